// MFMA GEMMs for the Fock build (gfx950).
//
//   C(m, n) = sum_k A(m, k) B(n, k)
//
// A(m, k) lives at A[m*lda + k] (K-contiguous, "AK") or A[k*lda + m]
// (M-contiguous); B(n, k) likewise. C is written row-major, C[m*ldc + n].
// The same kernel covers the four products of an iteration:
//   K  = T T^T   (A = B = T, K-contiguous, symmetric: upper tiles only)
//   T  = B^T C   (A M-contiguous over the fitted tensor, B K-contiguous)
//   Z  = Phi P   (both K-contiguous)
//   V  = Phi^T W (both M-contiguous, K = grid points)
//
// Arithmetic:
//   kF64       FP64 operands, v_mfma_f64_16x16x4f64, FP64 accumulate
//   kBF16x3    FP32 operands split on the fly into bf16 hi + lo while staging
//              to LDS; per k-step three v_mfma_f32_16x16x32_bf16 products
//              (hi*hi + hi*lo + lo*hi), FP32 accumulate
//
// Tiling: 256 threads = 4 wavefronts in a 2x2 arrangement over a 128x128
// output tile, each wavefront 64x64 = 4x4 MFMA 16x16 blocks. K is split over
// gridDim.z into a workspace of per-split partial tiles, then a deterministic
// reduction kernel sums the splits in order (and mirrors the symmetric case).
#include <cstdint>

#include "dfks_common.h"

namespace dfks {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

enum GemmPrec { kF64 = 0, kBF16x3 = 1 };

constexpr int kBM = 128;
constexpr int kBN = 128;

template <int PREC>
struct GemmTraits;

template <>
struct GemmTraits<kF64> {
  using In = double;
  static constexpr int BK = 16;
};

template <>
struct GemmTraits<kBF16x3> {
  using In = float;
  static constexpr int BK = 32;
};

// Map a linear tile id to (tm, tn); symmetric mode enumerates tm <= tn only.
__device__ __forceinline__ void tile_of(int tid, int ntn, bool sym, int& tm, int& tn) {
  if (!sym) {
    tm = tid / ntn;
    tn = tid % ntn;
    return;
  }
  // row r holds tiles (r, r..ntn-1): count ntn - r
  int r = 0, base = 0;
  while (base + (ntn - r) <= tid) {
    base += ntn - r;
    ++r;
  }
  tm = r;
  tn = r + (tid - base);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8). Give each XCD a contiguous range of tile ids so tiles that share an
// operand panel (consecutive ids) share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int per = (ntiles + 7) / 8;
  return (bid % 8) * per + bid / 8;
}

// Per-thread slice of a 128-row operand tile: row r, k offset kq, VEC values.
template <bool KC, int VEC>
__device__ __forceinline__ void slice_of(int tid, int& r, int& kq) {
  if (KC) {  // rows take 256/128 = 2 threads each, VEC consecutive k per thread
    r = tid >> 1;
    kq = (tid & 1) * VEC;
  } else {   // 128 consecutive rows per k-group: coalesced along the row index
    r = tid & 127;
    kq = (tid >> 7) * VEC;
  }
}

// Load VEC values X(row, k0+kq .. k0+kq+VEC-1) into regs (zeros outside).
template <bool KC, int VEC, typename T>
__device__ __forceinline__ void load_slice(const T* __restrict__ X, int64_t ld, int rows, int row0, int r, int64_t k0,
                                           int kq, int64_t k_end, bool vec_ok, T* regs) {
  const int row = row0 + r;
  const int64_t kb = k0 + kq;
  if (row >= rows) {
#pragma unroll
    for (int j = 0; j < VEC; ++j) regs[j] = (T)0;
    return;
  }
  if (KC) {
    const T* src = X + (int64_t)row * ld + kb;
    if (vec_ok && kb + VEC <= k_end) {
      constexpr int PER = 16 / sizeof(T);
      typedef T vecT __attribute__((ext_vector_type(PER)));
#pragma unroll
      for (int v = 0; v < VEC / PER; ++v) {
        vecT x = *reinterpret_cast<const vecT*>(src + v * PER);
#pragma unroll
        for (int j = 0; j < PER; ++j) regs[v * PER + j] = x[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < VEC; ++j) regs[j] = (kb + j < k_end) ? src[j] : (T)0;
    }
  } else {
#pragma unroll
    for (int j = 0; j < VEC; ++j) regs[j] = (kb + j < k_end) ? X[(kb + j) * ld + row] : (T)0;
  }
}

// NJ: 16-wide MFMA blocks per wavefront along N (tile width BN = 32 NJ: 4 ->
// 128, 5 -> 160 for N = 160-column exchange products); NJ = 5 only without the
// symmetric mode.
// Two workgroups per CU (the register budget of 2 waves per SIMD) where the
// tile fits it without spilling; the FP64 160-wide tile and the 160-wide
// split-BF16 tiles with an N-contiguous B keep one.
template <int PREC, bool BKC, int NJ>
constexpr int gemm_blocks_per_cu() {
  return (NJ == 5 && (PREC == kF64 || !BKC)) ? 1 : 2;
}

template <int PREC, bool AK, bool BKC, typename Out, int NJ = 4>
__global__ __launch_bounds__(256, (gemm_blocks_per_cu<PREC, BKC, NJ>())) void k_gemm(int M, int N, int64_t K, const typename GemmTraits<PREC>::In* __restrict__ A,
                                              int64_t lda, const typename GemmTraits<PREC>::In* __restrict__ B,
                                              int64_t ldb, Out* __restrict__ C, int64_t ldc, int64_t split_stride,
                                              int64_t kchunk, int sym, int ntiles, int vec_a, int vec_b, int64_t bs_a,
                                              int64_t bs_b, int64_t bs_c) {
  using In = typename GemmTraits<PREC>::In;
  // batch index on gridDim.y (strided operands; a zero stride shares the operand)
  A += blockIdx.y * bs_a;
  B += blockIdx.y * bs_b;
  C += blockIdx.y * bs_c;
  constexpr int BK = GemmTraits<PREC>::BK;
  constexpr int VEC = (128 * BK) / 256;  // values per thread per operand tile
  constexpr int BN = 32 * NJ;
  static_assert(NJ == 4 || NJ == 5, "128- or 160-wide tiles");
  // split-by-XCD mode (sym bit 1): block x runs on XCD x % 8, which takes the
  // splits z = x % 8 (mod 8) for every tile, so the tiles on one XCD walk the same
  // k-window and share it through that XCD's L2 (the operand is read about once)
  const bool xs = (sym & 2) != 0;
  const int tile = xs ? (int)(blockIdx.x >> 3) : xcd_tile(blockIdx.x, ntiles);
  const int split = xs ? (int)((blockIdx.x & 7) + 8 * blockIdx.z) : (int)blockIdx.z;
  if (tile >= ntiles) return;
  const int ntn = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(tile, ntn, (sym & 1) != 0, tm, tn);
  const int m0 = tm * kBM, n0 = tn * BN;
  const int64_t k_begin = (int64_t)split * kchunk;
  const int64_t k_end = k_begin + kchunk < K ? k_begin + kchunk : K;
  Out* Cz = C + (int64_t)split * split_stride;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * (16 * NJ);
  const int li = lane & 15, lk = lane >> 4;
  int ra_r, ra_k, rb_r, rb_k;
  slice_of<AK, VEC>(tid, ra_r, ra_k);
  slice_of<BKC, VEC>(tid, rb_r, rb_k);
  In ra[VEC], rb[VEC];
  // rows 128 .. BN-1 of the B tile (NJ = 5): threads 0..63, XV = 32 BK / 64 values each
  constexpr int XR = BN - 128;
  constexpr int XV = BK / 2;
  int rx_r = 0, rx_k = 0;
  if (BKC) { rx_r = 128 + (tid >> 1); rx_k = (tid & 1) * XV; }
  else { rx_r = 128 + (tid & 31); rx_k = (tid >> 5) * XV; }
  const bool has_x = XR > 0 && tid < 64;
  In rbx[XV];
  if (k_begin < k_end) {
    load_slice<AK, VEC>(A, lda, M, m0, ra_r, k_begin, ra_k, k_end, vec_a != 0, ra);
    load_slice<BKC, VEC>(B, ldb, N, n0, rb_r, k_begin, rb_k, k_end, vec_b != 0, rb);
    if (has_x) load_slice<BKC, XV>(B, ldb, N, n0, rx_r, k_begin, rx_k, k_end, vec_b != 0, rbx);
  }

  if constexpr (PREC == kF64) {
    __shared__ double sA[kBM][BK + 1];
    __shared__ double sB[BN][BK + 1];
    dbl4 acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[i][j] = dbl4{0, 0, 0, 0};
    for (int64_t k0 = k_begin; k0 < k_end; k0 += BK) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < VEC; ++j) sA[ra_r][ra_k + j] = ra[j];
#pragma unroll
      for (int j = 0; j < VEC; ++j) sB[rb_r][rb_k + j] = rb[j];
      if (has_x) {
#pragma unroll
        for (int j = 0; j < XV; ++j) sB[rx_r][rx_k + j] = rbx[j];
      }
      __syncthreads();
      if (k0 + BK < k_end) {  // prefetch the next tile while this one computes
        load_slice<AK, VEC>(A, lda, M, m0, ra_r, k0 + BK, ra_k, k_end, vec_a != 0, ra);
        load_slice<BKC, VEC>(B, ldb, N, n0, rb_r, k0 + BK, rb_k, k_end, vec_b != 0, rb);
        if (has_x) load_slice<BKC, XV>(B, ldb, N, n0, rx_r, k0 + BK, rx_k, k_end, vec_b != 0, rbx);
      }
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        double a[4], b[NJ];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = sA[wm + 16 * i + li][kk + lk];
#pragma unroll
        for (int j = 0; j < NJ; ++j) b[j] = sB[wn + 16 * j + li][kk + lk];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < NJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    // v_mfma_f64_16x16x4f64 result: lane l, register r holds row (l/16) + 4r, column l%16
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int gm = m0 + wm + 16 * i + lk + 4 * r;
          int gn = n0 + wn + 16 * j + li;
          if (gm < M && gn < N) Cz[(int64_t)gm * ldc + gn] = (Out)acc[i][j][r];
        }
  } else {
    constexpr int PADK = BK + 8;  // bf16 per LDS row; 80-byte rows keep 16-byte reads conflict-free
    __shared__ __attribute__((aligned(16))) __bf16 sAh[kBM][PADK];
    __shared__ __attribute__((aligned(16))) __bf16 sAl[kBM][PADK];
    __shared__ __attribute__((aligned(16))) __bf16 sBh[BN][PADK];
    __shared__ __attribute__((aligned(16))) __bf16 sBl[BN][PADK];
    flt4 acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j) acc[i][j] = flt4{0, 0, 0, 0};
    for (int64_t k0 = k_begin; k0 < k_end; k0 += BK) {
      __syncthreads();
      // split into bf16 hi + lo and store 16-byte vectors
#pragma unroll
      for (int v = 0; v < VEC / 8; ++v) {
        bf16x8 h, l, hb, lb;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          float x = ra[8 * v + j];
          __bf16 xh = (__bf16)x;
          h[j] = xh;
          l[j] = (__bf16)(x - (float)xh);
          float y = rb[8 * v + j];
          __bf16 yh = (__bf16)y;
          hb[j] = yh;
          lb[j] = (__bf16)(y - (float)yh);
        }
        *reinterpret_cast<bf16x8*>(&sAh[ra_r][ra_k + 8 * v]) = h;
        *reinterpret_cast<bf16x8*>(&sAl[ra_r][ra_k + 8 * v]) = l;
        *reinterpret_cast<bf16x8*>(&sBh[rb_r][rb_k + 8 * v]) = hb;
        *reinterpret_cast<bf16x8*>(&sBl[rb_r][rb_k + 8 * v]) = lb;
      }
      if (has_x) {
#pragma unroll
        for (int v = 0; v < XV / 8; ++v) {
          bf16x8 hb, lb;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            float y = rbx[8 * v + j];
            __bf16 yh = (__bf16)y;
            hb[j] = yh;
            lb[j] = (__bf16)(y - (float)yh);
          }
          *reinterpret_cast<bf16x8*>(&sBh[rx_r][rx_k + 8 * v]) = hb;
          *reinterpret_cast<bf16x8*>(&sBl[rx_r][rx_k + 8 * v]) = lb;
        }
      }
      __syncthreads();
      if (k0 + BK < k_end) {
        load_slice<AK, VEC>(A, lda, M, m0, ra_r, k0 + BK, ra_k, k_end, vec_a != 0, ra);
        load_slice<BKC, VEC>(B, ldb, N, n0, rb_r, k0 + BK, rb_k, k_end, vec_b != 0, rb);
        if (has_x) load_slice<BKC, XV>(B, ldb, N, n0, rx_r, k0 + BK, rx_k, k_end, vec_b != 0, rbx);
      }
      bf16x8 ah[4], al[4], bh[NJ], bl[NJ];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ah[i] = *reinterpret_cast<const bf16x8*>(&sAh[wm + 16 * i + li][8 * lk]);
        al[i] = *reinterpret_cast<const bf16x8*>(&sAl[wm + 16 * i + li][8 * lk]);
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        bh[j] = *reinterpret_cast<const bf16x8*>(&sBh[wn + 16 * j + li][8 * lk]);
        bl[j] = *reinterpret_cast<const bf16x8*>(&sBl[wn + 16 * j + li][8 * lk]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    // v_mfma_f32_16x16x32_bf16 result: lane l, register r holds row 4(l/16) + r, column l%16
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int gm = m0 + wm + 16 * i + 4 * lk + r;
          int gn = n0 + wn + 16 * j + li;
          if (gm < M && gn < N) Cz[(int64_t)gm * ldc + gn] = (Out)acc[i][j][r];
        }
  }
}

// FP64 product C = A B^T with B K-contiguous (B(n, k) = B[n*ldb + k]) and A
// either M-contiguous (AK = false: A(m, k) = A[k*lda + m], the exchange
// product T = B^T C over the fitted tensor) or K-contiguous (AK = true: the
// symmetric exchange product K = T T^T). One 128 x 32NJ tile per workgroup of
// 512 threads = 8 wavefronts (4 along M x 2 along N, each 32 x 16NJ = 2 x NJ
// MFMA 16x16 blocks), so every SIMD holds two wavefronts and one wavefront's
// MFMAs cover the other's barrier and LDS waits. The LDS tiles are
// double-buffered: the next k-step's global loads are issued before the
// current step's MFMAs and stored to the other buffer after them, with one
// barrier per k-step. Tiles, the symmetric mode (upper tiles only) and the
// split of K over gridDim.z (partial tiles at C + z * split_stride) follow
// k_gemm.
constexpr int kW8BM = 128, kW8BK = 16;

template <bool AK, int NJ>
__global__ __launch_bounds__(512, 1) void k_gemm_w8(int64_t M, int N, int64_t K, const double* __restrict__ A,
                                                   int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                   double* __restrict__ C, int64_t ldc, int64_t split_stride,
                                                   int64_t kchunk, int sym, int ntiles, int vec_a, int vec_b,
                                                   RowDot rd) {
  constexpr int BN = 32 * NJ;
  __shared__ double sA[2][kW8BM][kW8BK + 1];
  __shared__ double sB[2][BN][kW8BK + 1];
  // split-by-XCD mode (sym bit 1) as in k_gemm
  const bool xs = (sym & 2) != 0;
  const int tile = xs ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  const int split = xs ? (int)((blockIdx.x & 7) + 8 * blockIdx.z) : (int)blockIdx.z;
  if (tile >= ntiles) return;
  const int ntn = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(tile, ntn, (sym & 1) != 0, tm, tn);
  const int64_t m0 = (int64_t)tm * kW8BM;
  const int n0 = tn * BN;
  const int64_t k_begin = (int64_t)split * kchunk;
  const int64_t k_end = k_begin + kchunk < K ? k_begin + kchunk : K;
  double* __restrict__ Cz = C + (int64_t)split * split_stride;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * (16 * NJ);
  const int li = lane & 15, lk = lane >> 4;
  // A slice, 4 values per thread: M-contiguous row tid % 128, k 4 (tid / 128) .. +3
  // (coalesced along m); K-contiguous row tid / 4, k 4 (tid % 4) .. +3
  const int ar = AK ? tid >> 2 : tid & 127, ak = AK ? (tid & 3) * 4 : (tid >> 7) * 4;
  // B slice (threads < 2 BN): row tid / 2, k 8 (tid % 2) .. +7
  const int br = tid >> 1, bk = (tid & 1) * 8;
  const bool has_b = tid < 2 * BN;
  const bool a_in = m0 + ar < M;
  const bool b_in = has_b && n0 + br < N;
  const double* __restrict__ pa = AK ? A + (a_in ? m0 + ar : 0) * lda : A + (a_in ? m0 + ar : 0);
  const double* __restrict__ pb = B + (int64_t)(b_in ? n0 + br : 0) * ldb;
  // gamma row terms: the tile's rows span at most two m (nq >= 128), so the coefficients
  // sg[i] cf[i, m] are 2 x BN values, loaded here and stored to LDS after the main loop
  double c_sg = 0.0, c_cf = 0.0;
  if (rd.out && tid < 2 * BN) {
    const int gn = n0 + tid % BN;
    const int64_t mm = m0 / rd.nq + tid / BN;
    if (gn < N && mm < rd.nao) {
      c_sg = rd.sg[gn];
      c_cf = rd.cf[(int64_t)gn * rd.nao + mm];
    }
  }
  // A and B are loaded one k-step ahead into registers.
  double ra[4], rb[8];
  auto load_a = [&](int64_t k0) {
    const int64_t ka = k0 + ak;
    if (AK && a_in && vec_a && ka + 4 <= k_end) {
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const double2 x = *reinterpret_cast<const double2*>(pa + ka + 2 * v);
        ra[2 * v] = x.x;
        ra[2 * v + 1] = x.y;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) ra[j] = (a_in && ka + j < k_end) ? (AK ? pa[ka + j] : pa[(ka + j) * lda]) : 0.0;
    }
  };
  auto load_b = [&](int64_t k0) {
    if (!has_b) return;
    const int64_t kb = k0 + bk;
    if (b_in && vec_b && kb + 8 <= k_end) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double2 x = *reinterpret_cast<const double2*>(pb + kb + 2 * v);
        rb[2 * v] = x.x;
        rb[2 * v + 1] = x.y;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) rb[j] = (b_in && kb + j < k_end) ? pb[kb + j] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) sA[buf][ar][ak + j] = ra[j];
    if (has_b) {
#pragma unroll
      for (int j = 0; j < 8; ++j) sB[buf][br][bk + j] = rb[j];
    }
  };
  dbl4 acc[2][NJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = dbl4{0, 0, 0, 0};
  auto compute = [&](int cur) {
#pragma unroll
    for (int kk = 0; kk < kW8BK; kk += 4) {
      double a[2], b[NJ];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = sA[cur][wm + 16 * i + li][kk + lk];
#pragma unroll
      for (int j = 0; j < NJ; ++j) b[j] = sB[cur][wn + 16 * j + li][kk + lk];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  };
  const int64_t nk = k_end > k_begin ? (k_end - k_begin + kW8BK - 1) / kW8BK : 0;
  load_a(k_begin);
  load_b(k_begin);
  store(0);
  __syncthreads();
  for (int64_t t = 0; t < nk; ++t) {
    const int cur = (int)(t & 1);
    if (t + 1 < nk) {
      load_a(k_begin + (t + 1) * kW8BK);
      load_b(k_begin + (t + 1) * kW8BK);
    }
    compute(cur);
    if (t + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
  // v_mfma_f64_16x16x4f64 result: lane l, register r holds row (l/16) + 4r, column l%16
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + wm + 16 * i + lk + 4 * r;
        const int gn = n0 + wn + 16 * j + li;
        if (gm < M && gn < N) Cz[gm * ldc + gn] = acc[i][j][r];
      }
  if (rd.out) {
    // row sums over this tile's columns: 16 lanes of a row group, then the two N wavefronts through LDS
    __shared__ double s_row[2][kW8BM];
    __shared__ double s_cs[2][BN];
    if (tid < 2 * BN) s_cs[tid / BN][tid % BN] = c_sg * c_cf;
    __syncthreads();
    const int64_t m_first = m0 / rd.nq;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = wm + 16 * i + lk + 4 * r;
        const int64_t gm = m0 + lr;
        const int which = gm < M ? (int)(gm / rd.nq - m_first) : 0;
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int c = wn + 16 * j + li;
          v += acc[i][j][r] * s_cs[which][c];
        }
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        if (li == 0) s_row[wave & 1][lr] = v;
      }
    __syncthreads();
    if (tid < kW8BM && m0 + tid < M) rd.out[m0 + tid] = s_row[0][tid] + s_row[1][tid];
  }
}

// Split-BF16 (three-term) product C = A B^T with FP32 accumulate for A
// M-contiguous (A(m, k) = A[k*lda + m]) and B K-contiguous, C row-major FP32
// with ldc = N: the mixed-precision exchange product T = B^T C. One 256 x 160
// tile per workgroup of 512 threads = 8 wavefronts (4 along M x 2 along N, each
// 64 x 80 = 4 x 5 MFMA 16x16x32 blocks). Each operand value is split into bf16
// hi + lo once, while it is staged to LDS; the LDS tiles are double-buffered
// with one barrier per k-step. The 256-row tile halves the re-reads of the
// small operand per row of the large one against a 128-row tile.
//
// The global loads run DEPTH k-steps ahead of the MFMAs (DEPTH = 2: two
// register sets in flight). Every load is unconditional, so that none has to
// complete before the next is issued: addresses are clamped into the operands
// (rows past M or N load finite values of the last row and give output rows or
// columns that are never stored, and row terms weighted by zero; k past K
// loads finite values of the last k, and B's are replaced by zeros when
// staged, so those k contribute nothing), the k-row offsets of A are
// wave-uniform (scalar) and only the row offset within the tile is per lane.
// VEC: A in 8-byte pairs of rows (M even, 8-byte aligned rows), B in 16-byte
// groups along k (16-byte aligned rows); otherwise single floats.
constexpr int kB3BM = 256, kB3BN = 160, kB3BK = 32, kB3PAD = kB3BK + 8;

__device__ __forceinline__ void split_bf16x8(const float* x, bf16x8& h, bf16x8& l) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 xh = (__bf16)x[j];
    h[j] = xh;
    l[j] = (__bf16)(x[j] - (float)xh);
  }
}

template <int DEPTH, bool VEC>
__global__ __launch_bounds__(512, 1) void k_gemm_b3w8(int64_t M, int N, int K, const float* __restrict__ A,
                                                     int64_t lda, const float* __restrict__ B, int64_t ldb,
                                                     float* __restrict__ C, RowDot rd) {
  __shared__ __attribute__((aligned(16))) __bf16 sAh[2][kB3BM][kB3PAD];
  __shared__ __attribute__((aligned(16))) __bf16 sAl[2][kB3BM][kB3PAD];
  __shared__ __attribute__((aligned(16))) __bf16 sBh[2][kB3BN][kB3PAD];
  __shared__ __attribute__((aligned(16))) __bf16 sBl[2][kB3BN][kB3PAD];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t m0 = (int64_t)blockIdx.x * kB3BM;
  const int n0 = blockIdx.y * kB3BN;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 80;
  const int li = lane & 15, lk = lane >> 4;
  // A slice: rows ar, ar + 1 with ar = 2 (lane) + 128 (wave % 2) (8 bytes per
  // lane, 512 contiguous bytes per wavefront and k) at k ak .. ak + 7 with ak =
  // 8 (wave / 2), so that each row goes to LDS as one 16-byte chunk per plane:
  // a wavefront's 64 rows of one chunk then fall on 16 of the 32 banks
  const int ar = 2 * lane + 128 * (wave & 1), ak = 8 * (wave >> 1);
  // row offsets within the tile, clamped to the last row(s) of A
  const int64_t m_room = M - m0;   // rows of A from m0 on (> 0)
  const int a_off = VEC ? (int)(ar < m_room - 2 ? ar : m_room - 2) : 0;
  const int a_off0 = (int)(ar < m_room - 1 ? ar : m_room - 1);
  const int a_off1 = (int)(ar + 1 < m_room - 1 ? ar + 1 : m_room - 1);
  const float* __restrict__ Am = A + m0;
  // B slice (wavefronts 0-4): row tid / 2, k 16 (tid % 2) .. +15
  const int br = tid >> 1, bk = (tid & 1) * 16;
  const bool has_b = wave < 5;
  const float* __restrict__ pb = B + (int64_t)(n0 + br < N ? n0 + br : N - 1) * ldb;
  const int k_last4 = ((K - 1) / 4) * 4;   // start of the last 4-float group along k of B
  // gamma row terms: the tile's rows span at most two m (nq >= 256), so the coefficients
  // sg[i] cf[i, m] are 2 x 160 values, loaded here and stored to LDS after the main loop
  double c_sg = 0.0, c_cf = 0.0;
  if (rd.out && tid < 2 * kB3BN) {
    const int gn = n0 + tid % kB3BN;
    const int64_t mm = m0 / rd.nq + tid / kB3BN;
    if (gn < N && mm < rd.nao) {
      c_sg = rd.sg[gn];
      c_cf = rd.cf[(int64_t)gn * rd.nao + mm];
    }
  }
  // register sets: ra[2 q + c] row ar + c, k ak + q; rb[j] row br, k bk + j
  float ra0[16], rb0[16], ra1[16], rb1[16];
  auto load_a = [&](float (&ra)[16], int k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + ak + q;
      const float* __restrict__ row = Am + (int64_t)(k < K ? k : K - 1) * lda;
      if (VEC) {
        const float2 x = *reinterpret_cast<const float2*>(row + a_off);
        ra[2 * q] = x.x;
        ra[2 * q + 1] = x.y;
      } else {
        ra[2 * q] = row[a_off0];
        ra[2 * q + 1] = row[a_off1];
      }
    }
  };
  auto load_b = [&](float (&rb)[16], int k0) {
    if (!has_b) return;
    const int kb = k0 + bk;
    if (VEC) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int kv = kb + 4 * v;
        const float4 x = *reinterpret_cast<const float4*>(pb + (kv < k_last4 ? kv : k_last4));
        rb[4 * v] = x.x;
        rb[4 * v + 1] = x.y;
        rb[4 * v + 2] = x.z;
        rb[4 * v + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; ++j) rb[j] = pb[kb + j < K ? kb + j : K - 1];
    }
  };
  auto load = [&](float (&ra)[16], float (&rb)[16], int k0) {
    load_a(ra, k0);
    load_b(rb, k0);
  };
  // kvalid = K - (first k of the step): B's values at k >= K become zeros (the last step only)
  auto store = [&](const float (&ra)[16], float (&rb)[16], int buf, int kvalid) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      bf16x8 h, l;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float x = ra[2 * q + c];
        const __bf16 xh = (__bf16)x;
        h[q] = xh;
        l[q] = (__bf16)(x - (float)xh);
      }
      *reinterpret_cast<bf16x8*>(&sAh[buf][ar + c][ak]) = h;
      *reinterpret_cast<bf16x8*>(&sAl[buf][ar + c][ak]) = l;
    }
    if (has_b) {
      if (kvalid < kB3BK) {
#pragma unroll
        for (int j = 0; j < 16; ++j) rb[j] = bk + j < kvalid ? rb[j] : 0.f;
      }
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        bf16x8 h, l;
        split_bf16x8(rb + 8 * v, h, l);
        *reinterpret_cast<bf16x8*>(&sBh[buf][br][bk + 8 * v]) = h;
        *reinterpret_cast<bf16x8*>(&sBl[buf][br][bk + 8 * v]) = l;
      }
    }
  };
  flt4 acc[4][5];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = flt4{0, 0, 0, 0};
  auto compute = [&](int cur) {
    bf16x8 bh[5], bl[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      bh[j] = *reinterpret_cast<const bf16x8*>(&sBh[cur][wn + 16 * j + li][8 * lk]);
      bl[j] = *reinterpret_cast<const bf16x8*>(&sBl[cur][wn + 16 * j + li][8 * lk]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&sAh[cur][wm + 16 * i + li][8 * lk]);
      const bf16x8 al = *reinterpret_cast<const bf16x8*>(&sAl[cur][wm + 16 * i + li][8 * lk]);
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
      }
    }
  };
  const int nk = (K + kB3BK - 1) / kB3BK;
  if (DEPTH == 1) {
    load(ra0, rb0, 0);
    store(ra0, rb0, 0, K);
    __syncthreads();
    for (int t = 0; t < nk; ++t) {
      load(ra0, rb0, (t + 1) * kB3BK);   // past the last step: clamped loads, never stored
      compute(t & 1);
      if (t + 1 < nk) store(ra0, rb0, (t & 1) ^ 1, K - (t + 1) * kB3BK);
      __syncthreads();
    }
  } else {
    // at the top of an even step t, LDS buffer 0 holds step t, (ra1, rb1) step
    // t + 1 and (ra0, rb0) step t + 2 (loads past the last step are clamped and
    // never stored); the body covers steps t and t + 1 so that each register set
    // keeps a fixed name
    load(ra0, rb0, 0);
    load(ra1, rb1, kB3BK);
    store(ra0, rb0, 0, K);
    __syncthreads();
    load(ra0, rb0, 2 * kB3BK);
    for (int t = 0; t < nk; t += 2) {
      compute(0);
      if (t + 1 < nk) store(ra1, rb1, 1, K - (t + 1) * kB3BK);
      __syncthreads();
      if (t + 1 >= nk) break;
      load(ra1, rb1, (t + 3) * kB3BK);
      compute(1);
      if (t + 2 < nk) store(ra0, rb0, 0, K - (t + 2) * kB3BK);
      __syncthreads();
      load(ra0, rb0, (t + 4) * kB3BK);
    }
  }
  // v_mfma_f32_16x16x32_bf16 result: lane l, register r holds row 4(l/16) + r, column l%16
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + wm + 16 * i + 4 * lk + r;
        const int gn = n0 + wn + 16 * j + li;
        if (gm < M && gn < N) C[gm * N + gn] = acc[i][j][r];
      }
  if (rd.out) {
    // row sums of the FP32 products in FP64 (as the gamma kernel sums the stored T): 16 lanes of
    // a row group, then the two N wavefronts through LDS
    __shared__ double s_row[2][kB3BM];
    __shared__ double s_cs[2][kB3BN];
    if (tid < 2 * kB3BN) s_cs[tid / kB3BN][tid % kB3BN] = c_sg * c_cf;
    __syncthreads();
    const int64_t m_first = m0 / rd.nq;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = wm + 16 * i + 4 * lk + r;
        const int64_t gm = m0 + lr;
        const int which = gm < M ? (int)(gm / rd.nq - m_first) : 0;
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
          const int c = wn + 16 * j + li;
          v += (double)acc[i][j][r] * s_cs[which][c];
        }
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        if (li == 0) s_row[wave & 1][lr] = v;
      }
    __syncthreads();
    if (tid < kB3BM && m0 + tid < M) rd.out[m0 + tid] = s_row[0][tid] + s_row[1][tid];
  }
}

// out[m*ldo + n] = alpha * sum_z ws[z][m][n] (+ beta * out); symmetric mode reads
// the upper triangle (m <= n) and mirrors it. Fixed summation order.
template <typename W, typename Out>
__global__ void k_gemm_reduce(int M, int N, int nsplit, const W* __restrict__ ws, int64_t split_stride, int64_t ldw,
                              Out* __restrict__ out, int64_t ldo, double alpha, double beta, int sym) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * N) return;
  int m = (int)(idx / N), n = (int)(idx % N);
  int sm = m, sn = n;
  if (sym && m > n) { sm = n; sn = m; }
  double s = 0.0;
  for (int z = 0; z < nsplit; ++z) s += (double)ws[z * split_stride + (int64_t)sm * ldw + sn];
  double v = alpha * s;
  if (beta != 0.0) v += beta * (double)out[(int64_t)m * ldo + n];
  out[(int64_t)m * ldo + n] = (Out)v;
}

// Symmetric mode of the reduce for N x N: one workgroup per upper 64 x 64 tile
// (ti <= tj) sums the partials of the tile's upper elements (coalesced, fixed
// order), writes them, and writes the mirror tile through LDS, so both stores
// are row-contiguous (the elementwise kernel read the lower half's partials
// down columns). out = alpha * sum (+ beta * out) for every element, the
// lower one from its upper partner's sum, as in the elementwise kernel.
template <typename W, typename Out>
__global__ __launch_bounds__(256) void k_gemm_reduce_sym(int N, int nsplit, const W* __restrict__ ws,
                                                         int64_t split_stride, int64_t ldw, Out* __restrict__ out,
                                                         int64_t ldo, double alpha, double beta) {
  __shared__ double t[64][65];
  const int ntn = (N + 63) / 64;
  int ti = 0, b = (int)blockIdx.x;
  while (b >= ntn - ti) {
    b -= ntn - ti;
    ++ti;
  }
  const int tj = ti + b;
  if (ti >= ntn) return;
  const int m0 = 64 * ti, n0 = 64 * tj, c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
  for (int j = 0; j < 16; ++j) {
    const int r = r0 + 4 * j, m = m0 + r, n = n0 + c;
    if (m < N && n < N && (ti != tj || r <= c)) {
      double s = 0.0;
      for (int z = 0; z < nsplit; ++z) s += (double)ws[z * split_stride + (int64_t)m * ldw + n];
      const double v = alpha * s;
      t[r][c] = v;
      out[(int64_t)m * ldo + n] = (Out)(beta != 0.0 ? v + beta * (double)out[(int64_t)m * ldo + n] : v);
    }
  }
  __syncthreads();
  for (int j = 0; j < 16; ++j) {
    const int r = r0 + 4 * j, m = n0 + r, n = m0 + c;   // mirror element (n0 + r, m0 + c) of tile (c, r)
    if (m < N && n < N && (ti != tj || c < r)) {
      const double v = t[c][r];
      out[(int64_t)m * ldo + n] = (Out)(beta != 0.0 ? v + beta * (double)out[(int64_t)m * ldo + n] : v);
    }
  }
}

struct GemmArgs {
  int M, N;
  int64_t K;
  const void* A;
  int64_t lda;
  bool a_kcontig;
  const void* B;
  int64_t ldb;
  bool b_kcontig;
};

template <int PREC, typename Out, int NJ = 4>
static hipError_t launch_gemm_t(const GemmArgs& g, Out* C, int64_t ldc, int nsplit, int sym, hipStream_t st,
                                int batch = 1, int64_t bs_a = 0, int64_t bs_b = 0, int64_t bs_c = 0) {
  using In = typename GemmTraits<PREC>::In;
  constexpr int BK = GemmTraits<PREC>::BK;
  constexpr int BN = 32 * NJ;
  if ((sym & 1) && NJ != 4) return hipErrorInvalidValue;
  const bool xs = (sym & 2) != 0;
  if (xs && (nsplit % 8 != 0 || batch != 1)) return hipErrorInvalidValue;
  const int ntm = (g.M + kBM - 1) / kBM, ntn = (g.N + BN - 1) / BN;
  const int ntiles = (sym & 1) ? ntn * (ntn + 1) / 2 : ntm * ntn;
  int64_t kchunk = (g.K + nsplit - 1) / nsplit;
  kchunk = (kchunk + BK - 1) / BK * BK;
  const int64_t split_stride = (int64_t)g.M * ldc;
  const int nblk = (ntiles + 7) / 8 * 8;
  const dim3 grid = xs ? dim3(8 * ntiles, 1, nsplit / 8) : dim3(nblk, batch, nsplit);
  const In* A = static_cast<const In*>(g.A);
  const In* B = static_cast<const In*>(g.B);
  // 16-byte vector loads need a 16-byte aligned base and row pitch (K-contiguous operands)
  constexpr int PER = 16 / sizeof(In);
  const int vec_a = ((uintptr_t)A % 16 == 0) && (g.lda % PER == 0);
  const int vec_b = ((uintptr_t)B % 16 == 0) && (g.ldb % PER == 0);
  const auto run = [&](auto kern) {
    return launch_k(kern, grid, dim3(256), st, g.M, g.N, g.K, A, g.lda, B, g.ldb, C, ldc, split_stride, kchunk, sym,
                    ntiles, vec_a, vec_b, bs_a, bs_b, bs_c);
  };
  if (g.a_kcontig && g.b_kcontig) return run(k_gemm<PREC, true, true, Out, NJ>);
  if (g.a_kcontig && !g.b_kcontig) return run(k_gemm<PREC, true, false, Out, NJ>);
  if (!g.a_kcontig && g.b_kcontig) return run(k_gemm<PREC, false, true, Out, NJ>);
  return run(k_gemm<PREC, false, false, Out, NJ>);
}

// k_gemm_w8 launch: ldc, split and symmetric mode as launch_gemm_t (NJ = 5
// only without the symmetric mode); B must be K-contiguous.
template <int NJ>
static hipError_t launch_w8_t(const GemmArgs& g, double* C, int64_t ldc, int nsplit, int sym, hipStream_t st,
                              const RowDot& rd = RowDot()) {
  constexpr int BN = 32 * NJ;
  if (!g.b_kcontig || ((sym & 1) && NJ != 4)) return hipErrorInvalidValue;
  if (rd.out && (g.N > BN || nsplit != 1 || sym || rd.nq < kW8BM)) return hipErrorInvalidValue;
  const bool xs = (sym & 2) != 0;
  if (xs && nsplit % 8 != 0) return hipErrorInvalidValue;
  const int64_t ntm = ((int64_t)g.M + kW8BM - 1) / kW8BM, ntn = (g.N + BN - 1) / BN;
  const int64_t ntl = (sym & 1) ? ntn * (ntn + 1) / 2 : ntm * ntn;
  if (8 * ntl > 0x7fffffff || nsplit > 65535) return hipErrorInvalidValue;
  int64_t kchunk = (g.K + nsplit - 1) / nsplit;
  kchunk = (kchunk + kW8BK - 1) / kW8BK * kW8BK;
  const double* A = static_cast<const double*>(g.A);
  const double* B = static_cast<const double*>(g.B);
  const int vec_a = ((uintptr_t)A % 16 == 0) && (g.lda % 2 == 0);
  const int vec_b = ((uintptr_t)B % 16 == 0) && (g.ldb % 2 == 0);
  const dim3 grid = xs ? dim3((unsigned)(8 * ntl), 1, (unsigned)(nsplit / 8)) : dim3((unsigned)ntl, 1, (unsigned)nsplit);
  const int64_t split_stride = (int64_t)g.M * ldc;
  if (g.a_kcontig)
    return launch_k(k_gemm_w8<true, NJ>, grid, dim3(512), st, (int64_t)g.M, g.N, g.K, A, g.lda, B, g.ldb, C, ldc,
                    split_stride, kchunk, sym, (int)ntl, vec_a, vec_b, rd);
  return launch_k(k_gemm_w8<false, NJ>, grid, dim3(512), st, (int64_t)g.M, g.N, g.K, A, g.lda, B, g.ldb, C, ldc,
                  split_stride, kchunk, sym, (int)ntl, vec_a, vec_b, rd);
}

// C = A B^T in FP64 by k_gemm_w8, no split (ldc = N): 128- or 160-wide N
// tiles, whichever pads N less. B must be K-contiguous. rd.out (optional,
// N <= 160): the row terms of gamma.
hipError_t launch_gemm_w8(const GemmArgs& g, double* C, const RowDot& rd, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const int t128 = (g.N + 127) / 128, t160 = (g.N + 159) / 160;
  if (t160 < t128 && (int64_t)t160 * 160 - g.N < (int64_t)t128 * 128 - g.N)
    return launch_w8_t<5>(g, C, g.N, 1, 0, st, rd);
  return launch_w8_t<4>(g, C, g.N, 1, 0, st, rd);
}

// Split-K partial tiles of A B^T into ws (nsplit x M x N, FP64 for kF64 and
// FP32 for kBF16x3; upper tiles only in the symmetric mode). variant 1 takes
// the 8-wavefront FP64 tile where it applies (B K-contiguous), 0 k_gemm.
hipError_t launch_gemm_partials(int prec, const GemmArgs& g, int nsplit, int sym, void* ws, int variant,
                                hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (!ws || nsplit < 1) return hipErrorInvalidValue;
  if (prec == kF64) {
    if (variant == 1 && g.b_kcontig) return launch_w8_t<4>(g, static_cast<double*>(ws), g.N, nsplit, sym, st);
    return launch_gemm_t<kF64, double>(g, static_cast<double*>(ws), g.N, nsplit, sym, st);
  }
  return launch_gemm_t<kBF16x3, float>(g, static_cast<float*>(ws), g.N, nsplit, sym, st);
}

// C = alpha * sum of the nsplit partial tiles in ws (+ beta C), mirrored in the
// symmetric mode; fixed summation order.
hipError_t launch_gemm_reduce(int prec, const GemmArgs& g, double* C, double alpha, double beta, int nsplit, int sym,
                              const void* ws, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const int64_t stride = (int64_t)g.M * g.N;
  if ((sym & 1) && g.M == g.N) {
    const int ntn = (g.N + 63) / 64;
    const unsigned nt = (unsigned)((int64_t)ntn * (ntn + 1) / 2);
    if (prec == kF64)
      return launch_k(k_gemm_reduce_sym<double, double>, dim3(nt), dim3(256), st, g.N, nsplit,
                      static_cast<const double*>(ws), stride, (int64_t)g.N, C, (int64_t)g.N, alpha, beta);
    return launch_k(k_gemm_reduce_sym<float, double>, dim3(nt), dim3(256), st, g.N, nsplit,
                    static_cast<const float*>(ws), stride, (int64_t)g.N, C, (int64_t)g.N, alpha, beta);
  }
  const unsigned nb = (unsigned)((stride + 255) / 256);
  if (prec == kF64)
    return launch_k(k_gemm_reduce<double, double>, dim3(nb), dim3(256), st, g.M, g.N, nsplit,
                    static_cast<const double*>(ws), stride, (int64_t)g.N, C, (int64_t)g.N, alpha, beta, sym & 1);
  return launch_k(k_gemm_reduce<float, double>, dim3(nb), dim3(256), st, g.M, g.N, nsplit,
                  static_cast<const float*>(ws), stride, (int64_t)g.N, C, (int64_t)g.N, alpha, beta, sym & 1);
}

// C = alpha * A B^T (+ beta C), C row-major M x N (ldc = N), FP64 output.
// ws: workspace of nsplit * M * N doubles (FP64) or floats (BF16x3), used when
// nsplit > 1 or sym; may be null otherwise.
hipError_t launch_gemm(int prec, const GemmArgs& g, double* C, double alpha, double beta, int nsplit, int sym,
                       void* ws, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (nsplit < 1) nsplit = 1;
  const bool direct = nsplit == 1 && !sym && alpha == 1.0 && beta == 0.0 && prec == kF64;
  if ((sym & 2) && nsplit % 8 != 0) return hipErrorInvalidValue;
  if (direct) {
    const int t128 = (g.N + 127) / 128, t160 = (g.N + 159) / 160;
    if (t160 < t128 && (int64_t)t160 * 160 - g.N < (int64_t)t128 * 128 - g.N)
      return launch_gemm_t<kF64, double, 5>(g, C, g.N, 1, 0, st);
    return launch_gemm_t<kF64, double>(g, C, g.N, 1, 0, st);
  }
  const hipError_t e = launch_gemm_partials(prec, g, nsplit, sym, ws, 0, st);
  if (e != hipSuccess) return e;
  return launch_gemm_reduce(prec, g, C, alpha, beta, nsplit, sym, ws, st);
}

// FP32 output variant (mixed path intermediates): C = A B^T, no split, no sym.
// N of 129..160 (the exchange product at 160 occupied orbitals) or a multiple
// of 160 takes 160-wide tiles: one tile per row block instead of a padded pair.
hipError_t launch_gemm_f32out(const GemmArgs& g, float* C, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  const int t128 = (g.N + 127) / 128, t160 = (g.N + 159) / 160;
  if (t160 < t128 && (int64_t)t160 * 160 - g.N < (int64_t)t128 * 128 - g.N)
    return launch_gemm_t<kBF16x3, float, 5>(g, C, g.N, 1, 0, st);
  return launch_gemm_t<kBF16x3, float>(g, C, g.N, 1, 0, st);
}

// Batched FP32-output variant: for b < batch, C + b*bs_c = (A + b*bs_a)(B + b*bs_b)^T.
hipError_t launch_gemm_f32out_batched(const GemmArgs& g, float* C, int batch, int64_t bs_a, int64_t bs_b,
                                      int64_t bs_c, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0 || batch <= 0) return hipSuccess;
  if (batch > 65535) return hipErrorInvalidValue;
  const int t128 = (g.N + 127) / 128, t160 = (g.N + 159) / 160;
  if (t160 < t128 && (int64_t)t160 * 160 - g.N < (int64_t)t128 * 128 - g.N)
    return launch_gemm_t<kBF16x3, float, 5>(g, C, g.N, 1, 0, st, batch, bs_a, bs_b, bs_c);
  return launch_gemm_t<kBF16x3, float>(g, C, g.N, 1, 0, st, batch, bs_a, bs_b, bs_c);
}

// C = A B^T in FP32 by k_gemm_b3w8 (split-BF16, A M-contiguous, B K-contiguous, ldc = N);
// rd.out (optional, N <= 160): the row terms of gamma; depth: k-steps of operands in flight (1, 2).
hipError_t launch_gemm_b3w8(const GemmArgs& g, float* C, const RowDot& rd, hipStream_t st, int depth) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (g.a_kcontig || !g.b_kcontig || g.K <= 0 || g.K > 0x7fffffff) return hipErrorInvalidValue;
  if (rd.out && (g.N > kB3BN || rd.nq < kB3BM)) return hipErrorInvalidValue;
  const float* A = static_cast<const float*>(g.A);
  const float* B = static_cast<const float*>(g.B);
  // 8-byte loads of 2 consecutive rows of A (M even), 16-byte loads along k of B
  const bool vec = ((uintptr_t)A % 8 == 0) && (g.lda % 2 == 0) && (g.M % 2 == 0) && ((uintptr_t)B % 16 == 0) &&
                   (g.ldb % 4 == 0);
  const dim3 grid((unsigned)(((int64_t)g.M + kB3BM - 1) / kB3BM), (unsigned)((g.N + kB3BN - 1) / kB3BN));
  const auto run = [&](auto kern) {
    return launch_k(kern, grid, dim3(512), st, (int64_t)g.M, g.N, (int)g.K, A, g.lda, B, g.ldb, C, rd);
  };
  if (depth == 2) return vec ? run(k_gemm_b3w8<2, true>) : run(k_gemm_b3w8<2, false>);
  return vec ? run(k_gemm_b3w8<1, true>) : run(k_gemm_b3w8<1, false>);
}

// Batched FP64 variant (no split, FP64 accumulate): for b < batch,
// C + b*bs_c = (A + b*bs_a)(B + b*bs_b)^T with ldc = N; 128- or 160-wide N
// tiles, whichever pads N less (the per-block products of the compact XC grid).
hipError_t launch_gemm_f64_batched(const GemmArgs& g, double* C, int batch, int64_t bs_a, int64_t bs_b,
                                   int64_t bs_c, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0 || batch <= 0) return hipSuccess;
  if (batch > 65535) return hipErrorInvalidValue;
  const int t128 = (g.N + 127) / 128, t160 = (g.N + 159) / 160;
  if (t160 < t128 && (int64_t)t160 * 160 - g.N < (int64_t)t128 * 128 - g.N)
    return launch_gemm_t<kF64, double, 5>(g, C, g.N, 1, 0, st, batch, bs_a, bs_b, bs_c);
  return launch_gemm_t<kF64, double>(g, C, g.N, 1, 0, st, batch, bs_a, bs_b, bs_c);
}

}  // namespace dfks
