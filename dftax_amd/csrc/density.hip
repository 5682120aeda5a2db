// Aufbau density of an orthonormal-basis Fock matrix without an eigensolver.
//
// The SCF step needs the projector onto the nocc lowest eigenvectors of
// F' = X^T F X and a factor of it for the exchange build. A dense FP64
// eigensolve is launch-latency bound on the GPU (tens of thousands of small
// kernels for n ~ 10^3), so the projector comes from trace-correcting second-
// order spectral projection (SP2; Niklasson, Phys. Rev. B 66, 155115 (2002)):
//
//   X_0     = (e_max I - F') / (e_max - e_min)        spectrum mapped into [0, 1]
//   X_{k+1} = X_k^2          if |tr X_k^2 - nocc| < |2 tr X_k - tr X_k^2 - nocc|
//           = 2 X_k - X_k^2  otherwise
//
// with e_min, e_max from 24 Lanczos steps (Ritz values widened by their
// residuals, inside the Gershgorin bounds). Both maps keep the spectrum in [0, 1],
// so e_k = tr X_k - tr X_k^2 = sum lambda (1 - lambda) >= 0 bounds every
// eigenvalue's distance from {0, 1}. Once e_k < 1e-9 two forced steps follow
// (X^2, then 2X - X^2), which square the remaining deviations on both sides,
// and the iteration freezes. For n <= 2015 one kernel per iteration
// (k_sp2_step) forms the upper 64 x 64 tiles of X_k^2 on FP64 MFMA, applies
// the branch, writes X_{k+1} (tile and mirror) into the other of two buffers
// and accumulates tr X_{k+1} and sum_ij (X_{k+1})_ij^2 = tr X_{k+1}^2, which
// decide the next branch before its product; larger n take split-K partial
// tiles of the product and a kernel that sums them and does the same update.
// The stop rule runs on the device, so the host only reads the state between
// chunks of iterations.
//
// The occupied factor L (X = L L^T, n x nocc) comes from a pivoted Cholesky of
// the converged projector in panels of kPanel columns: one workgroup picks the
// panel's pivots (largest residual diagonal) and forms its columns, then a
// GEMM removes the panel from the trailing residual. (The SCF driver uses it
// on the first step; later steps project the previous factor.)
#include <cstdint>

#include "dfks_common.h"

namespace dfks {

constexpr double kSp2Tol = 1e-9;

__device__ __forceinline__ double block_sum_1024(double v, double* s_red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if (lane == 0) s_red[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_red[w];
  __syncthreads();
  return t;  // valid in thread 0
}

// lohi[i] = F_ii - r_i, lohi[n + i] = F_ii + r_i with r_i = sum_{j != i} |F_ij|
__global__ __launch_bounds__(256) void k_gersh(int n, const double* __restrict__ F, double* __restrict__ lohi) {
  __shared__ double s[256];
  const int i = blockIdx.x;
  double r = 0.0;
  for (int j = threadIdx.x; j < n; j += 256)
    if (j != i) r += fabs(F[(int64_t)i * n + j]);
  s[threadIdx.x] = r;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double d = F[(int64_t)i * n + i];
    lohi[i] = d - s[0];
    lohi[n + i] = d + s[0];
  }
}

// one block: spectral bounds and a fresh state
__global__ __launch_bounds__(256) void k_sp2_start(int n, const double* __restrict__ lohi, double* __restrict__ st) {
  __shared__ double smin[256], smax[256];
  double lo = 1e300, hi = -1e300;
  for (int i = threadIdx.x; i < n; i += 256) {
    lo = fmin(lo, lohi[i]);
    hi = fmax(hi, lohi[n + i]);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    lo = smin[0];
    hi = smax[0];
    // a zero-width interval only for F' = c I; widen it so the map stays defined
    if (!(hi - lo > 1e-12 * fmax(1.0, fabs(hi)))) { lo -= 1.0; hi += 1.0; }
    st[kSpEmin] = lo;
    st[kSpEmax] = hi;
    st[kSpFlag] = 0.0;
    st[kSpCount] = 0.0;
    st[kSpIter] = 0.0;
    st[kSpErr] = 1e300;
    st[kSpTrace] = 0.0;
    st[kSpResid] = 0.0;
    st[kSpMinPiv] = 1e300;
  }
}

// X = (e_max I - F) / (e_max - e_min)
__global__ void k_sp2_init(int n, const double* __restrict__ F, const double* __restrict__ st, double* __restrict__ X) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  if (idx >= nn) return;
  const double lo = st[kSpEmin], hi = st[kSpEmax];
  const int i = (int)(idx / n), j = (int)(idx % n);
  X[idx] = ((i == j ? hi : 0.0) - F[idx]) / (hi - lo);
}

// The branch (or the freeze) of one iteration from t1 = tr X and t2 = tr X^2
// (thread 0 of k_sp2_decide).
__device__ void sp2_decide(double t1, double t2, int nocc, double* __restrict__ st) {
  if (st[kSpFlag] < 0.0) return;  // frozen
  const int it = (int)st[kSpIter];
  const double e = t1 - t2;
  if (it < kSp2MaxIter) st[kSpHist + it] = e;
  st[kSpIter] = it + 1;
  st[kSpErr] = e;
  st[kSpTrace] = t1;
  const double cnt = st[kSpCount];
  double flag;
  if (cnt >= 2.0) {
    flag = -1.0;  // converged: X is final
  } else if (cnt == 1.0) {
    flag = 0.0;   // second forced step: 2X - X^2
    st[kSpCount] = 2.0;
  } else if (fabs(e) < kSp2Tol) {
    flag = 1.0;   // first forced step: X^2
    st[kSpCount] = 1.0;
  } else {
    flag = fabs(t2 - nocc) < fabs(2.0 * t1 - t2 - nocc) ? 1.0 : 0.0;
  }
  st[kSpFlag] = flag;
}

// Per-workgroup partial sums of tr X and of sum_ij X_ij^2 (= tr X^2 for the
// symmetric X) into st[kSpPart + 2 b], st[kSpPart + 2 b + 1], from the NT
// threads of workgroup b (kSp2Groups workgroups).
template <int NT = 256>
__device__ void sp2_partials(double a, double b, double* __restrict__ st) {
  __shared__ double s_a[NT], s_b[NT];
  s_a[threadIdx.x] = a;
  s_b[threadIdx.x] = b;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_a[threadIdx.x] += s_a[threadIdx.x + w];
      s_b[threadIdx.x] += s_b[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    st[kSpPart + 2 * blockIdx.x] = s_a[0];
    st[kSpPart + 2 * blockIdx.x + 1] = s_b[0];
  }
}

__global__ __launch_bounds__(256) void k_sp2_norms(int n, const double* __restrict__ X, double* __restrict__ st) {
  const int64_t nn = (int64_t)n * n;
  double a = 0.0, b = 0.0;
  for (int64_t idx = blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)kSp2Groups * 256) {
    const double x = X[idx];
    if (idx / n == idx % n) a += x;
    b += x * x;
  }
  sp2_partials(a, b, st);
}

// One block: t1, t2 from the kSp2Groups partial pairs (fixed order), then the branch.
__global__ __launch_bounds__(kSp2Groups) void k_sp2_decide(int nocc, double* __restrict__ st) {
  __shared__ double s_a[kSp2Groups], s_b[kSp2Groups];
  s_a[threadIdx.x] = st[kSpPart + 2 * threadIdx.x];
  s_b[threadIdx.x] = st[kSpPart + 2 * threadIdx.x + 1];
  __syncthreads();
  for (int w = kSp2Groups / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_a[threadIdx.x] += s_a[threadIdx.x + w];
      s_b[threadIdx.x] += s_b[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sp2_decide(s_a[0], s_b[0], nocc, st);
}

// X <- X^2 or 2X - X^2 by the decided branch, X^2 summed from the nsplit
// partial upper tiles of the symmetric product (mirrored), with the partial
// sums of the new X's trace and squared norm for the next decision. Nothing
// changes once the iteration is frozen.
__global__ __launch_bounds__(256) void k_sp2_reduce_update(int n, int nsplit, const double* __restrict__ ws,
                                                           int64_t split_stride, double* __restrict__ X,
                                                           double* __restrict__ st) {
  const double flag = st[kSpFlag];
  if (flag < 0.0) return;
  const int64_t nn = (int64_t)n * n;
  double a = 0.0, b = 0.0;
  for (int64_t idx = blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)kSp2Groups * 256) {
    const int m = (int)(idx / n), c = (int)(idx % n);
    const int64_t off = m <= c ? (int64_t)m * n + c : (int64_t)c * n + m;
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += ws[z * split_stride + off];
    const double x = flag > 0.5 ? s : 2.0 * X[idx] - s;
    X[idx] = x;
    if (m == c) a += x;
    b += x * x;
  }
  sp2_partials(a, b, st);
}

// One purification iteration in one kernel, for n with at most kSp2Groups upper
// 64 x 64 tiles (n <= 2015). Workgroup b < ntiles takes upper tile (ti, tj),
// ti <= tj, of S = X X (FP64 MFMA over the full K; both operands are rows of
// the symmetric X), applies the decided branch, Xn = S or 2X - S, writes the
// tile and its mirror (through LDS, so both stores are row-contiguous), and
// its share of tr Xn and sum_ij Xn_ij^2 (off-diagonal tiles count twice) to
// the partial slot b; workgroups ntiles..kSp2Groups-1 write zero partials.
// Once frozen, Xn = X. X and Xn alternate between iterations.
//
// 16 wavefronts: two groups of 8 (two per SIMD each) take the two halves of K
// in stages of 64, each wavefront a 32 x 16 sub-tile; the second group's sums
// join the first's through LDS at the end. Half the stages of one group over
// the whole K, so half the barriers and global round trips per tile.
typedef double sq4 __attribute__((ext_vector_type(4)));
constexpr int kSqT = 64;   // tile edge
constexpr int kSqK = 64;   // k per LDS stage
constexpr int kSqN = 1024; // threads

__global__ __launch_bounds__(kSqN) void k_sp2_step(int n, int ntiles, const double* __restrict__ X,
                                                   double* __restrict__ Xn, double* __restrict__ st) {
  // operand stages per K half, rows padded to 66 doubles (the MFMA operand reads of a
  // half-wavefront, 16 rows x 2 k, fall on distinct bank pairs); after the k loop the
  // first one holds the second half's sums, then the updated tile for its mirror
  constexpr int kLd = kSqK + 2;
  __shared__ double sA[2][kSqT][kLd];
  __shared__ double sB[2][kSqT][kLd];
  static_assert(kSqK == kSqT, "the mirror tile reuses sA");
  const int b = blockIdx.x;
  if (b >= ntiles) {
    sp2_partials<kSqN>(0.0, 0.0, st);
    return;
  }
  const int ntn = (n + kSqT - 1) / kSqT;
  int ti = 0, t = b;
  while (t >= ntn - ti) {
    t -= ntn - ti;
    ++ti;
  }
  const int tj = ti + t;
  const int m0 = ti * kSqT, n0 = tj * kSqT;
  const double flag = st[kSpFlag];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kh = wave >> 3, wg = wave & 7;                 // K half, wavefront in its group
  const int wm = (wg >> 2) * 32, wn = (wg & 3) * 16;
  const int li = lane & 15, lk = lane >> 4;
  double (*hA)[kLd] = sA[kh];
  double (*hB)[kLd] = sB[kh];
  sq4 acc[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) acc[i] = sq4{0, 0, 0, 0};
  if (flag >= 0.0) {
    // stages of each half: both groups run the same count (barriers are shared); stages
    // past n load zeros
    const int nst = ((n + kSqK - 1) / kSqK + 1) / 2;
    const int kbeg = kh * nst * kSqK;
    // staging: a wavefront loads 64 consecutive k of one row per instruction (512
    // contiguous bytes); thread t of a group takes k = t % 64 of rows t / 64 + 8 j,
    // j < 8, clamped to element 0 outside X and masked
    const int gt = tid & 511, lq = gt & 63, lr = gt >> 6;
    double xa[8], xb[8];
    auto load = [&](int k0) {
      const int k = k0 + lq;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int ra = m0 + lr + 8 * j, rb = n0 + lr + 8 * j;
        const bool oka = ra < n && k < n, okb = rb < n && k < n;
        const double va = X[oka ? (int64_t)ra * n + k : 0];
        const double vb = X[okb ? (int64_t)rb * n + k : 0];
        xa[j] = oka ? va : 0.0;
        xb[j] = okb ? vb : 0.0;
      }
    };
    load(kbeg);
    for (int s2 = 0; s2 < nst; ++s2) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        hA[lr + 8 * j][lq] = xa[j];
        hB[lr + 8 * j][lq] = xb[j];
      }
      __syncthreads();
      if (s2 + 1 < nst) load(kbeg + (s2 + 1) * kSqK);
#pragma unroll
      for (int kk = 0; kk < kSqK; kk += 4) {
        const double bb = hB[wn + li][kk + lk];
#pragma unroll
        for (int i = 0; i < 2; ++i)
          acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(hA[wm + 16 * i + li][kk + lk], bb, acc[i], 0, 0, 0);
      }
    }
  }
  // v_mfma_f64_16x16x4f64 result: lane l, register r holds row (l/16) + 4r, column l%16
  double (*sT)[kLd] = sA[0];
  __syncthreads();   // every wavefront is done reading the stages
  if (kh == 1) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) sT[wm + 16 * i + lk + 4 * r][wn + li] = acc[i][r];
  }
  __syncthreads();
  double ta = 0.0, tb = 0.0;
  if (kh == 0) {
    const double wsq = ti == tj ? 1.0 : 2.0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lm = wm + 16 * i + lk + 4 * r, ln = wn + li;
        const int gm = m0 + lm, gn = n0 + ln;
        double x = 0.0;
        if (gm < n && gn < n) {
          const double x0 = X[(int64_t)gm * n + gn];
          const double sq = acc[i][r] + sT[lm][ln];
          x = flag < 0.0 ? x0 : flag > 0.5 ? sq : 2.0 * x0 - sq;
          Xn[(int64_t)gm * n + gn] = x;
          if (gm == gn) ta += x;
          tb += wsq * x * x;
        }
        sT[lm][ln] = x;   // each element is read above by the same thread only
      }
  }
  if (ti != tj) {
    __syncthreads();
    // mirror: Xn[n0 + c][m0 + r] = tile[r][c], 64 consecutive columns per wavefront row
    for (int e = tid; e < kSqT * kSqT; e += kSqN) {
      const int c = e >> 6, r = e & 63;
      const int gm = n0 + c, gn = m0 + r;
      if (gm < n && gn < n) Xn[(int64_t)gm * n + gn] = sT[r][c];
    }
  }
  sp2_partials<kSqN>(ta, tb, st);
}

// d[i] = R_ii
__global__ void k_diag(int n, const double* __restrict__ R, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = R[(int64_t)i * n + i];
}

// Pivoted Cholesky panel: columns k0 .. k0+b-1 of L (column-major, L[k*n + i]).
// R is the residual after the previous panels (row-major, symmetric), d its
// diagonal minus the columns of this panel formed so far. One workgroup.
__global__ __launch_bounds__(1024) void k_pchol_panel(int n, int k0, int b, const double* __restrict__ R,
                                                      double* __restrict__ d, double* __restrict__ L,
                                                      int32_t* __restrict__ piv, double* __restrict__ st, int last) {
  __shared__ double s_v[16];
  __shared__ int s_i[16];
  __shared__ double s_lp[kPanel];
  __shared__ int s_p;
  __shared__ double s_dp;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nw = blockDim.x >> 6;
  for (int j = 0; j < b; ++j) {
    const int k = k0 + j;
    // pivot: largest residual diagonal, lowest index among equals
    double bv = -1e300;
    int bi = n;
    for (int i = tid; i < n; i += blockDim.x) {
      const double v = d[i];
      if (v > bv) { bv = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_down(bv, off);
      const int oi = __shfl_down(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { s_v[wid] = bv; s_i[wid] = bi; }
    __syncthreads();
    if (tid == 0) {
      double v = s_v[0];
      int p = s_i[0];
      for (int w = 1; w < nw; ++w)
        if (s_v[w] > v || (s_v[w] == v && s_i[w] < p)) { v = s_v[w]; p = s_i[w]; }
      if (p >= n) p = 0;
      s_p = p;
      s_dp = v;
      piv[k] = p;
      st[kSpMinPiv] = fmin(st[kSpMinPiv], v);
    }
    __syncthreads();
    const int p = s_p;
    const double dp = s_dp;
    if (tid < j) s_lp[tid] = L[(int64_t)(k0 + tid) * n + p];
    __syncthreads();
    const double inv = dp > 0.0 ? 1.0 / sqrt(dp) : 0.0;
    for (int i = tid; i < n; i += blockDim.x) {
      double c = R[(int64_t)p * n + i];
      for (int t = 0; t < j; ++t) c -= L[(int64_t)(k0 + t) * n + i] * s_lp[t];
      c *= inv;
      L[(int64_t)k * n + i] = c;
      d[i] = (i == p) ? 0.0 : d[i] - c * c;
    }
    __syncthreads();
  }
  if (last) {
    __shared__ double s_red[16];
    double a = 0.0;
    for (int i = tid; i < n; i += blockDim.x) a += d[i];
    const double r = block_sum_1024(a, s_red);
    if (tid == 0) st[kSpResid] = r;
  }
}

// ---------------------------------------------------------------------------
// Spectral bounds for X_0 by Lanczos. Gershgorin intervals of F' overestimate
// its spectrum several-fold (for (H2O)_32 def2-TZVP: [-371, 400] Ha against
// [-19.3, 43.7]), and SP2 needs about 2 log2(range / gap) iterations, so kLzSteps
// Lanczos steps from a fixed pseudo-random start give the extreme Ritz values
// theta and their residuals rho = beta_k |y_k| (an eigenvalue lies within rho
// of theta); the bounds become theta_min - m, theta_max + m with m = max(rho,
// 1e-3 of the Ritz range), intersected with Gershgorin's. Lanczos finds the
// extreme eigenvalues first; should a bound still fall inside the spectrum,
// the eigenvalues beyond it map slightly outside [0, 1] and the iteration
// still takes them to 0 or 1, and the step's idempotency and trace checks
// stand behind it as before.
// Work vectors: lz[0..n) v, lz[n..2n) previous v, lz[2n..3n) w, then
// kLzParts partial dot products, then alpha_k, beta_k (2 kLzSteps).
// ---------------------------------------------------------------------------

// v_0: a fixed +-(1 + small) pattern, normalised; previous v = 0; coefficients cleared.
__global__ __launch_bounds__(1024) void k_lz_start(int n, double* __restrict__ lz) {
  __shared__ double s_red[16];
  double* v = lz;
  double* vp = lz + n;
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t hsh = (uint32_t)i * 2654435761u;
    const double x = ((hsh >> 15) & 1u ? 1.0 : -1.0) * (1.0 + (double)((hsh >> 20) & 255u) / 1024.0);
    v[i] = x;
    vp[i] = 0.0;
    a += x * x;
  }
  const double t = block_sum_1024(a, s_red);
  __shared__ double s_inv;
  if (threadIdx.x == 0) s_inv = 1.0 / sqrt(t);
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] *= s_inv;
  double* ab = lz + 3 * (int64_t)n + kLzParts;
  for (int i = threadIdx.x; i < 2 * kLzSteps; i += blockDim.x) ab[i] = 0.0;
}

// w = F v (one wavefront per row of the symmetric F), and per workgroup the partial v . w
__global__ __launch_bounds__(256) void k_lz_matvec(int n, const double* __restrict__ F, double* __restrict__ lz) {
  __shared__ double s_p[4];
  const double* v = lz;
  double* w = lz + 2 * (int64_t)n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = blockIdx.x * 4 + wv;
  double acc = 0.0;
  if (i < n) {
    const double* row = F + (int64_t)i * n;
    for (int j = lane; j < n; j += 64) acc += row[j] * v[j];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) {
    if (i < n) w[i] = acc;
    s_p[wv] = i < n ? acc * v[i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) lz[3 * (int64_t)n + blockIdx.x] = s_p[0] + s_p[1] + s_p[2] + s_p[3];
}

// alpha = v . w (the partials in order); w -= alpha v + beta_{k-1} v_prev; beta = |w|;
// v_prev = v; v = w / beta (a zero beta ends the recurrence: v stays 0).
__global__ __launch_bounds__(1024) void k_lz_update(int n, int nparts, int k, double* __restrict__ lz) {
  __shared__ double s_red[16];
  __shared__ double s_alpha, s_beta;
  double* v = lz;
  double* vp = lz + n;
  double* w = lz + 2 * (int64_t)n;
  const double* part = lz + 3 * (int64_t)n;
  double* ab = lz + 3 * (int64_t)n + kLzParts;
  {
    // alpha from the matrix-vector kernel's partials: strided per thread, then the tree
    double pa = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) pa += part[i];
    const double a = block_sum_1024(pa, s_red);
    if (threadIdx.x == 0) s_alpha = a;
    __syncthreads();
  }
  const double alpha = s_alpha, bprev = k > 0 ? ab[2 * (k - 1) + 1] : 0.0;
  double q = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double x = w[i] - alpha * v[i] - bprev * vp[i];
    w[i] = x;
    q += x * x;
  }
  const double t = block_sum_1024(q, s_red);
  if (threadIdx.x == 0) {
    const double b = sqrt(t);
    s_beta = b > 1e-12 * (fabs(alpha) + 1.0) ? b : 0.0;
    ab[2 * k] = alpha;
    ab[2 * k + 1] = s_beta;
  }
  __syncthreads();
  const double beta = s_beta;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double vi = v[i];
    vp[i] = vi;
    v[i] = beta > 0.0 ? w[i] / beta : 0.0;
  }
}

// Eigenvalues of the Lanczos tridiagonal below x (Sturm count).
__device__ int lz_count_below(const double* ab, int m, double x) {
  int c = 0;
  double d = 1.0;
  for (int i = 0; i < m; ++i) {
    const double b2 = i > 0 ? ab[2 * (i - 1) + 1] * ab[2 * (i - 1) + 1] : 0.0;
    d = (ab[2 * i] - x) - (i > 0 ? b2 / d : 0.0);
    if (d == 0.0) d = -1e-300;
    if (d < 0.0) ++c;
  }
  return c;
}

// Residual beta_m |y_m| of the Ritz value theta (y the normalised eigenvector of T).
__device__ double lz_residual(const double* ab, int m, double theta) {
  double y0 = 1.0, y1 = 0.0, nrm = 1.0;
  double last = 1.0;
  for (int i = 0; i + 1 < m; ++i) {
    const double b = ab[2 * i + 1];
    const double bp = i > 0 ? ab[2 * (i - 1) + 1] : 0.0;
    const double y2 = ((theta - ab[2 * i]) * y0 - bp * y1) / b;
    y1 = y0;
    y0 = y2;
    nrm += y2 * y2;
    last = y2;
    if (nrm > 1e200) {   // rescale
      const double sc = 1.0 / sqrt(nrm);
      y0 *= sc; y1 *= sc; last *= sc; nrm = 1.0;
    }
  }
  return fabs(ab[2 * (m - 1) + 1] * last) / sqrt(nrm);
}

// One wavefront: the extreme Ritz values by multisection (lanes 0-31 close in on
// the lowest, lanes 32-63 on the highest: each round tests 32 points of the
// current interval by Sturm counts and keeps the sub-interval where the count
// steps), their residuals, and the bounds in st (intersected with the
// Gershgorin bounds already there).
__global__ __launch_bounds__(64) void k_lz_bounds(int n, const double* __restrict__ lz, double* __restrict__ st) {
  __shared__ double s_ab[2 * kLzSteps];
  if (blockIdx.x != 0) return;
  const int lane = threadIdx.x;
  const double* ab_g = lz + 3 * (int64_t)n + kLzParts;
  for (int i = lane; i < 2 * kLzSteps; i += 64) s_ab[i] = ab_g[i];
  __syncthreads();
  const double* ab = s_ab;
  int m = 0;
  while (m < kLzSteps) {
    ++m;
    if (ab[2 * (m - 1) + 1] == 0.0) break;
  }
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? fabs(ab[2 * (i - 1) + 1]) : 0.0) + (i + 1 < m ? fabs(ab[2 * i + 1]) : 0.0);
    lo = fmin(lo, ab[2 * i] - r);
    hi = fmax(hi, ab[2 * i] + r);
  }
  // half 0 looks for the point where the count of eigenvalues below reaches 1,
  // half 1 where it reaches m
  const int half = lane >> 5, j = lane & 31;
  const int target = half == 0 ? 1 : m;
  double a = lo, b = hi;
  for (int round = 0; round < 16; ++round) {
    if (!(b - a > 1e-13 * fmax(1.0, fabs(a) + fabs(b)))) break;   // uniform within the half
    const double x = a + (b - a) * (double)(j + 1) / 33.0;
    const bool reached = lz_count_below(ab, m, x) >= target;
    const unsigned long long bal = __ballot(reached);
    const unsigned mask = (unsigned)(half == 0 ? bal : bal >> 32);
    // the count is monotone in x: the first point that reaches the target closes the interval
    const int first = mask ? __ffs(mask) - 1 : 32;
    const double na = first == 0 ? a : a + (b - a) * (double)first / 33.0;
    const double nb = first == 32 ? b : a + (b - a) * (double)(first + 1) / 33.0;
    a = na;
    b = nb;
  }
  __shared__ double s_t[2];
  if (j == 0) s_t[half] = 0.5 * (a + b);
  __syncthreads();
  if (lane != 0) return;
  const double tmin = s_t[0], tmax = s_t[1];
  const bool exact = ab[2 * (m - 1) + 1] == 0.0;   // invariant subspace: the Ritz values are eigenvalues
  const double rmin = exact ? 0.0 : lz_residual(ab, m, tmin);
  const double rmax = exact ? 0.0 : lz_residual(ab, m, tmax);
  const double pad = 1e-3 * (tmax - tmin);
  const double emin = tmin - fmax(rmin, pad), emax = tmax + fmax(rmax, pad);
  if (!(emin < emax) || !isfinite(emin) || !isfinite(emax)) return;   // keep Gershgorin's
  st[kSpEmin] = fmax(st[kSpEmin], emin);
  st[kSpEmax] = fmin(st[kSpEmax], emax);
}

// The Lanczos bounds after launch_gersh: kLzSteps steps, kernels through launch_k
// (so that they can be recorded as a graph). lz: 3 n + kLzParts + 2 kLzSteps doubles.
hipError_t launch_sp2_lanczos(int n, const double* F, double* lz, double* st, hipStream_t s) {
  const int nparts = (n + 3) / 4;
  if (nparts > kLzParts) return hipErrorInvalidValue;
  hipError_t e = launch_k(k_lz_start, dim3(1), dim3(1024), s, n, lz);
  for (int k = 0; k < kLzSteps && e == hipSuccess; ++k) {
    e = launch_k(k_lz_matvec, dim3((unsigned)nparts), dim3(256), s, n, F, lz);
    if (e == hipSuccess) e = launch_k(k_lz_update, dim3(1), dim3(1024), s, n, nparts, k, lz);
  }
  if (e == hipSuccess) e = launch_k(k_lz_bounds, dim3(1), dim3(64), s, n, (const double*)lz, st);
  return e;
}

// Spectral bounds of F from those of the previous step's F' (Fp) without a
// Lanczos run: by Weyl's inequality every eigenvalue moves by at most
// ||F - Fp||_2 <= ||F - Fp||_F, so [lo_p - d, hi_p + d] with d = ||F - Fp||_F
// contains the spectrum of F whenever [lo_p, hi_p] contained that of Fp.
// k_weyl_part: per-workgroup partial sums of the squared differences (kSp2Groups
// workgroups of 256 threads), and Fp <- F in the same pass.
__global__ __launch_bounds__(256) void k_weyl_part(int n, const double* __restrict__ F, double* __restrict__ Fp,
                                                   double* __restrict__ st) {
  const int64_t nn = (int64_t)n * n;
  double a = 0.0;
  for (int64_t idx = blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)kSp2Groups * 256) {
    const double f = F[idx], d = f - Fp[idx];
    a += d * d;
    Fp[idx] = f;
  }
  sp2_partials(a, 0.0, st);
}

// One block: d from the partials (fixed order), then the widened previous bounds
// pb intersected with the Gershgorin bounds already in st.
__global__ __launch_bounds__(kSp2Groups) void k_weyl_bounds(const double* __restrict__ pb, double* __restrict__ st) {
  __shared__ double s_a[kSp2Groups];
  s_a[threadIdx.x] = st[kSpPart + 2 * threadIdx.x];
  __syncthreads();
  for (int w = kSp2Groups / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s_a[threadIdx.x] += s_a[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double d = sqrt(s_a[0]);
    const double lo = pb[0] - d, hi = pb[1] + d;
    if (isfinite(lo) && isfinite(hi) && lo < hi) {
      st[kSpEmin] = fmax(st[kSpEmin], lo);
      st[kSpEmax] = fmin(st[kSpEmax], hi);
    }
  }
}

// pb <- the bounds the purification used
__global__ void k_save_bounds(const double* __restrict__ st, double* __restrict__ pb) {
  if (threadIdx.x == 0) {
    pb[0] = st[kSpEmin];
    pb[1] = st[kSpEmax];
  }
}

hipError_t launch_sp2_weyl(int n, const double* F, double* Fp, const double* pb, double* st, hipStream_t s) {
  hipError_t e = launch_k(k_weyl_part, dim3(kSp2Groups), dim3(256), s, n, F, Fp, st);
  if (e == hipSuccess) e = launch_k(k_weyl_bounds, dim3(1), dim3(kSp2Groups), s, pb, st);
  return e;
}

hipError_t launch_save_bounds(const double* st, double* pb, hipStream_t s) {
  return launch_k(k_save_bounds, dim3(1), dim3(64), s, st, pb);
}

hipError_t launch_gersh(int n, const double* F, double* lohi, double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_gersh, dim3(n), dim3(256), 0, s, n, F, lohi);
  hipLaunchKernelGGL(k_sp2_start, dim3(1), dim3(256), 0, s, n, lohi, st);
  return hipGetLastError();
}

hipError_t launch_sp2_init(int n, const double* F, const double* st, double* X, hipStream_t s) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(k_sp2_init, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, n, F, st, X);
  return hipGetLastError();
}

// The first decision, from X_0 (after launch_sp2_init).
hipError_t launch_sp2_begin(int n, const double* X, int nocc, double* st, hipStream_t s) {
  const hipError_t e = launch_k(k_sp2_norms, dim3(kSp2Groups), dim3(256), s, n, X, st);
  if (e != hipSuccess) return e;
  return launch_k(k_sp2_decide, dim3(1), dim3(kSp2Groups), s, nocc, st);
}

// One iteration after the product's split-K partials are in ws: update X, decide the next branch.
hipError_t launch_sp2_update(int n, int nsplit, const double* ws, double* X, int nocc, double* st, hipStream_t s) {
  const hipError_t e = launch_k(k_sp2_reduce_update, dim3(kSp2Groups), dim3(256), s, n, nsplit, ws,
                                (int64_t)n * n, X, st);
  if (e != hipSuccess) return e;
  return launch_k(k_sp2_decide, dim3(1), dim3(kSp2Groups), s, nocc, st);
}

// Upper 64 x 64 tiles of an n x n purification step; k_sp2_step takes n when this is
// at most kSp2Groups.
int sp2_step_tiles(int n) {
  const int ntn = (n + kSqT - 1) / kSqT;
  return ntn * (ntn + 1) / 2;
}

// One fused iteration, X -> Xn, then the next decision.
hipError_t launch_sp2_step(int n, const double* X, double* Xn, int nocc, double* st, hipStream_t s) {
  const int ntiles = sp2_step_tiles(n);
  if (ntiles > kSp2Groups) return hipErrorInvalidValue;
  const hipError_t e = launch_k(k_sp2_step, dim3(kSp2Groups), dim3(kSqN), s, n, ntiles, X, Xn, st);
  if (e != hipSuccess) return e;
  return launch_k(k_sp2_decide, dim3(1), dim3(kSp2Groups), s, nocc, st);
}

hipError_t launch_diag(int n, const double* R, double* d, hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, R, d);
  return hipGetLastError();
}

hipError_t launch_pchol_panel(int n, int k0, int b, const double* R, double* d, double* L, int32_t* piv, double* st,
                              int last, hipStream_t s) {
  if (b < 1 || b > kPanel) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pchol_panel, dim3(1), dim3(1024), 0, s, n, k0, b, R, d, L, piv, st, last);
  return hipGetLastError();
}

}  // namespace dfks
