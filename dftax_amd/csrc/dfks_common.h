// Shared device-side definitions for the DF-KS library (gfx950, FP64 path).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <tuple>
#include <utility>

#include "c2s_tables.h"

namespace dfks {

constexpr int kLMax = 4;              // largest l of an orbital shell (def2-TZVP f; aux g)
constexpr int kBoysNg = 401;          // table nodes t = 0, 0.1, ..., 40
constexpr int kBoysN = 25;            // F_0 .. F_24 tabulated
constexpr int kBoysOrder = 6;         // local Taylor degree
constexpr double kBoysDt = 0.1;
constexpr double kBoysTmax = 40.0;
constexpr double kPi = 3.14159265358979323846;

// Sparse cart->sph lists: for l and spherical index s, up to 12 (cart, coef).
struct C2SList {
  int n[7][13];
  int cart[7][13][12];
  double coef[7][13][12];
};

// SP2 purification state vector (doubles) and limits, shared by host and kernels
enum Sp2State {
  kSpEmin = 0, kSpEmax = 1, kSpFlag = 2, kSpCount = 3, kSpIter = 4, kSpErr = 5, kSpTrace = 6,
  kSpResid = 7, kSpMinPiv = 8, kSpHist = 16
};
constexpr int kSp2MaxIter = 128;
constexpr int kSp2Groups = 512;       // workgroups of the SP2 update (partial trace sums)
constexpr int kSpPart = kSpHist + kSp2MaxIter;  // their partial sums: 2 kSp2Groups doubles
constexpr int kSp2State = kSpPart + 2 * kSp2Groups;
constexpr int kSp2Block = 4;          // purification iterations per graph launch (divides kSp2MaxIter)
// widest compact-grid block (functions) the fused mixed-precision XC kernels take
constexpr int kXfMaxN = 2048;
// One bin of the compact XC grid inside a single fused mixed-precision launch: its
// width n (functions per block), blocks, Z-column tiles and K chunks, the offsets of
// its AO rows (floats), column lists (ints), block products (floats), points, partial
// sums (doubles) and W coefficients (floats), and the first workgroup id of its
// blocks in each of the three kernels (Z tiles, per-point pass, Vh tiles).
struct XcfBin {
  int32_t n, nb, nt, nks;
  int64_t ao_off, col_off, q_off, pt0, part_off, coef_off;
  int32_t wz0, wp0, wv0, pad;
};
constexpr int kPanel = 32;            // pivoted Cholesky panel width
constexpr int kLzSteps = 24;          // Lanczos steps for the SP2 spectral bounds
constexpr int kLzParts = 4096;        // partial dot products of its matrix-vector kernel (n <= 4 kLzParts)

// Graph recording without stream capture. While a thread has a GraphRec
// installed in g_rec, launch_k appends each kernel as a node of g_rec->g that
// depends on the previous one, instead of launching it. Building a graph this
// way puts no stream into capture mode, so other threads' work on other
// streams (including library calls that touch the legacy stream) cannot
// collide with it.
struct GraphRec {
  hipGraph_t g = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;
};
extern thread_local GraphRec* g_rec;

template <typename... P, typename... A>
inline hipError_t launch_k(void (*k)(P...), dim3 grid, dim3 block, hipStream_t st, A&&... a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  if (!g_rec) {
    hipLaunchKernelGGL(k, grid, block, 0, st, static_cast<P>(std::forward<A>(a))...);
    return hipGetLastError();
  }
  std::tuple<P...> vals(static_cast<P>(std::forward<A>(a))...);
  void* args[sizeof...(P)];
  std::apply([&](auto&... v) {
    int i = 0;
    ((args[i++] = static_cast<void*>(&v)), ...);
  }, vals);
  hipKernelNodeParams p{};
  p.func = reinterpret_cast<void*>(k);
  p.gridDim = grid;
  p.blockDim = block;
  p.sharedMemBytes = 0;
  p.kernelParams = args;
  p.extra = nullptr;
  hipGraphNode_t node = nullptr;
  const hipError_t e = hipGraphAddKernelNode(&node, g_rec->g, g_rec->last ? &g_rec->last : nullptr,
                                             g_rec->last ? 1 : 0, &p);
  if (e == hipSuccess) g_rec->last = node;
  else g_rec->err = e;
  return e;
}

// Optional epilogue of the exchange product T (rows m' = m * nq + q, columns
// i): out[m'] = sum_i sg[i] cf[i * nao + m] T[m', i], the per-row terms of the
// fitted density gamma_q = sum_m out[m * nq + q], computed from the tile in
// registers so T is not read again for it. Needs the whole N range in one tile
// and nq at least the tile's rows (a tile then spans at most two m).
struct RowDot {
  const double* cf = nullptr;
  const double* sg = nullptr;
  int nao = 0, nq = 1;
  double* out = nullptr;
};

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nsph(int l) { return 2 * l + 1; }
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Shell-major basis on the device (structure of arrays).
struct ShellSet {
  int32_t n;              // shells
  const int32_t* l;
  const int32_t* nprim;
  const int32_t* prim0;
  const int32_t* ao0;     // first function in the final (spherical or cartesian) span
  const double* centre;   // [3n]
  const double* exps;
  const double* coefs;    // primitive norm (and the l<=1 contracted norm) folded in
  int32_t spherical;
  // operator of the two- and three-centre integrals taken over this set (the
  // auxiliary side): 0 the Coulomb operator 1/r, > 0 the long-range part
  // erf(omega r)/r of a range-separated hybrid
  double omega = 0.0;
};

// Final-span size of a shell.
__host__ __device__ inline int nfun(int l, int spherical) {
  return (spherical && l >= 2) ? nsph(l) : ncart(l);
}

// Compile-time cartesian component table.
template <int L>
__device__ __forceinline__ constexpr int cart_exp(int c, int x) {
  if constexpr (L == 0) return kCart0[c][x];
  else if constexpr (L == 1) return kCart1[c][x];
  else if constexpr (L == 2) return kCart2[c][x];
  else if constexpr (L == 3) return kCart3[c][x];
  else if constexpr (L == 4) return kCart4[c][x];
  else if constexpr (L == 5) return kCart5[c][x];
  else return kCart6[c][x];
}

template <int L>
__device__ __forceinline__ constexpr double c2s_coef(int c, int s) {
  if constexpr (L == 0) return kC2S0[c][s];
  else if constexpr (L == 1) return kC2S1[c][s];
  else if constexpr (L == 2) return kC2S2[c][s];
  else if constexpr (L == 3) return kC2S3[c][s];
  else if constexpr (L == 4) return kC2S4[c][s];
  else if constexpr (L == 5) return kC2S5[c][s];
  else return kC2S6[c][s];
}

// F_m(T) for m = 0..NMAX. Top order by the tabulated degree-6 Taylor expansion
// (nearest node of a 0.1-spaced grid), lower orders by the stable downward
// recursion F_{m-1} = (2T F_m + e^{-T}) / (2m - 1); beyond T = 40 the
// asymptote Gamma(m+1/2) / (2 T^{m+1/2}) with its upward recursion.
template <int NMAX>
__device__ __forceinline__ void boys_array(double T, double* F, const double* __restrict__ tbl) {
  if (T < kBoysTmax) {
    int idx = __double2int_rn(T * 10.0);
    double dx = idx * kBoysDt - T;
    const double* row = tbl + idx * kBoysN + NMAX;
    double acc = row[kBoysOrder];
#pragma unroll
    for (int k = kBoysOrder - 1; k >= 0; --k) acc = row[k] + acc * dx * (1.0 / (k + 1));
    F[NMAX] = acc;
    if constexpr (NMAX > 0) {
      double e = exp(-T);
      double t2 = 2.0 * T;
#pragma unroll
      for (int m = NMAX; m > 0; --m) F[m - 1] = (t2 * F[m] + e) * (1.0 / (2 * m - 1));
    }
  } else {
    double it = 1.0 / T;
    F[0] = 0.88622692545275801365 * sqrt(it);  // sqrt(pi)/2 / sqrt(T)
#pragma unroll
    for (int m = 0; m < NMAX; ++m) F[m + 1] = F[m] * (2 * m + 1) * 0.5 * it;
  }
}

// Runtime-order variant (same algorithm) for kernels without a fixed order.
__device__ __forceinline__ void boys_array_rt(int nmax, double T, double* F, const double* __restrict__ tbl) {
  if (T < kBoysTmax) {
    int idx = __double2int_rn(T * 10.0);
    double dx = idx * kBoysDt - T;
    const double* row = tbl + idx * kBoysN + nmax;
    double acc = row[kBoysOrder];
    for (int k = kBoysOrder - 1; k >= 0; --k) acc = row[k] + acc * dx / (k + 1);
    F[nmax] = acc;
    double e = exp(-T);
    for (int m = nmax; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + e) / (2 * m - 1);
  } else {
    double it = 1.0 / T;
    F[0] = 0.88622692545275801365 * sqrt(it);
    for (int m = 0; m < nmax; ++m) F[m + 1] = F[m] * (2 * m + 1) * 0.5 * it;
  }
}

// Boys values of the operator of an integral: omega = 0 the Coulomb operator,
// F_m(T); omega > 0 the attenuated erf(omega r)/r, s^(m+1/2) F_m(s T) with
// s = omega^2 / (omega^2 + rho) and rho the reduced exponent of the two charge
// distributions (dftax/integrals/eri3c_bucketed.py:340-346).
template <int NMAX>
__device__ __forceinline__ void boys_op(double T, double rho, double omega, double* F, const double* __restrict__ tbl) {
  if (omega > 0.0) {
    const double s = omega * omega / (omega * omega + rho);
    boys_array<NMAX>(s * T, F, tbl);
    double f = sqrt(s);
#pragma unroll
    for (int m = 0; m <= NMAX; ++m) {
      F[m] *= f;
      f *= s;
    }
  } else {
    boys_array<NMAX>(T, F, tbl);
  }
}

__device__ __forceinline__ void boys_op_rt(int nmax, double T, double rho, double omega, double* F,
                                           const double* __restrict__ tbl) {
  if (omega > 0.0) {
    const double s = omega * omega / (omega * omega + rho);
    boys_array_rt(nmax, s * T, F, tbl);
    double f = sqrt(s);
    for (int m = 0; m <= nmax; ++m) {
      F[m] *= f;
      f *= s;
    }
  } else {
    boys_array_rt(nmax, T, F, tbl);
  }
}

}  // namespace dfks
