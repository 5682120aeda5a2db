// Bandwidth-bound helpers of the Fock build and the SCF step (FP64).
//
//  * k_gamma_part + k_colsum
//               gamma_Q = sum_{m,i} s_i Cf[m,i] T[m,Q,i]   (fitted density, from the
//               exchange intermediate T = B^T Cf, so B is not re-read for it)
//  * k_jvec     J[mn] = sum_Q B[mn,Q] gamma_Q              (one read of B, HBM bound)
//  * k_fock     F = hcore + J + cK sym(K) + sym(Vxc)        (upper-triangle inputs; hcore may be absent)
//  * deterministic two-pass reductions for the energy terms
//  * DIIS: Gram row, bordered solve (one lane, FP64, partial pivoting),
//    extrapolation (dftax/ks/scf.py:252-276)
#include <cstdint>

#include "dfks_common.h"

namespace dfks {

constexpr int kRedBlocks = 512;

// partial[b] = sum over a grid-stride slice of x[i]*y[i] (y may be null -> x only)
__global__ __launch_bounds__(256) void k_dot_partial(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += y ? x[i] * y[i] : x[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// out[0] = scale * sum(part[0..np))  (+= if accumulate)
__global__ __launch_bounds__(256) void k_sum_final(int np, const double* __restrict__ part, double scale,
                                                   double* __restrict__ out, int accumulate) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = accumulate ? out[0] + scale * s[0] : scale * s[0];
}

hipError_t launch_dot(int64_t n, const double* x, const double* y, double* part, double scale, double* out,
                      int accumulate, hipStream_t st) {
  int nb = (int)((n + 255) / 256);
  if (nb > kRedBlocks) nb = kRedBlocks;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, st, n, x, y, part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, nb, part, scale, out, accumulate);
  return hipGetLastError();
}

// gamma_q = sum_{m,i} s_i Cf[m,i] T[m,q,i] in two deterministic passes.
// Pass 1: one wavefront per (m, block of 8 q): lanes over i, wave reduction;
// partial[m*nq + q] written. Pass 2: column sums over m, fixed order.
// T layout: T[(m*nq + q)*nf + i]; Cf column-major: Cf[i*nao + m].
template <typename TT>
__global__ __launch_bounds__(256) void k_gamma_part(int nao, int nq, int nf, const TT* __restrict__ T,
                                                    const double* __restrict__ Cf, const double* __restrict__ sgn,
                                                    double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int qblocks = (nq + 7) / 8;
  if (w >= (int64_t)nao * qblocks) return;
  const int m = (int)(w / qblocks);
  const int q0 = (int)(w % qblocks) * 8;
  // this lane's coefficients s_i Cf[m, i] for i = lane, lane + 64, ...
  double c[8];
  int ni = 0;
  for (int i = lane; i < nf && ni < 8; i += 64) c[ni++] = sgn[i] * Cf[(int64_t)i * nao + m];
  for (int q = q0; q < q0 + 8 && q < nq; ++q) {
    const TT* row = T + ((int64_t)m * nq + q) * nf;
    double a = 0.0;
    int k = 0;
    for (int i = lane; i < nf && k < 8; i += 64, ++k) a += c[k] * (double)row[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) part[(int64_t)m * nq + q] = a;
  }
}

// out[q] (+)= sum_m part[m][q]: a workgroup per 16 columns, its 16 row groups
// (m = g, g + 16, ...) summed in a fixed order through LDS.
__global__ __launch_bounds__(256) void k_colsum(int nrows, int ncols, const double* __restrict__ part,
                                                double* __restrict__ out, int accumulate) {
  __shared__ double s_part[16][17];
  const int c = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int q = blockIdx.x * 16 + c;
  double s = 0.0;
  if (q < ncols)
    for (int m = g; m < nrows; m += 16) s += part[(int64_t)m * ncols + q];
  s_part[g][c] = s;
  __syncthreads();
  if (g == 0 && q < ncols) {
    double t = 0.0;
    for (int k = 0; k < 16; ++k) t += s_part[k][c];
    out[q] = accumulate ? out[q] + t : t;
  }
}

template <typename TT>
static hipError_t gamma_impl(int nao, int nq, int nf, const TT* T, const double* Cf, const double* sgn, double* gamma,
                             double* part, int accumulate, hipStream_t st) {
  if (nf > 512) return hipErrorInvalidValue;  // eight coefficients per lane
  const int64_t waves = (int64_t)nao * ((nq + 7) / 8);
  hipLaunchKernelGGL((k_gamma_part<TT>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, nao, nq, nf, T, Cf,
                     sgn, part);
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)((nq + 15) / 16)), dim3(256), 0, st, nao, nq, part, gamma,
                     accumulate);
  return hipGetLastError();
}

hipError_t launch_colsum(int nrows, int ncols, const double* part, double* out, int accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)((ncols + 15) / 16)), dim3(256), 0, st, nrows, ncols, part, out,
                     accumulate);
  return hipGetLastError();
}

hipError_t launch_gamma(int nao, int nq, int nf, const double* T, const double* Cf, const double* sgn, double* gamma,
                        double* part, int accumulate, hipStream_t st) {
  return gamma_impl<double>(nao, nq, nf, T, Cf, sgn, gamma, part, accumulate, st);
}

hipError_t launch_gamma_f32(int nao, int nq, int nf, const float* T, const double* Cf, const double* sgn,
                            double* gamma, double* part, int accumulate, hipStream_t st) {
  return gamma_impl<float>(nao, nq, nf, T, Cf, sgn, gamma, part, accumulate, st);
}

// J[r] = sum_q B[r*nq + q] g[q]; one wavefront per row, 4 rows per block per step.
// sym_n > 0: the rows are the pairs r = l sym_n + m of a B symmetric in (l, m); only
// the rows m >= l are read (half of B), and each sum goes to J[l][m] and J[m][l].
template <typename TB, bool VEC>
__global__ __launch_bounds__(256) void k_jvec(int64_t nrows, int nq, const TB* __restrict__ B,
                                              const double* __restrict__ g, double* __restrict__ J, int accumulate,
                                              int sym_n) {
  extern __shared__ double s_g[];
  for (int i = threadIdx.x; i < nq; i += blockDim.x) s_g[i] = g[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  constexpr int PER = 16 / sizeof(TB);
  typedef TB vecT __attribute__((ext_vector_type(PER)));
  for (int64_t r = blockIdx.x * 4 + (threadIdx.x >> 6); r < nrows; r += (int64_t)gridDim.x * 4) {
    int64_t rt = r;   // the mirrored row
    if (sym_n > 0) {
      const int64_t l = r / sym_n, m = r % sym_n;
      if (m < l) continue;
      rt = m * sym_n + l;
    }
    const TB* row = B + r * nq;
    double acc = 0.0;
    if (VEC) {
      // four 16-byte loads in flight per lane, four partial sums (fixed order)
      const vecT* rv = reinterpret_cast<const vecT*>(row);
      const int nv = nq / PER;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int q = lane;
      for (; q + 192 < nv; q += 256) {
        const vecT x0 = rv[q], x1 = rv[q + 64], x2 = rv[q + 128], x3 = rv[q + 192];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
          a0 += (double)x0[j] * s_g[q * PER + j];
          a1 += (double)x1[j] * s_g[(q + 64) * PER + j];
          a2 += (double)x2[j] * s_g[(q + 128) * PER + j];
          a3 += (double)x3[j] * s_g[(q + 192) * PER + j];
        }
      }
      for (; q < nv; q += 64) {
        const vecT x = rv[q];
#pragma unroll
        for (int j = 0; j < PER; ++j) a0 += (double)x[j] * s_g[q * PER + j];
      }
      acc = (a0 + a1) + (a2 + a3);
    } else {
      for (int q = lane; q < nq; q += 64) acc += (double)row[q] * s_g[q];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) {
      J[r] = accumulate ? J[r] + acc : acc;
      if (rt != r) J[rt] = accumulate ? J[rt] + acc : acc;
    }
  }
}

// sym_n > 0: B (nrows = sym_n^2 rows) is symmetric in its two orbital indices, and only
// the rows of the upper triangle are read (see k_jvec).
hipError_t launch_jvec(int64_t nrows, int nq, const double* B, const double* g, double* J, int sym_n, hipStream_t st) {
  if (sym_n > 0 && (int64_t)sym_n * sym_n != nrows) return hipErrorInvalidValue;
  int64_t nb = (nrows + 3) / 4;
  if (nb > 65536) nb = 65536;
  const bool vec = (nq % 2 == 0) && ((uintptr_t)B % 16 == 0);
  if (vec)
    hipLaunchKernelGGL((k_jvec<double, true>), dim3((unsigned)nb), dim3(256), nq * sizeof(double), st, nrows, nq, B, g,
                       J, 0, sym_n);
  else
    hipLaunchKernelGGL((k_jvec<double, false>), dim3((unsigned)nb), dim3(256), nq * sizeof(double), st, nrows, nq, B,
                       g, J, 0, sym_n);
  return hipGetLastError();
}

// J[r] += sum_q B[r, q] g[q] (one auxiliary slab of a recomputed 3-centre tensor)
hipError_t launch_jvec_acc(int64_t nrows, int nq, const double* B, const double* g, double* J, hipStream_t st) {
  int64_t nb = (nrows + 3) / 4;
  if (nb > 65536) nb = 65536;
  const bool vec = (nq % 2 == 0) && ((uintptr_t)B % 16 == 0);
  if (vec)
    hipLaunchKernelGGL((k_jvec<double, true>), dim3((unsigned)nb), dim3(256), nq * sizeof(double), st, nrows, nq, B, g,
                       J, 1, 0);
  else
    hipLaunchKernelGGL((k_jvec<double, false>), dim3((unsigned)nb), dim3(256), nq * sizeof(double), st, nrows, nq, B,
                       g, J, 1, 0);
  return hipGetLastError();
}

hipError_t launch_jvec_f32(int64_t nrows, int nq, const float* B, const double* g, double* J, int sym_n,
                           hipStream_t st) {
  if (sym_n > 0 && (int64_t)sym_n * sym_n != nrows) return hipErrorInvalidValue;
  int64_t nb = (nrows + 3) / 4;
  if (nb > 65536) nb = 65536;
  const bool vec = (nq % 4 == 0) && ((uintptr_t)B % 16 == 0);
  if (vec)
    hipLaunchKernelGGL((k_jvec<float, true>), dim3((unsigned)nb), dim3(256), nq * sizeof(double), st, nrows, nq, B, g,
                       J, 0, sym_n);
  else
    hipLaunchKernelGGL((k_jvec<float, false>), dim3((unsigned)nb), dim3(256), nq * sizeof(double), st, nrows, nq, B,
                       g, J, 0, sym_n);
  return hipGetLastError();
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

__global__ void k_to_f64(int64_t n, const float* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}

hipError_t launch_to_f64(int64_t n, const float* x, double* y, hipStream_t st) {
  if (n == 0) return hipSuccess;
  int64_t nb = (n + 255) / 256;
  if (nb > 131072) nb = 131072;
  hipLaunchKernelGGL(k_to_f64, dim3((unsigned)nb), dim3(256), 0, st, n, x, y);
  return hipGetLastError();
}

hipError_t launch_to_f32(int64_t n, const double* x, float* y, hipStream_t st) {
  if (n == 0) return hipSuccess;
  int64_t nb = (n + 255) / 256;
  if (nb > 131072) nb = 131072;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nb), dim3(256), 0, st, n, x, y);
  return hipGetLastError();
}

// F = H + J + cK * K + (Vh + Vh^T) with K a full symmetric matrix and Vh the
// half product Phi^T W. Either K or Vh may be null.
__global__ void k_fock(int n, const double* __restrict__ H, const double* __restrict__ J, const double* __restrict__ K,
                       double cK, const double* __restrict__ Vh, double* __restrict__ F) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  int r = (int)(idx / n), c = (int)(idx % n);
  double f = (H ? H[idx] : 0.0) + J[idx];
  if (K) f += cK * 0.5 * (K[idx] + K[(int64_t)c * n + r]);
  if (Vh) f += Vh[idx] + Vh[(int64_t)c * n + r];
  F[idx] = f;
}

hipError_t launch_fock(int n, const double* H, const double* J, const double* Ku, double cK, const double* Vu,
                       double* F, hipStream_t st) {
  int64_t tot = (int64_t)n * n;
  hipLaunchKernelGGL(k_fock, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, H, J, Ku, cK, Vu, F);
  return hipGetLastError();
}

// Fill the lower part of a column-major upper-triangular symmetric matrix.
__global__ void k_symfill(int n, double* __restrict__ A) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  int col = (int)(idx / n), row = (int)(idx % n);
  if (row > col) A[idx] = A[(int64_t)row * n + col];
}

hipError_t launch_symfill(int n, double* A, hipStream_t st) {
  int64_t tot = (int64_t)n * n;
  hipLaunchKernelGGL(k_symfill, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, A);
  return hipGetLastError();
}

// Metric factor: L_cm[k + q*nk] = U[q + (k0+k)*nq] / sqrt(w[k0+k]) for the kept
// (largest) eigenpairs; the B GEMM then consumes L^T directly.
__global__ void k_metric_factor(int nq, int k0, const double* __restrict__ U, const double* __restrict__ w,
                                double* __restrict__ Lt) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int nk = nq - k0;
  if (idx >= (int64_t)nk * nq) return;
  int q = (int)(idx / nk), k = (int)(idx % nk);
  Lt[idx] = U[q + (int64_t)(k0 + k) * nq] / sqrt(w[k0 + k]);
}

hipError_t launch_metric_factor(int nq, int k0, const double* U, const double* w, double* Lt, hipStream_t st) {
  int64_t tot = (int64_t)(nq - k0) * nq;
  hipLaunchKernelGGL(k_metric_factor, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, nq, k0, U, w, Lt);
  return hipGetLastError();
}

// Cf[:, i] = sqrt(|lam_i|) V[:, sel_i]; sgn_i = sign(lam_i). sel lists kept columns.
__global__ void k_factor_cols(int n, int nf, const double* __restrict__ V, const double* __restrict__ lam,
                              const int32_t* __restrict__ sel, double* __restrict__ Cf, double* __restrict__ sgn) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nf) return;
  int i = (int)(idx / n), m = (int)(idx % n);
  int c = sel[i];
  double l = lam[c];
  Cf[idx] = sqrt(fabs(l)) * V[(int64_t)c * n + m];
  if (m == 0) sgn[i] = l >= 0 ? 1.0 : -1.0;
}

hipError_t launch_factor_cols(int n, int nf, const double* V, const double* lam, const int32_t* sel, double* Cf,
                              double* sgn, hipStream_t st) {
  int64_t tot = (int64_t)n * nf;
  if (tot == 0) return hipSuccess;
  hipLaunchKernelGGL(k_factor_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, nf, V, lam, sel, Cf,
                     sgn);
  return hipGetLastError();
}

// y = a*x (+ y if acc)
__global__ void k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y, int acc) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = acc ? y[i] + a * x[i] : a * x[i];
}

hipError_t launch_axpy(int64_t n, double a, const double* x, double* y, int acc, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, x, y, acc);
  return hipGetLastError();
}

// E = A - A^T  (n x n)
__global__ void k_antisym(int n, const double* __restrict__ A, double* __restrict__ E) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  int r = (int)(idx / n), c = (int)(idx % n);
  E[idx] = A[idx] - A[(int64_t)c * n + r];
}

// out[0] = max_ij |G_ij - delta_ij| and out[1] = ||G - I||_F of the n x n matrix G
// (one workgroup, fixed-order tree; NaN counts as infinite).
__global__ __launch_bounds__(256) void k_eye_dev(int n, const double* __restrict__ G, double* __restrict__ out) {
  __shared__ double s[256], q[256];
  double m = 0.0, f = 0.0;
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = threadIdx.x; idx < nn; idx += 256) {
    const double d = G[idx] - (idx / n == idx % n ? 1.0 : 0.0);
    m = d == d ? fmax(m, fabs(d)) : HUGE_VAL;
    f += d * d;
  }
  s[threadIdx.x] = m;
  q[threadIdx.x] = f;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
      q[threadIdx.x] += q[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s[0];
    out[1] = q[0] == q[0] ? sqrt(q[0]) : HUGE_VAL;
  }
}

// dst[0] = a, dst[1] = b (a stream-ordered store of two values, no host copy)
__global__ void k_store2(double* __restrict__ dst, double a, double b) {
  if (threadIdx.x == 0) {
    dst[0] = a;
    dst[1] = b;
  }
}

hipError_t launch_store2(double* dst, double a, double b, hipStream_t st) {
  hipLaunchKernelGGL(k_store2, dim3(1), dim3(64), 0, st, dst, a, b);
  return hipGetLastError();
}

hipError_t launch_eye_dev(int n, const double* G, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_eye_dev, dim3(1), dim3(256), 0, st, n, G, out);
  return hipGetLastError();
}

hipError_t launch_antisym(int n, const double* A, double* E, hipStream_t st) {
  int64_t tot = (int64_t)n * n;
  hipLaunchKernelGGL(k_antisym, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, A, E);
  return hipGetLastError();
}

// DIIS bordered system: gram holds err_i . err_j for filled slots (m x m, row-major).
// Builds A, solves A c = (0,..,0,-1) by Gaussian elimination with partial pivoting
// on one lane and writes c[0..m). Mirrors dftax/ks/scf.py:252-276 (identity on
// empty slots, 1e-10 ridge on filled ones).
// DIIS coefficients from the bordered system [[G + 1e-10 I, -1], [-1^T, 0]] c = [0, -1]
// (unused slots become identity rows with c = 0), by Gaussian elimination with
// partial pivoting. One wavefront: lane i holds row i (m + 1 <= 17 rows, 18
// columns with the right-hand side) in registers; pivot search, row swap and the
// pivot row go through shuffles. The pivot (first largest |a_ik|), the updates
// and the back substitution run in the same order as a serial elimination.
__global__ __launch_bounds__(64) void k_diis_solve(int m, int count, const double* __restrict__ gram,
                                                   double* __restrict__ coef) {
  constexpr int kR = 17, kC = 18;
  const int lane = threadIdx.x;
  const int n = m + 1;
  double r[kC];
#pragma unroll
  for (int j = 0; j < kC; ++j) {
    double v = 0.0;
    if (lane < m) {
      if (j < m) {
        const bool vi = lane < count, vj = j < count;
        v = (vi && vj) ? gram[lane * m + j] : (lane == j ? 1.0 : 0.0);
        if (lane == j && vi) v += 1e-10;
      } else if (j == m) {
        v = lane < count ? -1.0 : 0.0;
      }
    } else if (lane == m) {
      if (j < m) v = j < count ? -1.0 : 0.0;
      else if (j == n) v = -1.0;   // right-hand side
    }
    r[j] = v;
  }
  // k stays a runtime value: every register array is indexed by unrolled j only
  for (int k = 0; k < n; ++k) {
    double rk = 0.0;
#pragma unroll
    for (int j = 0; j < kC; ++j)
      if (j == k) rk = r[j];
    double best = (lane >= k && lane < n) ? fabs(rk) : -1.0;
    int bi = lane;
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    const int piv = bi;
    if (piv != k) {
#pragma unroll
      for (int j = 0; j < kC; ++j) {
        const double vk = __shfl(r[j], k), vp = __shfl(r[j], piv);
        if (lane == k) r[j] = vp;
        else if (lane == piv) r[j] = vk;
      }
    }
    double p[kC], pk = 0.0;
    rk = 0.0;
#pragma unroll
    for (int j = 0; j < kC; ++j) {
      p[j] = __shfl(r[j], k);
      if (j == k) {
        pk = p[j];
        rk = r[j];
      }
    }
    if (lane > k && lane < n) {
      const double f = rk / pk;
      if (f != 0.0) {
#pragma unroll
        for (int j = 0; j < kC; ++j)
          if (j >= k) r[j] -= f * p[j];
      }
    }
  }
  double rhs = 0.0;
#pragma unroll
  for (int j = 0; j < kC; ++j)
    if (j == n) rhs = r[j];
  double x[kR];
#pragma unroll
  for (int i = kR - 1; i >= 0; --i) {
    x[i] = 0.0;
    if (i >= n) continue;
    double sum = rhs;
#pragma unroll
    for (int j = i + 1; j < kR; ++j)
      if (j < n) sum -= r[j] * x[j];
    x[i] = __shfl(sum / r[i], i);
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < kR; ++i)
      if (i < m) coef[i] = x[i];
  }
}

hipError_t launch_diis_solve(int m, int count, const double* gram, double* coef, hipStream_t st) {
  hipLaunchKernelGGL(k_diis_solve, dim3(1), dim3(64), 0, st, m, count, gram, coef);
  return hipGetLastError();
}

// Fx = sum_i coef[i] F_i  (F slots contiguous, n elements each)
__global__ void k_diis_extrap(int m, int64_t n, const double* __restrict__ coef, const double* __restrict__ Fs,
                              double* __restrict__ out) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  double s = 0.0;
  for (int i = 0; i < m; ++i) s += coef[i] * Fs[(int64_t)i * n + idx];
  out[idx] = s;
}

hipError_t launch_diis_extrap(int m, int64_t n, const double* coef, const double* Fs, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_diis_extrap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, m, n, coef, Fs, out);
  return hipGetLastError();
}

// gram row/column update for the new slot: gram[slot*m+i] = gram[i*m+slot] = dots[i]
__global__ void k_gram_update(int m, int slot, int count, const double* __restrict__ dots, double* __restrict__ gram) {
  int i = threadIdx.x;
  if (i >= count) return;
  gram[slot * m + i] = dots[i];
  gram[i * m + slot] = dots[i];
}

hipError_t launch_gram_update(int m, int slot, int count, const double* dots, double* gram, hipStream_t st) {
  hipLaunchKernelGGL(k_gram_update, dim3(1), dim3(64), 0, st, m, slot, count, dots, gram);
  return hipGetLastError();
}

}  // namespace dfks

namespace dfks {
// X[r + k*n] = U[r + (k0+k)*n] / sqrt(w[k0+k])  (column-major n x (n-k0))
__global__ void k_scaled_cols(int n, int k0, const double* __restrict__ U, const double* __restrict__ w,
                              double* __restrict__ X) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int nk = n - k0;
  if (idx >= (int64_t)n * nk) return;
  int k = (int)(idx / n), r = (int)(idx % n);
  X[idx] = U[r + (int64_t)(k0 + k) * n] / sqrt(w[k0 + k]);
}

hipError_t launch_scaled_cols(int n, int k0, const double* U, const double* w, double* X, hipStream_t st) {
  int64_t tot = (int64_t)n * (n - k0);
  if (tot == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scaled_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, k0, U, w, X);
  return hipGetLastError();
}
}  // namespace dfks
