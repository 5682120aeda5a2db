// Grid kernels: Becke fuzzy-cell partition and AO values/gradients (FP64).
//
// Becke partition (dftax/grid/becke.py:68-108): for a point p and atoms A, B
//   mu_AB = (|p-A| - |p-B|) / R_AB,  nu = mu + a_AB (1 - mu^2),
//   f = three iterations of f -> 1.5 f - 0.5 f^3,  s = (1 - f)/2,
//   P_A = prod_{B != A} s(nu_AB),   w = w_raw * P_owner / sum_A P_A.
// AO values (dftax/energy/gto.py:371-412): phi = x^lx y^ly z^lz sum_k c_k e^{-a_k r^2},
// followed by the per-shell cart->sph block; gradients are analytic.
// Layout of the AO block: ao[4][ng][nao] (value, d/dx, d/dy, d/dz), row-major
// per point so a wavefront reading one point's row is contiguous.
#include <algorithm>

#include "dfks_common.h"

namespace dfks {

__global__ __launch_bounds__(256) void k_becke(int64_t n, const double* __restrict__ pts, const double* __restrict__ raw,
                                                const int32_t* __restrict__ owner, int natm,
                                                const double* __restrict__ atm /*[natm*3]*/,
                                                const double* __restrict__ invR /*[natm*natm]*/,
                                                const double* __restrict__ aAB /*[natm*natm]*/, double* __restrict__ w) {
  extern __shared__ double s_atm[];  // 3*natm
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x) s_atm[i] = atm[i];
  __syncthreads();
  int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= n) return;
  double px = pts[3 * g], py = pts[3 * g + 1], pz = pts[3 * g + 2];
  const int own = owner[g];
  double total = 0.0, mine = 0.0;
  for (int a = 0; a < natm; ++a) {
    double dx = px - s_atm[3 * a], dy = py - s_atm[3 * a + 1], dz = pz - s_atm[3 * a + 2];
    double ra = sqrt(dx * dx + dy * dy + dz * dz + 1e-30);
    double prod = 1.0;
    for (int b = 0; b < natm; ++b) {
      if (b == a) continue;
      double ex = px - s_atm[3 * b], ey = py - s_atm[3 * b + 1], ez = pz - s_atm[3 * b + 2];
      double rb = sqrt(ex * ex + ey * ey + ez * ez + 1e-30);
      double mu = (ra - rb) * invR[a * natm + b];
      double f = mu + aAB[a * natm + b] * (1.0 - mu * mu);
      f = 1.5 * f - 0.5 * f * f * f;
      f = 1.5 * f - 0.5 * f * f * f;
      f = 1.5 * f - 0.5 * f * f * f;
      prod *= 0.5 * (1.0 - f);
    }
    total += prod;
    if (a == own) mine = prod;
  }
  w[g] = raw[g] * mine / total;
}

template <int L>
__device__ __forceinline__ double ipow(double x) {
  if constexpr (L == 0) return 1.0;
  else return x * ipow<L - 1>(x);
}

// One thread per (point, shell). Shell index fastest so a wavefront's writes
// land in one point's AO row.
template <int L>
__device__ __forceinline__ void ao_shell(double dx, double dy, double dz, double R, double dR,
                                          int sph, double* __restrict__ row0, double* __restrict__ row1,
                                          double* __restrict__ row2, double* __restrict__ row3) {
  constexpr int NC = ncart(L);
  double v[NC], gx[NC], gy[NC], gz[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    constexpr int dummy = 0;
    (void)dummy;
    const int lx = cart_exp<L>(c, 0), ly = cart_exp<L>(c, 1), lz = cart_exp<L>(c, 2);
    double px = 1.0, py = 1.0, pz = 1.0, qx = 0.0, qy = 0.0, qz = 0.0;
    // x^l and l x^{l-1}
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < lx) { qx = qx * dx + px; px *= dx; }
      if (k < ly) { qy = qy * dy + py; py *= dy; }
      if (k < lz) { qz = qz * dz + pz; pz *= dz; }
    }
    // qx accumulated sum_{k} x^k * x^{lx-1-k} = lx x^{lx-1}
    double ang = px * py * pz;
    v[c] = ang * R;
    gx[c] = qx * py * pz * R + ang * dR * dx;
    gy[c] = px * qy * pz * R + ang * dR * dy;
    gz[c] = px * py * qz * R + ang * dR * dz;
  }
  if (sph && L >= 2) {
    constexpr int NS = nsph(L);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double w = c2s_coef<L>(c, s);
        if (w != 0.0) {
          a0 += w * v[c]; a1 += w * gx[c]; a2 += w * gy[c]; a3 += w * gz[c];
        }
      }
      row0[s] = a0; row1[s] = a1; row2[s] = a2; row3[s] = a3;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) { row0[c] = v[c]; row1[c] = gx[c]; row2[c] = gy[c]; row3[c] = gz[c]; }
  }
}

__global__ __launch_bounds__(256) void k_ao(int64_t npts, int64_t g0, int64_t ng, const double* __restrict__ pts,
                                             ShellSet S, int64_t nao, double* __restrict__ ao) {
  int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= npts * S.n) return;
  const int64_t g = g0 + gid / S.n;
  const int ish = (int)(gid % S.n);
  double dx = pts[3 * g] - S.centre[3 * ish];
  double dy = pts[3 * g + 1] - S.centre[3 * ish + 1];
  double dz = pts[3 * g + 2] - S.centre[3 * ish + 2];
  double r2 = dx * dx + dy * dy + dz * dz;
  double R = 0.0, dR = 0.0;
  const int p0 = S.prim0[ish], np = S.nprim[ish];
  for (int k = 0; k < np; ++k) {
    double a = S.exps[p0 + k];
    double e = S.coefs[p0 + k] * exp(-a * r2);
    R += e;
    dR += a * e;
  }
  dR *= -2.0;
  const int64_t plane = ng * nao;
  double* r0 = ao + g * nao + S.ao0[ish];
  switch (S.l[ish]) {
    case 0: ao_shell<0>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
    case 1: ao_shell<1>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
    case 2: ao_shell<2>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
    case 3: ao_shell<3>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
    default: ao_shell<4>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
  }
}

// Compacted AO values for blocks of grid points (one bin: blocks of equal
// padded width n). Block b holds G points, rows b*G .. b*G+G-1 of the bin's
// planes ao[4][nblk*G][n]; its significant shells are shl[off[b] .. off[b+1])
// as (shell, first compact column). One workgroup per block; padded columns
// are left as they are (zeroed once by the caller).
__global__ __launch_bounds__(256) void k_ao_cmp(int nblk, int G, int n, const double* __restrict__ pts,
                                                 const int2* __restrict__ shl, const int64_t* __restrict__ off,
                                                 ShellSet S, double* __restrict__ ao) {
  const int b = blockIdx.x;
  if (b >= nblk) return;
  const int64_t s0 = off[b], ns = off[b + 1] - s0;
  const int64_t plane = (int64_t)nblk * G * n;
  for (int64_t idx = threadIdx.x; idx < ns * G; idx += blockDim.x) {
    const int p = (int)(idx % G);
    const int2 e = shl[s0 + idx / G];
    const int ish = e.x;
    const int64_t row = (int64_t)b * G + p;
    const double dx = pts[3 * row] - S.centre[3 * ish];
    const double dy = pts[3 * row + 1] - S.centre[3 * ish + 1];
    const double dz = pts[3 * row + 2] - S.centre[3 * ish + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    double R = 0.0, dR = 0.0;
    const int p0 = S.prim0[ish], np = S.nprim[ish];
    for (int k = 0; k < np; ++k) {
      const double a = S.exps[p0 + k];
      const double ev = S.coefs[p0 + k] * exp(-a * r2);
      R += ev;
      dR += a * ev;
    }
    dR *= -2.0;
    double* r0 = ao + row * n + e.y;
    switch (S.l[ish]) {
      case 0: ao_shell<0>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
      case 1: ao_shell<1>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
      case 2: ao_shell<2>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
      case 3: ao_shell<3>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
      default: ao_shell<4>(dx, dy, dz, R, dR, S.spherical, r0, r0 + plane, r0 + 2 * plane, r0 + 3 * plane); break;
    }
  }
}

hipError_t launch_ao_cmp(int nblk, int G, int n, const double* pts, const int2* shl, const int64_t* off,
                         const ShellSet& S, double* ao, hipStream_t st) {
  if (nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ao_cmp, dim3((unsigned)nblk), dim3(256), 0, st, nblk, G, n, pts, shl, off, S, ao);
  return hipGetLastError();
}

// Pg[b][k][j] = P[c_k][c_j] for the compact columns c = cols[b][0..n) of each
// block of a bin (0 where either column is padding, c < 0). One wavefront per
// row (b, k), grid-stride over the rows.
template <typename TP, typename T>
__global__ __launch_bounds__(256) void k_gather_sym(int64_t nblk, int n, const int32_t* __restrict__ cols,
                                                     const TP* __restrict__ P, int64_t nao, T* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t nrow = nblk * n;
  for (int64_t row = blockIdx.x * 4 + (threadIdx.x >> 6); row < nrow; row += (int64_t)gridDim.x * 4) {
    const int64_t b = row / n;
    const int32_t* cb = cols + b * n;
    const int32_t cr = cols[row];
    T* o = out + row * n;
    const TP* pr = P + (int64_t)(cr >= 0 ? cr : 0) * nao;
    for (int c = lane; c < n; c += 64) {
      const int32_t cc = cb[c];
      o[c] = (cr >= 0 && cc >= 0) ? (T)pr[cc] : (T)0;
    }
  }
}

template <typename TP, typename T>
hipError_t launch_gather_sym(int64_t nblk, int n, const int32_t* cols, const TP* P, int64_t nao, T* out,
                             hipStream_t st) {
  const int64_t nrow = nblk * n;
  if (nrow == 0) return hipSuccess;
  const int64_t nb = std::min<int64_t>((nrow + 3) / 4, 65536);
  hipLaunchKernelGGL((k_gather_sym<TP, T>), dim3((unsigned)nb), dim3(256), 0, st, nblk, n, cols, P, nao, out);
  return hipGetLastError();
}
template hipError_t launch_gather_sym<double, double>(int64_t, int, const int32_t*, const double*, int64_t, double*,
                                                     hipStream_t);
template hipError_t launch_gather_sym<float, float>(int64_t, int, const int32_t*, const float*, int64_t, float*,
                                                   hipStream_t);

// Vh[m][:] = sum over the (block, local column i) entries of row m of the
// block products prod_b[i][:], scattered to the blocks' columns. Workgroup
// (m, s) takes slice s of row m's entries, one entry per wavefront at a time,
// and accumulates the row in LDS (FP64 LDS atomics: wavefronts may hit the
// same column); slice s writes out + s * nao^2 (the slices are summed after).
template <typename T>
__global__ __launch_bounds__(256) void k_vassemble(int nao, int nslice, const int64_t* __restrict__ rowptr,
                                                    const int64_t* __restrict__ prow, const int64_t* __restrict__ pcol,
                                                    const int32_t* __restrict__ plen, const T* __restrict__ q,
                                                    const int32_t* __restrict__ cols, double* __restrict__ out) {
  extern __shared__ double acc[];
  const int m = blockIdx.x, s = blockIdx.y;
  if (m >= nao || s >= nslice) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = threadIdx.x; j < nao; j += blockDim.x) acc[j] = 0.0;
  __syncthreads();
  const int64_t r0 = rowptr[m], cnt = rowptr[m + 1] - r0;
  const int64_t e0 = r0 + cnt * s / nslice, e1 = r0 + cnt * (s + 1) / nslice;
  // two entries per wavefront in flight
  int64_t e = e0 + wave;
  for (; e + 4 < e1; e += 8) {
    const T* s0 = q + prow[e];
    const T* s1 = q + prow[e + 4];
    const int32_t* d0 = cols + pcol[e];
    const int32_t* d1 = cols + pcol[e + 4];
    const int n0 = plen[e], n1 = plen[e + 4];
    const int nmax = n0 > n1 ? n0 : n1;
    for (int j = lane; j < nmax; j += 64) {
      const int32_t c0 = j < n0 ? d0[j] : -1;
      const int32_t c1 = j < n1 ? d1[j] : -1;
      const double v0 = c0 >= 0 ? (double)s0[j] : 0.0;
      const double v1 = c1 >= 0 ? (double)s1[j] : 0.0;
      if (c0 >= 0) atomicAdd(&acc[c0], v0);
      if (c1 >= 0) atomicAdd(&acc[c1], v1);
    }
  }
  for (; e < e1; e += 4) {
    const T* src = q + prow[e];
    const int32_t* dst = cols + pcol[e];
    const int n = plen[e];
    for (int j = lane; j < n; j += 64) {
      const int32_t c = dst[j];
      if (c >= 0) atomicAdd(&acc[c], (double)src[j]);
    }
  }
  __syncthreads();
  double* o = out + (int64_t)s * nao * nao + (int64_t)m * nao;
  for (int j = threadIdx.x; j < nao; j += blockDim.x) o[j] = acc[j];
}

// V[i] = sum_s ws[s][i]
__global__ __launch_bounds__(256) void k_sum_slices(int64_t nn, int nslice, const double* __restrict__ ws,
                                                     double* __restrict__ V) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  double v = 0.0;
  for (int s = 0; s < nslice; ++s) v += ws[(int64_t)s * nn + i];
  V[i] = v;
}

// ws: nslice * nao^2 doubles when nslice > 1 (unused otherwise).
template <typename T>
hipError_t launch_vassemble(int nao, const int64_t* rowptr, const int64_t* prow, const int64_t* pcol,
                            const int32_t* plen, const T* q, const int32_t* cols, double* V, double* ws, int nslice,
                            hipStream_t st) {
  if (nao <= 0) return hipSuccess;
  if ((size_t)nao * sizeof(double) > 64 * 1024 || nslice < 1 || (nslice > 1 && !ws)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_vassemble<T>, dim3((unsigned)nao, (unsigned)nslice), dim3(256), nao * sizeof(double), st, nao,
                     nslice, rowptr, prow, pcol, plen, q, cols, nslice > 1 ? ws : V);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nslice == 1) return e;
  const int64_t nn = (int64_t)nao * nao;
  hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, nn, nslice, ws, V);
  return hipGetLastError();
}
template hipError_t launch_vassemble<double>(int, const int64_t*, const int64_t*, const int64_t*, const int32_t*,
                                            const double*, const int32_t*, double*, double*, int, hipStream_t);
template hipError_t launch_vassemble<float>(int, const int64_t*, const int64_t*, const int64_t*, const int32_t*,
                                           const float*, const int32_t*, double*, double*, int, hipStream_t);

hipError_t launch_becke(int64_t n, const double* pts, const double* raw, const int32_t* owner, int natm,
                        const double* atm, const double* invR, const double* aAB, double* w, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL(k_becke, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 3 * natm * sizeof(double), st, n, pts,
                     raw, owner, natm, atm, invR, aAB, w);
  return hipGetLastError();
}

hipError_t launch_ao(int64_t ng, const double* pts, const ShellSet& S, int64_t nao, double* ao, hipStream_t st) {
  const int bs = 256;
  const int64_t total = ng * S.n;
  const int64_t chunk_pts = ((1LL << 31) / S.n / bs) * bs;  // keep each launch under 2^31 threads
  for (int64_t g0 = 0; g0 < ng; g0 += chunk_pts) {
    int64_t n = ng - g0 < chunk_pts ? ng - g0 : chunk_pts;
    int64_t th = n * S.n;
    hipLaunchKernelGGL(k_ao, dim3((unsigned)((th + bs - 1) / bs)), dim3(bs), 0, st, n, g0, ng, pts, S, nao, ao);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  (void)total;
  return hipSuccess;
}

}  // namespace dfks
