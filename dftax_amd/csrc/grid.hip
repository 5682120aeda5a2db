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
// (m, s) takes slice s of row m's entries, two entries per wavefront at a
// time, and accumulates the row in LDS (FP64 LDS atomics: wavefronts may hit
// the same column); slice s writes out + s * nao^2 (the slices are summed
// after). Each lane reads 16 bytes of a product row (4 floats or 2 doubles)
// and the matching column indices per load; the block widths are multiples
// of 32, so every row and column list starts 16-byte aligned. Loads past a
// row's end are clamped to its start and masked, so every lane loads.
template <typename T>
__global__ __launch_bounds__(256) void k_vassemble(int nao, int nslice, const int64_t* __restrict__ rowptr,
                                                    const int64_t* __restrict__ prow, const int64_t* __restrict__ pcol,
                                                    const int32_t* __restrict__ plen, const T* __restrict__ q,
                                                    const int32_t* __restrict__ cols, double* __restrict__ out) {
  constexpr int V = 16 / sizeof(T);   // elements per 16-byte load
  struct alignas(16) Tv { T x[V]; };
  struct alignas(4 * V) Cv { int32_t c[V]; };
  extern __shared__ double acc[];
  const int m = blockIdx.x, s = blockIdx.y;
  if (m >= nao || s >= nslice) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = threadIdx.x; j < nao; j += blockDim.x) acc[j] = 0.0;
  __syncthreads();
  const int64_t r0 = rowptr[m], cnt = rowptr[m + 1] - r0;
  const int64_t e0 = r0 + cnt * s / nslice, e1 = r0 + cnt * (s + 1) / nslice;
  auto add = [&](const Tv& x, const Cv& c, bool ok) {
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (ok && c.c[v] >= 0) atomicAdd(&acc[c.c[v]], (double)x.x[v]);
  };
  int64_t e = e0 + wave;
  for (; e + 4 < e1; e += 8) {
    const Tv* s0 = reinterpret_cast<const Tv*>(q + prow[e]);
    const Tv* s1 = reinterpret_cast<const Tv*>(q + prow[e + 4]);
    const Cv* d0 = reinterpret_cast<const Cv*>(cols + pcol[e]);
    const Cv* d1 = reinterpret_cast<const Cv*>(cols + pcol[e + 4]);
    const int n0 = plen[e] / V, n1 = plen[e + 4] / V;
    const int nmax = n0 > n1 ? n0 : n1;
    for (int j = lane; j < nmax; j += 64) {
      const bool ok0 = j < n0, ok1 = j < n1;
      const Tv x0 = s0[ok0 ? j : 0], x1 = s1[ok1 ? j : 0];
      const Cv c0 = d0[ok0 ? j : 0], c1 = d1[ok1 ? j : 0];
      add(x0, c0, ok0);
      add(x1, c1, ok1);
    }
  }
  for (; e < e1; e += 4) {
    const Tv* src = reinterpret_cast<const Tv*>(q + prow[e]);
    const Cv* dst = reinterpret_cast<const Cv*>(cols + pcol[e]);
    const int n = plen[e] / V;
    for (int j = lane; j < n; j += 64) add(src[j], dst[j], true);
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
  // 16-byte loads: the product rows and column lists start at multiples of 32 elements
  if ((uintptr_t)q % 16 != 0 || (uintptr_t)cols % 16 != 0) return hipErrorInvalidValue;
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


// ---------------------------------------------------------------------------
// Nuclear gradient of E_xc on the moving Becke grid (dftax/ks/forces.py:245-270:
// the basis centres and the grid points follow their atoms, the Becke weights
// are differentiated). With Z = Phi P, X = (grad rho . grad Phi) P and, per point g
// and function m,
//   q_{g m j} = v_rho d_j phi Z + 2 v_sigma ( (H_j . grad rho) Z + d_j phi X ),
// the basis part is -2 sum_{m on A} sum_g w_g q_{g m j}, the point part (points
// owned by A move with it) is 2 sum_{g owned by A} w_g sum_m q_{g m j}, and the
// weight part is sum_g f_g dw_g/dR.
// ---------------------------------------------------------------------------

// AO values, gradients and Hessians: planes [10][npts][nao] in the order
// value, x, y, z, xx, xy, xz, yy, yz, zz.
template <int L>
__device__ __forceinline__ void ao_shell_hess(double dx, double dy, double dz, double R, double dR, double ddR,
                                               int sph, double* __restrict__ r0, int64_t plane) {
  constexpr int NC = ncart(L);
  double o[10][NC];
  const double d[3] = {dx, dy, dz};
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    int l[3] = {cart_exp<L>(c, 0), cart_exp<L>(c, 1), cart_exp<L>(c, 2)};
    // per axis: x^l, l x^(l-1), l (l-1) x^(l-2)
    double p0[3], p1[3], p2[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double a = 1.0, b = 0.0, e = 0.0;   // value, first, second derivative of x^l
      for (int k = 0; k < l[x]; ++k) {
        e = e * d[x] + 2.0 * b;
        b = b * d[x] + a;
        a *= d[x];
      }
      p0[x] = a; p1[x] = b; p2[x] = e;
    }
    const double A = p0[0] * p0[1] * p0[2];
    const double Ag[3] = {p1[0] * p0[1] * p0[2], p0[0] * p1[1] * p0[2], p0[0] * p0[1] * p1[2]};
    o[0][c] = A * R;
#pragma unroll
    for (int x = 0; x < 3; ++x) o[1 + x][c] = Ag[x] * R + A * dR * d[x];
    int h = 4;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        double Aij;
        if (i == j) Aij = (i == 0 ? p2[0] * p0[1] * p0[2] : (i == 1 ? p0[0] * p2[1] * p0[2] : p0[0] * p0[1] * p2[2]));
        else {
          double f[3] = {p0[0], p0[1], p0[2]};
          f[i] = p1[i];
          f[j] = p1[j];
          Aij = f[0] * f[1] * f[2];
        }
        o[h][c] = Aij * R + (Ag[i] * d[j] + Ag[j] * d[i]) * dR + A * (ddR * d[i] * d[j] + (i == j ? dR : 0.0));
        ++h;
      }
  }
  if (sph && L >= 2) {
    constexpr int NS = nsph(L);
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const double w = c2s_coef<L>(c, s);
          if (w != 0.0) acc += w * o[k][c];
        }
        r0[k * plane + s] = acc;
      }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 10; ++k) r0[k * plane + c] = o[k][c];
  }
}

__global__ __launch_bounds__(256) void k_ao_hess(int64_t npts, const double* __restrict__ pts, ShellSet S, int64_t nao,
                                                  double* __restrict__ ao) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= npts * S.n) return;
  const int64_t g = gid / S.n;
  const int ish = (int)(gid % S.n);
  const double dx = pts[3 * g] - S.centre[3 * ish];
  const double dy = pts[3 * g + 1] - S.centre[3 * ish + 1];
  const double dz = pts[3 * g + 2] - S.centre[3 * ish + 2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  double R = 0.0, dR = 0.0, ddR = 0.0;
  for (int k = S.prim0[ish]; k < S.prim0[ish] + S.nprim[ish]; ++k) {
    const double a = S.exps[k];
    const double e = S.coefs[k] * exp(-a * r2);
    R += e;
    dR += a * e;
    ddR += a * a * e;
  }
  dR *= -2.0;    // (1/r) dR/dr
  ddR *= 4.0;    // its derivative over r, per unit x_i x_j
  const int64_t plane = npts * nao;
  double* r0 = ao + g * nao + S.ao0[ish];
  switch (S.l[ish]) {
    case 0: ao_shell_hess<0>(dx, dy, dz, R, dR, ddR, S.spherical, r0, plane); break;
    case 1: ao_shell_hess<1>(dx, dy, dz, R, dR, ddR, S.spherical, r0, plane); break;
    case 2: ao_shell_hess<2>(dx, dy, dz, R, dR, ddR, S.spherical, r0, plane); break;
    default: ao_shell_hess<3>(dx, dy, dz, R, dR, ddR, S.spherical, r0, plane); break;
  }
}

// rho, grad rho and sigma at each point from the AO planes and Z = Phi P
// (one wavefront per point, lanes over the functions).
__global__ __launch_bounds__(64) void k_xcg_rho(int64_t npts, int64_t nao, const double* __restrict__ ao,
                                                const double* __restrict__ Z, double* __restrict__ rho,
                                                double* __restrict__ grad, double* __restrict__ sigma) {
  const int64_t g = blockIdx.x;
  if (g >= npts) return;
  const int64_t plane = npts * nao;
  double a[4] = {0, 0, 0, 0};
  for (int64_t m = threadIdx.x; m < nao; m += 64) {
    const double z = Z[g * nao + m];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] += ao[k * plane + g * nao + m] * z;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (int off = 32; off > 0; off >>= 1) a[k] += __shfl_xor(a[k], off, 64);
  if (threadIdx.x == 0) {
    rho[g] = a[0];
    const double gx = 2.0 * a[1], gy = 2.0 * a[2], gz = 2.0 * a[3];
    grad[3 * g] = gx; grad[3 * g + 1] = gy; grad[3 * g + 2] = gz;
    sigma[g] = gx * gx + gy * gy + gz * gz;
  }
}

// Gd[g][m] = grad rho_g . grad phi_m(g)
__global__ __launch_bounds__(256) void k_xcg_gdot(int64_t npts, int64_t nao, const double* __restrict__ ao,
                                                   const double* __restrict__ grad, double* __restrict__ Gd) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= npts * nao) return;
  const int64_t g = i / nao, plane = npts * nao;
  Gd[i] = grad[3 * g] * ao[plane + i] + grad[3 * g + 1] * ao[2 * plane + i] + grad[3 * g + 2] * ao[3 * plane + i];
}

// q_{g m j} (above) into q[3][npts][nao]; the point part 2 w_g sum_m q to the
// point's owner atom. xc[3 g ..] = {f, v_rho, v_sigma}; points with
// rho <= thresh contribute nothing (the energy's density mask).
__global__ __launch_bounds__(64) void k_xcg_q(int64_t npts, int64_t nao, const double* __restrict__ ao,
                                              const double* __restrict__ Z, const double* __restrict__ X,
                                              const double* __restrict__ rho, const double* __restrict__ grad,
                                              const double* __restrict__ xc, const double* __restrict__ w,
                                              const int32_t* __restrict__ owner, double thresh, double* __restrict__ q,
                                              double* __restrict__ gout, double* __restrict__ fm) {
  const int64_t g = blockIdx.x;
  if (g >= npts) return;
  const int64_t plane = npts * nao;
  const bool on = rho[g] > thresh;
  if (threadIdx.x == 0) fm[g] = on ? xc[3 * g] : 0.0;   // energy density f of the masked functional
  const double vr = on ? xc[3 * g + 1] : 0.0, vs = on ? xc[3 * g + 2] : 0.0;
  const double gr[3] = {grad[3 * g], grad[3 * g + 1], grad[3 * g + 2]};
  double acc[3] = {0, 0, 0};
  for (int64_t m = threadIdx.x; m < nao; m += 64) {
    const int64_t i = g * nao + m;
    const double z = Z[i], x = X[i];
    const double d[3] = {ao[plane + i], ao[2 * plane + i], ao[3 * plane + i]};
    const double hxx = ao[4 * plane + i], hxy = ao[5 * plane + i], hxz = ao[6 * plane + i];
    const double hyy = ao[7 * plane + i], hyz = ao[8 * plane + i], hzz = ao[9 * plane + i];
    const double hg[3] = {hxx * gr[0] + hxy * gr[1] + hxz * gr[2], hxy * gr[0] + hyy * gr[1] + hyz * gr[2],
                          hxz * gr[0] + hyz * gr[1] + hzz * gr[2]};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double v = vr * d[j] * z + 2.0 * vs * (hg[j] * z + d[j] * x);
      q[j * plane + i] = v;
      acc[j] += v;
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
    for (int off = 32; off > 0; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
  if (threadIdx.x == 0 && on) {
    const int a = owner[g];
#pragma unroll
    for (int j = 0; j < 3; ++j) atomicAdd(&gout[3 * a + j], 2.0 * w[g] * acc[j]);
  }
}

// Becke weight derivatives: sum_g f_g dw_g/dR_D (becke.py:68-108 differentiated),
// with w = w_raw P_O / sum_A P_A, P_A = prod_{B != A} s(nu_AB), and the point
// moving with its owner O:
//   dw/dR_D = w sum_A (delta_AO - P_A / Z) sum_{B != A} (s'/s)(nu_AB) (1 - 2 a_AB mu_AB) dmu_AB/dR_D,
//   dmu_AB/dR_A = -u_A / R_AB - mu_AB (R_A - R_B) / R_AB^2,
//   dmu_AB/dR_B =  u_B / R_AB + mu_AB (R_A - R_B) / R_AB^2,
//   dmu_AB/dr   = (u_A - u_B) / R_AB  (to the owner), u_X = (r - R_X)/|r - R_X|.
// One thread per point; per-atom sums in LDS, added to gout once per workgroup.
__device__ __forceinline__ void becke_s(double nu, double& s, double& ds) {
  const double f1 = 1.5 * nu - 0.5 * nu * nu * nu;
  const double f2 = 1.5 * f1 - 0.5 * f1 * f1 * f1;
  const double f3 = 1.5 * f2 - 0.5 * f2 * f2 * f2;
  s = 0.5 * (1.0 - f3);
  ds = -0.5 * (1.5 - 1.5 * f2 * f2) * (1.5 - 1.5 * f1 * f1) * (1.5 - 1.5 * nu * nu);
}

__global__ __launch_bounds__(256) void k_becke_grad(int64_t n, const double* __restrict__ pts,
                                                     const double* __restrict__ raw, const int32_t* __restrict__ owner,
                                                     const double* __restrict__ f, int natm,
                                                     const double* __restrict__ atm /*[3 natm]*/,
                                                     const double* __restrict__ invR, const double* __restrict__ aAB,
                                                     double* __restrict__ gout) {
  extern __shared__ double s_g[];   // 3 natm accumulators
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x) s_g[i] = 0.0;
  __syncthreads();
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < n && f[g] != 0.0) {
    const double px = pts[3 * g], py = pts[3 * g + 1], pz = pts[3 * g + 2];
    const int own = owner[g];
    auto dist = [&](int a, double* u) {
      const double dx = px - atm[3 * a], dy = py - atm[3 * a + 1], dz = pz - atm[3 * a + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz + 1e-30);
      u[0] = dx / r; u[1] = dy / r; u[2] = dz / r;
      return r;
    };
    // pass 1: Z = sum_A P_A and P_owner
    double Zs = 0.0, Po = 0.0, u[3];
    for (int a = 0; a < natm; ++a) {
      const double ra = dist(a, u);
      double prod = 1.0;
      for (int b = 0; b < natm; ++b) {
        if (b == a) continue;
        const double rb = dist(b, u);
        const double mu = (ra - rb) * invR[a * natm + b];
        double s, ds;
        becke_s(mu + aAB[a * natm + b] * (1.0 - mu * mu), s, ds);
        prod *= s;
      }
      Zs += prod;
      if (a == own) Po = prod;
    }
    if (Zs > 0.0) {
      const double wg = raw[g] * Po / Zs;
      const double fw = f[g] * wg;
      // pass 2: the derivative terms
      for (int a = 0; a < natm; ++a) {
        double ua[3];
        const double ra = dist(a, ua);
        double prod = 1.0;
        for (int b = 0; b < natm; ++b) {
          if (b == a) continue;
          double ub[3];
          const double rb = dist(b, ub);
          const double mu = (ra - rb) * invR[a * natm + b];
          double s, ds;
          becke_s(mu + aAB[a * natm + b] * (1.0 - mu * mu), s, ds);
          prod *= s;
        }
        const double Ka = fw * ((a == own ? 1.0 : 0.0) - prod / Zs);
        if (Ka == 0.0) continue;
        for (int b = 0; b < natm; ++b) {
          if (b == a) continue;
          double ub[3];
          const double rb = dist(b, ub);
          const double iR = invR[a * natm + b];
          const double mu = (ra - rb) * iR;
          const double ab = aAB[a * natm + b];
          double s, ds;
          becke_s(mu + ab * (1.0 - mu * mu), s, ds);
          if (s < 1e-300) continue;
          const double c = Ka * ds / s * (1.0 - 2.0 * ab * mu);
          for (int x = 0; x < 3; ++x) {
            const double Rab = atm[3 * a + x] - atm[3 * b + x];
            const double da = -ua[x] * iR - mu * Rab * iR * iR;
            const double db = ub[x] * iR + mu * Rab * iR * iR;
            const double dr = (ua[x] - ub[x]) * iR;
            atomicAdd(&s_g[3 * a + x], c * da);
            atomicAdd(&s_g[3 * b + x], c * db);
            atomicAdd(&s_g[3 * own + x], c * dr);
          }
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * natm; i += blockDim.x)
    if (s_g[i] != 0.0) atomicAdd(&gout[i], s_g[i]);
}

hipError_t launch_ao_hess(int64_t npts, const double* pts, const ShellSet& S, int64_t nao, double* ao,
                          hipStream_t st) {
  const int64_t total = npts * S.n;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ao_hess, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, npts, pts, S, nao, ao);
  return hipGetLastError();
}

hipError_t launch_xcg_rho(int64_t npts, int64_t nao, const double* ao, const double* Z, double* rho, double* grad,
                          double* sigma, hipStream_t st) {
  if (npts == 0) return hipSuccess;
  hipLaunchKernelGGL(k_xcg_rho, dim3((unsigned)npts), dim3(64), 0, st, npts, nao, ao, Z, rho, grad, sigma);
  return hipGetLastError();
}

hipError_t launch_xcg_gdot(int64_t npts, int64_t nao, const double* ao, const double* grad, double* Gd,
                           hipStream_t st) {
  const int64_t total = npts * nao;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_xcg_gdot, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, npts, nao, ao, grad, Gd);
  return hipGetLastError();
}

hipError_t launch_xcg_q(int64_t npts, int64_t nao, const double* ao, const double* Z, const double* X,
                        const double* rho, const double* grad, const double* xc, const double* w,
                        const int32_t* owner, double thresh, double* q, double* gout, double* fm, hipStream_t st) {
  if (npts == 0) return hipSuccess;
  hipLaunchKernelGGL(k_xcg_q, dim3((unsigned)npts), dim3(64), 0, st, npts, nao, ao, Z, X, rho, grad, xc, w, owner,
                     thresh, q, gout, fm);
  return hipGetLastError();
}

hipError_t launch_becke_grad(int64_t n, const double* pts, const double* raw, const int32_t* owner, const double* f,
                             int natm, const double* atm, const double* invR, const double* aAB, double* gout,
                             hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_becke_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 3 * natm * sizeof(double), st, n,
                     pts, raw, owner, f, natm, atm, invR, aAB, gout);
  return hipGetLastError();
}

}  // namespace dfks
