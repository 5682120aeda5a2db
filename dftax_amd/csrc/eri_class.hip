// Three-centre (ab|c) for the higher angular-momentum classes: one 64-lane
// wavefront per task (bra shell pair x auxiliary shell), one template instance
// per class.
//
// Same scheme as the generic wavefront kernel in eri.hip (E tables, R ladder
// and the cartesian block in LDS, lanes over the cartesian outputs, sparse
// cart->sph at the end), with every size a compile-time constant: the LDS
// footprint is the class's own (a (pd|p) task needs ~3 KB where the generic
// kernel reserves ~20 KB for (ff|g)), so several times more wavefronts fit on
// a CU, and the R-ladder index decode is a table lookup.
// Formula: dftax/integrals/eri3c_bucketed.py:290-426.
#include <cstdint>

#include "dfks_common.h"

namespace dfks {

namespace {

struct TuvTab {
  int t[286], u[286], v[286];
};
__host__ __device__ constexpr TuvTab make_tab() {
  TuvTab d{};
  int e = 0;
  for (int s = 0; s <= 10; ++s)
    for (int r = 0; r <= s; ++r)
      for (int v = 0; v <= r; ++v) {
        d.t[e] = s - r;
        d.u[e] = r - v;
        d.v[e] = v;
        ++e;
      }
  return d;
}
__constant__ TuvTab c_tuv = make_tab();

__host__ __device__ constexpr int nr_tet(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__device__ __forceinline__ int rx(int t, int u, int v) {
  const int s = t + u + v, r = u + v;
  return s * (s + 1) * (s + 2) / 6 + r * (r + 1) / 2 + v;
}

}  // namespace

// cartesian exponents and sparse cart->sph lists (filled by eri.hip's table upload)
__constant__ int g_cart_cls[7][28][3];
__constant__ C2SList g_c2s_cls;

void upload_cls_tables(const C2SList& h, const int (&cart)[7][28][3]) {
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_c2s_cls), &h, sizeof(h));
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_cart_cls), cart, sizeof(cart));
}

template <int LA, int LB, int LC>
__global__ __launch_bounds__(64) void k_eri_cls(const int2* __restrict__ pairs, const int32_t* __restrict__ aux,
                                                int npair, int naux_l, ShellSet A, ShellSet C,
                                                const double* __restrict__ boys, double* __restrict__ out, int64_t sa,
                                                int64_t sb, int64_t c0, int tr, int64_t blk0) {
  constexpr int L = LA + LB + LC, LAB = LA + LB, MT = L + 1, MTAB = LAB + 1;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), NOUT = NA * NB * NC;
  constexpr int NR = nr_tet(L);
  __shared__ double s_E[3][(LA + 1) * (LB + 1) * MTAB];
  __shared__ double s_Ec[(LC + 1) * MT];
  __shared__ double s_F[MT];
  __shared__ double s_R[2][NR];
  __shared__ double s_cart[NOUT];

  const int lane = threadIdx.x;
  const int64_t task = blk0 + blockIdx.x;
  const int k = (int)(task / naux_l);
  const int j = (int)(task % naux_l);
  if (k >= npair) return;
  const int2 pr = pairs[k];
  const int ia = pr.x, ib = pr.y, ic = aux[j];
  const double Ax = A.centre[3 * ia], Ay = A.centre[3 * ia + 1], Az = A.centre[3 * ia + 2];
  const double Bx = A.centre[3 * ib], By = A.centre[3 * ib + 1], Bz = A.centre[3 * ib + 2];
  const double Cx = C.centre[3 * ic], Cy = C.centre[3 * ic + 1], Cz = C.centre[3 * ic + 2];
  const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
  const double ab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  const int npa = A.nprim[ia], npb = A.nprim[ib], p0a = A.prim0[ia], p0b = A.prim0[ib];

  for (int o = lane; o < NOUT; o += 64) s_cart[o] = 0.0;

  for (int pa = 0; pa < npa; ++pa)
    for (int pb = 0; pb < npb; ++pb) {
      const double al = A.exps[p0a + pa], be = A.exps[p0b + pb];
      const double cab = A.coefs[p0a + pa] * A.coefs[p0b + pb];
      const double p = al + be;
      const double Px = (al * Ax + be * Bx) / p, Py = (al * Ay + be * By) / p, Pz = (al * Az + be * Bz) / p;
      const double Kab = exp(-al * be / p * ab2);
      __syncthreads();
      if (lane < 3) {
        // E^{ij}_t of one axis (same recurrences as eri.hip e_table_1d)
        const double xab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
        const double i2g = 0.5 / p, xpa = -be * xab / p, xpb = al * xab / p;
        double* E = s_E[lane];
        for (int q = 0; q < (LA + 1) * (LB + 1) * MTAB; ++q) E[q] = 0.0;
        E[0] = 1.0;
        for (int i = 0; i < LA; ++i)
          for (int t = 0; t <= i + 1 && t < MTAB; ++t) {
            double v = xpa * E[(i * (LB + 1)) * MTAB + t];
            if (t > 0) v += i2g * E[(i * (LB + 1)) * MTAB + t - 1];
            if (t + 1 < MTAB) v += (t + 1) * E[(i * (LB + 1)) * MTAB + t + 1];
            E[((i + 1) * (LB + 1)) * MTAB + t] = v;
          }
        for (int jj = 0; jj < LB; ++jj)
          for (int i = 0; i <= LA; ++i)
            for (int t = 0; t <= i + jj + 1 && t < MTAB; ++t) {
              double v = xpb * E[(i * (LB + 1) + jj) * MTAB + t];
              if (t > 0) v += i2g * E[(i * (LB + 1) + jj) * MTAB + t - 1];
              if (t + 1 < MTAB) v += (t + 1) * E[(i * (LB + 1) + jj) * MTAB + t + 1];
              E[(i * (LB + 1) + jj + 1) * MTAB + t] = v;
            }
      }
      for (int pc = 0; pc < C.nprim[ic]; ++pc) {
        const double ga = C.exps[C.prim0[ic] + pc];
        const double cc = C.coefs[C.prim0[ic] + pc];
        const double rho = p * ga / (p + ga);
        double pref = cab * cc * Kab * 2.0 * 17.493418327624862 / (p * ga * sqrt(p + ga));  // 2 pi^2.5
        if (LC & 1) pref = -pref;
        const double pcx = Px - Cx, pcy = Py - Cy, pcz = Pz - Cz;
        const double T = rho * (pcx * pcx + pcy * pcy + pcz * pcz);
        __syncthreads();
        if (lane == 0) {
          boys_op<L>(T, rho, C.omega, s_F, boys);
        } else if (lane == 1) {
          const double i2g = 0.5 / ga;
          for (int q = 0; q < (LC + 1) * MT; ++q) s_Ec[q] = 0.0;
          s_Ec[0] = 1.0;
          for (int kk = 0; kk < LC; ++kk)
            for (int t = 0; t < MT; ++t) {
              double v = 0.0;
              if (t > 0) v += i2g * s_Ec[kk * MT + t - 1];
              if (t + 1 < MT) v += (t + 1) * s_Ec[kk * MT + t + 1];
              s_Ec[(kk + 1) * MT + t] = v;
            }
        }
        __syncthreads();
        // R ladder: level L holds R_000 only; each level below from the one above
        const double n2r = -2.0 * rho;
        double pw[MT];
        pw[0] = 1.0;
#pragma unroll
        for (int q = 0; q < L; ++q) pw[q + 1] = pw[q] * n2r;
        if (lane == 0) s_R[0][0] = pw[L] * s_F[L];
        __syncthreads();
        int cur = 0;
#pragma unroll
        for (int m = L - 1; m >= 0; --m) {
          const double* prev = s_R[cur];
          double* nxt = s_R[cur ^ 1];
          const int count = nr_tet(L - m);
          for (int e = lane; e < count; e += 64) {
            const int t = c_tuv.t[e], u = c_tuv.u[e], v = c_tuv.v[e];
            double val;
            if (e == 0) {
              val = pw[m] * s_F[m];
            } else if (t > 0) {
              val = pcx * prev[rx(t - 1, u, v)] + (t > 1 ? (t - 1) * prev[rx(t - 2, u, v)] : 0.0);
            } else if (u > 0) {
              val = pcy * prev[rx(t, u - 1, v)] + (u > 1 ? (u - 1) * prev[rx(t, u - 2, v)] : 0.0);
            } else {
              val = pcz * prev[rx(t, u, v - 1)] + (v > 1 ? (v - 1) * prev[rx(t, u, v - 2)] : 0.0);
            }
            nxt[e] = val;
          }
          __syncthreads();
          cur ^= 1;
        }
        const double* R = s_R[cur];
        for (int o = lane; o < NOUT; o += 64) {
          const int ci = o % NC, bi = (o / NC) % NB, ai = o / (NC * NB);
          const int a0 = g_cart_cls[LA][ai][0], a1 = g_cart_cls[LA][ai][1], a2 = g_cart_cls[LA][ai][2];
          const int b0 = g_cart_cls[LB][bi][0], b1 = g_cart_cls[LB][bi][1], b2 = g_cart_cls[LB][bi][2];
          const int e0 = g_cart_cls[LC][ci][0], e1 = g_cart_cls[LC][ci][1], e2 = g_cart_cls[LC][ci][2];
          const double* Ex = s_E[0] + (a0 * (LB + 1) + b0) * MTAB;
          const double* Ey = s_E[1] + (a1 * (LB + 1) + b1) * MTAB;
          const double* Ez = s_E[2] + (a2 * (LB + 1) + b2) * MTAB;
          const double* Ecx = s_Ec + e0 * MT;
          const double* Ecy = s_Ec + e1 * MT;
          const double* Ecz = s_Ec + e2 * MT;
          double sum = 0.0;
          for (int t = 0; t <= a0 + b0; ++t) {
            const double et = Ex[t];
            if (et == 0.0) continue;
            for (int u = 0; u <= a1 + b1; ++u) {
              const double eu = et * Ey[u];
              if (eu == 0.0) continue;
              for (int v = 0; v <= a2 + b2; ++v) {
                const double ev = eu * Ez[v];
                if (ev == 0.0) continue;
                double inner = 0.0;
                for (int ta = e0 & 1; ta <= e0; ta += 2) {
                  const double gx = Ecx[ta];
                  for (int nu = e1 & 1; nu <= e1; nu += 2) {
                    const double gy = gx * Ecy[nu];
                    for (int ph = e2 & 1; ph <= e2; ph += 2) inner += gy * Ecz[ph] * R[rx(t + ta, u + nu, v + ph)];
                  }
                }
                sum += ev * inner;
              }
            }
          }
          s_cart[o] += pref * sum;
        }
      }
    }

  // cart -> final span (spherical aux; orbital span per A.spherical) and store
  __syncthreads();
  const int sphA = A.spherical && LA >= 2, sphB = A.spherical && LB >= 2, sphC = C.spherical && LC >= 2;
  const int fa = sphA ? nsph(LA) : NA, fb = sphB ? nsph(LB) : NB, fc = sphC ? nsph(LC) : NC;
  const int64_t ao_a = A.ao0[ia], ao_b = A.ao0[ib], ao_c = C.ao0[ic] - c0;
  for (int o = lane; o < fa * fb * fc; o += 64) {
    const int Ci = o % fc, Bi = (o / fc) % fb, Ai = o / (fc * fb);
    const int nA = sphA ? g_c2s_cls.n[LA][Ai] : 1, nB = sphB ? g_c2s_cls.n[LB][Bi] : 1, nC = sphC ? g_c2s_cls.n[LC][Ci] : 1;
    double v = 0.0;
    for (int x = 0; x < nA; ++x) {
      const int ai = sphA ? g_c2s_cls.cart[LA][Ai][x] : Ai;
      const double wa = sphA ? g_c2s_cls.coef[LA][Ai][x] : 1.0;
      for (int y = 0; y < nB; ++y) {
        const int bi = sphB ? g_c2s_cls.cart[LB][Bi][y] : Bi;
        const double wb = wa * (sphB ? g_c2s_cls.coef[LB][Bi][y] : 1.0);
        for (int z = 0; z < nC; ++z) {
          const int ci = sphC ? g_c2s_cls.cart[LC][Ci][z] : Ci;
          const double wc = wb * (sphC ? g_c2s_cls.coef[LC][Ci][z] : 1.0);
          v += wc * s_cart[(ai * NB + bi) * NC + ci];
        }
      }
    }
    out[(ao_a + Ai) * sa + (ao_b + Bi) * sb + ao_c + Ci] = v;
    if (tr && ia != ib) out[(ao_b + Bi) * sa + (ao_a + Ai) * sb + ao_c + Ci] = v;
  }
}

template <int LA, int LB, int LC>
static hipError_t go_cls(const int2* pairs, int npair, const int32_t* aux, int naux_l, const ShellSet& A,
                         const ShellSet& C, const double* boys, double* out, int64_t sa, int64_t sb, int64_t c0,
                         int tr, hipStream_t st) {
  const int64_t nblk = (int64_t)npair * naux_l;
  const int64_t chunk = 1LL << 24;
  for (int64_t b0 = 0; b0 < nblk; b0 += chunk) {
    const int64_t nb = nblk - b0 < chunk ? nblk - b0 : chunk;
    hipLaunchKernelGGL((k_eri_cls<LA, LB, LC>), dim3((unsigned)nb), dim3(64), 0, st, pairs, aux, npair, naux_l, A, C,
                       boys, out, sa, sb, c0, tr, b0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Every 3-centre class with l_a >= l_b, l_a <= 3, l_c <= 4.
hipError_t launch_eri_cls(int la, int lb, int lc, const int2* pairs, int npair, const int32_t* aux, int naux_l,
                          const ShellSet& A, const ShellSet& C, const double* boys, double* out, int64_t sa,
                          int64_t sb, int64_t c0, int tr, hipStream_t st) {
  if (npair == 0 || naux_l == 0) return hipSuccess;
#define DFKS_CLS(a, b) \
  if (la == a && lb == b) { \
    switch (lc) { \
      case 0: return go_cls<a, b, 0>(pairs, npair, aux, naux_l, A, C, boys, out, sa, sb, c0, tr, st); \
      case 1: return go_cls<a, b, 1>(pairs, npair, aux, naux_l, A, C, boys, out, sa, sb, c0, tr, st); \
      case 2: return go_cls<a, b, 2>(pairs, npair, aux, naux_l, A, C, boys, out, sa, sb, c0, tr, st); \
      case 3: return go_cls<a, b, 3>(pairs, npair, aux, naux_l, A, C, boys, out, sa, sb, c0, tr, st); \
      case 4: return go_cls<a, b, 4>(pairs, npair, aux, naux_l, A, C, boys, out, sa, sb, c0, tr, st); \
      default: return hipErrorInvalidValue; \
    } \
  }
  DFKS_CLS(0, 0) DFKS_CLS(1, 0) DFKS_CLS(1, 1) DFKS_CLS(2, 0) DFKS_CLS(2, 1) DFKS_CLS(2, 2)
  DFKS_CLS(3, 0) DFKS_CLS(3, 1) DFKS_CLS(3, 2) DFKS_CLS(3, 3)
#undef DFKS_CLS
  return hipErrorInvalidValue;
}

}  // namespace dfks
