// Three-centre (ab|c) for the low angular-momentum classes: one thread per
// task (bra shell pair x auxiliary shell), the whole McMurchie-Davidson
// evaluation in registers.
//
// The wavefront-per-task kernel (eri.hip) leaves most lanes idle on classes
// such as (ss|s) or (ps|p), which hold most of the tasks of a def2 basis with
// an even-tempered fitting set. Here each class is a template instance, so
// the Hermite E tables, the R ladder from (-2 rho)^m F_m, the auxiliary
// single-centre expansion and the cartesian->spherical transforms are fully
// unrolled with compile-time indices. Same formula as eri.hip
// (dftax/integrals/eri3c_bucketed.py:290-426):
//   (ab|c) = sum_prims c_a c_b c_c K_AB 2 pi^2.5 / (p q sqrt(p+q)) (-1)^{l_c}
//            sum_{tuv} E^{ab}_t E^{ab}_u E^{ab}_v
//            sum_{tau nu phi} E^c_tau E^c_nu E^c_phi R_{t+tau, u+nu, v+phi}
// Adjacent threads take adjacent auxiliary shells of one bra pair, so a wave
// shares the bra pair's data and writes neighbouring output rows.
#include <cstdint>

#include "dfks_common.h"

namespace dfks {

// compact index of (t,u,v) in the t+u+v <= L tetrahedron (same order as eri.hip)
__host__ __device__ constexpr int rid(int t, int u, int v) {
  return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
__host__ __device__ constexpr int nr_of(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Decoded compact index: (t, u, v) of entry e of the tetrahedron, built at
// compile time so every loop below has a constant trip count and unrolls to
// straight-line code with register-resident tables.
struct Tuv {
  int t[84], u[84], v[84];
};
__host__ __device__ constexpr Tuv make_tuv() {
  Tuv d{};
  int e = 0;
  for (int s = 0; s <= 6; ++s)
    for (int r = 0; r <= s; ++r)
      for (int v = 0; v <= r; ++v) {
        d.t[e] = s - r;
        d.u[e] = r - v;
        d.v[e] = v;
        ++e;
      }
  return d;
}
constexpr Tuv kTuv = make_tuv();

// One level of the ladder: cur[e] for t+u+v <= TOP from prev (level m+1).
template <int TOP>
__device__ __forceinline__ void r_level(const double* prev, double* cur, double x, double y, double z, double base) {
#pragma unroll
  for (int e = 0; e < nr_of(TOP); ++e) {
    const int t = kTuv.t[e], u = kTuv.u[e], v = kTuv.v[e];
    double val;
    if (e == 0) {
      val = base;
    } else if (t > 0) {
      val = x * prev[rid(t - 1, u, v)];
      if (t > 1) val += (t - 1) * prev[rid(t - 2, u, v)];
    } else if (u > 0) {
      val = y * prev[rid(t, u - 1, v)];
      if (u > 1) val += (u - 1) * prev[rid(t, u - 2, v)];
    } else {
      val = z * prev[rid(t, u, v - 1)];
      if (v > 1) val += (v - 1) * prev[rid(t, u, v - 2)];
    }
    cur[e] = val;
  }
}

template <int L, int TOP>
__device__ __forceinline__ void r_levels(double* a, double* b, double x, double y, double z, const double* pw,
                                         const double* F) {
  if constexpr (TOP <= L) {
    r_level<TOP>(a, b, x, y, z, pw[L - TOP] * F[L - TOP]);
    r_levels<L, TOP + 1>(b, a, x, y, z, pw, F);
  }
}

// R^{(0)}_{tuv}, t+u+v <= L, by the descending ladder; the result ends in
// R (L even) or in the scratch half (L odd), returned.
template <int L>
__device__ __forceinline__ const double* r_build(double n2r, double x, double y, double z, const double* F,
                                                 double* R, double* tmp) {
  double pw[L + 1];
  pw[0] = 1.0;
#pragma unroll
  for (int k = 0; k < L; ++k) pw[k + 1] = pw[k] * n2r;
  R[0] = pw[L] * F[L];
  r_levels<L, 1>(R, tmp, x, y, z, pw, F);
  return (L & 1) ? tmp : R;
}

// Two-centre Hermite coefficients of one axis: E[i][j][t], i <= LA, j <= LB.
template <int LA, int LB>
__device__ __forceinline__ void e_build(double xpa, double xpb, double i2p, double (&E)[LA + 1][LB + 1][LA + LB + 1]) {
  constexpr int MT = LA + LB + 1;
#pragma unroll
  for (int i = 0; i <= LA; ++i)
#pragma unroll
    for (int j = 0; j <= LB; ++j)
#pragma unroll
      for (int t = 0; t < MT; ++t) E[i][j][t] = 0.0;
  E[0][0][0] = 1.0;
#pragma unroll
  for (int i = 0; i < LA; ++i)
#pragma unroll
    for (int t = 0; t <= i + 1; ++t) {
      double v = xpa * E[i][0][t];
      if (t > 0) v += i2p * E[i][0][t - 1];
      if (t + 1 <= i) v += (t + 1) * E[i][0][t + 1];
      E[i + 1][0][t] = v;
    }
#pragma unroll
  for (int j = 0; j < LB; ++j)
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = xpb * E[i][j][t];
        if (t > 0) v += i2p * E[i][j][t - 1];
        if (t + 1 <= i + j) v += (t + 1) * E[i][j][t + 1];
        E[i][j + 1][t] = v;
      }
}

template <int L>
__device__ __forceinline__ constexpr int nfun_t(int sph) { return (sph && L >= 2) ? nsph(L) : ncart(L); }

template <int LA, int LB, int LC, bool SAB>
__global__ __launch_bounds__(256) void k_eri_thread(const int2* __restrict__ pairs, const int32_t* __restrict__ aux,
                                                    int npair, int naux_l, ShellSet A, ShellSet C,
                                                    const double* __restrict__ boys, double* __restrict__ out,
                                                    int64_t sa, int64_t sb, int64_t c0, int tr, int64_t ntask) {
  constexpr int L = LA + LB + LC;
  constexpr int LAB = LA + LB;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC);
  const int64_t task = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (task >= ntask) return;
  const int k = (int)(task / naux_l);
  const int jx = (int)(task % naux_l);
  const int2 pr = pairs[k];
  const int ia = pr.x, ib = pr.y, ic = aux[jx];

  const double Ax = A.centre[3 * ia], Ay = A.centre[3 * ia + 1], Az = A.centre[3 * ia + 2];
  const double Bx = A.centre[3 * ib], By = A.centre[3 * ib + 1], Bz = A.centre[3 * ib + 2];
  const double Cx = C.centre[3 * ic], Cy = C.centre[3 * ic + 1], Cz = C.centre[3 * ic + 2];
  const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
  const double ab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  const int npa = A.nprim[ia], npb = A.nprim[ib], npc = C.nprim[ic];
  const int p0a = A.prim0[ia], p0b = A.prim0[ib], p0c = C.prim0[ic];

  double acc[NA * NB * NC];
#pragma unroll
  for (int o = 0; o < NA * NB * NC; ++o) acc[o] = 0.0;

  for (int pa = 0; pa < npa; ++pa) {
    const double al = A.exps[p0a + pa], ca = A.coefs[p0a + pa];
    for (int pb = 0; pb < npb; ++pb) {
      const double be = A.exps[p0b + pb];
      const double p = al + be, ip = 1.0 / p;
      const double Px = (al * Ax + be * Bx) * ip, Py = (al * Ay + be * By) * ip, Pz = (al * Az + be * Bz) * ip;
      const double cab = ca * A.coefs[p0b + pb] * exp(-al * be * ip * ab2);
      double Ex[LA + 1][LB + 1][LAB + 1], Ey[LA + 1][LB + 1][LAB + 1], Ez[LA + 1][LB + 1][LAB + 1];
      const double i2p = 0.5 * ip;
      e_build<LA, LB>(Px - Ax, Px - Bx, i2p, Ex);
      e_build<LA, LB>(Py - Ay, Py - By, i2p, Ey);
      e_build<LA, LB>(Pz - Az, Pz - Bz, i2p, Ez);
      for (int pc = 0; pc < npc; ++pc) {
        const double ga = C.exps[p0c + pc];
        const double rho = p * ga / (p + ga);
        double pref = cab * C.coefs[p0c + pc] * 2.0 * 17.493418327624862 / (p * ga * sqrt(p + ga));
        if (LC & 1) pref = -pref;
        const double pcx = Px - Cx, pcy = Py - Cy, pcz = Pz - Cz;
        double F[L + 1];
        boys_op<L>(rho * (pcx * pcx + pcy * pcy + pcz * pcz), rho, C.omega, F, boys);
        double Ra[nr_of(L)], Rb[nr_of(L)];
        const double* R = r_build<L>(-2.0 * rho, pcx, pcy, pcz, F, Ra, Rb);
        // single-centre expansion of the auxiliary Gaussian: Ec[k][tau], tau <= k
        double Ec[LC + 1][LC + 1];
#pragma unroll
        for (int a = 0; a <= LC; ++a)
#pragma unroll
          for (int b = 0; b <= LC; ++b) Ec[a][b] = 0.0;
        Ec[0][0] = 1.0;
        const double i2g = 0.5 / ga;
#pragma unroll
        for (int kk = 0; kk < LC; ++kk)
#pragma unroll
          for (int t = 0; t <= kk + 1; ++t) {
            double v = 0.0;
            if (t > 0) v += i2g * Ec[kk][t - 1];
            if (t + 1 <= kk) v += (t + 1) * Ec[kk][t + 1];
            Ec[kk + 1][t] = v;
          }
#pragma unroll
        for (int ci = 0; ci < NC; ++ci) {
          const int c0x = cart_exp<LC>(ci, 0), c1x = cart_exp<LC>(ci, 1), c2x = cart_exp<LC>(ci, 2);
          // G[tuv] = sum_{tau nu phi} Ec Ec Ec R[t+tau, u+nu, v+phi], t+u+v <= LAB
          double G[nr_of(LAB)];
#pragma unroll
          for (int e = 0; e < nr_of(LAB); ++e) {
            const int t = kTuv.t[e], u = kTuv.u[e], v = kTuv.v[e];
            {
                double s = 0.0;
#pragma unroll
                for (int ta = 0; ta <= LC; ++ta) {
                  if (ta > c0x || ((c0x - ta) & 1)) continue;
#pragma unroll
                  for (int nu = 0; nu <= LC; ++nu) {
                    if (nu > c1x || ((c1x - nu) & 1)) continue;
#pragma unroll
                    for (int ph = 0; ph <= LC; ++ph) {
                      if (ph > c2x || ((c2x - ph) & 1)) continue;
                      s += Ec[c0x][ta] * Ec[c1x][nu] * Ec[c2x][ph] * R[rid(t + ta, u + nu, v + ph)];
                    }
                  }
                }
                G[rid(t, u, v)] = s;
            }
          }
#pragma unroll
          for (int ai = 0; ai < NA; ++ai) {
            const int a0 = cart_exp<LA>(ai, 0), a1 = cart_exp<LA>(ai, 1), a2 = cart_exp<LA>(ai, 2);
#pragma unroll
            for (int bi = 0; bi < NB; ++bi) {
              const int b0 = cart_exp<LB>(bi, 0), b1 = cart_exp<LB>(bi, 1), b2 = cart_exp<LB>(bi, 2);
              double s = 0.0;
#pragma unroll
              for (int t = 0; t <= LAB; ++t) {
                if (t > a0 + b0) continue;
#pragma unroll
                for (int u = 0; u <= LAB; ++u) {
                  if (u > a1 + b1) continue;
#pragma unroll
                  for (int v = 0; v <= LAB; ++v) {
                    if (v > a2 + b2) continue;
                    s += Ex[a0][b0][t] * Ey[a1][b1][u] * Ez[a2][b2][v] * G[rid(t, u, v)];
                  }
                }
              }
              acc[(ai * NB + bi) * NC + ci] += pref * s;
            }
          }
        }
      }
    }
  }

  // cartesian -> final span (spherical aux; orbital span per SAB) and store (both bra triangles)
  const int64_t ao_a = A.ao0[ia], ao_b = A.ao0[ib], ao_c = C.ao0[ic] - c0;
  constexpr bool SA = SAB && LA >= 2, SB = SAB && LB >= 2, SC = LC >= 2;
  constexpr int FA = SA ? nsph(LA) : NA, FB = SB ? nsph(LB) : NB, FC = SC ? nsph(LC) : NC;
#pragma unroll
  for (int A_ = 0; A_ < FA; ++A_)
#pragma unroll
    for (int B_ = 0; B_ < FB; ++B_)
#pragma unroll
      for (int C_ = 0; C_ < FC; ++C_) {
        double v = 0.0;
#pragma unroll
        for (int ai = 0; ai < NA; ++ai) {
          const double wa = SA ? c2s_coef<LA>(ai, A_) : (ai == A_ ? 1.0 : 0.0);
          if (wa == 0.0) continue;
#pragma unroll
          for (int bi = 0; bi < NB; ++bi) {
            const double wb = SB ? c2s_coef<LB>(bi, B_) : (bi == B_ ? 1.0 : 0.0);
            if (wb == 0.0) continue;
#pragma unroll
            for (int ci = 0; ci < NC; ++ci) {
              const double wc = SC ? c2s_coef<LC>(ci, C_) : (ci == C_ ? 1.0 : 0.0);
              if (wc == 0.0) continue;
              v += wa * wb * wc * acc[(ai * NB + bi) * NC + ci];
            }
          }
        }
        out[(ao_a + A_) * sa + (ao_b + B_) * sb + ao_c + C_] = v;
        if (tr && ia != ib) out[(ao_b + B_) * sa + (ao_a + A_) * sb + ao_c + C_] = v;
      }
}

template <int LA, int LB, int LC>
static hipError_t go(const int2* pairs, int npair, const int32_t* aux, int naux_l, const ShellSet& A,
                     const ShellSet& C, const double* boys, double* out, int64_t sa, int64_t sb, int64_t c0, int tr,
                     hipStream_t st) {
  const int64_t ntask = (int64_t)npair * naux_l;
  const int64_t nb = (ntask + 255) / 256;
  if (A.spherical)
    hipLaunchKernelGGL((k_eri_thread<LA, LB, LC, true>), dim3((unsigned)nb), dim3(256), 0, st, pairs, aux, npair,
                       naux_l, A, C, boys, out, sa, sb, c0, tr, ntask);
  else
    hipLaunchKernelGGL((k_eri_thread<LA, LB, LC, false>), dim3((unsigned)nb), dim3(256), 0, st, pairs, aux, npair,
                       naux_l, A, C, boys, out, sa, sb, c0, tr, ntask);
  return hipGetLastError();
}

// Classes handled here (la >= lb, spherical auxiliary span): those whose
// unrolled evaluation stays in registers on gfx950 (no scratch); the rest run
// on the wavefront kernel.
#define DFKS_ERI_THREAD_CLASSES(X)                                                   \
  X(0, 0, 0) X(0, 0, 1) X(0, 0, 2) X(0, 0, 3) X(0, 0, 4)                             \
  X(1, 0, 0) X(1, 0, 1) X(1, 0, 2) X(1, 0, 3)                                        \
  X(1, 1, 0) X(1, 1, 1) X(1, 1, 2)                                                   \
  X(2, 0, 0) X(2, 0, 1) X(2, 0, 2) X(2, 0, 3)                                        \
  X(2, 1, 0)                                                                         \
  X(3, 0, 0) X(3, 0, 1)

bool eri_thread_has(int la, int lb, int lc) {
#define DFKS_HAS(a, b, c) if (la == a && lb == b && lc == c) return true;
  DFKS_ERI_THREAD_CLASSES(DFKS_HAS)
#undef DFKS_HAS
  return false;
}

hipError_t launch_eri_thread(int la, int lb, int lc, const int2* pairs, int npair, const int32_t* aux, int naux_l,
                             const ShellSet& A, const ShellSet& C, const double* boys, double* out, int64_t sa,
                             int64_t sb, int64_t c0, int tr, hipStream_t st) {
  if (npair == 0 || naux_l == 0) return hipSuccess;
  if ((int64_t)npair * naux_l > (int64_t)INT32_MAX * 128) return hipErrorInvalidValue;
  if (!C.spherical) return hipErrorInvalidValue;
#define DFKS_GO(a, b, c) \
  if (la == a && lb == b && lc == c) return go<a, b, c>(pairs, npair, aux, naux_l, A, C, boys, out, sa, sb, c0, tr, st);
  DFKS_ERI_THREAD_CLASSES(DFKS_GO)
#undef DFKS_GO
  return hipErrorInvalidValue;
}

}  // namespace dfks
