// McMurchie-Davidson 1-, 2- and 3-centre integrals on gfx950 (FP64).
//
// Algorithm (dftax/integrals/eri3c_bucketed.py:290-426, 697-734):
//   (ab|c) = K_AB 2 pi^2.5 / (p q sqrt(p+q)) *
//            sum_{tuv} E^{ab}_t E^{ab}_u E^{ab}_v
//            sum_{tau nu phi} (-1)^{tau+nu+phi} E^c_tau E^c_nu E^c_phi R_{t+tau,u+nu,v+phi}
// with the Hermite E tables of the bra pair, the single-centre E table of the
// auxiliary Gaussian, and R built by the descending m-ladder from
// (-2 rho)^m F_m(rho |PC|^2).
//
// Work decomposition: one 64-lane wavefront per task (bra shell pair x aux
// shell). The R ladder is built cooperatively in LDS in compact (t+u+v <= L)
// indexing, each lane accumulates a strided subset of the cartesian outputs
// in an LDS block, and the cart->sph transform reads that block with sparse
// coefficient lists before the result is written (both bra triangles for 3c).
// Limits: orbital l <= 3, auxiliary l <= 4 (LDS sized for (ff|g)).
#include <cstring>

#include "dfks_common.h"

namespace dfks {

constexpr int kWave = 64;
constexpr int kMaxL3 = 3 + 3 + 4;                       // (ff|g)
constexpr int kNR = (kMaxL3 + 1) * (kMaxL3 + 2) * (kMaxL3 + 3) / 6;  // 286
constexpr int kMaxCartBlk = 10 * 10 * 15;
constexpr int kMaxOutPerLane = (kMaxCartBlk + kWave - 1) / kWave;     // 24

__device__ __forceinline__ int ridx(int t, int u, int v) {
  int s = t + u + v;
  int r = s - t;
  return s * (s + 1) * (s + 2) / 6 + r * (r + 1) / 2 + v;
}

__constant__ C2SList g_c2s;

// cartesian exponent table for runtime l (l <= 6)
__constant__ int g_cart[7][28][3];

__device__ __forceinline__ int cexp(int l, int c, int x) { return g_cart[l][c][x]; }

struct Task3 {
  const int2* pairs;     // (ia, ib) per bra pair of this class
  const int32_t* aux;    // aux shell ids of this lc
  int npair, naux_l;
};

// E^{ij}_t for one axis into a [(la+1)][(lb+1)][mt] LDS block (lanes 0..2 call it).
__device__ void e_table_1d(int la, int lb, double al, double be, double xab, int mt, double* E) {
  double p = al + be, i2g = 0.5 / p, xpa = -be * xab / p, xpb = al * xab / p;
  int nb = lb + 1;
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j)
      for (int t = 0; t < mt; ++t) E[(i * nb + j) * mt + t] = 0.0;
  E[0] = 1.0;
  for (int i = 0; i < la; ++i)
    for (int t = 0; t <= i + 1 && t < mt; ++t) {
      double v = xpa * E[(i * nb) * mt + t];
      if (t > 0) v += i2g * E[(i * nb) * mt + t - 1];
      if (t + 1 < mt) v += (t + 1) * E[(i * nb) * mt + t + 1];
      E[((i + 1) * nb) * mt + t] = v;
    }
  for (int j = 0; j < lb; ++j)
    for (int i = 0; i <= la; ++i)
      for (int t = 0; t <= i + j + 1 && t < mt; ++t) {
        double v = xpb * E[(i * nb + j) * mt + t];
        if (t > 0) v += i2g * E[(i * nb + j) * mt + t - 1];
        if (t + 1 < mt) v += (t + 1) * E[(i * nb + j) * mt + t + 1];
        E[(i * nb + j + 1) * mt + t] = v;
      }
}

// Build R^{(0)}_{tuv}, t+u+v <= L, in LDS (compact). buf0/buf1 ping-pong.
// Returns a pointer to the buffer that holds level 0.
__device__ double* r_ladder(int L, double rho, double pcx, double pcy, double pcz, const double* F,
                            double* buf0, double* buf1) {
  const int lane = threadIdx.x;
  double* prev = buf0;
  double* cur = buf1;
  double n2r = -2.0 * rho;
  // level L: only R_000
  if (lane == 0) {
    double pw = 1.0;
    for (int k = 0; k < L; ++k) pw *= n2r;
    prev[0] = pw * F[L];
  }
  __syncthreads();
  for (int m = L - 1; m >= 0; --m) {
    int top = L - m;                           // highest t+u+v at this level
    int count = (top + 1) * (top + 2) * (top + 3) / 6;
    for (int e = lane; e < count; e += kWave) {
      // decode compact index e -> (t,u,v)
      int s = 0;
      while ((s + 1) * (s + 2) * (s + 3) / 6 <= e) ++s;
      int rem = e - s * (s + 1) * (s + 2) / 6;
      int r = 0;
      while ((r + 1) * (r + 2) / 2 <= rem) ++r;
      int v = rem - r * (r + 1) / 2;
      int t = s - r, u = r - v;
      double val;
      if (s == 0) {
        double pw = 1.0;
        for (int k = 0; k < m; ++k) pw *= n2r;
        val = pw * F[m];
      } else if (t > 0) {
        val = pcx * prev[ridx(t - 1, u, v)] + (t > 1 ? (t - 1) * prev[ridx(t - 2, u, v)] : 0.0);
      } else if (u > 0) {
        val = pcy * prev[ridx(t, u - 1, v)] + (u > 1 ? (u - 1) * prev[ridx(t, u - 2, v)] : 0.0);
      } else {
        val = pcz * prev[ridx(t, u, v - 1)] + (v > 1 ? (v - 1) * prev[ridx(t, u, v - 2)] : 0.0);
      }
      cur[e] = val;
    }
    __syncthreads();
    double* tmp = prev;
    prev = cur;
    cur = tmp;
  }
  return prev;
}

// Generic 2-/3-centre kernel. For the 2-centre metric the bra partner b is a
// unit s function of zero exponent on A (b_unit != 0), which makes (a b|c)
// the two-centre (a|c).
// Output index: ao_a*sa + ao_b*sb + ao_c (+ transposed bra write if tr).
__global__ __launch_bounds__(64) void k_eri_wave(
    int la, int lb, int lc, Task3 task, ShellSet A, ShellSet C, int b_unit,
    const double* __restrict__ boys, double* __restrict__ out, int64_t sa, int64_t sb, int tr, int64_t blk0,
    int64_t c0) {
  __shared__ double s_E[3][4 * 4 * 8];
  __shared__ double s_Ec[5 * 11];
  __shared__ double s_F[12];
  __shared__ double s_R[2][kNR];
  __shared__ double s_cart[kMaxCartBlk];

  const int lane = threadIdx.x;
  const int64_t tid = blk0 + blockIdx.x;
  const int k = (int)(tid / task.naux_l);
  const int j = (int)(tid % task.naux_l);
  if (k >= task.npair) return;
  int2 pr = task.pairs[k];
  const int ia = pr.x, ib = pr.y, ic = task.aux[j];
  const int L = la + lb + lc;
  const int mt = L + 1;
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc);
  const int nout = na * nb * nc;

  double Ax = A.centre[3 * ia], Ay = A.centre[3 * ia + 1], Az = A.centre[3 * ia + 2];
  double Bx = Ax, By = Ay, Bz = Az;
  int npa = A.nprim[ia], npb = 1, p0a = A.prim0[ia], p0b = 0;
  if (!b_unit) {
    Bx = A.centre[3 * ib]; By = A.centre[3 * ib + 1]; Bz = A.centre[3 * ib + 2];
    npb = A.nprim[ib]; p0b = A.prim0[ib];
  }
  double Cx = C.centre[3 * ic], Cy = C.centre[3 * ic + 1], Cz = C.centre[3 * ic + 2];
  double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
  double ab2 = ABx * ABx + ABy * ABy + ABz * ABz;

  for (int o = lane; o < nout; o += kWave) s_cart[o] = 0.0;

  const int lab = la + lb;
  for (int pa = 0; pa < npa; ++pa)
    for (int pb = 0; pb < npb; ++pb) {
      double al = A.exps[p0a + pa];
      double be = b_unit ? 0.0 : A.exps[p0b + pb];
      double cab = A.coefs[p0a + pa] * (b_unit ? 1.0 : A.coefs[p0b + pb]);
      double p = al + be;
      double Px = (al * Ax + be * Bx) / p, Py = (al * Ay + be * By) / p, Pz = (al * Az + be * Bz) / p;
      double Kab = exp(-al * be / p * ab2);
      __syncthreads();
      if (lane < 3) {
        double xab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
        e_table_1d(la, lb, al, be, xab, lab + 1, s_E[lane]);
      }
      for (int pc = 0; pc < C.nprim[ic]; ++pc) {
        double ga = C.exps[C.prim0[ic] + pc];
        double cc = C.coefs[C.prim0[ic] + pc];
        double rho = p * ga / (p + ga);
        double pref = cab * cc * Kab * 2.0 * 17.493418327624862 / (p * ga * sqrt(p + ga));  // 2 pi^2.5
        double pcx = Px - Cx, pcy = Py - Cy, pcz = Pz - Cz;
        double T = rho * (pcx * pcx + pcy * pcy + pcz * pcz);
        __syncthreads();
        if (lane == 0) {
          boys_op_rt(L, T, rho, C.omega, s_F, boys);
          // single-centre E^k_tau for the aux Gaussian
          double i2g = 0.5 / ga;
          for (int q = 0; q < (lc + 1) * mt; ++q) s_Ec[q] = 0.0;
          s_Ec[0] = 1.0;
          for (int kk = 0; kk < lc; ++kk)
            for (int t = 0; t < mt; ++t) {
              double v = 0.0;
              if (t > 0) v += i2g * s_Ec[kk * mt + t - 1];
              if (t + 1 < mt) v += (t + 1) * s_Ec[kk * mt + t + 1];
              s_Ec[(kk + 1) * mt + t] = v;
            }
        }
        __syncthreads();
        const double* R = r_ladder(L, rho, pcx, pcy, pcz, s_F, s_R[0], s_R[1]);
        // contraction
        const int nbb = lb + 1;
        const int mtab = lab + 1;
        for (int o = lane; o < nout; o += kWave) {
          int ci = o % nc, bi = (o / nc) % nb, ai = o / (nc * nb);
          int a0 = cexp(la, ai, 0), a1 = cexp(la, ai, 1), a2 = cexp(la, ai, 2);
          int b0 = cexp(lb, bi, 0), b1 = cexp(lb, bi, 1), b2 = cexp(lb, bi, 2);
          int c0 = cexp(lc, ci, 0), c1 = cexp(lc, ci, 1), c2 = cexp(lc, ci, 2);
          const double* Ex = s_E[0] + (a0 * nbb + b0) * mtab;
          const double* Ey = s_E[1] + (a1 * nbb + b1) * mtab;
          const double* Ez = s_E[2] + (a2 * nbb + b2) * mtab;
          const double* Ecx = s_Ec + c0 * mt;
          const double* Ecy = s_Ec + c1 * mt;
          const double* Ecz = s_Ec + c2 * mt;
          double sum = 0.0;
          for (int t = 0; t <= a0 + b0; ++t) {
            double et = Ex[t];
            if (et == 0.0) continue;
            for (int u = 0; u <= a1 + b1; ++u) {
              double eu = et * Ey[u];
              if (eu == 0.0) continue;
              for (int v = 0; v <= a2 + b2; ++v) {
                double ev = eu * Ez[v];
                if (ev == 0.0) continue;
                double inner = 0.0;
                for (int ta = c0 & 1; ta <= c0; ta += 2) {
                  double gx = Ecx[ta];
                  for (int nu = c1 & 1; nu <= c1; nu += 2) {
                    double gy = gx * Ecy[nu];
                    for (int ph = c2 & 1; ph <= c2; ph += 2) {
                      double g = gy * Ecz[ph];
                      inner += g * R[ridx(t + ta, u + nu, v + ph)];
                    }
                  }
                }
                // (-1)^{tau+nu+phi} = (-1)^{c0+c1+c2} since tau == c0 mod 2 etc.
                sum += ev * inner;
              }
            }
          }
          s_cart[o] += ((lc & 1) ? -pref : pref) * sum;
        }
      }
    }

  // cart -> final span via LDS
  __syncthreads();
  const int sphA = A.spherical && la >= 2, sphB = A.spherical && lb >= 2 && !b_unit, sphC = C.spherical && lc >= 2;
  const int fa = sphA ? nsph(la) : na, fb = sphB ? nsph(lb) : nb, fc = sphC ? nsph(lc) : nc;
  const int nf = fa * fb * fc;
  const int64_t ao_a = A.ao0[ia], ao_b = b_unit ? 0 : A.ao0[ib], ao_c = C.ao0[ic] - c0;
  for (int o = lane; o < nf; o += kWave) {
    int Ci = o % fc, Bi = (o / fc) % fb, Ai = o / (fc * fb);
    int nA = sphA ? g_c2s.n[la][Ai] : 1;
    int nB = sphB ? g_c2s.n[lb][Bi] : 1;
    int nC = sphC ? g_c2s.n[lc][Ci] : 1;
    double v = 0.0;
    for (int x = 0; x < nA; ++x) {
      int ai = sphA ? g_c2s.cart[la][Ai][x] : Ai;
      double wa = sphA ? g_c2s.coef[la][Ai][x] : 1.0;
      for (int y = 0; y < nB; ++y) {
        int bi = sphB ? g_c2s.cart[lb][Bi][y] : Bi;
        double wb = wa * (sphB ? g_c2s.coef[lb][Bi][y] : 1.0);
        for (int z = 0; z < nC; ++z) {
          int ci = sphC ? g_c2s.cart[lc][Ci][z] : Ci;
          double wc = wb * (sphC ? g_c2s.coef[lc][Ci][z] : 1.0);
          v += wc * s_cart[(ai * nb + bi) * nc + ci];
        }
      }
    }
    out[(ao_a + Ai) * sa + (ao_b + Bi) * sb + ao_c + Ci] = v;
    if (tr && ia != ib) out[(ao_b + Bi) * sa + (ao_a + Ai) * sb + ao_c + Ci] = v;
  }
}

// ---------------------------------------------------------------------------
// One-electron S, T, V: one wave per ordered shell pair, lanes over the
// cartesian outputs; the nuclear R ladder per (primitive pair, atom) in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_int1e(ShellSet A, int natm, const double* __restrict__ atm,
                                              const double* __restrict__ boys, double* __restrict__ S,
                                              double* __restrict__ Tm, double* __restrict__ V, int64_t nao) {
  __shared__ double s_E[3][4 * 6 * 10];
  __shared__ double s_F[12];
  __shared__ double s_R[2][84];
  __shared__ double s_cart[3][100];
  const int lane = threadIdx.x;
  const int ia = blockIdx.x / A.n, ib = blockIdx.x % A.n;
  const int la = A.l[ia], lb = A.l[ib];
  const int na = ncart(la), nb = ncart(lb);
  const int L = la + lb;
  const int mtS = L + 3;   // kinetic needs j + 2
  const int nbb = lb + 3;
  double Ax = A.centre[3 * ia], Ay = A.centre[3 * ia + 1], Az = A.centre[3 * ia + 2];
  double Bx = A.centre[3 * ib], By = A.centre[3 * ib + 1], Bz = A.centre[3 * ib + 2];
  double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;
  double ab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  const int nout = na * nb;
  for (int o = lane; o < nout; o += kWave) { s_cart[0][o] = 0.0; s_cart[1][o] = 0.0; s_cart[2][o] = 0.0; }
  for (int pa = 0; pa < A.nprim[ia]; ++pa)
    for (int pb = 0; pb < A.nprim[ib]; ++pb) {
      double al = A.exps[A.prim0[ia] + pa], be = A.exps[A.prim0[ib] + pb];
      double cab = A.coefs[A.prim0[ia] + pa] * A.coefs[A.prim0[ib] + pb];
      double p = al + be;
      double Px = (al * Ax + be * Bx) / p, Py = (al * Ay + be * By) / p, Pz = (al * Az + be * Bz) / p;
      double Kab = exp(-al * be / p * ab2);
      __syncthreads();
      if (lane < 3) {
        double xab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
        e_table_1d(la, lb + 2, al, be, xab, mtS, s_E[lane]);
      }
      __syncthreads();
      double s0 = pow(kPi / p, 1.5) * Kab * cab;
      for (int o = lane; o < nout; o += kWave) {
        int ai = o / nb, bi = o % nb;
        double s1[3], t1[3];
        for (int x = 0; x < 3; ++x) {
          int i = cexp(la, ai, x), jj = cexp(lb, bi, x);
          const double* E = s_E[x];
          double sij = E[(i * nbb + jj) * mtS];
          double tt = -2.0 * be * be * E[(i * nbb + jj + 2) * mtS] + be * (2 * jj + 1) * sij;
          if (jj >= 2) tt -= 0.5 * jj * (jj - 1) * E[(i * nbb + jj - 2) * mtS];
          s1[x] = sij;
          t1[x] = tt;
        }
        s_cart[0][o] += s0 * s1[0] * s1[1] * s1[2];
        s_cart[1][o] += s0 * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
      }
      for (int c = 0; c < natm; ++c) {
        double pcx = Px - atm[4 * c], pcy = Py - atm[4 * c + 1], pcz = Pz - atm[4 * c + 2];
        double Z = atm[4 * c + 3];
        __syncthreads();
        if (lane == 0) boys_array_rt(L, p * (pcx * pcx + pcy * pcy + pcz * pcz), s_F, boys);
        __syncthreads();
        const double* R = r_ladder(L, p, pcx, pcy, pcz, s_F, s_R[0], s_R[1]);
        double pref = -Z * cab * Kab * 2.0 * kPi / p;
        for (int o = lane; o < nout; o += kWave) {
          int ai = o / nb, bi = o % nb;
          int a0 = cexp(la, ai, 0), a1 = cexp(la, ai, 1), a2 = cexp(la, ai, 2);
          int b0 = cexp(lb, bi, 0), b1 = cexp(lb, bi, 1), b2 = cexp(lb, bi, 2);
          const double* Ex = s_E[0] + (a0 * nbb + b0) * mtS;
          const double* Ey = s_E[1] + (a1 * nbb + b1) * mtS;
          const double* Ez = s_E[2] + (a2 * nbb + b2) * mtS;
          double sum = 0.0;
          for (int t = 0; t <= a0 + b0; ++t)
            for (int u = 0; u <= a1 + b1; ++u)
              for (int v = 0; v <= a2 + b2; ++v) sum += Ex[t] * Ey[u] * Ez[v] * R[ridx(t, u, v)];
          s_cart[2][o] += pref * sum;
        }
        __syncthreads();
      }
    }
  __syncthreads();
  const int sphA = A.spherical && la >= 2, sphB = A.spherical && lb >= 2;
  const int fa = sphA ? nsph(la) : na, fb = sphB ? nsph(lb) : nb;
  for (int o = lane; o < fa * fb; o += kWave) {
    int Ai = o / fb, Bi = o % fb;
    int nA = sphA ? g_c2s.n[la][Ai] : 1, nB = sphB ? g_c2s.n[lb][Bi] : 1;
    double v[3] = {0, 0, 0};
    for (int x = 0; x < nA; ++x) {
      int ai = sphA ? g_c2s.cart[la][Ai][x] : Ai;
      double wa = sphA ? g_c2s.coef[la][Ai][x] : 1.0;
      for (int y = 0; y < nB; ++y) {
        int bi = sphB ? g_c2s.cart[lb][Bi][y] : Bi;
        double w = wa * (sphB ? g_c2s.coef[lb][Bi][y] : 1.0);
        for (int z = 0; z < 3; ++z) v[z] += w * s_cart[z][ai * nb + bi];
      }
    }
    int64_t idx = (int64_t)(A.ao0[ia] + Ai) * nao + A.ao0[ib] + Bi;
    S[idx] = v[0];
    Tm[idx] = v[1];
    V[idx] = v[2];
  }
}

// ---------------------------------------------------------------------------
// Schwarz factors of bra shell pairs: Q_ab = max over cartesian components
// (mu nu) of sqrt((mu nu|mu nu)) (dftax/integrals/eri4c.py:206-220, taken per
// shell pair as in eri3c_bucketed.py:140-169). One wavefront per pair, lanes
// over the component pairs, primitive quartets in sequence:
//   (ab|ab) = sum_{p,q} K_p K_q 2 pi^2.5 / (g_p g_q sqrt(g_p + g_q))
//             sum_{tuv} E^p_{tuv} sum_{tau nu phi} (-1)^{tau+nu+phi} E^q_{tau nu phi} R_{t+tau,u+nu,v+phi}
// ---------------------------------------------------------------------------
constexpr int kMaxLQ = 2 * (3 + 3);                                     // (ff|ff)
constexpr int kNRQ = (kMaxLQ + 1) * (kMaxLQ + 2) * (kMaxLQ + 3) / 6;     // 455

__global__ __launch_bounds__(64) void k_schwarz(int la, int lb, const int2* __restrict__ pairs, int npair, ShellSet A,
                                                const double* __restrict__ boys, double* __restrict__ q_out) {
  __shared__ double s_Ep[3][4 * 4 * 7];
  __shared__ double s_Eq[3][4 * 4 * 7];
  __shared__ double s_F[kMaxLQ + 1];
  __shared__ double s_R[2][kNRQ];
  const int lane = threadIdx.x;
  const int k = blockIdx.x;
  if (k >= npair) return;
  const int2 pr = pairs[k];
  const int ia = pr.x, ib = pr.y;
  const int lab = la + lb, L = 2 * lab, mtab = lab + 1, nbb = lb + 1;
  const int na = ncart(la), nb = ncart(lb);
  const double Ax = A.centre[3 * ia], Ay = A.centre[3 * ia + 1], Az = A.centre[3 * ia + 2];
  const double Bx = A.centre[3 * ib], By = A.centre[3 * ib + 1], Bz = A.centre[3 * ib + 2];
  const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz, ab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  const int npa = A.nprim[ia], npb = A.nprim[ib], p0a = A.prim0[ia], p0b = A.prim0[ib];
  const int npp = npa * npb;
  double acc[2] = {0.0, 0.0};  // components lane and lane + 64 (na * nb <= 100)
  for (int p = 0; p < npp; ++p) {
    const double al = A.exps[p0a + p / npb], be = A.exps[p0b + p % npb];
    const double gp = al + be;
    const double Px = (al * Ax + be * Bx) / gp, Py = (al * Ay + be * By) / gp, Pz = (al * Az + be * Bz) / gp;
    const double Kp = A.coefs[p0a + p / npb] * A.coefs[p0b + p % npb] * exp(-al * be / gp * ab2);
    __syncthreads();
    if (lane < 3) {
      const double xab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
      e_table_1d(la, lb, al, be, xab, mtab, s_Ep[lane]);
    }
    for (int q = 0; q < npp; ++q) {
      const double ga = A.exps[p0a + q / npb], de = A.exps[p0b + q % npb];
      const double gq = ga + de;
      const double Qx = (ga * Ax + de * Bx) / gq, Qy = (ga * Ay + de * By) / gq, Qz = (ga * Az + de * Bz) / gq;
      const double Kq = A.coefs[p0a + q / npb] * A.coefs[p0b + q % npb] * exp(-ga * de / gq * ab2);
      const double rho = gp * gq / (gp + gq);
      const double pqx = Px - Qx, pqy = Py - Qy, pqz = Pz - Qz;
      const double pref = Kp * Kq * 2.0 * 17.493418327624862 / (gp * gq * sqrt(gp + gq));
      __syncthreads();
      if (lane < 3) {
        const double xab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
        e_table_1d(la, lb, ga, de, xab, mtab, s_Eq[lane]);
      }
      if (lane == 0) boys_array_rt(L, rho * (pqx * pqx + pqy * pqy + pqz * pqz), s_F, boys);
      __syncthreads();
      const double* R = r_ladder(L, rho, pqx, pqy, pqz, s_F, s_R[0], s_R[1]);
      for (int h = 0; h < 2; ++h) {
        const int o = lane + 64 * h;
        if (o >= na * nb) continue;
        const int ai = o / nb, bi = o % nb;
        const int a0 = cexp(la, ai, 0), a1 = cexp(la, ai, 1), a2 = cexp(la, ai, 2);
        const int b0 = cexp(lb, bi, 0), b1 = cexp(lb, bi, 1), b2 = cexp(lb, bi, 2);
        const double* Epx = s_Ep[0] + (a0 * nbb + b0) * mtab;
        const double* Epy = s_Ep[1] + (a1 * nbb + b1) * mtab;
        const double* Epz = s_Ep[2] + (a2 * nbb + b2) * mtab;
        const double* Eqx = s_Eq[0] + (a0 * nbb + b0) * mtab;
        const double* Eqy = s_Eq[1] + (a1 * nbb + b1) * mtab;
        const double* Eqz = s_Eq[2] + (a2 * nbb + b2) * mtab;
        double sum = 0.0;
        for (int t = 0; t <= a0 + b0; ++t)
          for (int u = 0; u <= a1 + b1; ++u)
            for (int v = 0; v <= a2 + b2; ++v) {
              const double e1 = Epx[t] * Epy[u] * Epz[v];
              if (e1 == 0.0) continue;
              double inner = 0.0;
              for (int ta = 0; ta <= a0 + b0; ++ta)
                for (int nu = 0; nu <= a1 + b1; ++nu)
                  for (int ph = 0; ph <= a2 + b2; ++ph) {
                    const double e2 = Eqx[ta] * Eqy[nu] * Eqz[ph];
                    inner += (((ta + nu + ph) & 1) ? -e2 : e2) * R[ridx(t + ta, u + nu, v + ph)];
                  }
              sum += e1 * inner;
            }
        acc[h] += pref * sum;
      }
      __syncthreads();
    }
  }
  double m = fmax(acc[0], acc[1]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off));
  if (lane == 0) q_out[k] = sqrt(fmax(m, 0.0));
}

hipError_t launch_schwarz(int la, int lb, const int2* pairs, int npair, const ShellSet& A, const double* boys,
                          double* q_out, hipStream_t st) {
  if (npair == 0) return hipSuccess;
  if (la > 3 || lb > 3 || ncart(la) * ncart(lb) > 128) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_schwarz, dim3((unsigned)npair), dim3(64), 0, st, la, lb, pairs, npair, A, boys, q_out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
void upload_cls_tables(const C2SList& h, const int (&cart)[7][28][3]);

void upload_c2s_tables() {
  static C2SList h;
  static int cart[7][28][3];
  memset(&h, 0, sizeof(h));
  auto fill = [&](int l, auto blk) {
    int nc = ncart(l), ns = nsph(l);
    for (int s = 0; s < ns; ++s) {
      int k = 0;
      for (int c = 0; c < nc; ++c) {
        double v = blk[c][s];
        if (v != 0.0 && k < 12) {
          h.cart[l][s][k] = c;
          h.coef[l][s][k] = v;
          ++k;
        }
      }
      h.n[l][s] = k;
    }
  };
  fill(0, kC2S0); fill(1, kC2S1); fill(2, kC2S2); fill(3, kC2S3); fill(4, kC2S4); fill(5, kC2S5); fill(6, kC2S6);
  auto fc = [&](int l, auto tab) {
    for (int c = 0; c < ncart(l); ++c)
      for (int x = 0; x < 3; ++x) cart[l][c][x] = tab[c][x];
  };
  fc(0, kCart0); fc(1, kCart1); fc(2, kCart2); fc(3, kCart3); fc(4, kCart4); fc(5, kCart5); fc(6, kCart6);
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_c2s), &h, sizeof(h));
  (void)hipMemcpyToSymbol(HIP_SYMBOL(g_cart), cart, sizeof(cart));
  upload_cls_tables(h, cart);
}

hipError_t launch_eri_wave(int la, int lb, int lc, const int2* pairs, int npair, const int32_t* aux, int naux_l,
                           const ShellSet& A, const ShellSet& C, int b_unit, const double* boys, double* out,
                           int64_t sa, int64_t sb, int tr, int64_t c0, hipStream_t stream) {
  if (npair == 0 || naux_l == 0) return hipSuccess;
  Task3 t{pairs, aux, npair, naux_l};
  const int64_t nblk = (int64_t)npair * naux_l;
  // keep gridDim.x * blockDim.x below 2^32 threads per launch
  const int64_t chunk = 1LL << 24;
  for (int64_t b0 = 0; b0 < nblk; b0 += chunk) {
    int64_t nb = nblk - b0 < chunk ? nblk - b0 : chunk;
    hipLaunchKernelGGL(k_eri_wave, dim3((unsigned)nb), dim3(kWave), 0, stream, la, lb, lc, t, A, C, b_unit, boys,
                       out, sa, sb, tr, b0, c0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_int1e(const ShellSet& A, int natm, const double* atm, const double* boys, double* S, double* T,
                        double* V, int64_t nao, hipStream_t stream) {
  hipLaunchKernelGGL(k_int1e, dim3((unsigned)(A.n * A.n)), dim3(kWave), 0, stream, A, natm, atm, boys, S, T, V, nao);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Nuclear-coordinate derivatives of the integrals, contracted on the fly with
// the densities they meet in dE/dR (dftax/ks/forces.py:245-270, where the
// energy is differentiated with the basis centres following their atoms).
//
// A cartesian Gaussian's centre derivative raises and lowers its exponent,
// d/dA_x x_A^i e^{-a x_A^2} = 2a x_A^{i+1} e^{..} - i x_A^{i-1} e^{..}, so in
// the Hermite picture the x-axis table of the bra pair is replaced by
//   DA^{ij}_t = 2a E^{i+1,j}_t - i E^{i-1,j}_t   (derivative in A)
//   DB^{ij}_t = 2b E^{i,j+1}_t - j E^{i,j-1}_t   (derivative in B)
// and the R table needs one more order. The auxiliary centre follows from
// translational invariance, d/dC = -(d/dA + d/dB). The spherical density
// blocks are taken to the cartesian components first, so each lane weights
// its cartesian derivative integrals directly; the wavefront's six sums go to
// the atoms with one atomic add each.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

constexpr int kMaxLG = kMaxL3 + 1;                                   // (ff|g) plus one derivative order
constexpr int kNRG = (kMaxLG + 1) * (kMaxLG + 2) * (kMaxLG + 3) / 6;  // 364

// Spherical block of a symmetric density weight -> cartesian components:
// G_cart[a][b][c] = sum_{m n p} T[a][m] T[b][n] T[c][p] G[m][n][p] with the sparse
// cart->sph lists (the transpose of the integral transform). Layout of the
// spherical block: element (m, n, p) at G[p * sp + m * sm + n * sn].
__device__ void weight_to_cart(int la, int lb, int lc, int sphA, int sphB, int sphC, const double* __restrict__ G,
                               int64_t sp, int64_t sm, int64_t sn, int64_t m0, int64_t n0, int64_t p0,
                               double* s_out /*[na][nb][nc]*/) {
  const int lane = threadIdx.x;
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc);
  const int fa = sphA ? nsph(la) : na, fb = sphB ? nsph(lb) : nb, fc = sphC ? nsph(lc) : nc;
  for (int o = lane; o < na * nb * nc; o += kWave) s_out[o] = 0.0;
  __syncthreads();
  // one lane per spherical element scatters into the cartesian block (LDS atomics)
  for (int o = lane; o < fa * fb * fc; o += kWave) {
    const int C = o % fc, B = (o / fc) % fb, A = o / (fc * fb);
    const double g = G[(p0 + C) * sp + (m0 + A) * sm + (n0 + B) * sn];
    const int nA = sphA ? g_c2s.n[la][A] : 1, nB = sphB ? g_c2s.n[lb][B] : 1, nC = sphC ? g_c2s.n[lc][C] : 1;
    for (int x = 0; x < nA; ++x) {
      const int ai = sphA ? g_c2s.cart[la][A][x] : A;
      const double wa = g * (sphA ? g_c2s.coef[la][A][x] : 1.0);
      for (int y = 0; y < nB; ++y) {
        const int bi = sphB ? g_c2s.cart[lb][B][y] : B;
        const double wb = wa * (sphB ? g_c2s.coef[lb][B][y] : 1.0);
        for (int z = 0; z < nC; ++z) {
          const int ci = sphC ? g_c2s.cart[lc][C][z] : C;
          atomicAdd(&s_out[(ai * nb + bi) * nc + ci], wb * (sphC ? g_c2s.coef[lc][C][z] : 1.0));
        }
      }
    }
  }
  __syncthreads();
}

// sum over the 3-/2-centre derivative integrals weighted by Gamma:
//   3c: tasks (bra pair ia >= ib, aux shell ic); Gamma[P][m][n] (naux x nao x nao),
//       factor 2 for ia != ib (both bra orders);
//   2c (b_unit): bra = aux shell p with a unit s partner, ket = aux shell q, all
//       ordered (p, q); Gamma2[p][q] (naux x naux).
// grad[3 * atom + x] receives dE/dR.
__global__ __launch_bounds__(64) void k_eri_grad(const int2* __restrict__ pairs, int npair, int nket, ShellSet A,
                                                 ShellSet C, const int32_t* __restrict__ atom_a,
                                                 const int32_t* __restrict__ atom_c, int b_unit,
                                                 const double* __restrict__ boys, const double* __restrict__ G,
                                                 int64_t nao_g, double* __restrict__ grad, int64_t blk0) {
  __shared__ double s_E[3][5 * 5 * 10];
  __shared__ double s_Ec[5 * (kMaxLG + 1)];
  __shared__ double s_F[kMaxLG + 1];
  __shared__ double s_R[2][kNRG];
  __shared__ double s_G[kMaxCartBlk];
  const int lane = threadIdx.x;
  const int64_t task = blk0 + blockIdx.x;
  const int k = (int)(task / nket), ic = (int)(task % nket);
  if (k >= npair) return;
  const int2 pr = pairs[k];
  const int ia = pr.x, ib = b_unit ? -1 : pr.y;
  const int la = A.l[ia], lb = b_unit ? 0 : A.l[ib], lc = C.l[ic];
  const int L = la + lb + lc + 1;           // one order above the integrals
  const int mt = L + 1, lab = la + lb, mtab = lab + 3, nbb = lb + 2;
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc);
  const int sphA = A.spherical && la >= 2, sphB = !b_unit && A.spherical && lb >= 2, sphC = C.spherical && lc >= 2;
  if (b_unit)   // Gamma2[p][q]: rows p (bra), columns q (ket)
    weight_to_cart(la, 0, lc, sphA, 0, sphC, G, 1, nao_g, 0, A.ao0[ia], 0, C.ao0[ic], s_G);
  else          // Gamma[P][m][n]
    weight_to_cart(la, lb, lc, sphA, sphB, sphC, G, nao_g * nao_g, nao_g, 1, A.ao0[ia], A.ao0[ib], C.ao0[ic], s_G);

  const double Ax = A.centre[3 * ia], Ay = A.centre[3 * ia + 1], Az = A.centre[3 * ia + 2];
  double Bx = Ax, By = Ay, Bz = Az;
  int npb = 1, p0b = 0;
  if (!b_unit) {
    Bx = A.centre[3 * ib]; By = A.centre[3 * ib + 1]; Bz = A.centre[3 * ib + 2];
    npb = A.nprim[ib]; p0b = A.prim0[ib];
  }
  const double Cx = C.centre[3 * ic], Cy = C.centre[3 * ic + 1], Cz = C.centre[3 * ic + 2];
  const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz, ab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
  for (int pa = 0; pa < A.nprim[ia]; ++pa)
    for (int pb = 0; pb < npb; ++pb) {
      const double al = A.exps[A.prim0[ia] + pa];
      const double be = b_unit ? 0.0 : A.exps[p0b + pb];
      const double cab = A.coefs[A.prim0[ia] + pa] * (b_unit ? 1.0 : A.coefs[p0b + pb]);
      const double p = al + be;
      const double Px = (al * Ax + be * Bx) / p, Py = (al * Ay + be * By) / p, Pz = (al * Az + be * Bz) / p;
      const double Kab = exp(-al * be / p * ab2);
      __syncthreads();
      if (lane < 3) {
        const double xab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
        e_table_1d(la + 1, lb + 1, al, be, xab, mtab, s_E[lane]);
      }
      for (int pc = 0; pc < C.nprim[ic]; ++pc) {
        const double ga = C.exps[C.prim0[ic] + pc];
        const double cc = C.coefs[C.prim0[ic] + pc];
        const double rho = p * ga / (p + ga);
        const double pref = cab * cc * Kab * 2.0 * 17.493418327624862 / (p * ga * sqrt(p + ga));  // 2 pi^2.5
        const double pcx = Px - Cx, pcy = Py - Cy, pcz = Pz - Cz;
        __syncthreads();
        if (lane == 0) {
          boys_array_rt(L, rho * (pcx * pcx + pcy * pcy + pcz * pcz), s_F, boys);
          const double i2g = 0.5 / ga;
          for (int q = 0; q < (lc + 1) * mt; ++q) s_Ec[q] = 0.0;
          s_Ec[0] = 1.0;
          for (int kk = 0; kk < lc; ++kk)
            for (int t = 0; t < mt; ++t) {
              double v = 0.0;
              if (t > 0) v += i2g * s_Ec[kk * mt + t - 1];
              if (t + 1 < mt) v += (t + 1) * s_Ec[kk * mt + t + 1];
              s_Ec[(kk + 1) * mt + t] = v;
            }
        }
        __syncthreads();
        const double* R = r_ladder(L, rho, pcx, pcy, pcz, s_F, s_R[0], s_R[1]);
        const double sgn = (lc & 1) ? -pref : pref;
        for (int o = lane; o < na * nb * nc; o += kWave) {
          const double w = s_G[o];
          if (w == 0.0) continue;
          const int ci = o % nc, bi = (o / nc) % nb, ai = o / (nc * nb);
          int ae[3], be3[3], ce[3];
          for (int x = 0; x < 3; ++x) {
            ae[x] = cexp(la, ai, x);
            be3[x] = b_unit ? 0 : cexp(lb, bi, x);
            ce[x] = cexp(lc, ci, x);
          }
          // per axis: value E, A-derivative DA and B-derivative DB tables (t <= i + j + 1)
          double Ev[3][8], DA[3][8], DB[3][8];
          for (int x = 0; x < 3; ++x) {
            const double* E = s_E[x];
            const int i = ae[x], j = be3[x];
            for (int t = 0; t <= i + j + 1; ++t) {
              const double e0 = (t <= i + j) ? E[(i * nbb + j) * mtab + t] : 0.0;
              double da = 2.0 * al * E[((i + 1) * nbb + j) * mtab + t];
              if (i > 0 && t <= i + j - 1) da -= i * E[((i - 1) * nbb + j) * mtab + t];
              double db = 2.0 * be * E[(i * nbb + j + 1) * mtab + t];
              if (j > 0 && t <= i + j - 1) db -= j * E[(i * nbb + j - 1) * mtab + t];
              Ev[x][t] = e0;
              DA[x][t] = da;
              DB[x][t] = db;
            }
          }
          const double* Ecx = s_Ec + ce[0] * mt;
          const double* Ecy = s_Ec + ce[1] * mt;
          const double* Ecz = s_Ec + ce[2] * mt;
          double sa[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
          for (int t = 0; t <= ae[0] + be3[0] + 1; ++t)
            for (int u = 0; u <= ae[1] + be3[1] + 1; ++u)
              for (int v = 0; v <= ae[2] + be3[2] + 1; ++v) {
                double inner = 0.0;
                for (int ta = ce[0] & 1; ta <= ce[0]; ta += 2)
                  for (int nu = ce[1] & 1; nu <= ce[1]; nu += 2)
                    for (int ph = ce[2] & 1; ph <= ce[2]; ph += 2)
                      inner += Ecx[ta] * Ecy[nu] * Ecz[ph] * R[ridx(t + ta, u + nu, v + ph)];
                sa[0] += DA[0][t] * Ev[1][u] * Ev[2][v] * inner;
                sa[1] += Ev[0][t] * DA[1][u] * Ev[2][v] * inner;
                sa[2] += Ev[0][t] * Ev[1][u] * DA[2][v] * inner;
                sb[0] += DB[0][t] * Ev[1][u] * Ev[2][v] * inner;
                sb[1] += Ev[0][t] * DB[1][u] * Ev[2][v] * inner;
                sb[2] += Ev[0][t] * Ev[1][u] * DB[2][v] * inner;
              }
          for (int x = 0; x < 3; ++x) {
            gA[x] += w * sgn * sa[x];
            gB[x] += w * sgn * sb[x];
          }
        }
      }
    }
  const double f = (!b_unit && ia != ib) ? 2.0 : 1.0;
  double tot[6];
  for (int x = 0; x < 3; ++x) {
    tot[x] = wave_sum(gA[x]) * f;
    tot[3 + x] = wave_sum(gB[x]) * f;
  }
  if (lane == 0) {
    const int at_a = atom_a[ia], at_c = atom_c[ic];
    for (int x = 0; x < 3; ++x) {
      atomicAdd(&grad[3 * at_a + x], tot[x]);
      if (!b_unit) atomicAdd(&grad[3 * atom_a[ib] + x], tot[3 + x]);
      atomicAdd(&grad[3 * at_c + x], -(tot[x] + tot[3 + x]));
    }
  }
}

// One-electron part of dE/dR over every ordered shell pair (a, b), derivative in
// the bra centre only (the ket-centre share equals it by symmetry, hence the 2):
//   2 sum_{mn in (a,b)} [ P_mn (dT_mn + sum_C dV^C_mn) - W_mn dS_mn ]  -> atom of a,
//   and the Hellmann-Feynman term of each nucleus C, d/dC V^C = -(d/dA + d/dB) V^C:
//   -2 sum_{mn} P_mn dV^C_mn -> atom C.
__global__ __launch_bounds__(64) void k_int1e_grad(ShellSet A, const int32_t* __restrict__ atom_a, int natm,
                                                   const double* __restrict__ atm, const double* __restrict__ boys,
                                                   const double* __restrict__ P, const double* __restrict__ W,
                                                   int64_t nao, double* __restrict__ grad) {
  __shared__ double s_E[3][5 * 6 * 11];
  __shared__ double s_F[12];
  __shared__ double s_R[2][120];
  __shared__ double s_P[100], s_W[100];
  const int lane = threadIdx.x;
  const int ia = blockIdx.x / A.n, ib = blockIdx.x % A.n;
  const int la = A.l[ia], lb = A.l[ib];
  const int na = ncart(la), nb = ncart(lb), nout = na * nb;
  const int L = la + lb + 1;
  const int mt = la + lb + 4;   // E up to (la+1, lb+2)
  const int nbb = lb + 3;
  const int sphA = A.spherical && la >= 2, sphB = A.spherical && lb >= 2;
  weight_to_cart(la, lb, 0, sphA, sphB, 0, P, 0, nao, 1, A.ao0[ia], A.ao0[ib], 0, s_P);
  weight_to_cart(la, lb, 0, sphA, sphB, 0, W, 0, nao, 1, A.ao0[ia], A.ao0[ib], 0, s_W);
  const double Ax = A.centre[3 * ia], Ay = A.centre[3 * ia + 1], Az = A.centre[3 * ia + 2];
  const double Bx = A.centre[3 * ib], By = A.centre[3 * ib + 1], Bz = A.centre[3 * ib + 2];
  const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz, ab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  double gA[3] = {0, 0, 0};   // this lane's share of the S and T terms
  double gV[3] = {0, 0, 0};   // the wavefront's nuclear-attraction sums (the same on every lane)
  for (int pa = 0; pa < A.nprim[ia]; ++pa)
    for (int pb = 0; pb < A.nprim[ib]; ++pb) {
      const double al = A.exps[A.prim0[ia] + pa], be = A.exps[A.prim0[ib] + pb];
      const double cab = A.coefs[A.prim0[ia] + pa] * A.coefs[A.prim0[ib] + pb];
      const double p = al + be;
      const double Px = (al * Ax + be * Bx) / p, Py = (al * Ay + be * By) / p, Pz = (al * Az + be * Bz) / p;
      const double Kab = exp(-al * be / p * ab2);
      __syncthreads();
      if (lane < 3) {
        const double xab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
        e_table_1d(la + 1, lb + 2, al, be, xab, mt, s_E[lane]);
      }
      __syncthreads();
      const double s0 = pow(kPi / p, 1.5) * Kab * cab;
      // overlap and kinetic: per-axis factors s_ij = E^{ij}_0, t_ij from E^{i,j+2}, E^{ij}, E^{i,j-2};
      // their bra derivatives 2a f_{i+1,j} - i f_{i-1,j}
      for (int o = lane; o < nout; o += kWave) {
        const int ai = o / nb, bi = o % nb;
        double s1[3], t1[3], ds[3], dt[3];
        for (int x = 0; x < 3; ++x) {
          const int i = cexp(la, ai, x), j = cexp(lb, bi, x);
          const double* E = s_E[x];
          auto sv = [&](int ii, int jj) { return E[(ii * nbb + jj) * mt]; };
          auto tv = [&](int ii, int jj) {
            double r = -2.0 * be * be * sv(ii, jj + 2) + be * (2 * jj + 1) * sv(ii, jj);
            if (jj >= 2) r -= 0.5 * jj * (jj - 1) * sv(ii, jj - 2);
            return r;
          };
          s1[x] = sv(i, j);
          t1[x] = tv(i, j);
          ds[x] = 2.0 * al * sv(i + 1, j) - (i > 0 ? i * sv(i - 1, j) : 0.0);
          dt[x] = 2.0 * al * tv(i + 1, j) - (i > 0 ? i * tv(i - 1, j) : 0.0);
        }
        const double wP = s_P[o], wW = s_W[o];
        for (int x = 0; x < 3; ++x) {
          const int y = (x + 1) % 3, z = (x + 2) % 3;
          const double dS = s0 * ds[x] * s1[y] * s1[z];
          const double dT = s0 * (dt[x] * s1[y] * s1[z] + ds[x] * t1[y] * s1[z] + ds[x] * s1[y] * t1[z]);
          gA[x] += wP * dT - wW * dS;
        }
      }
      // nuclear attraction, nucleus by nucleus
      for (int c = 0; c < natm; ++c) {
        const double pcx = Px - atm[4 * c], pcy = Py - atm[4 * c + 1], pcz = Pz - atm[4 * c + 2];
        __syncthreads();
        if (lane == 0) boys_array_rt(L, p * (pcx * pcx + pcy * pcy + pcz * pcz), s_F, boys);
        __syncthreads();
        const double* R = r_ladder(L, p, pcx, pcy, pcz, s_F, s_R[0], s_R[1]);
        const double pref = -atm[4 * c + 3] * cab * Kab * 2.0 * kPi / p;
        double gc[3] = {0, 0, 0};
        for (int o = lane; o < nout; o += kWave) {
          const double wP = s_P[o];
          if (wP == 0.0) continue;
          const int ai = o / nb, bi = o % nb;
          double Ev[3][8], DA[3][8];
          int ij[3];
          for (int x = 0; x < 3; ++x) {
            const int i = cexp(la, ai, x), j = cexp(lb, bi, x);
            const double* E = s_E[x];
            ij[x] = i + j;
            for (int t = 0; t <= i + j + 1; ++t) {
              Ev[x][t] = t <= i + j ? E[(i * nbb + j) * mt + t] : 0.0;
              double da = 2.0 * al * E[((i + 1) * nbb + j) * mt + t];
              if (i > 0 && t <= i + j - 1) da -= i * E[((i - 1) * nbb + j) * mt + t];
              DA[x][t] = da;
            }
          }
          double d3[3] = {0, 0, 0};
          for (int t = 0; t <= ij[0] + 1; ++t)
            for (int u = 0; u <= ij[1] + 1; ++u)
              for (int v = 0; v <= ij[2] + 1; ++v) {
                const double r = R[ridx(t, u, v)];
                d3[0] += DA[0][t] * Ev[1][u] * Ev[2][v] * r;
                d3[1] += Ev[0][t] * DA[1][u] * Ev[2][v] * r;
                d3[2] += Ev[0][t] * Ev[1][u] * DA[2][v] * r;
              }
          for (int x = 0; x < 3; ++x) gc[x] += wP * pref * d3[x];
        }
        for (int x = 0; x < 3; ++x) {
          const double v = wave_sum(gc[x]);
          gV[x] += v;
          if (lane == 0) atomicAdd(&grad[3 * c + x], -2.0 * v);
        }
        __syncthreads();
      }
    }
  for (int x = 0; x < 3; ++x) {
    const double v = wave_sum(gA[x]) + gV[x];
    if (lane == 0) atomicAdd(&grad[3 * atom_a[ia] + x], 2.0 * v);
  }
}

hipError_t launch_eri_grad(const int2* pairs, int npair, int nket, const ShellSet& A, const ShellSet& C,
                           const int32_t* atom_a, const int32_t* atom_c, int b_unit, const double* boys,
                           const double* G, int64_t nao_g, double* grad, hipStream_t st) {
  const int64_t ntask = (int64_t)npair * nket;
  const int64_t chunk = 1 << 30;
  for (int64_t b0 = 0; b0 < ntask; b0 += chunk) {
    const int64_t nb = std::min(chunk, ntask - b0);
    hipLaunchKernelGGL(k_eri_grad, dim3((unsigned)nb), dim3(kWave), 0, st, pairs, npair, nket, A, C, atom_a, atom_c,
                       b_unit, boys, G, nao_g, grad, b0);
  }
  return hipGetLastError();
}

hipError_t launch_int1e_grad(const ShellSet& A, const int32_t* atom_a, int natm, const double* atm,
                             const double* boys, const double* P, const double* W, int64_t nao, double* grad,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_int1e_grad, dim3((unsigned)(A.n * A.n)), dim3(kWave), 0, st, A, atom_a, natm, atm, boys, P, W,
                     nao, grad);
  return hipGetLastError();
}

}  // namespace dfks
