/*
 * ORACLE -- test infrastructure only. Nothing in the product links this file;
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg call it as
 * the checker.
 *
 * CPU FP64 restatement of dftax's McMurchie-Davidson integral engine:
 *   - Boys function: table + degree-6 local Taylor, asymptote for t >= 40
 *     (dftax/energy/boys.py:33-36, 81-123, 130-176);
 *   - Hermite E tables (dftax/integrals/eri3c_bucketed.py:290-323);
 *   - Hermite Coulomb R table by the descending m-ladder
 *     (eri3c_bucketed.py:326-369);
 *   - 3-centre (ab|c) with prefactor K_AB 2 pi^2.5 / (p q sqrt(p+q)) and the
 *     (-1)^tau sign on the auxiliary side (eri3c_bucketed.py:376-426);
 *   - 2-centre (P|Q) (eri3c_bucketed.py:697-734);
 *   - the long-range operator erf(omega r)/r of range-separated hybrids by
 *     attenuated Boys values (eri3c_bucketed.py:340-346), oracle_set_omega;
 *   - overlap, kinetic and nuclear attraction by the same E/R tables
 *     (eri3c_bucketed.py:534-569, 655-694).
 * Output is in the final AO span: cartesian components are contracted with
 * the per-l cart->sph blocks passed in by the caller.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BOYS_ORDER 6
#define BOYS_DT 0.1
#define BOYS_TMAX 40.0
#define BOYS_NMAX 24
#define BOYS_NG 401
#define LMAX 6
#define MTMAX (3 * LMAX + 3)

static double boys_tbl[BOYS_NG][BOYS_NMAX + 1];
static int boys_ready = 0;

static void boys_build(void) {
    const int n_top = (BOYS_NMAX > (int)BOYS_TMAX ? BOYS_NMAX : (int)BOYS_TMAX) + 20;
    for (int j = 0; j < BOYS_NG; ++j) {
        double t = j * BOYS_DT, two_t = 2.0 * t, emt = exp(-t);
        double term = 1.0 / (2 * n_top + 1), total = term;
        for (int k = 1; k < 500; ++k) {
            term = term * two_t / (2 * n_top + 2 * k + 1);
            total += term;
            if (term <= 1e-18 * total) break;
        }
        double f = emt * total;
        for (int n = n_top; n > 0; --n) {
            if (n <= BOYS_NMAX) boys_tbl[j][n] = f;
            f = (two_t * f + emt) / (2 * n - 1);
        }
        boys_tbl[j][0] = f;
    }
    boys_ready = 1;
}

/* F_n(t) for n + BOYS_ORDER <= BOYS_NMAX (boys.py:130-176). */
double oracle_boys(int n, double t) {
    if (!boys_ready) boys_build();
    if (t < BOYS_TMAX) {
        int idx = (int)nearbyint(t / BOYS_DT);
        if (idx < 0) idx = 0;
        if (idx > BOYS_NG - 1) idx = BOYS_NG - 1;
        double dx = idx * BOYS_DT - t;
        double acc = boys_tbl[idx][n + BOYS_ORDER];
        for (int k = BOYS_ORDER - 1; k >= 0; --k) acc = boys_tbl[idx][n + k] + acc * dx / (k + 1);
        return acc;
    }
    double a = n + 0.5;
    return 0.5 * exp(lgamma(a)) * pow(t, -a);
}

void oracle_boys_table(double *out) {
    if (!boys_ready) boys_build();
    memcpy(out, boys_tbl, sizeof(boys_tbl));
}

/* --- basis description ---------------------------------------------------- */
typedef struct {
    int nshell;
    const int *l, *prim0, *nprim, *ao0;
    const double *centre, *exps, *coefs;
    int spherical;
} basis_t;

static const double *c2s_data;
static const int *c2s_off;

static inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static inline int nfun(const basis_t *b, int l) { return (b->spherical && l >= 2) ? 2 * l + 1 : ncart(l); }

static void cart_comps(int l, int (*c)[3]) {
    int i = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) { c[i][0] = lx; c[i][1] = ly; c[i][2] = l - lx - ly; ++i; }
}

/* E^{ij}_t for one axis, i<=la, j<=lb (eri3c_bucketed.py:290-310). */
static void e_table(int la, int lb, double a, double b, double xab, int mt, double *E /*[(la+1)(lb+1)mt]*/) {
    double p = a + b, i2g = 0.5 / p, xpa = -b * xab / p, xpb = a * xab / p;
#define EI(i, j, t) E[((i) * (lb + 1) + (j)) * mt + (t)]
    memset(E, 0, sizeof(double) * (la + 1) * (lb + 1) * mt);
    EI(0, 0, 0) = 1.0;
    for (int i = 0; i < la; ++i)
        for (int t = 0; t < mt; ++t) {
            double v = xpa * EI(i, 0, t);
            if (t > 0) v += i2g * EI(i, 0, t - 1);
            if (t + 1 < mt) v += (t + 1) * EI(i, 0, t + 1);
            EI(i + 1, 0, t) = v;
        }
    for (int j = 0; j < lb; ++j)
        for (int i = 0; i <= la; ++i)
            for (int t = 0; t < mt; ++t) {
                double v = xpb * EI(i, j, t);
                if (t > 0) v += i2g * EI(i, j, t - 1);
                if (t + 1 < mt) v += (t + 1) * EI(i, j, t + 1);
                EI(i, j + 1, t) = v;
            }
#undef EI
}

/* single-centre E^{k}_tau, k<=lc (eri3c_bucketed.py:313-323) */
static void ec_table(int lc, double g, int mt, double *E /*[(lc+1)mt]*/) {
    double i2g = 0.5 / g;
    memset(E, 0, sizeof(double) * (lc + 1) * mt);
    E[0] = 1.0;
    for (int k = 0; k < lc; ++k)
        for (int t = 0; t < mt; ++t) {
            double v = 0.0;
            if (t > 0) v += i2g * E[k * mt + t - 1];
            if (t + 1 < mt) v += (t + 1) * E[k * mt + t + 1];
            E[(k + 1) * mt + t] = v;
        }
}

/* Operator of the two-electron integrals: 0 the Coulomb operator 1/r, > 0 the
 * long-range erf(omega r)/r of a range-separated hybrid, whose Boys values are
 * s^(m+1/2) F_m(s T) with s = omega^2 / (omega^2 + rho) (eri3c_bucketed.py:340-346).
 * Set by oracle_set_omega before an integral call (the oracle is driven from
 * one thread). */
static double g_omega = 0.0;

void oracle_set_omega(double omega) { g_omega = omega; }

/* R^0_{tuv}, t+u+v < mt, stored [t][u][v] in an mt^3 cube (eri3c_bucketed.py:326-369). */
static void r_table(double rho, const double *pc, int mt, double *R) {
    double T = rho * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
    double s = 1.0;
    if (g_omega > 0.0) {
        s = g_omega * g_omega / (g_omega * g_omega + rho);
        T *= s;
    }
    static __thread double bufA[MTMAX * MTMAX * MTMAX], bufB[MTMAX * MTMAX * MTMAX];
    double *prev = bufA, *cur = bufB;
    memset(prev, 0, sizeof(double) * mt * mt * mt);
#define IX(t, u, v) (((t) * mt + (u)) * mt + (v))
    double n2r = -2.0 * rho;
    for (int m = mt - 1; m >= 0; --m) {
        memset(cur, 0, sizeof(double) * mt * mt * mt);
        cur[IX(0, 0, 0)] = pow(n2r, m) * oracle_boys(m, T) * (g_omega > 0.0 ? pow(s, m + 0.5) : 1.0);
        int lev = mt - 1 - m; /* highest t+u+v reachable at this level */
        for (int s = 1; s <= lev; ++s)
            for (int t = 0; t <= s; ++t)
                for (int u = 0; u <= s - t; ++u) {
                    int v = s - t - u;
                    double val;
                    if (t > 0) val = pc[0] * prev[IX(t - 1, u, v)] + (t > 1 ? (t - 1) * prev[IX(t - 2, u, v)] : 0.0);
                    else if (u > 0) val = pc[1] * prev[IX(t, u - 1, v)] + (u > 1 ? (u - 1) * prev[IX(t, u - 2, v)] : 0.0);
                    else val = pc[2] * prev[IX(t, u, v - 1)] + (v > 1 ? (v - 1) * prev[IX(t, u, v - 2)] : 0.0);
                    cur[IX(t, u, v)] = val;
                }
        double *tmp = prev; prev = cur; cur = tmp;
    }
    memcpy(R, prev, sizeof(double) * mt * mt * mt);
#undef IX
}

/* Transform a cartesian block [na_c][nb_c][nc_c] into the final span. */
static void c2s_apply3(const basis_t *ba, int la, const basis_t *bb, int lb, const basis_t *bc, int lc,
                       const double *in, double *out) {
    int nac = ncart(la), nbc = ncart(lb), ncc = ncart(lc);
    int naf = nfun(ba, la), nbf = nfun(bb, lb), ncf = nfun(bc, lc);
    const double *Ta = c2s_data + c2s_off[la], *Tb = c2s_data + c2s_off[lb], *Tc = c2s_data + c2s_off[lc];
    int sa = (ba->spherical && la >= 2), sb = (bb->spherical && lb >= 2), sc = (bc->spherical && lc >= 2);
    double *t1 = (double *)malloc(sizeof(double) * naf * nbc * ncc);
    double *t2 = (double *)malloc(sizeof(double) * naf * nbf * ncc);
    for (int A = 0; A < naf; ++A)
        for (int j = 0; j < nbc * ncc; ++j) {
            double s = 0;
            if (sa) for (int i = 0; i < nac; ++i) s += Ta[i * naf + A] * in[i * nbc * ncc + j];
            else s = in[A * nbc * ncc + j];
            t1[A * nbc * ncc + j] = s;
        }
    for (int A = 0; A < naf; ++A)
        for (int B = 0; B < nbf; ++B)
            for (int k = 0; k < ncc; ++k) {
                double s = 0;
                if (sb) for (int j = 0; j < nbc; ++j) s += Tb[j * nbf + B] * t1[(A * nbc + j) * ncc + k];
                else s = t1[(A * nbc + B) * ncc + k];
                t2[(A * nbf + B) * ncc + k] = s;
            }
    for (int A = 0; A < naf; ++A)
        for (int B = 0; B < nbf; ++B)
            for (int C = 0; C < ncf; ++C) {
                double s = 0;
                if (sc) for (int k = 0; k < ncc; ++k) s += Tc[k * ncf + C] * t2[(A * nbf + B) * ncc + k];
                else s = t2[(A * nbf + B) * ncc + C];
                out[(A * nbf + B) * ncf + C] = s;
            }
    free(t1);
    free(t2);
}

/* cartesian (ab|c) block for one shell triple */
static void eri3c_block(const basis_t *bo, int ia, int ib, const basis_t *bx, int ic, double *blk) {
    int la = bo->l[ia], lb = bo->l[ib], lc = bx->l[ic];
    int mt = la + lb + lc + 1;
    if (mt < 2) mt = 2;
    int ca[28][3], cb[28][3], cc[28][3];
    cart_comps(la, ca); cart_comps(lb, cb); cart_comps(lc, cc);
    int nac = ncart(la), nbc = ncart(lb), ncc = ncart(lc);
    memset(blk, 0, sizeof(double) * nac * nbc * ncc);
    const double *A = bo->centre + 3 * ia, *B = bo->centre + 3 * ib, *C = bx->centre + 3 * ic;
    double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    double *Et = (double *)malloc(sizeof(double) * 3 * (la + 1) * (lb + 1) * mt);
    double *Ec = (double *)malloc(sizeof(double) * (lc + 1) * mt);
    double *R = (double *)malloc(sizeof(double) * mt * mt * mt);
    double *G = (double *)malloc(sizeof(double) * 3 * (la + 1) * (lb + 1) * (lc + 1) * mt);
    int eab = (la + 1) * (lb + 1) * mt;
    for (int pa = 0; pa < bo->nprim[ia]; ++pa)
        for (int pb = 0; pb < bo->nprim[ib]; ++pb) {
            double al = bo->exps[bo->prim0[ia] + pa], be = bo->exps[bo->prim0[ib] + pb];
            double cab = bo->coefs[bo->prim0[ia] + pa] * bo->coefs[bo->prim0[ib] + pb];
            double p = al + be;
            double P[3] = {(al * A[0] + be * B[0]) / p, (al * A[1] + be * B[1]) / p, (al * A[2] + be * B[2]) / p};
            double K = exp(-al * be / p * ab2);
            for (int x = 0; x < 3; ++x) e_table(la, lb, al, be, AB[x], mt, Et + x * eab);
            for (int pc = 0; pc < bx->nprim[ic]; ++pc) {
                double ga = bx->exps[bx->prim0[ic] + pc];
                double cc_ = bx->coefs[bx->prim0[ic] + pc];
                double rho = p * ga / (p + ga);
                double pref = cab * cc_ * K * 2.0 * pow(M_PI, 2.5) / (p * ga * sqrt(p + ga));
                double PC[3] = {P[0] - C[0], P[1] - C[1], P[2] - C[2]};
                r_table(rho, PC, mt, R);
                ec_table(lc, ga, mt, Ec);
                /* G[x][i][j][k][s] = sum_{t+tau=s} E^{ij}_t (-1)^tau E^k_tau */
                for (int x = 0; x < 3; ++x)
                    for (int i = 0; i <= la; ++i)
                        for (int j = 0; j <= lb; ++j)
                            for (int k = 0; k <= lc; ++k)
                                for (int s = 0; s < mt; ++s) {
                                    double v = 0;
                                    for (int t = 0; t <= s; ++t) {
                                        int tau = s - t;
                                        double sg = (tau & 1) ? -1.0 : 1.0;
                                        v += Et[x * eab + (i * (lb + 1) + j) * mt + t] * sg * Ec[k * mt + tau];
                                    }
                                    G[((((x * (la + 1) + i) * (lb + 1) + j) * (lc + 1) + k) * mt) + s] = v;
                                }
#define GG(x, i, j, k, s) G[((((x) * (la + 1) + (i)) * (lb + 1) + (j)) * (lc + 1) + (k)) * mt + (s)]
                for (int a = 0; a < nac; ++a)
                    for (int b = 0; b < nbc; ++b)
                        for (int c = 0; c < ncc; ++c) {
                            double acc = 0;
                            for (int s = 0; s < mt; ++s) {
                                double gx = GG(0, ca[a][0], cb[b][0], cc[c][0], s);
                                if (gx == 0.0) continue;
                                for (int r = 0; r < mt - s; ++r) {
                                    double gy = GG(1, ca[a][1], cb[b][1], cc[c][1], r);
                                    if (gy == 0.0) continue;
                                    for (int q = 0; q < mt - s - r; ++q)
                                        acc += gx * gy * GG(2, ca[a][2], cb[b][2], cc[c][2], q) * R[(s * mt + r) * mt + q];
                                }
                            }
                            blk[(a * nbc + b) * ncc + c] += pref * acc;
                        }
#undef GG
            }
        }
    free(Et); free(Ec); free(R); free(G);
}

static void set_c2s(const double *data, const int *off) { c2s_data = data; c2s_off = off; }

/* (mn|P) in the final spans, out[(m*nao + n)*naux + P], both bra triangles. */
void oracle_int3c(int ns, const int *l, const int *prim0, const int *nprim, const int *ao0,
                  const double *centre, const double *exps, const double *coefs, int sph, int nao,
                  int nsx, const int *lx, const int *prim0x, const int *nprimx, const int *ao0x,
                  const double *centrex, const double *expsx, const double *coefsx, int sphx, int naux,
                  const double *c2s, const int *c2soff, double *out) {
    set_c2s(c2s, c2soff);
    if (!boys_ready) boys_build();
    basis_t bo = {ns, l, prim0, nprim, ao0, centre, exps, coefs, sph};
    basis_t bx = {nsx, lx, prim0x, nprimx, ao0x, centrex, expsx, coefsx, sphx};
    int npair = ns * (ns + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
    for (int pr = 0; pr < npair; ++pr) {
        int ia = (int)((sqrt(8.0 * pr + 1.0) - 1.0) / 2.0);
        while (ia * (ia + 1) / 2 > pr) --ia;
        while ((ia + 1) * (ia + 2) / 2 <= pr) ++ia;
        int ib = pr - ia * (ia + 1) / 2;
        int la = l[ia], lb = l[ib];
        double *cblk = (double *)malloc(sizeof(double) * 28 * 28 * 28);
        double *fblk = (double *)malloc(sizeof(double) * 28 * 28 * 28);
        for (int ic = 0; ic < nsx; ++ic) {
            int lc = lx[ic];
            eri3c_block(&bo, ia, ib, &bx, ic, cblk);
            c2s_apply3(&bo, la, &bo, lb, &bx, lc, cblk, fblk);
            int na = nfun(&bo, la), nb = nfun(&bo, lb), nc = nfun(&bx, lc);
            for (int a = 0; a < na; ++a)
                for (int b = 0; b < nb; ++b)
                    for (int c = 0; c < nc; ++c) {
                        double v = fblk[(a * nb + b) * nc + c];
                        int m = ao0[ia] + a, n = ao0[ib] + b, P = ao0x[ic] + c;
                        out[((size_t)m * nao + n) * naux + P] = v;
                        out[((size_t)n * nao + m) * naux + P] = v;
                    }
        }
        free(cblk);
        free(fblk);
    }
}

/* (P|Q) 2-centre metric in the final aux span. Built as (P s|Q) with a unit s
 * function of zero exponent folded in: single-centre E tables on both sides. */
void oracle_int2c(int ns, const int *l, const int *prim0, const int *nprim, const int *ao0,
                  const double *centre, const double *exps, const double *coefs, int sph, int naux,
                  const double *c2s, const int *c2soff, double *out) {
    set_c2s(c2s, c2soff);
    if (!boys_ready) boys_build();
    basis_t bx = {ns, l, prim0, nprim, ao0, centre, exps, coefs, sph};
#pragma omp parallel for schedule(dynamic, 1)
    for (int ip = 0; ip < ns; ++ip) {
        double *cblk = (double *)malloc(sizeof(double) * 28 * 28);
        double *fblk = (double *)malloc(sizeof(double) * 28 * 28);
        double *R = (double *)malloc(sizeof(double) * MTMAX * MTMAX * MTMAX);
        for (int iq = 0; iq < ns; ++iq) {
            int lp = l[ip], lq = l[iq];
            int mt = lp + lq + 1;
            if (mt < 2) mt = 2;
            int cp[28][3], cq[28][3];
            cart_comps(lp, cp); cart_comps(lq, cq);
            int npc = ncart(lp), nqc = ncart(lq);
            memset(cblk, 0, sizeof(double) * npc * nqc);
            const double *A = centre + 3 * ip, *C = centre + 3 * iq;
            double Ep[(LMAX + 1) * MTMAX], Eq[(LMAX + 1) * MTMAX];
            for (int a = 0; a < nprim[ip]; ++a)
                for (int c = 0; c < nprim[iq]; ++c) {
                    double p = exps[prim0[ip] + a], q = exps[prim0[iq] + c];
                    double cf = coefs[prim0[ip] + a] * coefs[prim0[iq] + c];
                    double rho = p * q / (p + q);
                    double pref = cf * 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q));
                    double PC[3] = {A[0] - C[0], A[1] - C[1], A[2] - C[2]};
                    r_table(rho, PC, mt, R);
                    ec_table(lp, p, mt, Ep);
                    ec_table(lq, q, mt, Eq);
                    for (int i = 0; i < npc; ++i)
                        for (int k = 0; k < nqc; ++k) {
                            double acc = 0;
                            for (int t1 = 0; t1 <= cp[i][0]; ++t1)
                                for (int u1 = 0; u1 <= cp[i][1]; ++u1)
                                    for (int v1 = 0; v1 <= cp[i][2]; ++v1) {
                                        double e1 = Ep[cp[i][0] * mt + t1] * Ep[cp[i][1] * mt + u1] * Ep[cp[i][2] * mt + v1];
                                        if (e1 == 0.0) continue;
                                        for (int t2 = 0; t2 <= cq[k][0]; ++t2)
                                            for (int u2 = 0; u2 <= cq[k][1]; ++u2)
                                                for (int v2 = 0; v2 <= cq[k][2]; ++v2) {
                                                    double e2 = Eq[cq[k][0] * mt + t2] * Eq[cq[k][1] * mt + u2] * Eq[cq[k][2] * mt + v2];
                                                    if (e2 == 0.0) continue;
                                                    double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                                                    acc += e1 * sg * e2 * R[((t1 + t2) * mt + (u1 + u2)) * mt + (v1 + v2)];
                                                }
                                    }
                            cblk[i * nqc + k] += pref * acc;
                        }
                }
            /* cart -> final span on both sides */
            int npf = nfun(&bx, lp), nqf = nfun(&bx, lq);
            int sp = (sph && lp >= 2), sq = (sph && lq >= 2);
            const double *Tp = c2s_data + c2s_off[lp], *Tq = c2s_data + c2s_off[lq];
            for (int I = 0; I < npf; ++I)
                for (int K = 0; K < nqf; ++K) {
                    double s = 0;
                    for (int i = 0; i < npc; ++i) {
                        double wi = sp ? Tp[i * npf + I] : (i == I ? 1.0 : 0.0);
                        if (wi == 0.0) continue;
                        for (int k = 0; k < nqc; ++k) {
                            double wk = sq ? Tq[k * nqf + K] : (k == K ? 1.0 : 0.0);
                            if (wk == 0.0) continue;
                            s += wi * wk * cblk[i * nqc + k];
                        }
                    }
                    fblk[I * nqf + K] = s;
                }
            for (int I = 0; I < npf; ++I)
                for (int K = 0; K < nqf; ++K)
                    out[(size_t)(ao0[ip] + I) * naux + ao0[iq] + K] = fblk[I * nqf + K];
        }
        free(cblk); free(fblk); free(R);
    }
}

/* S, T and V = -sum_C Z_C <a|1/r_C|b> in the final span (nao x nao each). */
void oracle_int1e(int ns, const int *l, const int *prim0, const int *nprim, const int *ao0,
                  const double *centre, const double *exps, const double *coefs, int sph, int nao,
                  int natm, const double *atm_xyz, const double *atm_z,
                  const double *c2s, const int *c2soff, double *S, double *T, double *V) {
    set_c2s(c2s, c2soff);
    if (!boys_ready) boys_build();
    basis_t bo = {ns, l, prim0, nprim, ao0, centre, exps, coefs, sph};
#pragma omp parallel for schedule(dynamic, 1)
    for (int ia = 0; ia < ns; ++ia) {
        double *R = (double *)malloc(sizeof(double) * MTMAX * MTMAX * MTMAX);
        double blkS[28 * 28], blkT[28 * 28], blkV[28 * 28];
        double one[1] = {1.0};
        for (int ib = 0; ib < ns; ++ib) {
            int la = l[ia], lb = l[ib];
            int ca[28][3], cb[28][3];
            cart_comps(la, ca); cart_comps(lb, cb);
            int nac = ncart(la), nbc = ncart(lb);
            memset(blkS, 0, sizeof(blkS)); memset(blkT, 0, sizeof(blkT)); memset(blkV, 0, sizeof(blkV));
            const double *A = centre + 3 * ia, *B = centre + 3 * ib;
            double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
            double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
            int mtS = la + lb + 3; /* kinetic needs j+2 */
            int mtV = la + lb + 1;
            double Es[3][(LMAX + 1) * (LMAX + 3) * (2 * LMAX + 3)];
            double Ev[3][(LMAX + 1) * (LMAX + 1) * (2 * LMAX + 1)];
            for (int pa = 0; pa < nprim[ia]; ++pa)
                for (int pb = 0; pb < nprim[ib]; ++pb) {
                    double al = exps[prim0[ia] + pa], be = exps[prim0[ib] + pb];
                    double cab = coefs[prim0[ia] + pa] * coefs[prim0[ib] + pb];
                    double p = al + be;
                    double P[3] = {(al * A[0] + be * B[0]) / p, (al * A[1] + be * B[1]) / p, (al * A[2] + be * B[2]) / p};
                    double K = exp(-al * be / p * ab2);
                    for (int x = 0; x < 3; ++x) {
                        e_table(la, lb + 2, al, be, AB[x], mtS, Es[x]);
                        e_table(la, lb, al, be, AB[x], mtV, Ev[x]);
                    }
                    double s0 = pow(M_PI / p, 1.5);
#define ES(x, i, j) Es[x][((i) * (lb + 3) + (j)) * mtS]
                    for (int a = 0; a < nac; ++a)
                        for (int b = 0; b < nbc; ++b) {
                            double s1[3], t1[3];
                            for (int x = 0; x < 3; ++x) {
                                int i = ca[a][x], j = cb[b][x];
                                s1[x] = ES(x, i, j);
                                double tt = -2.0 * be * be * ES(x, i, j + 2) + be * (2 * j + 1) * ES(x, i, j);
                                if (j >= 2) tt -= 0.5 * j * (j - 1) * ES(x, i, j - 2);
                                t1[x] = tt;
                            }
                            blkS[a * nbc + b] += cab * K * s0 * s1[0] * s1[1] * s1[2];
                            blkT[a * nbc + b] += cab * K * s0 * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                        }
#undef ES
                    for (int c = 0; c < natm; ++c) {
                        double PC[3] = {P[0] - atm_xyz[3 * c], P[1] - atm_xyz[3 * c + 1], P[2] - atm_xyz[3 * c + 2]};
                        int mt = mtV < 2 ? 2 : mtV;
                        r_table(p, PC, mt, R);
                        double pref = -atm_z[c] * cab * K * 2.0 * M_PI / p;
                        for (int a = 0; a < nac; ++a)
                            for (int b = 0; b < nbc; ++b) {
                                double acc = 0;
                                for (int t = 0; t <= ca[a][0] + cb[b][0]; ++t)
                                    for (int u = 0; u <= ca[a][1] + cb[b][1]; ++u)
                                        for (int v = 0; v <= ca[a][2] + cb[b][2]; ++v)
                                            acc += Ev[0][(ca[a][0] * (lb + 1) + cb[b][0]) * mtV + t] *
                                                   Ev[1][(ca[a][1] * (lb + 1) + cb[b][1]) * mtV + u] *
                                                   Ev[2][(ca[a][2] * (lb + 1) + cb[b][2]) * mtV + v] *
                                                   R[(t * mt + u) * mt + v];
                                blkV[a * nbc + b] += pref * acc;
                            }
                    }
                }
            /* to the final span */
            int naf = nfun(&bo, la), nbf = nfun(&bo, lb);
            int sa = (sph && la >= 2), sb = (sph && lb >= 2);
            const double *Ta = c2s_data + c2s_off[la], *Tb = c2s_data + c2s_off[lb];
            double *blks[3] = {blkS, blkT, blkV};
            double *outs[3] = {S, T, V};
            for (int w = 0; w < 3; ++w)
                for (int I = 0; I < naf; ++I)
                    for (int J = 0; J < nbf; ++J) {
                        double s = 0;
                        for (int i = 0; i < nac; ++i) {
                            double wi = sa ? Ta[i * naf + I] : (i == I ? 1.0 : 0.0);
                            if (wi == 0.0) continue;
                            for (int j = 0; j < nbc; ++j) {
                                double wj = sb ? Tb[j * nbf + J] : (j == J ? 1.0 : 0.0);
                                if (wj == 0.0) continue;
                                s += wi * wj * blks[w][i * nbc + j];
                            }
                        }
                        outs[w][(size_t)(ao0[ia] + I) * nao + ao0[ib] + J] = s;
                    }
        }
        free(R);
        (void)one;
    }
}

/* Four-centre (ab|cd) in the cartesian span, for orbital shells up to l = 2:
 *   (ab|cd) = K_ab K_cd 2 pi^2.5 / (p q sqrt(p+q))
 *             sum_{tuv} E^{ab}_{tuv} sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi}
 *             R_{t+tau, u+nu, v+phi}(pq/(p+q), P - Q)
 * (the exact-ERI path of dftax/integrals/eri4c.py in McMurchie-Davidson form,
 * with the same E and R tables as the 3-centre engine). Only the test pin of
 * the exact path uses it: out[((a*n + b)*n + c)*n + d], all n^4 entries. */
void oracle_int4c_cart(int ns, const int *l, const int *prim0, const int *nprim, const int *ao0,
                       const double *centre, const double *exps, const double *coefs, int n, double *out) {
    if (!boys_ready) boys_build();
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int ia = 0; ia < ns; ++ia)
        for (int ib = 0; ib < ns; ++ib) {
            double *R = (double *)malloc(sizeof(double) * MTMAX * MTMAX * MTMAX);
            double *Eab = (double *)malloc(sizeof(double) * 3 * 3 * 3 * MTMAX);
            double *Ecd = (double *)malloc(sizeof(double) * 3 * 3 * 3 * MTMAX);
            for (int ic = 0; ic < ns; ++ic)
                for (int id = 0; id < ns; ++id) {
                    int la = l[ia], lb = l[ib], lc = l[ic], ld = l[id];
                    int mt = la + lb + lc + ld + 1;
                    int ca[6][3], cb[6][3], cc[6][3], cd[6][3];
                    cart_comps(la, ca); cart_comps(lb, cb); cart_comps(lc, cc); cart_comps(ld, cd);
                    const double *A = centre + 3 * ia, *B = centre + 3 * ib, *C = centre + 3 * ic,
                                 *D = centre + 3 * id;
                    double AB[3], CD[3], ab2 = 0, cd2 = 0;
                    for (int x = 0; x < 3; ++x) {
                        AB[x] = A[x] - B[x]; CD[x] = C[x] - D[x];
                        ab2 += AB[x] * AB[x]; cd2 += CD[x] * CD[x];
                    }
                    int eab = (la + 1) * (lb + 1) * mt, ecd = (lc + 1) * (ld + 1) * mt;
                    for (int pa = 0; pa < nprim[ia]; ++pa)
                        for (int pb = 0; pb < nprim[ib]; ++pb) {
                            double al = exps[prim0[ia] + pa], be = exps[prim0[ib] + pb];
                            double p = al + be;
                            double cab = coefs[prim0[ia] + pa] * coefs[prim0[ib] + pb] * exp(-al * be / p * ab2);
                            double P[3];
                            for (int x = 0; x < 3; ++x) {
                                P[x] = (al * A[x] + be * B[x]) / p;
                                e_table(la, lb, al, be, AB[x], mt, Eab + x * eab);
                            }
                            for (int pc = 0; pc < nprim[ic]; ++pc)
                                for (int pd = 0; pd < nprim[id]; ++pd) {
                                    double ga = exps[prim0[ic] + pc], de = exps[prim0[id] + pd];
                                    double q = ga + de;
                                    double ccd = coefs[prim0[ic] + pc] * coefs[prim0[id] + pd] *
                                                 exp(-ga * de / q * cd2);
                                    double Q[3], PQ[3];
                                    for (int x = 0; x < 3; ++x) {
                                        Q[x] = (ga * C[x] + de * D[x]) / q;
                                        PQ[x] = P[x] - Q[x];
                                        e_table(lc, ld, ga, de, CD[x], mt, Ecd + x * ecd);
                                    }
                                    r_table(p * q / (p + q), PQ, mt, R);
                                    double pref = cab * ccd * 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q));
#define EAB(x, i, j, t) Eab[(x) * eab + ((i) * (lb + 1) + (j)) * mt + (t)]
#define ECD(x, i, j, t) Ecd[(x) * ecd + ((i) * (ld + 1) + (j)) * mt + (t)]
                                    for (int a = 0; a < ncart(la); ++a)
                                        for (int b = 0; b < ncart(lb); ++b)
                                            for (int c = 0; c < ncart(lc); ++c)
                                                for (int d = 0; d < ncart(ld); ++d) {
                                                    double acc = 0.0;
                                                    int ax = ca[a][0] , ay = ca[a][1], az = ca[a][2];
                                                    int bx = cb[b][0], by = cb[b][1], bz = cb[b][2];
                                                    int cx = cc[c][0], cy = cc[c][1], cz = cc[c][2];
                                                    int dx = cd[d][0], dy = cd[d][1], dz = cd[d][2];
                                                    for (int t = 0; t <= ax + bx; ++t)
                                                        for (int u = 0; u <= ay + by; ++u)
                                                            for (int v = 0; v <= az + bz; ++v) {
                                                                double e1 = EAB(0, ax, bx, t) * EAB(1, ay, by, u) *
                                                                            EAB(2, az, bz, v);
                                                                if (e1 == 0.0) continue;
                                                                for (int tt = 0; tt <= cx + dx; ++tt)
                                                                    for (int uu = 0; uu <= cy + dy; ++uu)
                                                                        for (int vv = 0; vv <= cz + dz; ++vv) {
                                                                            double sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                                                                            acc += e1 * sg * ECD(0, cx, dx, tt) *
                                                                                   ECD(1, cy, dy, uu) *
                                                                                   ECD(2, cz, dz, vv) *
                                                                                   R[((t + tt) * mt + (u + uu)) * mt +
                                                                                     (v + vv)];
                                                                        }
                                                            }
                                                    int m = ao0[ia] + a, nn = ao0[ib] + b, k = ao0[ic] + c,
                                                        s = ao0[id] + d;
                                                    out[(((size_t)m * n + nn) * n + k) * n + s] += pref * acc;
                                                }
#undef EAB
#undef ECD
                                }
                        }
                }
            free(R); free(Eab); free(Ecd);
        }
}

/* Selected shell triples (ia, ib | ic) of (mn|P) in the final spans, packed
 * block after block as [a][b][c] (for checks of a tensor too large to build
 * whole). trip[3 t + 0..2] = ia, ib, ic; out receives sum_t na nb nc values. */
void oracle_int3c_triples(int ns, const int *l, const int *prim0, const int *nprim, const int *ao0,
                          const double *centre, const double *exps, const double *coefs, int sph, int nao,
                          int nsx, const int *lx, const int *prim0x, const int *nprimx, const int *ao0x,
                          const double *centrex, const double *expsx, const double *coefsx, int sphx, int naux,
                          const double *c2s, const int *c2soff, int ntrip, const int *trip, const long *off,
                          double *out) {
    (void)nao; (void)naux;
    set_c2s(c2s, c2soff);
    if (!boys_ready) boys_build();
    basis_t bo = {ns, l, prim0, nprim, ao0, centre, exps, coefs, sph};
    basis_t bx = {nsx, lx, prim0x, nprimx, ao0x, centrex, expsx, coefsx, sphx};
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < ntrip; ++t) {
        int ia = trip[3 * t], ib = trip[3 * t + 1], ic = trip[3 * t + 2];
        double *cblk = (double *)malloc(sizeof(double) * 28 * 28 * 28);
        eri3c_block(&bo, ia, ib, &bx, ic, cblk);
        c2s_apply3(&bo, l[ia], &bo, l[ib], &bx, lx[ic], cblk, out + off[t]);
        free(cblk);
    }
}
