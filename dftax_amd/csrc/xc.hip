// Exchange-correlation on the grid (FP64), closed shell.
//
// Per point g:  rho = sum_n Z_gn phi_gn,  grad rho = 2 sum_n Z_gn grad phi_gn
// with Z = Phi P (a GEMM done before this kernel), then the pointwise
// functional f(rho, sigma) = eps_xc * rho with sigma = |grad rho|^2, masked to
// zero where rho <= rho_thresh (dftax/energy/grid.py:10-44). The Vxc matrix is
// V = Phi^T W + W^T Phi with W_gn = w_g (1/2 f_rho phi_gn + 2 f_sigma grad rho . grad phi_gn),
// i.e. sym(dE_xc/dP) of the quadrature (dftax/ks/terms.py:791-828 + jax.grad).
//
// The functionals are transcribed from dftax/energy/xc.py (PBEExchange 87-138,
// PBECorrelation 141-239, B88Exchange 255-294, VWN_RPACorrelation 376-419,
// LYPCorrelation 427-519, LDAExchange 64-84; PBE0 584-601, B3LYP 527-567),
// including the same density/gradient floors, and evaluated on a forward-mode
// dual number carrying d/drho and d/dsigma, so f_rho and f_sigma are exact to
// rounding.
#include <cstdint>

#include "dfks_common.h"

namespace dfks {

enum XcId { kXcPBE = 1, kXcPBE0 = 2, kXcB3LYP = 3, kXcCAMB3LYP = 4 };

// Forward-mode dual number over the two functional arguments: value, d/drho,
// d/dsigma. Dual<double> gives f, f_rho, f_sigma; Dual<Dual<double>> (a dual
// whose parts are duals) also carries the second derivatives f_rho rho,
// f_rho sigma, f_sigma sigma for the response kernel. The functionals below
// are written once, generic over the number type.
template <class T>
struct Dual {
  T v, r, s;  // value, d/drho, d/dsigma
};
using D = Dual<double>;
using D2 = Dual<Dual<double>>;

__device__ __forceinline__ double val(double x) { return x; }
template <class T>
__device__ __forceinline__ double val(const Dual<T>& a) { return val(a.v); }

template <class T>
__device__ __forceinline__ Dual<T> operator+(const Dual<T>& a, const Dual<T>& b) { return {a.v + b.v, a.r + b.r, a.s + b.s}; }
template <class T>
__device__ __forceinline__ Dual<T> operator-(const Dual<T>& a, const Dual<T>& b) { return {a.v - b.v, a.r - b.r, a.s - b.s}; }
template <class T>
__device__ __forceinline__ Dual<T> operator-(const Dual<T>& a) { return {-a.v, -a.r, -a.s}; }
template <class T>
__device__ __forceinline__ Dual<T> operator*(const Dual<T>& a, const Dual<T>& b) {
  return {a.v * b.v, a.r * b.v + a.v * b.r, a.s * b.v + a.v * b.s};
}
template <class T>
__device__ __forceinline__ Dual<T> operator/(const Dual<T>& a, const Dual<T>& b) {
  const T iv = 1.0 / b.v, q = a.v * iv;
  return {q, (a.r - q * b.r) * iv, (a.s - q * b.s) * iv};
}
template <class T>
__device__ __forceinline__ Dual<T> operator+(const Dual<T>& a, double c) { return {a.v + c, a.r, a.s}; }
template <class T>
__device__ __forceinline__ Dual<T> operator+(double c, const Dual<T>& a) { return {a.v + c, a.r, a.s}; }
template <class T>
__device__ __forceinline__ Dual<T> operator-(const Dual<T>& a, double c) { return {a.v - c, a.r, a.s}; }
template <class T>
__device__ __forceinline__ Dual<T> operator-(double c, const Dual<T>& a) { return {c - a.v, -a.r, -a.s}; }
template <class T>
__device__ __forceinline__ Dual<T> operator*(const Dual<T>& a, double c) { return {a.v * c, a.r * c, a.s * c}; }
template <class T>
__device__ __forceinline__ Dual<T> operator*(double c, const Dual<T>& a) { return {a.v * c, a.r * c, a.s * c}; }
template <class T>
__device__ __forceinline__ Dual<T> operator/(const Dual<T>& a, double c) { return {a.v / c, a.r / c, a.s / c}; }
template <class T>
__device__ __forceinline__ Dual<T> operator/(double c, const Dual<T>& a) {
  const T q = c / a.v;
  const T dq = -q / a.v;
  return {q, dq * a.r, dq * a.s};
}

// f(a) from f(a.v) and f'(a.v), both evaluated in the inner number type
template <class T>
__device__ __forceinline__ Dual<T> chain(const Dual<T>& a, const T& f, const T& df) { return {f, df * a.r, df * a.s}; }

__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ double dlog1p(double x) { return log1p(x); }
__device__ __forceinline__ double dexpm1(double x) { return expm1(x); }
__device__ __forceinline__ double datan(double x) { return atan(x); }
__device__ __forceinline__ double dasinh(double x) { return asinh(x); }
__device__ __forceinline__ double dpow(double x, double p) { return pow(x, p); }
__device__ __forceinline__ double derf(double x) { return erf(x); }

template <class T>
__device__ __forceinline__ Dual<T> dsqrt(const Dual<T>& a) { const T f = dsqrt(a.v); return chain(a, f, 0.5 / f); }
template <class T>
__device__ __forceinline__ Dual<T> dexp(const Dual<T>& a) { const T f = dexp(a.v); return chain(a, f, f); }
template <class T>
__device__ __forceinline__ Dual<T> dlog(const Dual<T>& a) { return chain(a, dlog(a.v), 1.0 / a.v); }
template <class T>
__device__ __forceinline__ Dual<T> dlog1p(const Dual<T>& a) { return chain(a, dlog1p(a.v), 1.0 / (1.0 + a.v)); }
template <class T>
__device__ __forceinline__ Dual<T> dexpm1(const Dual<T>& a) { return chain(a, dexpm1(a.v), dexp(a.v)); }
template <class T>
__device__ __forceinline__ Dual<T> datan(const Dual<T>& a) { return chain(a, datan(a.v), 1.0 / (1.0 + a.v * a.v)); }
template <class T>
__device__ __forceinline__ Dual<T> dasinh(const Dual<T>& a) {
  return chain(a, dasinh(a.v), 1.0 / dsqrt(1.0 + a.v * a.v));
}
template <class T>
__device__ __forceinline__ Dual<T> derf(const Dual<T>& a) {
  return chain(a, derf(a.v), 1.1283791670955126 * dexp(-(a.v * a.v)));   // 2/sqrt(pi) e^{-x^2}
}
template <class T>
__device__ __forceinline__ Dual<T> dpow(const Dual<T>& a, double p) {
  const T f = dpow(a.v, p);
  return chain(a, f, p * f / a.v);
}
// jnp.clip(x, lo): derivative passes where x >= lo
template <class N>
__device__ __forceinline__ N dclip(const N& a, double lo) { return val(a) >= lo ? a : N{} + lo; }

constexpr double kPi2 = kPi * kPi;

template <class N>
__device__ N lda_x(N rho) {
  N kF = dpow(3.0 * kPi2 * rho + 1e-30, 1.0 / 3.0);
  return -3.0 * kF / (4.0 * kPi);
}

// |grad + 1e-30| of dftax's PBE/B88 exchange, written on sigma
template <class N>
__device__ __forceinline__ N grad_norm_x(N sig) { return dsqrt(sig + 3e-60); }

template <class N>
__device__ N pbe_x(N rho, N sig) {
  const double kappa = 0.804, mu = 0.2195149727645171;
  N gn = grad_norm_x(sig);
  N kF = dpow(3.0 * kPi2 * rho + 1e-30, 1.0 / 3.0);
  N s = gn / dclip(2.0 * kF * rho, 1e-30);
  N eps_unif = -3.0 * kF / (4.0 * kPi);
  N f = 1.0 + kappa - kappa / (1.0 + mu * s * s / kappa);
  return eps_unif * f;
}

template <class N>
__device__ N pbe_c(N rho, N sig) {
  const double gamma = (1.0 - 0.69314718055994530942) / kPi2;
  const double beta = 0.06672455060314922;
  const double A = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  N rt = dclip(rho, 1e-20);
  N rs = dpow(3.0 / (4.0 * kPi * rt) + 1e-20, 1.0 / 3.0);
  N kf = dpow(3.0 * kPi2 * rt + 1e-20, 1.0 / 3.0);
  N ks = dsqrt(4.0 * kf / kPi);
  N gn = dsqrt(sig + 1e-30);
  N t = gn / dclip(2.0 * ks * rt, 1e-20);
  N rs_sqrt = dsqrt(rs);
  N denom = 2.0 * A * (b1 * rs_sqrt + b2 * rs + b3 * rs * rs_sqrt + b4 * rs * rs);
  // zeta = 0: the PW92 spin interpolation reduces to the paramagnetic branch
  N eps_pw = -2.0 * A * (1.0 + a1 * rs) * dlog1p(1.0 / dclip(denom, 1e-20));
  N denom_A = gamma * dexpm1(-eps_pw / gamma);
  N Ac = beta / dclip(denom_A, 1e-20);
  N t2 = t * t;
  N f1 = t2 + Ac * t2 * t2;
  N frac = (beta * f1) / dclip(gamma * (1.0 + Ac * f1), 1e-20);
  return eps_pw + gamma * dlog1p(frac);
}

template <class N>
__device__ N b88_x(N rho, N sig) {
  const double beta = 0.0042;
  N gn = grad_norm_x(sig);
  N r = dclip(rho * 0.5, 1e-20);
  N r13 = dpow(r, 1.0 / 3.0);
  const double C = 0.75 * 1.2407009817988000;  // (6/pi)^(1/3)
  N x = (gn * 0.5) / dclip(dpow(r, 4.0 / 3.0), 1e-20);
  N den = 1.0 + 6.0 * beta * x * dasinh(x);
  return -C * r13 - beta * r13 * x * x / dclip(den, 1e-20);
}

// Paramagnetic VWN correlation per electron (VWN-RPA and VWN5 differ in the
// constants of this one branch function, xc.py:302-320 and 376-419)
template <class N>
__device__ N vwn_para(N rho, double A, double b, double c, double x0) {
  N r = dclip(rho, 1e-20);
  N rs = dpow(3.0 / (4.0 * kPi * r), 1.0 / 3.0);
  N x = dsqrt(rs);
  N Xx = x * x + b * x + c;
  double Xx0 = x0 * x0 + b * x0 + c;
  double Q = sqrt(4.0 * c - b * b);
  N at = datan(Q / dclip(2.0 * x + b, 1e-20));
  N term1 = dlog(x * x / dclip(Xx, 1e-20));
  N term2 = (2.0 * b / Q) * at;
  N xm = x - x0;
  N term3 = -(b * x0 / Xx0) * (dlog(xm * xm / dclip(Xx, 1e-20)) + (2.0 * (b + 2.0 * x0) / Q) * at);
  return A * (term1 + term2 + term3);
}

template <class N>
__device__ N vwn_rpa_c(N rho) { return vwn_para(rho, 0.0310907, 13.0720, 42.7198, -0.409286); }

// VWN5 (xc.py:323-352, paramagnetic parameters)
template <class N>
__device__ N vwn5_c(N rho) { return vwn_para(rho, 0.0310907, 3.72744, 12.9352, -0.10498); }

template <class N>
__device__ N lyp_c(N rho, N sig) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double C_F = 0.3 * 9.5707800006273;  // (3 pi^2)^(2/3)
  N r = dclip(rho, 1e-20);
  N ri3 = dpow(r, -1.0 / 3.0);
  N g = 1.0 / (1.0 + d * ri3);
  N om = dexp(-c * ri3) * g * dpow(r, -11.0 / 3.0);
  N delta = c * ri3 + d * ri3 * g;
  N term_ff = C_F * dpow(r, 8.0 / 3.0) - sig * (3.0 + 7.0 * delta) / 72.0;
  return -a * g - a * b * om * r * term_ff;
}

// Short-range fraction F(a) of an erfc-attenuated exchange hole (xc.py:609-631):
// F(a) = 1 - (8/3) a [sqrt(pi) erf(1/2a) - 3a + 4a^3 + (2a - 4a^3) e^{-1/(4a^2)}],
// and its asymptote 1/(36 a^2) from a = 8 on.
template <class N>
__device__ N erf_attenuation(N a) {
  a = dclip(a, 1e-10);
  if (val(a) < 8.0) {
    N a2 = a * a;
    N a3 = a2 * a;
    return 1.0 - (8.0 / 3.0) * a *
                     (1.7724538509055160 * derf(1.0 / (2.0 * a)) - 3.0 * a + 4.0 * a3 +
                      (2.0 * a - 4.0 * a3) * dexp(-1.0 / (4.0 * a2)));
  }
  return 1.0 / (36.0 * a * a);
}

// Short-range B88 exchange by the ITYH scheme at omega = 0.33 (xc.py:664-709),
// closed shell: the per-spin B88 energy eps at (rho/2, |grad rho|/2) times
// F(a), a = omega sqrt(K) / (6 sqrt(pi) rho_s^{1/3}), K = -2 eps / rho_s^{1/3}.
template <class N>
__device__ N ityh_b88_x(N rho, N sig) {
  const double omega = 0.33;
  N eps = b88_x(rho, sig);
  N rs = dclip(rho * 0.5, 1e-20);
  N r13 = dpow(rs, 1.0 / 3.0);
  N K = dclip(-2.0 * eps / r13, 1e-20);
  N a = omega * dsqrt(K) / (6.0 * 1.7724538509055160 * r13);
  return eps * erf_attenuation(a);
}

template <class N>
__device__ N eps_xc(int id, N rho, N sig) {
  switch (id) {
    case kXcPBE: return pbe_x(rho, sig) + pbe_c(rho, sig);
    case kXcPBE0: return 0.75 * pbe_x(rho, sig) + pbe_c(rho, sig);
    case kXcCAMB3LYP:   // xc.py:711-750: its exact exchange is 0.19 K + 0.46 K_lr(omega = 0.33)
      return 0.35 * b88_x(rho, sig) + 0.46 * ityh_b88_x(rho, sig) + 0.81 * lyp_c(rho, sig) + 0.19 * vwn5_c(rho);
    default:
      return 0.08 * lda_x(rho) + 0.72 * b88_x(rho, sig) + 0.19 * vwn_rpa_c(rho) + 0.81 * lyp_c(rho, sig);
  }
}

// ---------------------------------------------------------------------------
// Spin-polarised functionals (dftax/energy/xc.py spin branches), evaluated on
// a dual number over the five arguments (rho_a, rho_b, sigma_aa, sigma_ab,
// sigma_bb), so one pass gives f and its five first partials. Each returns
// the energy per electron of the total density.
// ---------------------------------------------------------------------------
struct D5 {
  double v;
  double d[5];
};
__device__ __forceinline__ double val(const D5& a) { return a.v; }
__device__ __forceinline__ D5 d5_map(const D5& a, double f, double df) {
  D5 r;
  r.v = f;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = df * a.d[k];
  return r;
}
__device__ __forceinline__ D5 operator+(const D5& a, const D5& b) {
  D5 r; r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ D5 operator-(const D5& a, const D5& b) {
  D5 r; r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ D5 operator-(const D5& a) { return d5_map(a, -a.v, -1.0); }
__device__ __forceinline__ D5 operator*(const D5& a, const D5& b) {
  D5 r; r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ D5 operator/(const D5& a, const D5& b) {
  const double iv = 1.0 / b.v, q = a.v * iv;
  D5 r; r.v = q;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - q * b.d[k]) * iv;
  return r;
}
__device__ __forceinline__ D5 operator+(const D5& a, double c) { D5 r = a; r.v += c; return r; }
__device__ __forceinline__ D5 operator+(double c, const D5& a) { return a + c; }
__device__ __forceinline__ D5 operator-(const D5& a, double c) { D5 r = a; r.v -= c; return r; }
__device__ __forceinline__ D5 operator-(double c, const D5& a) { return d5_map(a, c - a.v, -1.0); }
__device__ __forceinline__ D5 operator*(const D5& a, double c) { return d5_map(a, a.v * c, c); }
__device__ __forceinline__ D5 operator*(double c, const D5& a) { return d5_map(a, a.v * c, c); }
__device__ __forceinline__ D5 operator/(const D5& a, double c) { return d5_map(a, a.v / c, 1.0 / c); }
__device__ __forceinline__ D5 operator/(double c, const D5& a) {
  const double q = c / a.v;
  return d5_map(a, q, -q / a.v);
}
__device__ __forceinline__ D5 dsqrt(const D5& a) { const double f = sqrt(a.v); return d5_map(a, f, 0.5 / f); }
__device__ __forceinline__ D5 dexp(const D5& a) { const double f = exp(a.v); return d5_map(a, f, f); }
__device__ __forceinline__ D5 dlog(const D5& a) { return d5_map(a, log(a.v), 1.0 / a.v); }
__device__ __forceinline__ D5 dlog1p(const D5& a) { return d5_map(a, log1p(a.v), 1.0 / (1.0 + a.v)); }
__device__ __forceinline__ D5 dexpm1(const D5& a) { return d5_map(a, expm1(a.v), exp(a.v)); }
__device__ __forceinline__ D5 datan(const D5& a) { return d5_map(a, atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
__device__ __forceinline__ D5 dasinh(const D5& a) { return d5_map(a, asinh(a.v), 1.0 / sqrt(1.0 + a.v * a.v)); }
__device__ __forceinline__ D5 dpow(const D5& a, double p) {
  const double f = pow(a.v, p);
  return d5_map(a, f, p * f / a.v);
}
// jnp.clip(x, lo, hi): derivative passes inside
template <class N>
__device__ __forceinline__ N dclip2(const N& a, double lo, double hi) {
  return val(a) < lo ? N{} + lo : (val(a) > hi ? N{} + hi : a);
}

// Spin scaling E_x[ra, rb] = (E_x[2 ra] + E_x[2 rb]) / 2 per electron (xc.py:76-83, 116-134):
// the unpolarised forms at (2 r_s, 4 sigma_ss), mixed by the channel populations.
template <class N>
__device__ N lda_x_pol(N ra, N rb) {
  N n = dclip(ra + rb, 1e-30);
  return (ra / n) * lda_x(2.0 * ra) + (rb / n) * lda_x(2.0 * rb);
}

template <class N>
__device__ N pbe_x_pol(N ra, N rb, N saa, N sbb) {
  N n = dclip(ra + rb, 1e-30);
  return (ra / n) * pbe_x(2.0 * ra, 4.0 * saa) + (rb / n) * pbe_x(2.0 * rb, 4.0 * sbb);
}

// B88 per channel (xc.py:271-292); b88_x(2 r_s, 4 sigma_ss) is the channel's form
template <class N>
__device__ N b88_x_pol(N ra, N rb, N saa, N sbb) {
  N n = dclip(ra + rb, 1e-20);
  return (ra / n) * b88_x(2.0 * ra, 4.0 * saa) + (rb / n) * b88_x(2.0 * rb, 4.0 * sbb);
}

// PW92 with its spin interpolation (xc.py:165-177); rows para, ferro, -alpha_c
template <class N>
__device__ N pw92_pol(N rs, N zeta) {
  const double P[3][6] = {{0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294},
                          {0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517},
                          {0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671}};
  const double fz20 = 1.709920934161365617563962776245;
  N rq = dsqrt(rs);
  N g[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double A = P[r][0], a1 = P[r][1], b1 = P[r][2], b2 = P[r][3], b3 = P[r][4], b4 = P[r][5];
    N denom = 2.0 * A * (b1 * rq + b2 * rs + b3 * rs * rq + b4 * rs * rs);
    g[r] = -2.0 * A * (1.0 + a1 * rs) * dlog1p(1.0 / dclip(denom, 1e-20));
  }
  N fz = (dpow(1.0 + zeta, 4.0 / 3.0) + dpow(1.0 - zeta, 4.0 / 3.0) - 2.0) / (2.5198420997897464 - 2.0);
  N z2 = zeta * zeta;
  return g[0] + z2 * z2 * fz * (g[1] - g[0] + g[2] / fz20) - fz * g[2] / fz20;
}

// PBE correlation of a polarised density (xc.py:179-237): stot = |grad ra + grad rb|^2
template <class N>
__device__ N pbe_c_pol(N ra, N rb, N stot) {
  const double gamma = (1.0 - 0.69314718055994530942) / kPi2;
  const double beta = 0.06672455060314922;
  N a = dclip(ra, 1e-20), b = dclip(rb, 1e-20);
  N n = a + b;
  N zeta = dclip2((a - b) / n, -1.0 + 1e-7, 1.0 - 1e-7);
  N rs = dpow(3.0 / (4.0 * kPi * n) + 1e-20, 1.0 / 3.0);
  N phi = dclip((dpow(1.0 + zeta, 2.0 / 3.0) + dpow(1.0 - zeta, 2.0 / 3.0)) * 0.5, 1e-20);
  N kf = dpow(3.0 * kPi2 * n + 1e-20, 1.0 / 3.0);
  N ks = dsqrt(4.0 * kf / kPi);
  N gn = dsqrt(stot + 1e-30);
  N t = gn / dclip(2.0 * phi * ks * n, 1e-20);
  N eps_pw = pw92_pol(rs, zeta);
  N gphi3 = gamma * phi * phi * phi;
  N Ac = beta / dclip(gamma * dexpm1(-eps_pw / gphi3), 1e-20);
  N t2 = t * t;
  N f1 = t2 + Ac * t2 * t2;
  N frac = (beta * f1) / dclip(gamma * (1.0 + Ac * f1), 1e-20);
  return eps_pw + gphi3 * dlog1p(frac);
}

template <class N>
__device__ N vwn_branch(N x, double A, double b, double c, double x0) {
  N Xx = x * x + b * x + c;
  const double Xx0 = x0 * x0 + b * x0 + c;
  const double Q = sqrt(4.0 * c - b * b);
  N at = datan(Q / dclip(2.0 * x + b, 1e-20));
  N xm = x - x0;
  return A * (dlog(x * x / dclip(Xx, 1e-20)) + (2.0 * b / Q) * at -
              (b * x0 / Xx0) * (dlog(xm * xm / dclip(Xx, 1e-20)) + (2.0 * (b + 2.0 * x0) / Q) * at));
}

// VWN RPA: para-to-ferro interpolation without spin stiffness (xc.py:403-417)
template <class N>
__device__ N vwn_rpa_c_pol(N ra, N rb) {
  N n = dclip(ra + rb, 1e-20);
  N zeta = dclip2((ra - rb) / n, -1.0 + 1e-7, 1.0 - 1e-7);
  N x = dsqrt(dpow(3.0 / (4.0 * kPi * n), 1.0 / 3.0));
  N ep = vwn_branch(x, 0.0310907, 13.0720, 42.7198, -0.409286);
  N ef = vwn_branch(x, 0.01554535, 20.1231, 101.578, -0.743294);
  N fz = (dpow(1.0 + zeta, 4.0 / 3.0) + dpow(1.0 - zeta, 4.0 / 3.0) - 2.0) / (2.5198420997897464 - 2.0);
  return ep + (ef - ep) * fz;
}

// Open-shell LYP, Miehlich et al. 1989 (xc.py:489-519), per electron
template <class N>
__device__ N lyp_c_pol(N ra, N rb, N gaa, N gab, N gbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double C_F = 0.3 * 9.5707800006273;  // (3 pi^2)^(2/3)
  N ca = dclip(ra, 1e-20), cb = dclip(rb, 1e-20);
  N rho = ca + cb;
  N gt = gaa + 2.0 * gab + gbb;
  N r13 = dpow(rho, -1.0 / 3.0);
  N g = 1.0 / (1.0 + d * r13);
  N om = dexp(-c * r13) * dpow(rho, -11.0 / 3.0) * g;
  N dl = c * r13 + d * r13 * g;
  N e1 = -a * 4.0 * ca * cb / rho * g;
  N br = ca * cb * (12.699208415745595 * C_F * (dpow(ca, 8.0 / 3.0) + dpow(cb, 8.0 / 3.0))
                    + (47.0 / 18.0 - 7.0 / 18.0 * dl) * gt - (2.5 - dl / 18.0) * (gaa + gbb)
                    - (dl - 11.0) / 9.0 * (ca / rho * gaa + cb / rho * gbb))
         - (2.0 / 3.0) * rho * rho * gt + ((2.0 / 3.0) * rho * rho - ca * ca) * gbb
         + ((2.0 / 3.0) * rho * rho - cb * cb) * gaa;
  return (e1 - a * b * om * br) / rho;
}

template <class N>
__device__ N eps_xc_pol(int id, N ra, N rb, N saa, N sab, N sbb) {
  switch (id) {
    case kXcPBE: return pbe_x_pol(ra, rb, saa, sbb) + pbe_c_pol(ra, rb, saa + 2.0 * sab + sbb);
    case kXcPBE0: return 0.75 * pbe_x_pol(ra, rb, saa, sbb) + pbe_c_pol(ra, rb, saa + 2.0 * sab + sbb);
    default:
      return 0.08 * lda_x_pol(ra, rb) + 0.72 * b88_x_pol(ra, rb, saa, sbb) + 0.19 * vwn_rpa_c_pol(ra, rb) +
             0.81 * lyp_c_pol(ra, rb, saa, sab, sbb);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One wavefront per grid point, 4 points per 256-thread block.
// zw: on entry Z = Phi P (ng x nao); on exit W (ng x nao), in place.
// TA / TZ: storage types of the AO planes and of Z/W (double, or float on the
// mixed-precision path); the reduction and the functional stay FP64.
template <typename TA, typename TZ, bool VEC>
__global__ __launch_bounds__(256) void k_xc(int64_t ng, int64_t nao, const TA* __restrict__ ao,
                                            const double* __restrict__ wts, TZ* __restrict__ zw, int xc_id,
                                            double thresh, double* __restrict__ exc, double* __restrict__ rho_out) {
  const int lane = threadIdx.x & 63;
  const int64_t g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;
  const int64_t plane = ng * nao;
  const TA* p0 = ao + g * nao;
  const TA* p1 = p0 + plane;
  const TA* p2 = p0 + 2 * plane;
  const TA* p3 = p0 + 3 * plane;
  TZ* zrow = zw + g * nao;
  double r = 0, gx = 0, gy = 0, gz = 0;
  constexpr int PER = 16 / sizeof(TA);
  typedef TA vecA __attribute__((ext_vector_type(PER)));
  typedef TZ vecZ __attribute__((ext_vector_type(PER)));
  if (VEC) {
    for (int64_t n = lane; n < nao / PER; n += 64) {
      vecZ z = reinterpret_cast<const vecZ*>(zrow)[n];
      vecA a0 = reinterpret_cast<const vecA*>(p0)[n], a1 = reinterpret_cast<const vecA*>(p1)[n];
      vecA a2 = reinterpret_cast<const vecA*>(p2)[n], a3 = reinterpret_cast<const vecA*>(p3)[n];
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        double zz = (double)z[j];
        r += zz * (double)a0[j];
        gx += zz * (double)a1[j];
        gy += zz * (double)a2[j];
        gz += zz * (double)a3[j];
      }
    }
  } else {
    for (int64_t n = lane; n < nao; n += 64) {
      double z = (double)zrow[n];
      r += z * (double)p0[n];
      gx += z * (double)p1[n];
      gy += z * (double)p2[n];
      gz += z * (double)p3[n];
    }
  }
  r = wave_sum(r);
  gx = 2.0 * wave_sum(gx);
  gy = 2.0 * wave_sum(gy);
  gz = 2.0 * wave_sum(gz);
  if (rho_out && lane == 0) {
    rho_out[4 * g] = r; rho_out[4 * g + 1] = gx; rho_out[4 * g + 2] = gy; rho_out[4 * g + 3] = gz;
  }
  double a = 0.0, bx = 0.0, by = 0.0, bz = 0.0, e = 0.0;
  if (r > thresh) {
    const double w = wts[g];
    D rho{r, 1.0, 0.0};
    D sig{gx * gx + gy * gy + gz * gz, 0.0, 1.0};
    D f = eps_xc(xc_id, rho, sig) * rho;
    e = w * f.v;
    a = 0.5 * w * f.r;
    const double t = 2.0 * w * f.s;
    bx = t * gx; by = t * gy; bz = t * gz;
  }
  if (lane == 0) exc[g] = e;
  if (VEC) {
    for (int64_t n = lane; n < nao / PER; n += 64) {
      vecA a0 = reinterpret_cast<const vecA*>(p0)[n], a1 = reinterpret_cast<const vecA*>(p1)[n];
      vecA a2 = reinterpret_cast<const vecA*>(p2)[n], a3 = reinterpret_cast<const vecA*>(p3)[n];
      vecZ o;
#pragma unroll
      for (int j = 0; j < PER; ++j)
        o[j] = (TZ)(a * (double)a0[j] + bx * (double)a1[j] + by * (double)a2[j] + bz * (double)a3[j]);
      reinterpret_cast<vecZ*>(zrow)[n] = o;
    }
  } else {
    for (int64_t n = lane; n < nao; n += 64)
      zrow[n] = (TZ)(a * (double)p0[n] + bx * (double)p1[n] + by * (double)p2[n] + bz * (double)p3[n]);
  }
}

// Single-read variant for the FP64 path: one 256-thread block per grid point;
// each thread keeps its slice of the four AO plane rows (J vectors of 16 bytes
// per plane) in registers between the rho / grad-rho reduction and the W
// update, so every AO value is read from HBM once per Fock build instead of
// twice. Needs the vector layout (nao a multiple of the vector width, 16-byte
// aligned rows) and nao <= 256 J PER. (The FP32 path keeps the
// one-wavefront-per-point kernel: its second read costs less than this
// variant's lower occupancy.)
template <typename TA, typename TZ, int J>
__global__ __launch_bounds__(256) void k_xc1(int64_t ng, int64_t nao, const TA* __restrict__ ao,
                                             const double* __restrict__ wts, TZ* __restrict__ zw, int xc_id,
                                             double thresh, double* __restrict__ exc) {
  __shared__ double s_red[4][4];
  __shared__ double s_coef[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t g = blockIdx.x;
  const int64_t plane = ng * nao;
  constexpr int PER = 16 / sizeof(TA);
  typedef TA vecA __attribute__((ext_vector_type(PER)));
  typedef TZ vecZ __attribute__((ext_vector_type(PER)));
  const vecA* p0 = reinterpret_cast<const vecA*>(ao + g * nao);
  const vecA* p1 = reinterpret_cast<const vecA*>(ao + plane + g * nao);
  const vecA* p2 = reinterpret_cast<const vecA*>(ao + 2 * plane + g * nao);
  const vecA* p3 = reinterpret_cast<const vecA*>(ao + 3 * plane + g * nao);
  vecZ* zrow = reinterpret_cast<vecZ*>(zw + g * nao);
  const int nv = (int)(nao / PER);
  vecA a0[J], a1[J], a2[J], a3[J];
  double r = 0, gx = 0, gy = 0, gz = 0;
#pragma unroll
  for (int jv = 0; jv < J; ++jv) {
    const int n = tid + 256 * jv;
    if (n < nv) {
      const vecZ z = zrow[n];
      a0[jv] = p0[n]; a1[jv] = p1[n]; a2[jv] = p2[n]; a3[jv] = p3[n];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const double zz = (double)z[q];
        r += zz * (double)a0[jv][q];
        gx += zz * (double)a1[jv][q];
        gy += zz * (double)a2[jv][q];
        gz += zz * (double)a3[jv][q];
      }
    }
  }
  r = wave_sum(r);
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  if (lane == 0) { s_red[wave][0] = r; s_red[wave][1] = gx; s_red[wave][2] = gy; s_red[wave][3] = gz; }
  __syncthreads();
  if (tid == 0) {
    r = s_red[0][0] + s_red[1][0] + s_red[2][0] + s_red[3][0];
    gx = 2.0 * (s_red[0][1] + s_red[1][1] + s_red[2][1] + s_red[3][1]);
    gy = 2.0 * (s_red[0][2] + s_red[1][2] + s_red[2][2] + s_red[3][2]);
    gz = 2.0 * (s_red[0][3] + s_red[1][3] + s_red[2][3] + s_red[3][3]);
    double a = 0.0, bx = 0.0, by = 0.0, bz = 0.0, e = 0.0;
    if (r > thresh) {
      const double w = wts[g];
      D rho{r, 1.0, 0.0};
      D sig{gx * gx + gy * gy + gz * gz, 0.0, 1.0};
      D f = eps_xc(xc_id, rho, sig) * rho;
      e = w * f.v;
      a = 0.5 * w * f.r;
      const double t = 2.0 * w * f.s;
      bx = t * gx; by = t * gy; bz = t * gz;
    }
    exc[g] = e;
    s_coef[0] = a; s_coef[1] = bx; s_coef[2] = by; s_coef[3] = bz;
  }
  __syncthreads();
  const double a = s_coef[0], bx = s_coef[1], by = s_coef[2], bz = s_coef[3];
#pragma unroll
  for (int jv = 0; jv < J; ++jv) {
    const int n = tid + 256 * jv;
    if (n < nv) {
      vecZ o;
#pragma unroll
      for (int q = 0; q < PER; ++q)
        o[q] = (TZ)(a * (double)a0[jv][q] + bx * (double)a1[jv][q] + by * (double)a2[jv][q] + bz * (double)a3[jv][q]);
      zrow[n] = o;
    }
  }
}

// Single-read variant for short rows (the compact grid's blocks): one
// wavefront per point, four points per block; each lane keeps its J vectors of
// the four AO plane rows in registers between the rho / grad-rho reduction and
// the W update, so every AO value is read once. Needs the vector layout and
// nao <= 64 J PER.
template <typename TA, typename TZ, int J>
__global__ __launch_bounds__(256) void k_xcw(int64_t ng, int64_t nao, const TA* __restrict__ ao,
                                             const double* __restrict__ wts, TZ* __restrict__ zw, int xc_id,
                                             double thresh, double* __restrict__ exc, double* __restrict__ rho_out) {
  const int lane = threadIdx.x & 63;
  const int64_t g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;
  constexpr int PER = 16 / sizeof(TA);
  typedef TA vecA __attribute__((ext_vector_type(PER)));
  typedef TZ vecZ __attribute__((ext_vector_type(PER)));
  const int64_t plane = ng * nao;
  const vecA* p0 = reinterpret_cast<const vecA*>(ao + g * nao);
  const vecA* p1 = reinterpret_cast<const vecA*>(ao + plane + g * nao);
  const vecA* p2 = reinterpret_cast<const vecA*>(ao + 2 * plane + g * nao);
  const vecA* p3 = reinterpret_cast<const vecA*>(ao + 3 * plane + g * nao);
  vecZ* zrow = reinterpret_cast<vecZ*>(zw + g * nao);
  const int nv = (int)(nao / PER);
  vecA a0[J], a1[J], a2[J], a3[J];
  double r = 0, gx = 0, gy = 0, gz = 0;
#pragma unroll
  for (int jv = 0; jv < J; ++jv) {
    const int n = lane + 64 * jv;
    if (n < nv) {
      const vecZ z = zrow[n];
      a0[jv] = p0[n]; a1[jv] = p1[n]; a2[jv] = p2[n]; a3[jv] = p3[n];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const double zz = (double)z[q];
        r += zz * (double)a0[jv][q];
        gx += zz * (double)a1[jv][q];
        gy += zz * (double)a2[jv][q];
        gz += zz * (double)a3[jv][q];
      }
    }
  }
  r = wave_sum(r);
  gx = 2.0 * wave_sum(gx);
  gy = 2.0 * wave_sum(gy);
  gz = 2.0 * wave_sum(gz);
  if (rho_out && lane == 0) {
    rho_out[4 * g] = r; rho_out[4 * g + 1] = gx; rho_out[4 * g + 2] = gy; rho_out[4 * g + 3] = gz;
  }
  double a = 0.0, bx = 0.0, by = 0.0, bz = 0.0, e = 0.0;
  if (r > thresh) {
    const double w = wts[g];
    D rho{r, 1.0, 0.0};
    D sig{gx * gx + gy * gy + gz * gz, 0.0, 1.0};
    D f = eps_xc(xc_id, rho, sig) * rho;
    e = w * f.v;
    a = 0.5 * w * f.r;
    const double t = 2.0 * w * f.s;
    bx = t * gx; by = t * gy; bz = t * gz;
  }
  if (lane == 0) exc[g] = e;
#pragma unroll
  for (int jv = 0; jv < J; ++jv) {
    const int n = lane + 64 * jv;
    if (n < nv) {
      vecZ o;
#pragma unroll
      for (int q = 0; q < PER; ++q)
        o[q] = (TZ)(a * (double)a0[jv][q] + bx * (double)a1[jv][q] + by * (double)a2[jv][q] + bz * (double)a3[jv][q]);
      zrow[n] = o;
    }
  }
}

template <typename TA, typename TZ>
static bool launch_xcw(int64_t ng, int64_t nao, const TA* ao, const double* w, TZ* zw, int xc_id, double thresh,
                       double* exc, double* rho_out, hipStream_t st) {
  constexpr int PER = 16 / sizeof(TA);
  const int64_t j = (nao / PER + 63) / 64;
  const dim3 grid((unsigned)((ng + 3) / 4)), blk(256);
#define DFKS_XCW(JJ)                                                                                               \
  hipLaunchKernelGGL((k_xcw<TA, TZ, JJ>), grid, blk, 0, st, ng, nao, ao, w, zw, xc_id, thresh, exc, rho_out)
  if (j <= 1) DFKS_XCW(1);
  else if (j <= 2) DFKS_XCW(2);
  else if (j <= 3) DFKS_XCW(3);
  else if (j <= 4) DFKS_XCW(4);
  else return false;
#undef DFKS_XCW
  return true;
}

template <typename TA, typename TZ>
static bool launch_xc1(int64_t ng, int64_t nao, const TA* ao, const double* w, TZ* zw, int xc_id, double thresh,
                       double* exc, hipStream_t st) {
  constexpr int PER = 16 / sizeof(TA);
  const int64_t j = (nao / PER + 255) / 256;
  const dim3 grid((unsigned)ng), blk(256);
  if (j <= 1) hipLaunchKernelGGL((k_xc1<TA, TZ, 1>), grid, blk, 0, st, ng, nao, ao, w, zw, xc_id, thresh, exc);
  else if (j <= 2) hipLaunchKernelGGL((k_xc1<TA, TZ, 2>), grid, blk, 0, st, ng, nao, ao, w, zw, xc_id, thresh, exc);
  else if (j <= 3) hipLaunchKernelGGL((k_xc1<TA, TZ, 3>), grid, blk, 0, st, ng, nao, ao, w, zw, xc_id, thresh, exc);
  else if (j <= 4) hipLaunchKernelGGL((k_xc1<TA, TZ, 4>), grid, blk, 0, st, ng, nao, ao, w, zw, xc_id, thresh, exc);
  else if (j <= 6) hipLaunchKernelGGL((k_xc1<TA, TZ, 6>), grid, blk, 0, st, ng, nao, ao, w, zw, xc_id, thresh, exc);
  else return false;
  return true;
}

hipError_t launch_xc(int64_t ng, int64_t nao, const double* ao, const double* w, double* zw, int xc_id, double thresh,
                     double* exc, double* rho_out, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  const bool vec = nao % 2 == 0 && (uintptr_t)ao % 16 == 0 && (uintptr_t)zw % 16 == 0;
  // every AO value read once: one block per point on long rows, one wavefront per
  // point with the row slices in registers on short rows (the compact grid's blocks)
  if (vec && !rho_out && nao >= 768 && ng < (1LL << 31) && launch_xc1(ng, nao, ao, w, zw, xc_id, thresh, exc, st))
    return hipGetLastError();
  if (vec && nao < 768 && launch_xcw(ng, nao, ao, w, zw, xc_id, thresh, exc, rho_out, st)) return hipGetLastError();
  if (vec)
    hipLaunchKernelGGL((k_xc<double, double, true>), dim3((unsigned)((ng + 3) / 4)), dim3(256), 0, st, ng, nao, ao, w,
                       zw, xc_id, thresh, exc, rho_out);
  else
    hipLaunchKernelGGL((k_xc<double, double, false>), dim3((unsigned)((ng + 3) / 4)), dim3(256), 0, st, ng, nao, ao,
                       w, zw, xc_id, thresh, exc, rho_out);
  return hipGetLastError();
}

hipError_t launch_xc_f32(int64_t ng, int64_t nao, const float* ao, const double* w, float* zw, int xc_id,
                         double thresh, double* exc, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  const bool vec = nao % 4 == 0 && (uintptr_t)ao % 16 == 0 && (uintptr_t)zw % 16 == 0;
  if (vec && launch_xcw(ng, nao, ao, w, zw, xc_id, thresh, exc, (double*)nullptr, st)) return hipGetLastError();
  if (vec)
    hipLaunchKernelGGL((k_xc<float, float, true>), dim3((unsigned)((ng + 3) / 4)), dim3(256), 0, st, ng, nao, ao, w,
                       zw, xc_id, thresh, exc, (double*)nullptr);
  else
    hipLaunchKernelGGL((k_xc<float, float, false>), dim3((unsigned)((ng + 3) / 4)), dim3(256), 0, st, ng, nao, ao, w,
                       zw, xc_id, thresh, exc, (double*)nullptr);
  return hipGetLastError();
}

// Pointwise functional only (for parity tests of f, f_rho, f_sigma).
__global__ void k_xc_pointwise(int64_t n, int xc_id, const double* __restrict__ rho, const double* __restrict__ sigma,
                               double* __restrict__ out /*[n][3]*/) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  D r{rho[i], 1.0, 0.0};
  D s{sigma[i], 0.0, 1.0};
  D f = eps_xc(xc_id, r, s) * r;
  out[3 * i] = f.v;
  out[3 * i + 1] = f.r;
  out[3 * i + 2] = f.s;
}

hipError_t launch_xc_pointwise(int64_t n, int xc_id, const double* rho, const double* sigma, double* out,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_xc_pointwise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, xc_id, rho, sigma, out);
  return hipGetLastError();
}

// f and its first and second partials at (rho, sigma) from a dual of duals:
// o = {f, f_r, f_s, f_rr, f_rs, f_ss}. The outer dual differentiates the inner
// one, so f.r = (f_r, f_rr, f_rs) and f.s = (f_s, f_sr, f_ss).
__device__ __forceinline__ void xc_second(int xc_id, double r, double s, double* o) {
  const D2 rho{D{r, 1.0, 0.0}, D{1.0, 0.0, 0.0}, D{0.0, 0.0, 0.0}};
  const D2 sig{D{s, 0.0, 1.0}, D{0.0, 0.0, 0.0}, D{1.0, 0.0, 0.0}};
  const D2 f = eps_xc(xc_id, rho, sig) * rho;
  o[0] = f.v.v; o[1] = f.v.r; o[2] = f.v.s; o[3] = f.r.r; o[4] = f.r.s; o[5] = f.s.s;
}

// Pointwise second derivatives (parity tests): out[6 i ..] = xc_second.
__global__ __launch_bounds__(128) void k_xc_pointwise2(int64_t n, int xc_id, const double* __restrict__ rho,
                                const double* __restrict__ sigma, double* __restrict__ out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double o[6];
  xc_second(xc_id, rho[i], sigma[i], o);
#pragma unroll
  for (int k = 0; k < 6; ++k) out[6 * i + k] = o[k];
}

hipError_t launch_xc_pointwise2(int64_t n, int xc_id, const double* rho, const double* sigma, double* out,
                                hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_xc_pointwise2, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, n, xc_id, rho, sigma, out);
  return hipGetLastError();
}

// XC response at a reference density P0 (the jvp of Vxc along dP):
//   rho1 = sum_mn phi_m dP_mn phi_n, grad rho1 = 2 sum phi_m dP_mn grad phi_n,
//   sigma1 = 2 grad rho0 . grad rho1,
//   W1_gn = w [1/2 (f_rr rho1 + f_rs sigma1) phi_gn
//              + 2 ((f_sr rho1 + f_ss sigma1) grad rho0 + f_s grad rho1) . grad phi_gn],
// V1 = Phi^T W1 + W1^T Phi: the derivative of the Vxc of dftax/ks/terms.py:791-828
// along dP, i.e. the XC part of jax.jvp(jax.grad(E)) in dftax/ks/newton.py:134-136.
//
// Per-point record at P0 (8 doubles): w {f_rr, f_rs, f_ss, f_s}, grad rho0, 0;
// all zero where rho0 <= thresh (the masked points carry no derivative).
__global__ __launch_bounds__(128) void k_fxc_coef(int64_t ng, const double* __restrict__ rho0, const double* __restrict__ wts, int xc_id,
                           double thresh, double* __restrict__ rec) {
  int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const double r = rho0[4 * g], gx = rho0[4 * g + 1], gy = rho0[4 * g + 2], gz = rho0[4 * g + 3];
  double c[8] = {};
  if (r > thresh) {
    double o[6];
    xc_second(xc_id, r, gx * gx + gy * gy + gz * gz, o);
    const double w = wts[g];
    c[0] = w * o[3]; c[1] = w * o[4]; c[2] = w * o[5]; c[3] = w * o[2];
    c[4] = gx; c[5] = gy; c[6] = gz;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) rec[8 * g + k] = c[k];
}

hipError_t launch_fxc_coef(int64_t ng, const double* rho0, const double* w, int xc_id, double thresh, double* rec,
                           hipStream_t st) {
  if (ng == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fxc_coef, dim3((unsigned)((ng + 127) / 128)), dim3(128), 0, st, ng, rho0, w, xc_id, thresh,
                     rec);
  return hipGetLastError();
}

// One wavefront per grid point: zw holds Z1 = Phi dP on entry and W1 on exit.
template <bool VEC>
__global__ __launch_bounds__(256) void k_fxc(int64_t ng, int64_t nao, const double* __restrict__ ao,
                                             double* __restrict__ zw, const double* __restrict__ rec) {
  const int lane = threadIdx.x & 63;
  const int64_t g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;
  const int64_t plane = ng * nao;
  const double* p0 = ao + g * nao;
  const double* p1 = p0 + plane;
  const double* p2 = p0 + 2 * plane;
  const double* p3 = p0 + 3 * plane;
  double* zrow = zw + g * nao;
  typedef double vec2 __attribute__((ext_vector_type(2)));
  double r = 0, gx = 0, gy = 0, gz = 0;
  if (VEC) {
    for (int64_t n = lane; n < nao / 2; n += 64) {
      const vec2 z = reinterpret_cast<const vec2*>(zrow)[n];
      const vec2 a0 = reinterpret_cast<const vec2*>(p0)[n], a1 = reinterpret_cast<const vec2*>(p1)[n];
      const vec2 a2 = reinterpret_cast<const vec2*>(p2)[n], a3 = reinterpret_cast<const vec2*>(p3)[n];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        r += z[j] * a0[j]; gx += z[j] * a1[j]; gy += z[j] * a2[j]; gz += z[j] * a3[j];
      }
    }
  } else {
    for (int64_t n = lane; n < nao; n += 64) {
      const double z = zrow[n];
      r += z * p0[n]; gx += z * p1[n]; gy += z * p2[n]; gz += z * p3[n];
    }
  }
  r = wave_sum(r);
  gx = 2.0 * wave_sum(gx);
  gy = 2.0 * wave_sum(gy);
  gz = 2.0 * wave_sum(gz);
  const double* c = rec + 8 * g;
  const double gx0 = c[4], gy0 = c[5], gz0 = c[6];
  const double s1 = 2.0 * (gx0 * gx + gy0 * gy + gz0 * gz);
  const double a = 0.5 * (c[0] * r + c[1] * s1);
  const double t = 2.0 * (c[1] * r + c[2] * s1);
  const double bx = t * gx0 + 2.0 * c[3] * gx, by = t * gy0 + 2.0 * c[3] * gy, bz = t * gz0 + 2.0 * c[3] * gz;
  if (VEC) {
    for (int64_t n = lane; n < nao / 2; n += 64) {
      const vec2 a0 = reinterpret_cast<const vec2*>(p0)[n], a1 = reinterpret_cast<const vec2*>(p1)[n];
      const vec2 a2 = reinterpret_cast<const vec2*>(p2)[n], a3 = reinterpret_cast<const vec2*>(p3)[n];
      vec2 o;
#pragma unroll
      for (int j = 0; j < 2; ++j) o[j] = a * a0[j] + bx * a1[j] + by * a2[j] + bz * a3[j];
      reinterpret_cast<vec2*>(zrow)[n] = o;
    }
  } else {
    for (int64_t n = lane; n < nao; n += 64) zrow[n] = a * p0[n] + bx * p1[n] + by * p2[n] + bz * p3[n];
  }
}

hipError_t launch_fxc(int64_t ng, int64_t nao, const double* ao, double* zw, const double* rec, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  const bool vec = nao % 2 == 0 && (uintptr_t)ao % 16 == 0 && (uintptr_t)zw % 16 == 0;
  const dim3 grid((unsigned)((ng + 3) / 4)), blk(256);
  if (vec) hipLaunchKernelGGL(k_fxc<true>, grid, blk, 0, st, ng, nao, ao, zw, rec);
  else hipLaunchKernelGGL(k_fxc<false>, grid, blk, 0, st, ng, nao, ao, zw, rec);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused compact-grid XC pass, mixed precision (split-BF16 MFMA, FP32
// accumulate; reductions and the functional in FP64).
//
// A bin holds nb blocks of kXfG = 128 grid points; block b touches the n
// functions cols[b][0..n) (n a multiple of 32, -1 marks padding) and its AO
// planes are ao[k][b 128 + g][i] (k = value, x, y, z; plane = nb 128 n floats).
//   k_xcf_rho: one workgroup per block. Z = Phi_b Pg_b in 128 x 128 tiles,
//     Pg_b[c][i] = P[cols[c]][cols[i]] gathered while staging; each Z tile is
//     reduced at once to the block's rho and grad rho (LDS), so Z is never
//     stored; then the functional per point, writing the W coefficients
//     {w f_rho / 2, 2 w f_sigma grad rho} and w f.
//   k_xcf_v: one workgroup per (block, 128 x 128 tile of the n x n block):
//     Vh_b[i][j] = sum_g Phi[g][i] W[g][j], W[g][j] = a_g phi_gj + b_g . grad phi_gj
//     formed while staging, written as the row assembly reads it.
// Operands are split into bf16 hi + lo when staged to LDS; every k-step is
// three v_mfma_f32_16x16x32_bf16 products (hi*hi + hi*lo + lo*hi).
// ---------------------------------------------------------------------------
typedef float xflt4 __attribute__((ext_vector_type(4)));
typedef __bf16 xbf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 xbf16x4 __attribute__((ext_vector_type(4)));

constexpr int kXfG = 128;      // points per block
constexpr int kXfPad = 40;     // bf16 per LDS row: 32 k values + 8 (16-byte reads stay conflict-free)

// Split 8 floats into bf16 hi and lo parts and store them as 16-byte vectors.
__device__ __forceinline__ void xf_split8(const float* x, __bf16* hi, __bf16* lo) {
  xbf16x8 h, l;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 xh = (__bf16)x[j];
    h[j] = xh;
    l[j] = (__bf16)(x[j] - (float)xh);
  }
  *reinterpret_cast<xbf16x8*>(hi) = h;
  *reinterpret_cast<xbf16x8*>(lo) = l;
}

// One k-step of the 2 x 2 wavefront arrangement over a 128 x 128 tile: wave
// (wm, wn) owns 64 x 64 = 4 x 4 MFMA blocks; A rows and B rows are LDS rows.
__device__ __forceinline__ void xf_mfma_step(const __bf16 (*sAh)[kXfPad], const __bf16 (*sAl)[kXfPad],
                                             const __bf16 (*sBh)[kXfPad], const __bf16 (*sBl)[kXfPad], int wm,
                                             int wn, int li, int lk, xflt4 (&acc)[4][4]) {
  // the B fragments stay in registers for the step; the A fragments are read one row block
  // at a time (40 operand registers live instead of 64)
  xbf16x8 bh[4], bl[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    bh[j] = *reinterpret_cast<const xbf16x8*>(&sBh[wn + 16 * j + li][8 * lk]);
    bl[j] = *reinterpret_cast<const xbf16x8*>(&sBl[wn + 16 * j + li][8 * lk]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const xbf16x8 ah = *reinterpret_cast<const xbf16x8*>(&sAh[wm + 16 * i + li][8 * lk]);
    const xbf16x8 al = *reinterpret_cast<const xbf16x8*>(&sAl[wm + 16 * i + li][8 * lk]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
      acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
      acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
    }
  }
}

// Z tile rows staged through LDS for the rho reduction: 64 rows of 128 floats
// (+4 so consecutive rows start in different banks), reusing the operand tiles' LDS.
constexpr int kXfZPad = 132;
static_assert(64 * kXfZPad * 4 <= 4 * 128 * kXfPad * 2, "Z half tile fits the staging LDS");

// Partial rho sums of one (block, 128-column Z tile, K chunk): rho = sum_c phi_gc Z_gc
// and Z = Phi Pg are linear in both, so the tiles and K chunks of a block run as
// separate workgroups and their sums add up (k_xcf_pt, in a fixed order).
// part[((b nt + t) nks + ks)][4][128]: sum_c phi Z, sum_c d_x phi Z, d_y, d_z.
__global__ __launch_bounds__(256, 3) void k_xcf_z(const XcfBin* __restrict__ tab, int nbins, int ztot, int kchunk,
                                                  const float* __restrict__ ao_base,
                                                  const int32_t* __restrict__ cols_base, const float* __restrict__ P,
                                                  int64_t nao, double* __restrict__ part_base) {
  // staging tiles (A: points x 32 functions, B: 128 Z columns x 32 functions, bf16 hi/lo),
  // reused after the K loop for half of the Z tile in FP32
  __shared__ __attribute__((aligned(16))) __bf16 s_stage[4][128][kXfPad];
  __shared__ int32_t s_col[kXfMaxN];
  __shared__ double s_r[4][kXfG];
  auto sAh = s_stage[0];
  auto sAl = s_stage[1];
  auto sBh = s_stage[2];
  auto sBl = s_stage[3];
  float (*sZ)[kXfZPad] = reinterpret_cast<float (*)[kXfZPad]>(&s_stage[0][0][0]);
  // workgroup x runs on XCD x % 8: each XCD takes a contiguous range of ids, so the
  // workgroups of one block share its AO rows and P's rows through one L2
  const int per = (ztot + 7) / 8;
  const int gid = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (gid >= ztot) return;
  int bi = 0;
  while (bi + 1 < nbins && tab[bi + 1].wz0 <= gid) ++bi;
  const int n = tab[bi].n, nt = tab[bi].nt, nks = tab[bi].nks;
  const int64_t plane = (int64_t)tab[bi].nb * kXfG * n;
  const float* ao = ao_base + tab[bi].ao_off;
  const int32_t* cols = cols_base + tab[bi].col_off;
  double* part = part_base + tab[bi].part_off;
  const int id = gid - tab[bi].wz0, per_blk = nt * nks;
  const int64_t blk = id / per_blk;
  const int t = (id % per_blk) / nks, ks = id % nks;
  const int c0 = 128 * t, kb = ks * kchunk, ke = min(n, kb + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t row0 = blk * kXfG;        // first AO row of the block
  for (int i = tid; i < n; i += 256) s_col[i] = cols[blk * n + i];
  for (int i = tid; i < 4 * kXfG; i += 256) (&s_r[0][0])[i] = 0.0;
  __syncthreads();
  // staging: row sr of the A tile (point), 16 consecutive k from sk (two threads per row);
  // row br of the B tile (Z column), 16 k from bk: a wavefront's lanes take 64
  // consecutive columns of one row of the symmetric P, Pg[c][k] = P[cols[k]][cols[c]]
  const int sr = tid >> 1, sk = (tid & 1) * 16;
  const int br = tid & 127, bk = (tid >> 7) * 16;
  const float* arow = ao + (row0 + sr) * n;
  // the block's rows of the four planes (uniform), read with 32-bit offsets
  const float* q0 = ao + row0 * n;
  const float* q1 = q0 + plane;
  const float* q2 = q0 + 2 * plane;
  const float* q3 = q0 + 3 * plane;
  xflt4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = xflt4{0.f, 0.f, 0.f, 0.f};
  const int32_t pc = c0 + br < n ? s_col[c0 + br] : -1;
  const int nao32 = (int)nao;
  float a[16], b[16];
  auto load = [&](int k0) {
    const float4* src = reinterpret_cast<const float4*>(arow + k0 + sk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float4 x = src[v];
      a[4 * v] = x.x; a[4 * v + 1] = x.y; a[4 * v + 2] = x.z; a[4 * v + 3] = x.w;
    }
    // 32-bit offsets from P (nao^2 < 2^31, checked at launch), clamped to element 0 for
    // padding so that every lane loads
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int32_t kc = s_col[k0 + bk + j];
      const bool ok = pc >= 0 && kc >= 0;
      const float v = P[ok ? (uint32_t)(kc * nao32 + pc) : 0u];
      b[j] = ok ? v : 0.f;
    }
  };
  load(kb);
  for (int k0 = kb; k0 < ke; k0 += 32) {
    __syncthreads();
    xf_split8(a, &sAh[sr][sk], &sAl[sr][sk]);
    xf_split8(a + 8, &sAh[sr][sk + 8], &sAl[sr][sk + 8]);
    xf_split8(b, &sBh[br][bk], &sBl[br][bk]);
    xf_split8(b + 8, &sBh[br][bk + 8], &sBl[br][bk + 8]);
    __syncthreads();
    if (k0 + 32 < ke) load(k0 + 32);     // next k-step's operands while this one computes
    xf_mfma_step(sAh, sAl, sBh, sBl, wm, wn, li, lk, acc);
  }
  // acc[i][j][r] = Z(point wm + 16 i + 4 lk + r, column c0 + wn + 16 j + li). Each half of
  // the tile's points goes to LDS, then four threads per point reduce it against the four
  // AO planes (float4 reads along the row), one writer per point, in a fixed order.
#pragma unroll 1
  for (int hf = 0; hf < 2; ++hf) {
    __syncthreads();
    if (wm == 64 * hf) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) sZ[16 * i + 4 * lk + r][wn + 16 * j + li] = acc[i][j][r];
    }
    __syncthreads();
    const int g = 64 * hf + (tid >> 2), pq = tid & 3;
    const int cc = c0 + 32 * pq;               // 32 columns: all in range or all out (n % 32 == 0)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (cc < n) {
      const int o = g * n + cc;
      const float4* z4 = reinterpret_cast<const float4*>(&sZ[tid >> 2][32 * pq]);
      const float4* v0 = reinterpret_cast<const float4*>(q0 + o);
      const float4* v1 = reinterpret_cast<const float4*>(q1 + o);
      const float4* v2 = reinterpret_cast<const float4*>(q2 + o);
      const float4* v3 = reinterpret_cast<const float4*>(q3 + o);
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const float4 z = z4[v], x0 = v0[v], x1 = v1[v], x2 = v2[v], x3 = v3[v];
        s0 += (double)z.x * x0.x + (double)z.y * x0.y + (double)z.z * x0.z + (double)z.w * x0.w;
        s1 += (double)z.x * x1.x + (double)z.y * x1.y + (double)z.z * x1.z + (double)z.w * x1.w;
        s2 += (double)z.x * x2.x + (double)z.y * x2.y + (double)z.z * x2.z + (double)z.w * x2.w;
        s3 += (double)z.x * x3.x + (double)z.y * x3.y + (double)z.z * x3.z + (double)z.w * x3.w;
      }
    }
#pragma unroll
    for (int off = 1; off < 4; off <<= 1) {
      s0 += __shfl_xor(s0, off, 64);
      s1 += __shfl_xor(s1, off, 64);
      s2 += __shfl_xor(s2, off, 64);
      s3 += __shfl_xor(s3, off, 64);
    }
    if (pq == 0) {
      s_r[0][g] = s0;
      s_r[1][g] = s1;
      s_r[2][g] = s2;
      s_r[3][g] = s3;
    }
  }
  __syncthreads();
  double* out = part + (int64_t)id * 4 * kXfG;
  for (int i = tid; i < 4 * kXfG; i += 256) out[i] = (&s_r[0][0])[i];
}

// Per point: the partial sums of its block in (tile, K chunk) order, then the
// functional: exc = w f, coef = {w f_rho / 2, 2 w f_sigma grad rho}.
__global__ __launch_bounds__(128) void k_xcf_pt(const XcfBin* __restrict__ tab, int nbins, int ptot,
                                                const double* __restrict__ part_base,
                                                const double* __restrict__ w_base, int xc_id, double thresh,
                                                double* __restrict__ exc_base, float* __restrict__ coef_base) {
  if (blockIdx.x >= (unsigned)ptot) return;       // one workgroup per block
  int bi = 0;
  while (bi + 1 < nbins && tab[bi + 1].wp0 <= (int)blockIdx.x) ++bi;
  const int lb = (int)blockIdx.x - tab[bi].wp0, nparts = tab[bi].nt * tab[bi].nks;
  const int64_t gg = lb * (int64_t)kXfG + threadIdx.x;
  const double* wts = w_base + tab[bi].pt0;
  double* exc = exc_base + tab[bi].pt0;
  float* coef = coef_base + tab[bi].coef_off;
  const double* src = part_base + tab[bi].part_off + (int64_t)lb * nparts * 4 * kXfG + threadIdx.x;
  double r = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  for (int pp = 0; pp < nparts; ++pp) {
    const double* s = src + (int64_t)pp * 4 * kXfG;
    r += s[0];
    gx += s[kXfG];
    gy += s[2 * kXfG];
    gz += s[3 * kXfG];
  }
  gx *= 2.0; gy *= 2.0; gz *= 2.0;
  double a = 0.0, bx = 0.0, by = 0.0, bz = 0.0, e = 0.0;
  if (r > thresh) {
    const double w = wts[gg];
    D rho{r, 1.0, 0.0};
    D sig{gx * gx + gy * gy + gz * gz, 0.0, 1.0};
    D f = eps_xc(xc_id, rho, sig) * rho;
    e = w * f.v;
    a = 0.5 * w * f.r;
    const double tt = 2.0 * w * f.s;
    bx = tt * gx; by = tt * gy; bz = tt * gz;
  }
  exc[gg] = e;
  reinterpret_cast<float4*>(coef)[gg] = make_float4((float)a, (float)bx, (float)by, (float)bz);
}

__global__ __launch_bounds__(256, 2) void k_xcf_v(const XcfBin* __restrict__ tab, int nbins, int vtot,
                                                  const float* __restrict__ ao_base,
                                                  const float* __restrict__ coef_base, float* __restrict__ q_base) {
  __shared__ __attribute__((aligned(16))) __bf16 sAh[128][kXfPad];
  __shared__ __attribute__((aligned(16))) __bf16 sAl[128][kXfPad];
  __shared__ __attribute__((aligned(16))) __bf16 sBh[128][kXfPad];
  __shared__ __attribute__((aligned(16))) __bf16 sBl[128][kXfPad];
  // workgroup x runs on XCD x % 8: each XCD takes a contiguous range of (block, tile)
  // ids, so the tiles of one block read its AO rows through one L2
  const int per = (vtot + 7) / 8;
  const int gid = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (gid >= vtot) return;
  int bi = 0;
  while (bi + 1 < nbins && tab[bi + 1].wv0 <= gid) ++bi;
  const int n = tab[bi].n, nt = tab[bi].nt, per_blk = nt * nt;
  const int64_t plane = (int64_t)tab[bi].nb * kXfG * n;
  const float* ao = ao_base + tab[bi].ao_off;
  const float* coef = coef_base + tab[bi].coef_off;
  float* q = q_base + tab[bi].q_off;
  const int id = gid - tab[bi].wv0;
  const int64_t blk = id / per_blk;
  const int t = id % per_blk, m0 = (t / nt) * 128, n0 = (t % nt) * 128;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t row0 = blk * kXfG;
  // staging: 4 consecutive functions (one float4) x 4 points per thread; 32 threads
  // span the tile's 128 functions, 8 groups its 32 points
  const int cg = 4 * (tid & 31), pg = 4 * (tid >> 5);
  const bool a_in = m0 + cg < n, b_in = n0 + cg < n;   // 4-column groups lie wholly inside or outside (n % 32 == 0)
  const float* p0 = ao + row0 * n;
  const float* p1 = p0 + plane;
  const float* p2 = p0 + 2 * plane;
  const float* p3 = p0 + 3 * plane;
  const float4* cb = reinterpret_cast<const float4*>(coef) + row0;
  xflt4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = xflt4{0.f, 0.f, 0.f, 0.f};
  float4 a[4], b[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int g = k0 + pg + p;
      const int o = g * n;
      a[p] = a_in ? *reinterpret_cast<const float4*>(p0 + o + m0 + cg) : make_float4(0.f, 0.f, 0.f, 0.f);
      if (b_in) {
        const float4 cf = cb[g];
        const int oc = o + n0 + cg;
        const float4 x0 = *reinterpret_cast<const float4*>(p0 + oc);
        const float4 x1 = *reinterpret_cast<const float4*>(p1 + oc);
        const float4 x2 = *reinterpret_cast<const float4*>(p2 + oc);
        const float4 x3 = *reinterpret_cast<const float4*>(p3 + oc);
        b[p] = make_float4(cf.x * x0.x + cf.y * x1.x + cf.z * x2.x + cf.w * x3.x,
                           cf.x * x0.y + cf.y * x1.y + cf.z * x2.y + cf.w * x3.y,
                           cf.x * x0.z + cf.y * x1.z + cf.z * x2.z + cf.w * x3.z,
                           cf.x * x0.w + cf.y * x1.w + cf.z * x2.w + cf.w * x3.w);
      } else {
        b[p] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  };
  // LDS row = function, 4 points (k) per 8-byte store, hi and lo
  auto store = [&](const float4 (&x)[4], __bf16 (*hi)[kXfPad], __bf16 (*lo)[kXfPad]) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      xbf16x4 h, l;
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const float v = e == 0 ? x[p].x : e == 1 ? x[p].y : e == 2 ? x[p].z : x[p].w;
        const __bf16 vh = (__bf16)v;
        h[p] = vh;
        l[p] = (__bf16)(v - (float)vh);
      }
      *reinterpret_cast<xbf16x4*>(&hi[cg + e][pg]) = h;
      *reinterpret_cast<xbf16x4*>(&lo[cg + e][pg]) = l;
    }
  };
  load(0);
  for (int k0 = 0; k0 < kXfG; k0 += 32) {
    __syncthreads();
    store(a, sAh, sAl);
    store(b, sBh, sBl);
    __syncthreads();
    if (k0 + 32 < kXfG) load(k0 + 32);
    xf_mfma_step(sAh, sAl, sBh, sBl, wm, wn, li, lk, acc);
  }
  float* qb = q + blk * (int64_t)n * n;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * i + 4 * lk + r;
        const int gn = n0 + wn + 16 * j + li;
        if (gm < n && gn < n) qb[(int64_t)gm * n + gn] = acc[i][j][r];
      }
}

// The fused pass over every bin of a table in one launch per kernel: bins of
// width n <= kXfMaxN (a multiple of 32); tab (device) as XcfBin above, with ztot,
// ptot, vtot the three kernels' workgroup counts; part the partial sums, coef the W
// coefficients (4 floats per point), q the block products (floats).
constexpr int kXfKChunk = 384;   // functions of K per k_xcf_z workgroup

hipError_t launch_xcf_all(const XcfBin* tab, int nbins, int ztot, int ptot, int vtot, const float* ao,
                          const int32_t* cols, const float* P, int64_t nao, const double* w, int xc_id, double thresh,
                          double* exc, double* part, float* coef, float* q, hipStream_t st) {
  if (nbins == 0) return hipSuccess;
  if ((uintptr_t)ao % 16 != 0 || (uintptr_t)coef % 16 != 0) return hipErrorInvalidValue;
  if (nao * nao > 0x7fffffff) return hipErrorInvalidValue;   // k_xcf_z addresses P with 32-bit offsets
  // 8 ceil(total / 8) workgroups: the XCD mappings produce every id < total only then
  hipLaunchKernelGGL(k_xcf_z, dim3((unsigned)(8 * (((int64_t)ztot + 7) / 8))), dim3(256), 0, st, tab, nbins, ztot,
                     kXfKChunk, ao, cols, P, nao, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_xcf_pt, dim3((unsigned)ptot), dim3(kXfG), 0, st, tab, nbins, ptot, (const double*)part, w,
                     xc_id, thresh, exc, coef);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_xcf_v, dim3((unsigned)(8 * (((int64_t)vtot + 7) / 8))), dim3(256), 0, st, tab, nbins, vtot,
                     ao, (const float*)coef, q);
  return hipGetLastError();
}

// Workgroup counts and workspace of a bin in the table: K chunks per Z tile, then
// the partial sums (doubles) and W coefficients (floats) it needs.
void xcf_bin_geometry(int n, int nb, int& nt, int& nks, int64_t& part_doubles, int64_t& coef_floats) {
  nt = (n + 127) / 128;
  nks = (n + kXfKChunk - 1) / kXfKChunk;
  part_doubles = (int64_t)nb * nt * nks * 4 * kXfG;
  coef_floats = (int64_t)nb * kXfG * 4;
}

// Spin-polarised pointwise pass, one wavefront per grid point: on entry zwa, zwb
// hold Z_s = Phi P_s (ng x nao); on exit the W_s of the half potential,
//   W_a = w (f_ra / 2 phi + (2 f_saa grad ra + f_sab grad rb) . grad phi),
//   W_b = w (f_rb / 2 phi + (2 f_sbb grad rb + f_sab grad ra) . grad phi),
// with f = eps_xc_pol * (ra + rb) masked as terms.py:838-874 (points with
// ra + rb <= thresh count nothing; a channel at or below thresh enters the
// functional as thresh with zero gradient).
__global__ __launch_bounds__(256) void k_xcs(int64_t ng, int64_t nao, const double* __restrict__ ao,
                                             const double* __restrict__ wts, double* __restrict__ zwa,
                                             double* __restrict__ zwb, int xc_id, double thresh,
                                             double* __restrict__ exc) {
  const int lane = threadIdx.x & 63;
  const int64_t g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;
  const int64_t plane = ng * nao;
  const double* p0 = ao + g * nao;
  const double* p1 = p0 + plane;
  const double* p2 = p0 + 2 * plane;
  const double* p3 = p0 + 3 * plane;
  double* za = zwa + g * nao;
  double* zb = zwb + g * nao;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t n = lane; n < nao; n += 64) {
    const double a = za[n], b = zb[n], v0 = p0[n], v1 = p1[n], v2 = p2[n], v3 = p3[n];
    s[0] += a * v0; s[1] += a * v1; s[2] += a * v2; s[3] += a * v3;
    s[4] += b * v0; s[5] += b * v1; s[6] += b * v2; s[7] += b * v3;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = wave_sum(s[k]);
  const double ra = s[0], rb = s[4];
  const double ga[3] = {2.0 * s[1], 2.0 * s[2], 2.0 * s[3]}, gb[3] = {2.0 * s[5], 2.0 * s[6], 2.0 * s[7]};
  double aa = 0.0, ab = 0.0, ba[3] = {0, 0, 0}, bb[3] = {0, 0, 0}, e = 0.0;
  if (ra + rb > thresh) {
    const bool ta = ra > thresh, tb = rb > thresh;
    D5 Ra{}, Rb{}, Saa{}, Sab{}, Sbb{}, Rt{};
    Ra.v = ta ? ra : 1e-10; Ra.d[0] = ta ? 1.0 : 0.0;
    Rb.v = tb ? rb : 1e-10; Rb.d[1] = tb ? 1.0 : 0.0;
    Saa.v = ta ? ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2] : 0.0; Saa.d[2] = ta ? 1.0 : 0.0;
    Sab.v = ta && tb ? ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2] : 0.0; Sab.d[3] = ta && tb ? 1.0 : 0.0;
    Sbb.v = tb ? gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2] : 0.0; Sbb.d[4] = tb ? 1.0 : 0.0;
    Rt.v = ra + rb; Rt.d[0] = 1.0; Rt.d[1] = 1.0;      // the unclamped total density
    const D5 f = eps_xc_pol(xc_id, Ra, Rb, Saa, Sab, Sbb) * Rt;
    const double w = wts[g];
    e = w * f.v;
    aa = 0.5 * w * f.d[0];
    ab = 0.5 * w * f.d[1];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      ba[x] = w * (2.0 * f.d[2] * ga[x] + f.d[3] * gb[x]);
      bb[x] = w * (2.0 * f.d[4] * gb[x] + f.d[3] * ga[x]);
    }
  }
  if (lane == 0) exc[g] = e;
  for (int64_t n = lane; n < nao; n += 64) {
    const double v0 = p0[n], v1 = p1[n], v2 = p2[n], v3 = p3[n];
    za[n] = aa * v0 + ba[0] * v1 + ba[1] * v2 + ba[2] * v3;
    zb[n] = ab * v0 + bb[0] * v1 + bb[1] * v2 + bb[2] * v3;
  }
}

hipError_t launch_xcs(int64_t ng, int64_t nao, const double* ao, const double* w, double* zwa, double* zwb,
                      int xc_id, double thresh, double* exc, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  hipLaunchKernelGGL(k_xcs, dim3((unsigned)((ng + 3) / 4)), dim3(256), 0, st, ng, nao, ao, w, zwa, zwb, xc_id,
                     thresh, exc);
  return hipGetLastError();
}

// Pointwise spin-polarised functional (parity tests): out[6 i ..] = {f, f_ra, f_rb,
// f_saa, f_sab, f_sbb} of f = eps_xc_pol (ra, rb, saa, sab, sbb) * (ra + rb), no masking.
__global__ void k_xc_pointwise_pol(int64_t n, int xc_id, const double* __restrict__ in5,
                                   double* __restrict__ out6) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  D5 x[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    x[k] = D5{};
    x[k].v = in5[5 * i + k];
    x[k].d[k] = 1.0;
  }
  const D5 f = eps_xc_pol(xc_id, x[0], x[1], x[2], x[3], x[4]) * (x[0] + x[1]);
  out6[6 * i] = f.v;
#pragma unroll
  for (int k = 0; k < 5; ++k) out6[6 * i + 1 + k] = f.d[k];
}

hipError_t launch_xc_pointwise_pol(int64_t n, int xc_id, const double* in5, double* out6, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_xc_pointwise_pol, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, n, xc_id, in5, out6);
  return hipGetLastError();
}

}  // namespace dfks
