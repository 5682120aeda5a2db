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

enum XcId { kXcPBE = 1, kXcPBE0 = 2, kXcB3LYP = 3 };

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

template <class N>
__device__ N vwn_rpa_c(N rho) {
  const double A = 0.0310907, b = 13.0720, c = 42.7198, x0 = -0.409286;
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

template <class N>
__device__ N eps_xc(int id, N rho, N sig) {
  switch (id) {
    case kXcPBE: return pbe_x(rho, sig) + pbe_c(rho, sig);
    case kXcPBE0: return 0.75 * pbe_x(rho, sig) + pbe_c(rho, sig);
    default:
      return 0.08 * lda_x(rho) + 0.72 * b88_x(rho, sig) + 0.19 * vwn_rpa_c(rho) + 0.81 * lyp_c(rho, sig);
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

}  // namespace dfks
