// C ABI of the DF-KS library: handle state, build pipeline, energy + Fock, SCF.
//
// Data layout in HBM (all FP64, row-major unless noted):
//   B      [nao][nao][nfit]   fitted three-centre tensor, B = (mn|P) L, V+ = L L^T
//   ao     [4][ng][nao]       AO values and x/y/z gradients on the grid
//   zw     [ng][nao]          Z = Phi P, overwritten in place by W (XC kernel)
//   T      [nao][nfit][nf]    exchange intermediate T[m,Q,i] = sum_l B[l,m,Q] Cf[l,i]
//   S, H, P, F, J, K, V       nao x nao (symmetric)
//   X, C   column-major       nao x nmo orthonormaliser / MO coefficients
// Plain GEMM/SYRK/SYR2K shapes go to rocBLAS and the symmetric eigensolves to
// rocSOLVER; every other step is a kernel of this library.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

#include "dfks.h"
#include "dfks_common.h"

namespace dfks {
void upload_c2s_tables();
hipError_t launch_eri_wave(int la, int lb, int lc, const int2* pairs, int npair, const int32_t* aux, int naux_l,
                           const ShellSet& A, const ShellSet& C, int b_unit, const double* boys, double* out,
                           int64_t sa, int64_t sb, int tr, int64_t c0, hipStream_t stream);
hipError_t launch_schwarz(int la, int lb, const int2* pairs, int npair, const ShellSet& A, const double* boys,
                          double* q_out, hipStream_t st);
hipError_t launch_jvec_acc(int64_t nrows, int nq, const double* B, const double* g, double* J, hipStream_t st);
bool eri_thread_has(int la, int lb, int lc);
hipError_t launch_eri_cls(int la, int lb, int lc, const int2* pairs, int npair, const int32_t* aux, int naux_l,
                          const ShellSet& A, const ShellSet& C, const double* boys, double* out, int64_t sa,
                          int64_t sb, int64_t c0, int tr, hipStream_t st);
hipError_t launch_eri_thread(int la, int lb, int lc, const int2* pairs, int npair, const int32_t* aux, int naux_l,
                             const ShellSet& A, const ShellSet& C, const double* boys, double* out, int64_t sa,
                             int64_t sb, int64_t c0, int tr, hipStream_t st);
hipError_t launch_int1e(const ShellSet& A, int natm, const double* atm, const double* boys, double* S, double* T,
                        double* V, int64_t nao, hipStream_t stream);
hipError_t launch_becke(int64_t n, const double* pts, const double* raw, const int32_t* owner, int natm,
                        const double* atm, const double* invR, const double* aAB, double* w, hipStream_t st);
hipError_t launch_ao(int64_t ng, const double* pts, const ShellSet& S, int64_t nao, double* ao, hipStream_t st);
hipError_t launch_ao_cmp(int nblk, int G, int n, const double* pts, const int2* shl, const int64_t* off,
                         const ShellSet& S, double* ao, hipStream_t st);
template <typename TP, typename T>
hipError_t launch_gather_sym(int64_t nblk, int n, const int32_t* cols, const TP* P, int64_t nao, T* out,
                             hipStream_t st);
template <typename T>
hipError_t launch_vassemble(int nao, const int64_t* rowptr, const int64_t* prow, const int64_t* pcol,
                            const int32_t* plen, const T* q, const int32_t* cols, double* V, double* ws, int nslice,
                            hipStream_t st);
hipError_t launch_xc(int64_t ng, int64_t nao, const double* ao, const double* w, double* zw, int xc_id, double thresh,
                     double* exc, double* rho_out, hipStream_t st);
hipError_t launch_xc_pointwise(int64_t n, int xc_id, const double* rho, const double* sigma, double* out,
                               hipStream_t st);
hipError_t launch_xcs(int64_t ng, int64_t nao, const double* ao, const double* w, double* zwa, double* zwb,
                      int xc_id, double thresh, double* exc, hipStream_t st);
hipError_t launch_xc_pointwise_pol(int64_t n, int xc_id, const double* in5, double* out6, hipStream_t st);
hipError_t launch_xc_pointwise2(int64_t n, int xc_id, const double* rho, const double* sigma, double* out,
                                hipStream_t st);
hipError_t launch_fxc_coef(int64_t ng, const double* rho0, const double* w, int xc_id, double thresh, double* rec,
                           hipStream_t st);
hipError_t launch_fxc(int64_t ng, int64_t nao, const double* ao, double* zw, const double* rec, hipStream_t st);
hipError_t launch_dot(int64_t n, const double* x, const double* y, double* part, double scale, double* out,
                      int accumulate, hipStream_t st);
hipError_t launch_gamma(int nao, int nq, int nf, const double* T, const double* Cf, const double* sgn, double* gamma,
                        double* part, int accumulate, hipStream_t st);
hipError_t launch_jvec(int64_t nrows, int nq, const double* B, const double* g, double* J, int sym_n, hipStream_t st);
hipError_t launch_fock(int n, const double* H, const double* J, const double* Ku, double cK, const double* Vu,
                       double* F, hipStream_t st);
hipError_t launch_symfill(int n, double* A, hipStream_t st);
hipError_t launch_metric_factor(int nq, int k0, const double* U, const double* w, double* Lt, hipStream_t st);
hipError_t launch_factor_cols(int n, int nf, const double* V, const double* lam, const int32_t* sel, double* Cf,
                              double* sgn, hipStream_t st);
hipError_t launch_axpy(int64_t n, double a, const double* x, double* y, int acc, hipStream_t st);
hipError_t launch_antisym(int n, const double* A, double* E, hipStream_t st);
hipError_t launch_eye_dev(int n, const double* G, double* out, hipStream_t st);
hipError_t launch_store2(double* dst, double a, double b, hipStream_t st);
hipError_t launch_diis_solve(int m, int count, const double* gram, double* coef, hipStream_t st);
hipError_t launch_diis_extrap(int m, int64_t n, const double* coef, const double* Fs, double* out, hipStream_t st);
hipError_t launch_gram_update(int m, int slot, int count, const double* dots, double* gram, hipStream_t st);
hipError_t launch_scaled_cols(int n, int k0, const double* U, const double* w, double* X, hipStream_t st);
hipError_t launch_xcf_all(const XcfBin* tab, int nbins, int ztot, int ptot, int vtot, const float* ao,
                          const int32_t* cols, const float* P, int64_t nao, const double* w, int xc_id, double thresh,
                          double* exc, double* part, float* coef, float* q, hipStream_t st);
void xcf_bin_geometry(int n, int nb, int& nt, int& nks, int64_t& part_doubles, int64_t& coef_floats);
hipError_t launch_xc_f32(int64_t ng, int64_t nao, const float* ao, const double* w, float* zw, int xc_id,
                         double thresh, double* exc, hipStream_t st);
hipError_t launch_gamma_f32(int nao, int nq, int nf, const float* T, const double* Cf, const double* sgn,
                            double* gamma, double* part, int accumulate, hipStream_t st);
hipError_t launch_jvec_f32(int64_t nrows, int nq, const float* B, const double* g, double* J, int sym_n,
                           hipStream_t st);
hipError_t launch_to_f32(int64_t n, const double* x, float* y, hipStream_t st);
hipError_t launch_to_f64(int64_t n, const float* x, double* y, hipStream_t st);
hipError_t launch_eri_grad(const int2* pairs, int npair, int nket, const ShellSet& A, const ShellSet& C,
                           const int32_t* atom_a, const int32_t* atom_c, int b_unit, const double* boys,
                           const double* G, int64_t nao_g, double* grad, hipStream_t st);
hipError_t launch_int1e_grad(const ShellSet& A, const int32_t* atom_a, int natm, const double* atm,
                             const double* boys, const double* P, const double* W, int64_t nao, double* grad,
                             hipStream_t st);
hipError_t launch_ao_hess(int64_t npts, const double* pts, const ShellSet& S, int64_t nao, double* ao,
                          hipStream_t st);
hipError_t launch_xcg_rho(int64_t npts, int64_t nao, const double* ao, const double* Z, double* rho, double* grad,
                          double* sigma, hipStream_t st);
hipError_t launch_xcg_gdot(int64_t npts, int64_t nao, const double* ao, const double* grad, double* Gd,
                           hipStream_t st);
hipError_t launch_xcg_q(int64_t npts, int64_t nao, const double* ao, const double* Z, const double* X,
                        const double* rho, const double* grad, const double* xc, const double* w,
                        const int32_t* owner, double thresh, double* q, double* gout, double* fm, hipStream_t st);
hipError_t launch_becke_grad(int64_t n, const double* pts, const double* raw, const int32_t* owner, const double* f,
                             int natm, const double* atm, const double* invR, const double* aAB, double* gout,
                             hipStream_t st);
enum GemmPrec { kF64 = 0, kBF16x3 = 1 };
struct GemmArgs {
  int M, N;
  int64_t K;
  const void* A;
  int64_t lda;
  bool a_kcontig;
  const void* B;
  int64_t ldb;
  bool b_kcontig;
};
hipError_t launch_gemm(int prec, const GemmArgs& g, double* C, double alpha, double beta, int nsplit, int sym,
                       void* ws, hipStream_t st);
hipError_t launch_gemm_f32out(const GemmArgs& g, float* C, hipStream_t st);
hipError_t launch_gemm_f32out_batched(const GemmArgs& g, float* C, int batch, int64_t bs_a, int64_t bs_b,
                                      int64_t bs_c, hipStream_t st);
hipError_t launch_gemm_f64_batched(const GemmArgs& g, double* C, int batch, int64_t bs_a, int64_t bs_b,
                                   int64_t bs_c, hipStream_t st);
hipError_t launch_gemm_w8(const GemmArgs& g, double* C, const RowDot& rd, hipStream_t st);
hipError_t launch_gemm_b3w8(const GemmArgs& g, float* C, const RowDot& rd, hipStream_t st, int depth);
hipError_t launch_colsum(int nrows, int ncols, const double* part, double* out, int accumulate, hipStream_t st);
hipError_t launch_gemm_partials(int prec, const GemmArgs& g, int nsplit, int sym, void* ws, int variant,
                                hipStream_t st);
hipError_t launch_gemm_reduce(int prec, const GemmArgs& g, double* C, double alpha, double beta, int nsplit, int sym,
                              const void* ws, hipStream_t st);
hipError_t launch_gersh(int n, const double* F, double* lohi, double* st, hipStream_t s);
hipError_t launch_sp2_lanczos(int n, const double* F, double* lz, double* st, hipStream_t s);
hipError_t launch_sp2_init(int n, const double* F, const double* st, double* X, hipStream_t s);
hipError_t launch_sp2_weyl(int n, const double* F, double* Fp, const double* pb, double* st, hipStream_t s);
hipError_t launch_save_bounds(const double* st, double* pb, hipStream_t s);
hipError_t launch_sp2_begin(int n, const double* X, int nocc, double* st, hipStream_t s);
hipError_t launch_sp2_update(int n, int nsplit, const double* ws, double* X, int nocc, double* st, hipStream_t s);
int sp2_step_tiles(int n);
hipError_t launch_sp2_step(int n, const double* X, double* Xn, int nocc, double* st, hipStream_t s);
hipError_t launch_diag(int n, const double* R, double* d, hipStream_t s);
hipError_t launch_pchol_panel(int n, int k0, int b, const double* R, double* d, double* L, int32_t* piv, double* st,
                              int last, hipStream_t s);
thread_local GraphRec* g_rec = nullptr;
}  // namespace dfks

using namespace dfks;

// Fock-build / SCF stages timed by dfks_stage_times (event i marks the end of stage i-1).
enum Stage {
  kStExchangeGemm = 0,  // T = B^T Cf
  kStExchangeSyrk = 1,  // gamma from T, and K = T^T T
  kStCoulomb = 2,       // E_J and J = B gamma
  kStXcGemm = 3,        // Z = Phi P
  kStXcKernel = 4,      // rho, functional, W
  kStXcSyr2k = 5,       // Vxc = Phi^T W + W^T Phi
  kStFock = 6,          // assembly and energies
  kStScfDiis = 7,       // commutator + DIIS
  kStScfEigh = 8,       // X^T F X, density (purification or eigensolve), factor, P
  kNumStages = 9
};

namespace {

thread_local std::string g_err;

struct Fail {
  int code;
};

void fail(int code, const std::string& msg) {
  g_err = msg;
  throw Fail{code};
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) fail(2, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)
#define RBCHK(expr)                                                                           \
  do {                                                                                        \
    rocblas_status s_ = (expr);                                                               \
    if (s_ != rocblas_status_success) fail(3, std::string(#expr) + ": rocblas status " +      \
                                              std::to_string((int)s_));                       \
  } while (0)

// Device memory is stream-ordered: every entry point that allocates or frees
// runs with its caller's stream as g_st, and allocations come from the
// device's default memory pool (hipMallocAsync / hipFreeAsync). Unlike
// hipFree, nothing here waits for the whole device, so handles of different
// molecules can be built and destroyed while others compute. The pool keeps
// released memory reserved for the next handle instead of returning it.
thread_local hipStream_t g_st = nullptr;

struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(g_st) { g_st = s; }
  ~StreamScope() { g_st = prev; }
};

std::mutex g_pool_mu;
std::vector<int> g_pool_ready;

void pool_setup() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (std::find(g_pool_ready.begin(), g_pool_ready.end(), dev) != g_pool_ready.end()) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  g_pool_ready.push_back(dev);
}

// Buffers of 8 MiB and more are kept by the library after they are freed. A
// fresh device allocation of a few GB can hold the calling thread for
// 0.1-1 s, and the device pool does not hand a freed block to a request of
// another size, so a molecule that follows one of a different size would pay
// that again for its (mn|P), B and grid buffers. Sizes are rounded up to eight
// classes per power of two (at most 12.5 % unused); a freed block keeps an
// event recorded on the freeing stream, and the stream that takes it next waits
// on that event before its first use. At most a third of the device's memory
// (and never more than 96 GB) is kept; any allocation that fails, large or
// small, returns every kept block of its device to the pool and tries again.
struct BigBlock {
  size_t bytes;
  int dev;
  hipEvent_t ev;
};
constexpr size_t kCacheMin = size_t(8) << 20;
constexpr size_t kCacheKeepMax = size_t(96) << 30;
std::mutex g_cache_mu;
std::map<void*, BigBlock> g_big_live;                // handed out
std::vector<std::pair<void*, BigBlock>> g_big_free;  // kept for reuse
std::map<int, size_t> g_big_free_bytes;              // kept bytes per device
std::map<int, size_t> g_cache_keep;                  // keep limit per device

size_t size_class(size_t b) {
  const int k = 63 - __builtin_clzll(b);
  const size_t step = size_t(1) << (k - 3);
  return (b + step - 1) / step * step;
}

// Keep limit of device dev (caller holds g_cache_mu; dev is current).
size_t cache_keep(int dev) {
  auto it = g_cache_keep.find(dev);
  if (it != g_cache_keep.end()) return it->second;
  size_t fr = 0, tot = 0;
  size_t keep = kCacheKeepMax;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess && tot > 0) keep = std::min(keep, tot / 3);
  g_cache_keep[dev] = keep;
  return keep;
}

// Returns every kept block of device dev to the pool. The blocks leave the
// list under the lock; the waits for their last users and the frees happen
// after it is dropped, so other threads' allocations do not queue behind a
// long kernel on another stream.
void cache_release(int dev) {
  std::vector<std::pair<void*, BigBlock>> out;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (size_t i = 0; i < g_big_free.size();) {
      if (g_big_free[i].second.dev != dev) { ++i; continue; }
      out.push_back(g_big_free[i]);
      g_big_free_bytes[dev] -= g_big_free[i].second.bytes;
      g_big_free[i] = g_big_free.back();
      g_big_free.pop_back();
    }
  }
  for (auto& f : out) {
    (void)hipEventSynchronize(f.second.ev);
    (void)hipEventDestroy(f.second.ev);
    (void)hipFreeAsync(f.first, g_st);
  }
}

// After a failed allocation on device dev: return the kept blocks and the
// pool's unused reserve, then let the caller try once more.
void release_for_retry(int dev) {
  (void)hipGetLastError();
  cache_release(dev);
  (void)hipStreamSynchronize(g_st);
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
}

void* big_alloc(size_t bytes) {
  const size_t cb = size_class(bytes);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (size_t i = 0; i < g_big_free.size(); ++i) {
      if (g_big_free[i].second.bytes != cb || g_big_free[i].second.dev != dev) continue;
      auto f = g_big_free[i];
      g_big_free[i] = g_big_free.back();
      g_big_free.pop_back();
      g_big_free_bytes[dev] -= cb;
      g_big_live[f.first] = f.second;
      HIPCHK(hipStreamWaitEvent(g_st, f.second.ev, 0));
      return f.first;
    }
  }
  // the lock is not held across the allocation, which can take a second
  void* p = nullptr;
  hipError_t e = hipMallocAsync(&p, cb, g_st);
  if (e != hipSuccess) {
    release_for_retry(dev);
    e = hipMallocAsync(&p, cb, g_st);
  }
  if (e != hipSuccess) fail(4, "hipMallocAsync of " + std::to_string(cb) + " bytes: " + hipGetErrorString(e));
  hipEvent_t ev = nullptr;
  e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipFreeAsync(p, g_st);
    fail(2, std::string("hipEventCreateWithFlags: ") + hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_big_live[p] = BigBlock{cb, dev, ev};
  return p;
}

// false when p is not a kept-size block
bool big_free(void* p) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = g_big_live.find(p);
  if (it == g_big_live.end()) return false;
  const BigBlock b = it->second;
  g_big_live.erase(it);
  size_t& kept = g_big_free_bytes[b.dev];
  if (kept + b.bytes <= cache_keep(b.dev) && hipEventRecord(b.ev, g_st) == hipSuccess) {
    g_big_free.emplace_back(p, b);
    kept += b.bytes;
  } else {
    (void)hipEventDestroy(b.ev);
    (void)hipFreeAsync(p, g_st);
  }
  return true;
}

template <class T>
T* dalloc(size_t n) {
  if (n == 0) n = 1;
  pool_setup();
  if (n * sizeof(T) >= kCacheMin) return static_cast<T*>(big_alloc(n * sizeof(T)));
  void* p = nullptr;
  hipError_t e = hipMallocAsync(&p, n * sizeof(T), g_st);
  if (e != hipSuccess) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    release_for_retry(dev);
    e = hipMallocAsync(&p, n * sizeof(T), g_st);
  }
  if (e != hipSuccess)
    fail(4, "hipMallocAsync of " + std::to_string(n * sizeof(T)) + " bytes: " + hipGetErrorString(e));
  return static_cast<T*>(p);
}

void dfree(void* p) {
  if (p && !big_free(p)) (void)hipFreeAsync(p, g_st);
}

// rocBLAS handles are pooled for the process: creating and destroying one per
// molecule would allocate and free its workspace (a device-wide wait) each time.
std::mutex g_rb_mu;
std::vector<rocblas_handle> g_rb_free;

rocblas_handle rb_acquire() {
  {
    std::lock_guard<std::mutex> lk(g_rb_mu);
    if (!g_rb_free.empty()) {
      rocblas_handle r = g_rb_free.back();
      g_rb_free.pop_back();
      return r;
    }
  }
  rocblas_handle r = nullptr;
  if (rocblas_create_handle(&r) != rocblas_status_success) fail(3, "rocblas_create_handle failed");
  return r;
}

void rb_release(rocblas_handle r) {
  if (!r) return;
  std::lock_guard<std::mutex> lk(g_rb_mu);
  g_rb_free.push_back(r);
}

struct Basis {
  std::vector<int32_t> l, nprim, prim0, ao0;
  std::vector<double> centre, exps, coefs;
  int32_t spherical = 1, nfunc = 0;
  int32_t *d_l = nullptr, *d_nprim = nullptr, *d_prim0 = nullptr, *d_ao0 = nullptr;
  double *d_centre = nullptr, *d_exps = nullptr, *d_coefs = nullptr;
  int n() const { return (int)l.size(); }
  int maxl() const { return l.empty() ? 0 : *std::max_element(l.begin(), l.end()); }
  ShellSet view() const {
    return ShellSet{(int32_t)l.size(), d_l, d_nprim, d_prim0, d_ao0, d_centre, d_exps, d_coefs, spherical};
  }
  void load(const dfks_basis_desc& d) {
    if (d.nshell <= 0) fail(1, "basis with no shells");
    l.assign(d.l, d.l + d.nshell);
    nprim.assign(d.nprim, d.nprim + d.nshell);
    prim0.assign(d.prim0, d.prim0 + d.nshell);
    ao0.assign(d.ao0, d.ao0 + d.nshell);
    centre.assign(d.centre, d.centre + 3 * d.nshell);
    exps.assign(d.exps, d.exps + d.nprim_total);
    coefs.assign(d.coefs, d.coefs + d.nprim_total);
    spherical = d.spherical;
    nfunc = d.nfunc;
    int expect = 0;
    for (int i = 0; i < d.nshell; ++i) {
      if (l[i] < 0 || l[i] > kLMax) fail(1, "shell angular momentum outside 0..4");
      if (ao0[i] != expect) fail(1, "shell ao0 offsets are not contiguous");
      if (prim0[i] < 0 || nprim[i] <= 0 || prim0[i] + nprim[i] > d.nprim_total) fail(1, "bad primitive range");
      expect += nfun(l[i], spherical);
    }
    if (expect != nfunc) fail(1, "nfunc does not match the shells");
  }
  // device copies on g_st; the host vectors stay alive with the Basis
  void upload() {
    auto up = [](auto& v) {
      using T = typename std::decay_t<decltype(v)>::value_type;
      T* p = dalloc<T>(v.size());
      HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, g_st));
      return p;
    };
    d_l = up(l); d_nprim = up(nprim); d_prim0 = up(prim0); d_ao0 = up(ao0);
    d_centre = up(centre); d_exps = up(exps); d_coefs = up(coefs);
  }
  void release() {
    dfree(d_l); dfree(d_nprim); dfree(d_prim0); dfree(d_ao0); dfree(d_centre); dfree(d_exps); dfree(d_coefs);
    d_l = d_nprim = d_prim0 = d_ao0 = nullptr;
    d_centre = d_exps = d_coefs = nullptr;
  }
};

std::vector<double> boys_table() {
  std::vector<double> tbl(kBoysNg * kBoysN);
  const int n_top = std::max(kBoysN - 1, (int)kBoysTmax) + 20;
  for (int j = 0; j < kBoysNg; ++j) {
    double t = j * kBoysDt, two_t = 2.0 * t, emt = std::exp(-t);
    double term = 1.0 / (2 * n_top + 1), total = term;
    for (int k = 1; k < 500; ++k) {
      term = term * two_t / (2 * n_top + 2 * k + 1);
      total += term;
      if (term <= 1e-18 * total) break;
    }
    double f = emt * total;
    for (int n = n_top; n > 0; --n) {
      if (n <= kBoysN - 1) tbl[j * kBoysN + n] = f;
      f = (two_t * f + emt) / (2 * n - 1);
    }
    tbl[j * kBoysN + 0] = f;
  }
  return tbl;
}

double hf_coeff(int xc) {
  switch (xc) {
    case 1: return 0.0;
    case 2: return 0.25;
    case 3: return 0.20;
    case 4: return 0.19;
    default: fail(1, "unknown xc_id " + std::to_string(xc));
  }
  return 0.0;
}

// Long-range exact-exchange fraction and range-separation parameter of a
// range-separated hybrid, K_hf = hf K + hf_lr K_lr(omega) (dftax/energy/xc.py:
// 49-55; CAM-B3LYP 711-750: 0.19 K + 0.46 K_lr(0.33)); zero for global hybrids.
double hf_lr_coeff(int xc) { return xc == 4 ? 0.46 : 0.0; }
double omega_of(int xc) { return xc == 4 ? 0.33 : 0.0; }

}  // namespace

struct dfks_ctx {
  int nao = 0, naux = 0, nfit = 0, nocc = 0, natom = 0, xc_id = 0, nspin = 1;
  // range-separated hybrids: K_hf = hf K + hf_lr K_lr with K_lr from the fitted
  // tensor of the erf(omega r)/r operator (nfit_lr columns of its own metric)
  double omega = 0.0, hf_lr = 0.0;
  int nfit_lr = 0;
  double *d_Blr = nullptr, *d_Klr = nullptr;
  int64_t ng = 0;
  double hf = 0.0, e_nn = 0.0;
  dfks_options opt{};
  Basis orb, aux;
  std::vector<double> atm_host;  // x, y, z, Z
  std::vector<double> boys_host, gxyz_host, gw_host;
  hipStream_t st = nullptr;      // stream of the last call (the destructor frees on it)
  std::vector<double> h_sgn;     // factor signs of the last Fock build
  // stage timers: events recorded on the work stream when timing is on
  bool timing = false;
  hipEvent_t scf_ev[3] = {};
  hipEvent_t ev[kNumStages + 1] = {};
  bool ev_used[kNumStages + 1] = {};
  double build_ms[10] = {};
  void mark(int i, hipStream_t st) {
    if (!timing) return;
    if (!ev[i]) (void)hipEventCreate(&ev[i]);
    (void)hipEventRecord(ev[i], st);
    ev_used[i] = true;
  }
  rocblas_handle rb = nullptr;
  bool built = false;

  double* d_boys = nullptr;
  double* d_atm = nullptr;
  double *d_S = nullptr, *d_T = nullptr, *d_V = nullptr, *d_H = nullptr;
  double *d_int2c = nullptr, *d_int3c = nullptr, *d_B = nullptr;
  double *d_gxyz = nullptr, *d_gw = nullptr, *d_ao = nullptr, *d_zw = nullptr, *d_exc = nullptr;
  // Fock-build scratch
  int nf_chunk = 0;
  double *d_Tx = nullptr, *d_gamma = nullptr, *d_J = nullptr, *d_K = nullptr, *d_Vxc = nullptr;
  double *d_part = nullptr, *d_Edev = nullptr, *d_Cf = nullptr, *d_sgn = nullptr;
  double *d_W1 = nullptr, *d_W2 = nullptr, *d_evals = nullptr, *d_offd = nullptr;
  double* d_gws = nullptr;    // split-K workspace (64 x nao^2 doubles)
  double* d_gpart = nullptr;  // per-row partial sums of gamma (nao x nfit)
  // mixed-precision operands (precision = 1)
  bool mixed_now = false;
  bool restart_diis = false;  // restart the DIIS history once the next step has run
  int diis_base = 0;          // iteration index where the DIIS history restarted
  int last_nf = 0;            // factor columns of the last Fock build (signs in h_sgn)
  float *d_B32 = nullptr, *d_ao32 = nullptr, *d_zw32 = nullptr, *d_T32 = nullptr, *d_Cf32 = nullptr, *d_P32 = nullptr;
  float *d_ws32 = nullptr, *d_eps32 = nullptr, *d_offd32 = nullptr;
  int32_t *d_info = nullptr, *d_sel = nullptr;
  // SCF state
  int nmo = 0, m = 8, it = 0;
  double level_shift = 0.0, e_prev = 0.0;
  // the orthonormal nmo x nocc factor the SCF density was formed from, P = 2 X D D^T X^T
  // (d_L after a purified step, the occupied eigenvectors in d_Ws after an eigensolve);
  // null while P is the caller's guess
  const double* p_fac = nullptr;
  double *d_X = nullptr, *d_P = nullptr, *d_F = nullptr, *d_Fs = nullptr, *d_Es = nullptr, *d_gram = nullptr;
  double *d_coef = nullptr, *d_dots = nullptr, *d_Fx = nullptr, *d_C = nullptr, *d_eps = nullptr, *d_Ws = nullptr;
  // purification: state, Gershgorin bounds, residual diagonal, occupied factor, pivots
  double *d_sp2 = nullptr, *d_lohi = nullptr, *d_dvec = nullptr, *d_L = nullptr;
  // the occupied factor's projection from the previous step: W = X L_prev (nmo x nocc),
  // S = W^T W (nocc x nocc); l_valid while d_L holds the previous step's factor
  double *d_Lw = nullptr, *d_Ls = nullptr, *d_Lq = nullptr;
  bool l_valid = false;
  double* d_lz = nullptr;   // Lanczos vectors and coefficients of the SP2 spectral bounds
  // the previous purified step's F' and spectral bounds (Weyl bounds for the next step)
  double *d_Fp = nullptr, *d_pb = nullptr;
  int pb_valid_n = -1;      // nmo of the step d_Fp, d_pb belong to (-1: none)
  double lz_range = 0.0;    // width of the bounds of the last Lanczos run
  hipGraphExec_t lz_exec = nullptr;  // the Lanczos bounds as one graph (per nmo)
  int lz_exec_n = -1;
  int32_t* d_piv = nullptr;
  // compact XC grid: bins of blocks of equal padded width (see plan_xc_blocks)
  struct XcBin {
    int n;                 // padded compact width (multiple of 32)
    int64_t nblk, blk0;    // blocks in the bin, first block (bin order)
    int64_t pt0;           // first padded point of the bin
    int64_t ao_off, q_off, col_off;
  };
  bool xcb = false;
  std::vector<XcBin> xbins;
  int64_t xc_npts = 0, xc_zw = 0, xc_nblk = 0, xc_qtot = 0;
  // the fused mixed-precision XC pass over all eligible bins at once: the bin table
  // (host copy kept for the upload), its device copy, the workspace, launch sizes
  std::vector<XcfBin> xcf_tab;
  std::vector<char> xcf_fused;  // per bin of xbins: taken by the fused launch
  XcfBin* d_xcf_tab = nullptr;
  double* d_xcf_ws = nullptr;
  int64_t xcf_part = 0;
  int xcf_ztot = 0, xcf_ptot = 0, xcf_vtot = 0;
  bool xcf_ready = false;
  // spin-polarised (nspin = 2) Fock builds: beta-channel XC buffers, the total density,
  // J of the total density, K per channel, gamma of the total density
  int nocc_b = 0, spin = 0;
  double *d_xc_zw2 = nullptr, *d_xc_q2 = nullptr, *d_Vxc2 = nullptr, *d_uP = nullptr, *d_uJ = nullptr;
  double *d_uK = nullptr, *d_ug = nullptr;
  double *d_xc_pts = nullptr, *d_xc_w = nullptr, *d_xc_exc = nullptr, *d_xc_ao = nullptr, *d_xc_zw = nullptr;
  double* d_xc_q = nullptr;
  float* d_xc_ao32 = nullptr;
  int32_t *d_xc_cols = nullptr, *d_xc_vlen = nullptr;
  int2* d_xc_shl = nullptr;
  int64_t *d_xc_shoff = nullptr, *d_xc_vrow = nullptr, *d_xc_vprow = nullptr, *d_xc_vpcol = nullptr;
  // response (jvp of the Fock matrix): per-point XC records at the reference
  // density (8 x ng), rho0 and grad rho0 (4 x ng), and the signed factor of dP
  double *d_fxc = nullptr, *d_rho0 = nullptr, *d_Rf = nullptr;
  int rf_cap = 0;
  bool resp_ready = false;
  int sp2_guess = 24;       // iterations the last purification needed (first chunk of the next)
  hipGraphExec_t sp2_exec = nullptr;  // kSp2Block purification iterations as one graph (per nmo)
  int sp2_exec_n = -1;
  int sp2_variant = 0;      // tile of the purification products (autotuned when the graph is built)
  bool c_stale = false;     // d_C / d_eps not yet computed for the last step (purified density)
  int n_eigh_steps = 0;     // steps that fell back to the eigensolver
  int n_pchol_steps = 0;    // purified steps whose occupied factor came from the pivoted Cholesky
  // recompute mode (options.aux_chunk > 0): screened bra pairs per (la, lb)
  // class, auxiliary slabs with per-lc shell lists, slab and exchange buffers
  bool recompute = false;
  bool slab_zero = false;   // screened pairs leave holes: clear a slab before filling it
  int2* d_pairs3 = nullptr;
  std::vector<size_t> pair3_off;
  std::vector<int> pair3_n;
  int64_t pairs_kept = 0, pairs_total = 0;
  int32_t* d_slab_aux = nullptr;
  std::vector<size_t> slab_aux_off;
  std::vector<int> slab_aux_n, slab_p0, slab_np;
  int max_np = 0;
  double *d_slab = nullptr, *d_Tslab = nullptr, *d_Ttil = nullptr, *d_Lt = nullptr, *d_fitd = nullptr;
  // AO grid streamed per block (options.ao_block > 0)
  int64_t ao_block = 0;
  double *d_aob = nullptr, *d_zwb = nullptr;
  // mixed steps of the recompute / streamed paths: FP32 slab, metric factor and
  // AO block; the FP32 exchange intermediates alias the FP64 ones (a Fock build
  // uses one precision), so they are not freed separately
  float *d_slab32 = nullptr, *d_Lt32 = nullptr, *d_Tslab32 = nullptr, *d_aob32 = nullptr, *d_zwb32 = nullptr;
  float* d_Ttil32 = nullptr;
  bool t32_alias = false;

  ~dfks_ctx() {
    orb.release();
    aux.release();
    void* ptrs[] = {d_boys, d_atm, d_S, d_T, d_V, d_H, d_int2c, d_int3c, d_B, d_gxyz, d_gw, d_ao, d_zw, d_exc,
                    d_Tx, d_gamma, d_J, d_K, d_Vxc, d_part, d_Edev, d_Cf, d_sgn, d_W1, d_W2, d_evals, d_offd,
                    d_info, d_sel, d_X, d_P, d_F, d_Fs, d_Es, d_gram, d_coef, d_dots, d_Fx, d_C, d_eps, d_Ws,
                    d_sp2, d_lz, d_Fp, d_pb, d_lohi, d_dvec, d_L, d_Lw, d_Ls, d_Lq, d_piv, d_pairs3, d_slab_aux, d_slab, d_Tslab, d_Ttil, d_Lt,
                    d_fitd, d_aob, d_zwb, d_fxc, d_rho0, d_Rf, d_xc_pts, d_xc_w, d_xc_exc, d_xc_ao,
                    d_xc_zw, d_xc_q, d_xc_ao32, d_xc_cols, d_xc_vlen, d_xc_shl, d_xc_shoff, d_xc_vrow,
                    d_xc_vprow, d_xc_vpcol, d_xc_zw2, d_xc_q2, d_Vxc2, d_uP, d_uJ, d_uK, d_ug, d_xcf_tab,
                    d_xcf_ws, d_Blr, d_Klr};
    for (void* p : ptrs) dfree(p);
    void* ptrs32[] = {d_gws, d_gpart, d_B32, d_ao32, d_zw32, t32_alias ? nullptr : d_T32, d_Cf32, d_P32, d_ws32,
                      d_eps32, d_offd32, d_slab32, d_Lt32, d_Tslab32, d_aob32, d_zwb32};
    for (void* p : ptrs32) dfree(p);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : scf_ev)
      if (e) (void)hipEventDestroy(e);
    if (sp2_exec) (void)hipGraphExecDestroy(sp2_exec);
    if (lz_exec) (void)hipGraphExecDestroy(lz_exec);
    rb_release(rb);
  }
};

namespace {

std::mutex g_mu;

hipStream_t S_(dfks_stream s) { return reinterpret_cast<hipStream_t>(s); }

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 9;
  }
}

void sync(hipStream_t st) { HIPCHK(hipStreamSynchronize(st)); }

// ---------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Bra pairs of the 3-centre builds: Schwarz factors of every shell pair and
// the pairs kept per (la, lb) class (all of them when screen = 0).
void setup_pairs(dfks_ctx* h, const std::vector<std::vector<int2>>& pairs, const int2* d_pairs,
                 const std::vector<size_t>& pair_off, hipStream_t st) {
  const int lo = h->orb.maxl();
  const ShellSet O = h->orb.view();
  size_t np_all = 0;
  for (auto& v : pairs) np_all += v.size();
  h->pair3_off.assign(pairs.size(), 0);
  h->pair3_n.assign(pairs.size(), 0);
  h->pairs_total = (int64_t)np_all;
  if (!(h->opt.screen > 0.0)) {
    // every pair: the class lists as they are
    for (size_t c = 0; c < pairs.size(); ++c) {
      h->pair3_off[c] = pair_off[c];
      h->pair3_n[c] = (int)pairs[c].size();
    }
    h->pairs_kept = (int64_t)np_all;
    h->slab_zero = false;
    h->d_pairs3 = dalloc<int2>(std::max<size_t>(np_all, 1));
    HIPCHK(hipMemcpyAsync(h->d_pairs3, d_pairs, sizeof(int2) * np_all, hipMemcpyDeviceToDevice, st));
    return;
  }
  double* d_q = dalloc<double>(np_all);
  for (int la = 0; la <= lo; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      const int cls = la * (lo + 1) + lb;
      HIPCHK(launch_schwarz(la, lb, d_pairs + pair_off[cls], (int)pairs[cls].size(), O, h->d_boys,
                            d_q + pair_off[cls], st));
    }
  std::vector<double> q(np_all);
  HIPCHK(hipMemcpyAsync(q.data(), d_q, sizeof(double) * np_all, hipMemcpyDeviceToHost, st));
  sync(st);
  dfree(d_q);
  double qmax = 0.0;
  for (double v : q) qmax = std::max(qmax, v);
  const double cut = h->opt.screen > 0.0 ? h->opt.screen * qmax : -1.0;
  std::vector<int2> kept;
  for (size_t c = 0; c < pairs.size(); ++c) {
    h->pair3_off[c] = kept.size();
    for (size_t i = 0; i < pairs[c].size(); ++i)
      if (q[pair_off[c] + i] >= cut) kept.push_back(pairs[c][i]);
    h->pair3_n[c] = (int)(kept.size() - h->pair3_off[c]);
  }
  h->pairs_kept = (int64_t)kept.size();
  h->slab_zero = h->pairs_kept < h->pairs_total;
  h->d_pairs3 = dalloc<int2>(std::max<size_t>(kept.size(), 1));
  if (!kept.empty())
    HIPCHK(hipMemcpyAsync(h->d_pairs3, kept.data(), kept.size() * sizeof(int2), hipMemcpyHostToDevice, st));
  sync(st);  // kept lives on the host until the copy is done
}

// Recompute mode: auxiliary slabs (runs of whole shells of at most aux_chunk
// functions) with their per-lc shell lists, and the slab-sized buffers.
void setup_slabs(dfks_ctx* h, hipStream_t st) {
  const int nao = h->nao, naux = h->naux;
  const int lx = h->aux.maxl();
  // slabs of whole shells in function order
  const int chunk = std::max(1, h->opt.aux_chunk);
  std::vector<int32_t> flat;
  int c = 0;
  while (c < h->aux.n()) {
    const int p0 = h->aux.ao0[c];
    int c1 = c, np = 0;
    while (c1 < h->aux.n()) {
      const int f = nfun(h->aux.l[c1], h->aux.spherical);
      if (np > 0 && np + f > chunk) break;
      np += f;
      ++c1;
    }
    for (int lc = 0; lc <= lx; ++lc) {
      h->slab_aux_off.push_back(flat.size());
      for (int k = c; k < c1; ++k)
        if (h->aux.l[k] == lc) flat.push_back(k);
      h->slab_aux_n.push_back((int)(flat.size() - h->slab_aux_off.back()));
    }
    h->slab_p0.push_back(p0);
    h->slab_np.push_back(np);
    h->max_np = std::max(h->max_np, np);
    c = c1;
  }
  h->d_slab_aux = dalloc<int32_t>(std::max<size_t>(flat.size(), 1));
  HIPCHK(hipMemcpyAsync(h->d_slab_aux, flat.data(), flat.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  h->d_slab = dalloc<double>((size_t)nao * nao * h->max_np);
  h->d_Tslab = dalloc<double>((size_t)nao * h->max_np * h->nf_chunk);
  h->d_Ttil = dalloc<double>((size_t)nao * naux * h->nf_chunk);
  sync(st);
}

// ---------------------------------------------------------------------------
// Compact XC grid
// ---------------------------------------------------------------------------
constexpr int kXcG = 128;          // points per block
constexpr double kXcTau = 1e-14;   // AO value / gradient bound below which a shell is dropped in a block

// max over r >= d of r^p exp(-a r^2)
double tail_max(int p, double d, double a) {
  if (p == 0) return std::exp(-a * d * d);
  const double rs = std::sqrt(p / (2.0 * a));
  const double r = std::max(d, rs);
  return std::pow(r, p) * std::exp(-a * r * r);
}

// Bound on |phi| and each |d phi / dx| of a shell's functions at distance >= d
// from its centre: sum_k |c_k| (r^l + l r^(l-1) + 2 a_k r^(l+1)) e^(-a_k r^2),
// times 4 for the cart->sph combination.
double shell_bound(const Basis& B, int s, double d) {
  const int l = B.l[s];
  double acc = 0.0;
  for (int k = B.prim0[s]; k < B.prim0[s] + B.nprim[s]; ++k) {
    const double a = B.exps[k], c = std::fabs(B.coefs[k]);
    double t = tail_max(l, d, a) + 2.0 * a * tail_max(l + 1, d, a);
    if (l > 0) t += l * tail_max(l - 1, d, a);
    acc += c * t;
  }
  return 4.0 * acc;
}

// Radius beyond which the shell's bound stays below kXcTau (-1: nowhere above).
double shell_radius(const Basis& B, int s) {
  if (shell_bound(B, s, 0.0) < kXcTau) return -1.0;
  double lo = 0.0, hi = 1.0;
  while (shell_bound(B, s, hi) >= kXcTau && hi < 1e4) hi *= 2.0;
  for (int it = 0; it < 60; ++it) {
    const double mid = 0.5 * (lo + hi);
    (shell_bound(B, s, mid) >= kXcTau ? lo : hi) = mid;
  }
  return hi;
}

uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000ff;
  v = (v | (v << 8)) & 0x0300f00f;
  v = (v | (v << 4)) & 0x030c30c3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}

// Blocks of kXcG grid points in Morton order; per block the shells whose bound
// reaches kXcTau anywhere in the block's bounding sphere; blocks binned by the
// padded count of their functions (multiples of 32). Uploads the binned points,
// weights, column maps, shell lists and the row lists of the Vxc assembly, and
// evaluates the compact AO values and gradients (resident).
void plan_xc_blocks(dfks_ctx* h, hipStream_t st) {
  const double tp0 = now_ms();
  const int64_t ng = h->ng;
  const int nao = h->nao;
  const Basis& O = h->orb;
  const int nsh = O.n();
  const double* X = h->gxyz_host.data();
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t g = 0; g < ng; ++g)
    for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], X[3 * g + d]); hi[d] = std::max(hi[d], X[3 * g + d]); }
  // points in Morton order of their 30-bit keys: a stable LSD radix sort (three
  // 10-bit passes), so equal keys keep the input order
  std::vector<std::pair<uint32_t, int64_t>> key((size_t)ng), tmp((size_t)ng);
  for (int64_t g = 0; g < ng; ++g) {
    uint32_t c = 0;
    for (int d = 0; d < 3; ++d) {
      const double t = (X[3 * g + d] - lo[d]) / (hi[d] - lo[d] + 1e-12);
      c |= spread3((uint32_t)std::min(1023.0, std::max(0.0, t * 1023.0))) << d;
    }
    key[g] = {c, g};
  }
  for (int pass = 0; pass < 3; ++pass) {
    const int sh = 10 * pass;
    std::vector<int64_t> cnt(1025, 0);
    for (int64_t g = 0; g < ng; ++g) ++cnt[((key[g].first >> sh) & 1023u) + 1];
    for (int b = 0; b < 1024; ++b) cnt[b + 1] += cnt[b];
    for (int64_t g = 0; g < ng; ++g) tmp[cnt[(key[g].first >> sh) & 1023u]++] = key[g];
    key.swap(tmp);
  }
  std::vector<std::pair<uint32_t, int64_t>>().swap(tmp);
  const double tp1 = now_ms();
  const int64_t nblk = (ng + kXcG - 1) / kXcG;
  std::vector<double> rcut(nsh);
  for (int s = 0; s < nsh; ++s) rcut[s] = shell_radius(O, s);
  std::vector<std::vector<int>> bsh((size_t)nblk);
  std::vector<int> npad((size_t)nblk);
  // blocks are independent: host threads take interleaved blocks
  auto scan_block = [&](int64_t b) {
    const int64_t p0 = b * kXcG, p1 = std::min(ng, p0 + kXcG);
    double c[3] = {0, 0, 0};
    for (int64_t p = p0; p < p1; ++p)
      for (int d = 0; d < 3; ++d) c[d] += X[3 * key[p].second + d];
    for (int d = 0; d < 3; ++d) c[d] /= (double)(p1 - p0);
    double rad = 0.0;
    for (int64_t p = p0; p < p1; ++p) {
      const double* x = X + 3 * key[p].second;
      rad = std::max(rad, std::sqrt((x[0] - c[0]) * (x[0] - c[0]) + (x[1] - c[1]) * (x[1] - c[1]) +
                                    (x[2] - c[2]) * (x[2] - c[2])));
    }
    int nb = 0;
    for (int s = 0; s < nsh; ++s) {
      if (rcut[s] < 0.0) continue;
      const double* A = &O.centre[3 * s];
      const double dist = std::sqrt((A[0] - c[0]) * (A[0] - c[0]) + (A[1] - c[1]) * (A[1] - c[1]) +
                                    (A[2] - c[2]) * (A[2] - c[2]));
      if (dist - rad <= rcut[s]) {
        bsh[b].push_back(s);
        nb += nfun(O.l[s], O.spherical);
      }
    }
    npad[b] = std::max(32, (nb + 31) / 32 * 32);
  };
  {
    // a few threads only: the process may run many builds at once (the serving loop) and
    // share a CPU quota with the host runtime's own worker threads
    const int nth = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)std::thread::hardware_concurrency(), 4,
                                                                   (nblk + 63) / 64}));
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; ++t)
      pool.emplace_back([&, t] { for (int64_t b = t; b < nblk; b += nth) scan_block(b); });
    for (int64_t b = 0; b < nblk; b += nth) scan_block(b);
    for (auto& th : pool) th.join();
  }
  const double tp2 = now_ms();
  // bins of equal padded width, blocks ordered by (width, position)
  std::vector<int64_t> order((size_t)nblk);
  for (int64_t b = 0; b < nblk; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return npad[a] < npad[b]; });
  h->xbins.clear();
  int64_t ao_tot = 0, q_tot = 0, col_tot = 0, zw_max = 0;
  for (int64_t i = 0; i < nblk;) {
    int64_t j = i;
    while (j < nblk && npad[order[j]] == npad[order[i]]) ++j;
    dfks_ctx::XcBin bn{npad[order[i]], j - i, i, i * kXcG, ao_tot, q_tot, col_tot};
    ao_tot += 4 * bn.nblk * kXcG * bn.n;
    q_tot += bn.nblk * (int64_t)bn.n * bn.n;
    col_tot += bn.nblk * bn.n;
    zw_max = std::max(zw_max, bn.nblk * kXcG * bn.n);
    h->xbins.push_back(bn);
    i = j;
  }
  const int64_t npts = nblk * kXcG;
  std::vector<double> pts((size_t)(3 * npts)), w((size_t)npts, 0.0);
  std::vector<int32_t> cols((size_t)col_tot, -1);
  std::vector<int64_t> shoff((size_t)nblk + 1, 0);
  std::vector<int2> shl;
  for (const auto& bn : h->xbins) {
    for (int64_t k = 0; k < bn.nblk; ++k) {
      const int64_t ob = bn.blk0 + k;             // block in bin order
      const int64_t b = order[ob];                // Morton block
      const int64_t p0 = b * kXcG, p1 = std::min(ng, p0 + kXcG);
      for (int64_t p = 0; p < kXcG; ++p) {
        const int64_t src = key[std::min(p0 + p, p1 - 1)].second;
        for (int d = 0; d < 3; ++d) pts[3 * (ob * kXcG + p) + d] = X[3 * src + d];
        w[ob * kXcG + p] = p0 + p < p1 ? h->gw_host[src] : 0.0;
      }
      int col = 0;
      int32_t* cb = cols.data() + bn.col_off + k * bn.n;
      for (int s : bsh[b]) {
        shl.push_back(make_int2(s, col));
        const int nf = nfun(O.l[s], O.spherical);
        for (int f = 0; f < nf; ++f) cb[col + f] = O.ao0[s] + f;
        col += nf;
      }
      shoff[ob + 1] = (int64_t)shl.size();
    }
  }
  // row lists of the Vxc assembly (CSR by AO row m): one entry per (block, local
  // column i) with cols = m, in bin and block order -- a counting pass, then the fill
  std::vector<int64_t> vrow((size_t)nao + 1, 0);
  for (const auto& bn : h->xbins)
    for (int64_t e = 0; e < bn.nblk * bn.n; ++e) {
      const int32_t m = cols[bn.col_off + e];
      if (m >= 0) ++vrow[m + 1];
    }
  for (int m = 0; m < nao; ++m) vrow[m + 1] += vrow[m];
  const int64_t nent = vrow[nao];
  std::vector<int64_t> vprow((size_t)std::max<int64_t>(nent, 1), 0), vpcol((size_t)std::max<int64_t>(nent, 1), 0);
  std::vector<int32_t> vlen((size_t)std::max<int64_t>(nent, 1), 0);
  {
    std::vector<int64_t> fill(vrow.begin(), vrow.end() - 1);
    for (const auto& bn : h->xbins)
      for (int64_t k = 0; k < bn.nblk; ++k) {
        const int64_t qb = bn.q_off + k * (int64_t)bn.n * bn.n, cbo = bn.col_off + k * bn.n;
        for (int i = 0; i < bn.n; ++i) {
          const int32_t m = cols[cbo + i];
          if (m < 0) continue;
          const int64_t e = fill[m]++;
          vprow[e] = qb + (int64_t)i * bn.n;
          vpcol[e] = cbo;
          vlen[e] = bn.n;
        }
      }
  }
  const double tp3 = now_ms();
  auto up = [&](auto& v) {
    using T = typename std::decay_t<decltype(v)>::value_type;
    T* p = dalloc<T>(v.size());
    HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return p;
  };
  h->d_xc_pts = up(pts);
  h->d_xc_w = up(w);
  h->d_xc_cols = up(cols);
  h->d_xc_shoff = up(shoff);
  if (shl.empty()) shl.push_back(make_int2(0, 0));
  h->d_xc_shl = up(shl);
  h->d_xc_vrow = up(vrow);
  h->d_xc_vprow = up(vprow);
  h->d_xc_vpcol = up(vpcol);
  h->d_xc_vlen = up(vlen);
  h->xc_npts = npts;
  h->xc_nblk = nblk;
  h->xc_zw = zw_max;
  h->xc_qtot = q_tot;
  h->d_xc_exc = dalloc<double>(npts);
  h->d_xc_zw = dalloc<double>(zw_max);
  h->d_xc_q = dalloc<double>(q_tot);
  h->d_xc_ao = dalloc<double>(ao_tot);
  const double tp4 = now_ms();
  HIPCHK(hipMemsetAsync(h->d_xc_ao, 0, sizeof(double) * ao_tot, st));
  const ShellSet S = O.view();
  for (const auto& bn : h->xbins)
    HIPCHK(launch_ao_cmp((int)bn.nblk, kXcG, bn.n, h->d_xc_pts + 3 * bn.pt0, h->d_xc_shl, h->d_xc_shoff + bn.blk0, S,
                         h->d_xc_ao + bn.ao_off, st));
  if (h->opt.precision == 1) {
    h->d_xc_ao32 = dalloc<float>(ao_tot);
    HIPCHK(launch_to_f32(ao_tot, h->d_xc_ao, h->d_xc_ao32, st));
  }
  sync(st);  // the uploads read host vectors that go out of scope here
  // sub-phases of the compact-grid build (ms): Morton sort, block shell scans, host tables,
  // uploads and allocations, AO evaluation to the final wait
  const double tp5 = now_ms();
  h->build_ms[5] = tp1 - tp0;
  h->build_ms[6] = tp2 - tp1;
  h->build_ms[7] = tp3 - tp2;
  h->build_ms[8] = tp4 - tp3;
  h->build_ms[9] = tp5 - tp4;
  h->xcb = true;
}

void build_all(dfks_ctx* h, hipStream_t st) {
  h->rb = rb_acquire();
  RBCHK(rocblas_set_stream(h->rb, st));
  double t0 = now_ms();
  // inputs to the device
  h->orb.upload();
  h->aux.upload();
  h->d_atm = dalloc<double>(h->atm_host.size());
  HIPCHK(hipMemcpyAsync(h->d_atm, h->atm_host.data(), h->atm_host.size() * sizeof(double), hipMemcpyHostToDevice, st));
  h->d_boys = dalloc<double>(h->boys_host.size());
  HIPCHK(hipMemcpyAsync(h->d_boys, h->boys_host.data(), h->boys_host.size() * sizeof(double), hipMemcpyHostToDevice,
                        st));
  if (h->ng > 0) {
    h->d_gxyz = dalloc<double>(3 * h->ng);
    h->d_gw = dalloc<double>(h->ng);
    HIPCHK(hipMemcpyAsync(h->d_gxyz, h->gxyz_host.data(), 3 * h->ng * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->d_gw, h->gw_host.data(), h->ng * sizeof(double), hipMemcpyHostToDevice, st));
  }
  h->d_part = dalloc<double>(1024);
  h->d_Edev = dalloc<double>(8);
  h->d_info = dalloc<int32_t>(1);
  const int nao = h->nao, naux = h->naux;
  const ShellSet O = h->orb.view(), A = h->aux.view();
  // one-electron
  h->d_S = dalloc<double>((size_t)nao * nao);
  h->d_T = dalloc<double>((size_t)nao * nao);
  h->d_V = dalloc<double>((size_t)nao * nao);
  h->d_H = dalloc<double>((size_t)nao * nao);
  HIPCHK(launch_int1e(O, h->natom, h->d_atm, h->d_boys, h->d_S, h->d_T, h->d_V, nao, st));
  HIPCHK(hipMemcpyAsync(h->d_H, h->d_T, sizeof(double) * nao * nao, hipMemcpyDeviceToDevice, st));
  HIPCHK(launch_axpy((int64_t)nao * nao, 1.0, h->d_V, h->d_H, 1, st));
  sync(st);
  h->build_ms[0] = now_ms() - t0;
  t0 = now_ms();

  h->nf_chunk = std::max(1, std::min(std::min(nao, 512), std::max(h->nocc, 1)));
  // class lists
  const int lo = h->orb.maxl(), lx = h->aux.maxl();
  std::vector<std::vector<int2>> pairs((lo + 1) * (lo + 1));
  for (int a = 0; a < h->orb.n(); ++a)
    for (int b = 0; b < h->orb.n(); ++b) {
      int la = h->orb.l[a], lb = h->orb.l[b];
      if (la < lb || (la == lb && b > a)) continue;
      pairs[la * (lo + 1) + lb].push_back(make_int2(a, b));
    }
  std::vector<std::vector<int32_t>> auxl(lx + 1);
  for (int c = 0; c < h->aux.n(); ++c) auxl[h->aux.l[c]].push_back(c);
  std::vector<std::vector<int2>> apairs(lx + 1);
  for (int c = 0; c < h->aux.n(); ++c) apairs[h->aux.l[c]].push_back(make_int2(c, c));
  // flatten to device
  std::vector<int2> all_pairs;
  std::vector<size_t> pair_off;
  for (auto& v : pairs) { pair_off.push_back(all_pairs.size()); all_pairs.insert(all_pairs.end(), v.begin(), v.end()); }
  std::vector<size_t> apair_off;
  for (auto& v : apairs) { apair_off.push_back(all_pairs.size()); all_pairs.insert(all_pairs.end(), v.begin(), v.end()); }
  std::vector<int32_t> all_aux;
  std::vector<size_t> aux_off;
  for (auto& v : auxl) { aux_off.push_back(all_aux.size()); all_aux.insert(all_aux.end(), v.begin(), v.end()); }
  int2* d_pairs = dalloc<int2>(all_pairs.size());
  int32_t* d_auxl = dalloc<int32_t>(all_aux.size());
  HIPCHK(hipMemcpyAsync(d_pairs, all_pairs.data(), all_pairs.size() * sizeof(int2), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_auxl, all_aux.data(), all_aux.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));

  // two-centre metric
  h->d_int2c = dalloc<double>((size_t)naux * naux);
  for (int la = 0; la <= lx; ++la)
    for (int lc = 0; lc <= lx; ++lc)
      HIPCHK(launch_eri_wave(la, 0, lc, d_pairs + apair_off[la], (int)apairs[la].size(), d_auxl + aux_off[lc],
                             (int)auxl[lc].size(), A, A, 1, h->d_boys, h->d_int2c, naux, 0, 0, 0, st));
  sync(st);
  h->build_ms[1] = now_ms() - t0;
  t0 = now_ms();
  // The metric's eigendecomposition (rocSOLVER, a long chain of small kernels
  // on an n = naux matrix) needs only (P|Q): it runs on a side stream with its
  // own rocBLAS handle while the three-centre build and the AO grid plan run on
  // st, and is collected before the metric factor. d_U, d_w, d_e, the side
  // stream and handle stay with this build until then.
  double* d_U = dalloc<double>((size_t)naux * naux);
  double* d_w = dalloc<double>(naux);
  double* d_e = dalloc<double>(naux);
  hipStream_t s_eig = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s_eig, hipStreamNonBlocking));
  rocblas_handle rb_eig = rb_acquire();
  struct SideStream {
    hipStream_t s;
    rocblas_handle rb;
    ~SideStream() {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
      if (rb) (void)rocblas_set_stream(rb, nullptr);
      rb_release(rb);
    }
  } side{s_eig, rb_eig};
  RBCHK(rocblas_set_stream(rb_eig, s_eig));
  const auto eig_launch = [&](const double* V, rocblas_handle rb, hipStream_t s) {
    HIPCHK(hipMemcpyAsync(d_U, V, sizeof(double) * naux * naux, hipMemcpyDeviceToDevice, s));
    RBCHK(rocsolver_dsyevd(rb, rocblas_evect_original, rocblas_fill_upper, naux, d_U, naux, d_w, d_e, h->d_info));
  };
  // metric pseudo-inverse factor L = U_keep w_keep^{-1/2} (nkeep columns kept by the relative cutoff) from
  // the eigenpairs in d_U, d_w once stream s has finished them
  // the side stream's eigensolve runs on its own host thread: rocSOLVER's divide and conquer
  // holds its calling thread for much of its run, which would otherwise delay the three-centre
  // launches queued behind it
  struct EigThread {
    std::thread t;
    int code = 0;
    std::string msg;
    void finish() {
      if (t.joinable()) t.join();
      if (code) fail(code, msg);
    }
    ~EigThread() {
      if (t.joinable()) t.join();
    }
  } eig;
  const auto eig_factor = [&](int& nkeep, hipStream_t s) {
    if (s == s_eig) eig.finish();
    std::vector<double> w(naux);
    HIPCHK(hipMemcpyAsync(w.data(), d_w, sizeof(double) * naux, hipMemcpyDeviceToHost, s));
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (info != 0) fail(5, "metric eigensolver failed, info=" + std::to_string(info));
    const double wmax = w[naux - 1];
    int k0 = 0;
    while (k0 < naux && !(w[k0] > h->opt.metric_cutoff * wmax)) ++k0;
    nkeep = naux - k0;
    double* Lt = dalloc<double>((size_t)nkeep * naux);  // column-major nkeep x naux
    HIPCHK(launch_metric_factor(naux, k0, d_U, d_w, Lt, st));
    return Lt;
  };
  {
    // the side stream starts after st's (P|Q) and the allocations above (stream-ordered on st, and a
    // reused buffer's wait for its previous user is queued on st)
    hipEvent_t ready = nullptr;
    HIPCHK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    const hipError_t e1 = hipEventRecord(ready, st);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(s_eig, ready, 0) : e1;
    (void)hipEventDestroy(ready);
    HIPCHK(e2);
  }
  {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    eig.t = std::thread([&eig, &eig_launch, h, dev, rb_eig, s_eig] {
      try {
        HIPCHK(hipSetDevice(dev));
        StreamScope ss(s_eig);
        eig_launch(h->d_int2c, rb_eig, s_eig);
        HIPCHK(hipStreamSynchronize(s_eig));
      } catch (const Fail& f) {
        eig.code = f.code;
        eig.msg = g_err;
      } catch (const std::exception& e) {
        eig.code = 9;
        eig.msg = e.what();
      }
    });
  }
  // bra pairs (Schwarz-screened when screen > 0)
  setup_pairs(h, pairs, d_pairs, pair_off, st);
  // three-centre (materialised mode only)
  double* d3 = nullptr;
  if (!h->recompute) {
    d3 = dalloc<double>((size_t)nao * nao * naux);
    if (h->slab_zero) HIPCHK(hipMemsetAsync(d3, 0, sizeof(double) * nao * nao * naux, st));
    for (int la = 0; la <= lo; ++la)
      for (int lb = 0; lb <= la; ++lb)
        for (int lc = 0; lc <= lx; ++lc) {
          const int cls = la * (lo + 1) + lb;
          const int2* pp = h->d_pairs3 + h->pair3_off[cls];
          const int npair = h->pair3_n[cls];
          if (eri_thread_has(la, lb, lc) && A.spherical)
            HIPCHK(launch_eri_thread(la, lb, lc, pp, npair, d_auxl + aux_off[lc], (int)auxl[lc].size(), O, A,
                                     h->d_boys, d3, (int64_t)nao * naux, naux, 0, 1, st));
          else
            HIPCHK(launch_eri_cls(la, lb, lc, pp, npair, d_auxl + aux_off[lc], (int)auxl[lc].size(), O, A,
                                  h->d_boys, d3, (int64_t)nao * naux, naux, 0, 1, st));
        }
    sync(st);
  } else {
    setup_slabs(h, st);
  }
  h->build_ms[2] = now_ms() - t0;
  t0 = now_ms();

  // AO values and gradients on the grid: compact blocks, resident, or per block in every Fock build
  if (h->ng > 0 && h->opt.xc_blocks) {
    plan_xc_blocks(h, st);
    h->ao_block = 0;
  } else if (h->ng > 0) {
    h->d_exc = dalloc<double>(h->ng);
    if (h->ao_block == 0) {
      h->d_ao = dalloc<double>((size_t)4 * h->ng * nao);
      h->d_zw = dalloc<double>((size_t)h->ng * nao);
      HIPCHK(launch_ao(h->ng, h->d_gxyz, O, nao, h->d_ao, st));
    } else {
      const int64_t gb = std::min<int64_t>(h->ao_block, h->ng);
      h->d_aob = dalloc<double>((size_t)4 * gb * nao);
      h->d_zwb = dalloc<double>((size_t)gb * nao);
    }
  }
  sync(st);
  h->build_ms[4] = now_ms() - t0;
  t0 = now_ms();

  // the metric factor from the side stream's eigenpairs, then B
  double* d_Lt = eig_factor(h->nfit, s_eig);
  if (!h->recompute) {
    // B_cm (nfit x nao^2) = Lt (nfit x naux) * int3c_cm (naux x nao^2)
    h->d_B = dalloc<double>((size_t)nao * nao * h->nfit);
    const double one = 1.0, zero = 0.0;
    RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_none, rocblas_operation_none, h->nfit, (int64_t)nao * nao, naux,
                           &one, d_Lt, h->nfit, d3, naux, &zero, h->d_B, h->nfit));
    sync(st);
    h->d_Lt = d_Lt;   // kept: the nuclear gradient maps fitted quantities back to (P|Q) (nfit naux words)
    if (h->opt.keep_int3c) h->d_int3c = d3;
    else dfree(d3);
  } else {
    h->d_Lt = d_Lt;
    h->d_fitd = dalloc<double>(naux);
  }
  if (h->hf_lr != 0.0) {
    // long-range operator erf(omega r)/r of a range-separated hybrid (terms.py:578-612,
    // energy.py:419-433): attenuated (P|Q) and (mn|P), the pseudo-inverse factor of the
    // attenuated metric with the same cutoff, B_lr = (mn|P) L_lr
    ShellSet Aw = A;
    Aw.omega = h->omega;
    double* d2lr = dalloc<double>((size_t)naux * naux);
    for (int la = 0; la <= lx; ++la)
      for (int lc = 0; lc <= lx; ++lc)
        HIPCHK(launch_eri_wave(la, 0, lc, d_pairs + apair_off[la], (int)apairs[la].size(), d_auxl + aux_off[lc],
                               (int)auxl[lc].size(), Aw, Aw, 1, h->d_boys, d2lr, naux, 0, 0, 0, st));
    double* d3lr = dalloc<double>((size_t)nao * nao * naux);
    if (h->slab_zero) HIPCHK(hipMemsetAsync(d3lr, 0, sizeof(double) * nao * nao * naux, st));
    for (int la = 0; la <= lo; ++la)
      for (int lb = 0; lb <= la; ++lb)
        for (int lc = 0; lc <= lx; ++lc) {
          const int cls = la * (lo + 1) + lb;
          const int2* pp = h->d_pairs3 + h->pair3_off[cls];
          const int npair = h->pair3_n[cls];
          if (eri_thread_has(la, lb, lc) && A.spherical)
            HIPCHK(launch_eri_thread(la, lb, lc, pp, npair, d_auxl + aux_off[lc], (int)auxl[lc].size(), O, Aw,
                                     h->d_boys, d3lr, (int64_t)nao * naux, naux, 0, 1, st));
          else
            HIPCHK(launch_eri_cls(la, lb, lc, pp, npair, d_auxl + aux_off[lc], (int)auxl[lc].size(), O, Aw,
                                  h->d_boys, d3lr, (int64_t)nao * naux, naux, 0, 1, st));
        }
    sync(st);
    eig_launch(d2lr, h->rb, st);
    double* d_Ltlr = eig_factor(h->nfit_lr, st);
    h->d_Blr = dalloc<double>((size_t)nao * nao * h->nfit_lr);
    const double one = 1.0, zero = 0.0;
    RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_none, rocblas_operation_none, h->nfit_lr, (int64_t)nao * nao,
                           naux, &one, d_Ltlr, h->nfit_lr, d3lr, naux, &zero, h->d_Blr, h->nfit_lr));
    sync(st);
    dfree(d3lr);
    dfree(d2lr);
    dfree(d_Ltlr);
    h->d_Klr = dalloc<double>((size_t)nao * nao);
  }
  dfree(d_pairs);
  dfree(d_auxl);
  sync(st);
  h->build_ms[3] = now_ms() - t0;
  t0 = now_ms();
  dfree(d_U); dfree(d_w); dfree(d_e);

  // Fock-build scratch: exchange intermediate for up to nf_chunk factor columns
  h->d_Tx = dalloc<double>((size_t)nao * std::max(h->nfit, h->nfit_lr) * h->nf_chunk);
  h->d_gamma = dalloc<double>(h->nfit);
  h->d_J = dalloc<double>((size_t)nao * nao);
  h->d_K = dalloc<double>((size_t)nao * nao);
  h->d_Vxc = dalloc<double>((size_t)nao * nao);
  h->d_Cf = dalloc<double>((size_t)nao * nao);
  h->d_sgn = dalloc<double>(2 * (size_t)nao);
  h->d_W1 = dalloc<double>((size_t)nao * nao);
  h->d_W2 = dalloc<double>((size_t)nao * nao);
  h->d_evals = dalloc<double>(nao);
  h->d_offd = dalloc<double>(nao);
  h->d_sel = dalloc<int32_t>(nao);
  h->d_gws = dalloc<double>((size_t)64 * nao * nao);
  // purification state: fixed size, allocated once, because the purification
  // graph keeps its address (as it keeps those of d_W1 and d_gws)
  h->d_sp2 = dalloc<double>(kSp2State);
  h->d_lz = dalloc<double>(3 * (size_t)nao + kLzParts + 2 * kLzSteps);
  h->d_gpart = dalloc<double>((size_t)nao * h->nfit);
  if (h->opt.precision == 1) {
    if (!h->recompute) {
      const size_t nB = (size_t)nao * nao * h->nfit;
      h->d_B32 = dalloc<float>(nB);
      HIPCHK(launch_to_f32((int64_t)nB, h->d_B, h->d_B32, st));
      h->d_T32 = dalloc<float>((size_t)nao * h->nfit * h->nf_chunk);
    } else {
      h->d_slab32 = dalloc<float>((size_t)nao * nao * h->max_np);
      h->d_Tslab32 = dalloc<float>((size_t)nao * h->max_np * h->nf_chunk);
      h->d_Lt32 = dalloc<float>((size_t)h->nfit * naux);
      HIPCHK(launch_to_f32((int64_t)h->nfit * naux, h->d_Lt, h->d_Lt32, st));
      h->d_Ttil32 = reinterpret_cast<float*>(h->d_Ttil);
      h->d_T32 = reinterpret_cast<float*>(h->d_Tx);
      h->t32_alias = true;
    }
    if (h->xcb) {
      // compact AO rows were converted in plan_xc_blocks
    } else if (h->ng > 0 && h->ao_block == 0) {
      h->d_ao32 = dalloc<float>((size_t)4 * h->ng * nao);
      HIPCHK(launch_to_f32((int64_t)4 * h->ng * nao, h->d_ao, h->d_ao32, st));
      h->d_zw32 = dalloc<float>((size_t)h->ng * nao);
    } else if (h->ng > 0) {
      const int64_t gb = std::min<int64_t>(h->ao_block, h->ng);
      h->d_aob32 = dalloc<float>((size_t)4 * gb * nao);
      h->d_zwb32 = dalloc<float>((size_t)gb * nao);
    }
    h->d_Cf32 = dalloc<float>((size_t)nao * nao);
    h->d_P32 = dalloc<float>((size_t)nao * nao);
    h->d_ws32 = dalloc<float>((size_t)nao * nao);
    h->d_eps32 = dalloc<float>(nao);
    h->d_offd32 = dalloc<float>(nao);
  }
  sync(st);
  h->build_ms[4] += now_ms() - t0;   // the AO grid plan above plus the Fock-build scratch
  h->built = true;
}

// ---------------------------------------------------------------------------
// energy + Fock for P = sum_i s_i c_i c_i^T (Cf column-major nao x nf)
// ---------------------------------------------------------------------------
// split count so a K-long product fills the chip (~2048 workgroups), k-chunks >= 512
int split_for(int64_t tiles, int64_t K) {
  int64_t s = 2048 / std::max<int64_t>(tiles, 1);
  s = std::min<int64_t>(s, std::max<int64_t>(1, K / 512));
  return (int)std::max<int64_t>(1, std::min<int64_t>(s, 64));
}

// (mn|P) for the functions of slab s into d_slab, layout [m][n][p] with
// p = P - p0 (kept bra pairs only; the others stay zero).
void eri_slab(dfks_ctx* h, int s, hipStream_t st) {
  const int nao = h->nao;
  const int lo = h->orb.maxl(), lx = h->aux.maxl();
  const int np = h->slab_np[s];
  const int64_t p0 = h->slab_p0[s];
  const ShellSet O = h->orb.view(), A = h->aux.view();
  if (h->slab_zero) HIPCHK(hipMemsetAsync(h->d_slab, 0, sizeof(double) * nao * nao * np, st));
  const int64_t sa = (int64_t)nao * np, sb = np;
  for (int la = 0; la <= lo; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      const int cls = la * (lo + 1) + lb;
      const int npair = h->pair3_n[cls];
      if (npair == 0) continue;
      const int2* pp = h->d_pairs3 + h->pair3_off[cls];
      for (int lc = 0; lc <= lx; ++lc) {
        const int sl = s * (lx + 1) + lc;
        const int nal = h->slab_aux_n[sl];
        if (nal == 0) continue;
        const int32_t* aa = h->d_slab_aux + h->slab_aux_off[sl];
        if (eri_thread_has(la, lb, lc) && A.spherical)
          HIPCHK(launch_eri_thread(la, lb, lc, pp, npair, aa, nal, O, A, h->d_boys, h->d_slab, sa, sb, p0, 1, st));
        else
          HIPCHK(launch_eri_cls(la, lb, lc, pp, npair, aa, nal, O, A, h->d_boys, h->d_slab, sa, sb, p0, 1, st));
      }
    }
}

// Exchange intermediate of nc factor columns from recomputed slabs:
// T~[m][P][i] = sum_l (l m|P) Cf[l,i] slab by slab, then the fitted
// T[m][Q][i] = sum_P T~[m][P][i] L[P,Q] (one strided-batched GEMM over m).
void exchange_recompute(dfks_ctx* h, const double* Cf, const float* Cf32, int nc, hipStream_t st) {
  const int nao = h->nao, naux = h->naux, nfit = h->nfit;
  const double one = 1.0, zero = 0.0;
  if (h->mixed_now) {
    // split-BF16 products: slab (FP32 copy) x Cf, then the metric factor per m (batched)
    for (size_t s = 0; s < h->slab_np.size(); ++s) {
      const int np = h->slab_np[s];
      eri_slab(h, (int)s, st);
      HIPCHK(launch_to_f32((int64_t)nao * nao * np, h->d_slab, h->d_slab32, st));
      const GemmArgs g{(int)((int64_t)nao * np), nc, nao, h->d_slab32, (int64_t)nao * np, false, Cf32, nao, true};
      HIPCHK(launch_gemm_f32out(g, h->d_Tslab32, st));
      HIPCHK(hipMemcpy2DAsync(h->d_Ttil32 + (size_t)h->slab_p0[s] * nc, sizeof(float) * naux * nc, h->d_Tslab32,
                              sizeof(float) * np * nc, sizeof(float) * np * nc, nao, hipMemcpyDeviceToDevice, st));
    }
    // T_m[Q][i] = sum_P L[P,Q] T~_m[P][i]: A(Q,P) = Lt column-major, B(i,P) = T~_m row-major [P][i]
    const GemmArgs gm{nfit, nc, naux, h->d_Lt32, nfit, false, h->d_Ttil32, nc, false};
    HIPCHK(launch_gemm_f32out_batched(gm, h->d_T32, nao, 0, (int64_t)naux * nc, (int64_t)nfit * nc, st));
    return;
  }
  for (size_t s = 0; s < h->slab_np.size(); ++s) {
    const int np = h->slab_np[s];
    eri_slab(h, (int)s, st);
    RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_transpose, rocblas_operation_transpose, nc, (int64_t)nao * np,
                           nao, &one, Cf, nao, h->d_slab, (int64_t)nao * np, &zero, h->d_Tslab, nc));
    HIPCHK(hipMemcpy2DAsync(h->d_Ttil + (size_t)h->slab_p0[s] * nc, sizeof(double) * naux * nc, h->d_Tslab,
                            sizeof(double) * np * nc, sizeof(double) * np * nc, nao, hipMemcpyDeviceToDevice, st));
  }
  RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_none, rocblas_operation_transpose, nc, nfit, naux, &one,
                                      h->d_Ttil, nc, (rocblas_stride)naux * nc, h->d_Lt, nfit, 0, &zero, h->d_Tx, nc,
                                      (rocblas_stride)nfit * nc, nao));
}

// J = sum_P (mn|P) d_P with d = L gamma, slab by slab.
void coulomb_recompute(dfks_ctx* h, hipStream_t st) {
  const int nao = h->nao, naux = h->naux, nfit = h->nfit;
  const double one = 1.0, zero = 0.0;
  RBCHK(rocblas_dgemv(h->rb, rocblas_operation_transpose, nfit, naux, &one, h->d_Lt, nfit, h->d_gamma, 1, &zero,
                      h->d_fitd, 1));
  for (size_t s = 0; s < h->slab_np.size(); ++s) {
    const int np = h->slab_np[s];
    eri_slab(h, (int)s, st);
    if (s == 0) HIPCHK(launch_jvec((int64_t)nao * nao, np, h->d_slab, h->d_fitd + h->slab_p0[s], h->d_J, 0, st));
    else HIPCHK(launch_jvec_acc((int64_t)nao * nao, np, h->d_slab, h->d_fitd + h->slab_p0[s], h->d_J, st));
  }
}

// FP64 exchange product T[m,Q,i] = sum_l B[l,m,Q] Cf[l,i], autotuned: the first
// product of each (device, M, N, K) shape runs both candidates -- the library
// kernel (128- or 160-wide N tiles, whichever pads less) and rocBLAS -- timed
// with HIP events on the work stream, and the process keeps the faster one for
// that shape. Both compute the same T to rounding.
// Per-shape choice among the variants of a product. Tags and variants:
//   0  FP64 exchange T = B^T C:      0 rocBLAS, 1 k_gemm, 2 k_gemm_w8
//   1  FP64 XC Z_b = Phi_b Pg_b:     0 rocBLAS, 1 k_gemm (batched)
//   2  FP64 XC Phi_b^T W_b:          0 rocBLAS, 1 k_gemm (batched)
//   3  FP64 exchange K partials:     0 k_gemm, 1 k_gemm_w8, 2 k_gemm_w8 split by XCD (base 1)
//   4  split-BF16 exchange T:        0 k_gemm, 1 k_gemm_b3w8 (base 1)
//   5  SP2 product partials X X:     0 k_gemm, 1 k_gemm_w8   (base 1)
//   6  split-BF16 exchange K:        0 k_gemm, 1 k_gemm split by XCD (base 1)
// The first call of a shape on a device times every variant with HIP events
// and keeps the fastest; the output of that call is recomputed by the kept
// variant when it is not the one that ran last, so every call of a shape
// returns the kept variant's result. options.gemm 1, 2 and 3 force variant
// gemm - 1 - base, clamped to the variants there are (gemm 2: k_gemm, 3: the
// 8-wavefront tiles).
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int64_t, int64_t, int64_t>, int> g_tune;

template <typename Run>
int tuned_product(dfks_ctx* h, int tag, int64_t d0, int64_t d1, int64_t d2, int ncand, int base, Run run,
                  hipStream_t st) {
  if (h->opt.gemm != 0) {
    const int c = std::max(0, std::min(h->opt.gemm - 1 - base, ncand - 1));
    run(c);
    return c;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, tag, d0, d1, d2);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      const int c = it->second;
      run(c);
      return c;
    }
  }
  hipEvent_t ev[5];
  for (int c = 0; c <= ncand; ++c) HIPCHK(hipEventCreate(&ev[c]));
  HIPCHK(hipEventRecord(ev[0], st));
  for (int c = 0; c < ncand; ++c) {
    run(c);
    HIPCHK(hipEventRecord(ev[c + 1], st));
  }
  HIPCHK(hipEventSynchronize(ev[ncand]));
  int best = 0;
  float t_best = 0.f, t_all[4] = {0.f, 0.f, 0.f, 0.f};
  for (int c = 0; c < ncand; ++c) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev[c], ev[c + 1]));
    t_all[c] = t;
    if (c == 0 || t < t_best) { best = c; t_best = t; }
  }
  static const bool verbose = std::getenv("DFKS_TUNE_LOG") != nullptr;
  if (verbose)
    std::fprintf(stderr, "dfks tune: product %d shape (%lld, %lld, %lld): %.3f / %.3f / %.3f / %.3f ms -> variant %d\n",
                 tag, (long long)d0, (long long)d1, (long long)d2, t_all[0], t_all[1], t_all[2], t_all[3], best);
  for (int c = 0; c <= ncand; ++c) (void)hipEventDestroy(ev[c]);
  int kept;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    kept = g_tune.emplace(key, best).first->second;
  }
  if (kept != ncand - 1) run(kept);
  return kept;
}

// T = B^T C (nao*nfit x nc), FP64
// Returns whether the gamma row terms of rd were produced with T (the
// 8-wavefront kernel computes them in its epilogue).
bool t_product(dfks_ctx* h, const double* C, int nc, const RowDot& rd, hipStream_t st) {
  const int nao = h->nao, nfit = h->nfit;
  const double one = 1.0, zero = 0.0;
  const bool fuse = nc <= 160 && nfit >= 128;
  const int v = tuned_product(h, 0, (int64_t)nao * nfit, nc, nao, 3, 0, [&](int choice) {
    const GemmArgs g{(int)((int64_t)nao * nfit), nc, nao, h->d_B, (int64_t)nao * nfit, false, C, nao, true};
    if (choice == 2) {
      HIPCHK(launch_gemm_w8(g, h->d_Tx, fuse ? rd : RowDot(), st));
    } else if (choice == 1) {
      HIPCHK(launch_gemm(kF64, g, h->d_Tx, 1.0, 0.0, 1, 0, nullptr, st));
    } else {
      RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_transpose, rocblas_operation_transpose, nc,
                             (int64_t)nao * nfit, nao, &one, C, nao, h->d_B, (int64_t)nao * nfit, &zero, h->d_Tx,
                             nc));
    }
  }, st);
  return fuse && v == 2;
}

// Kout = sg T^T T + beta Kout in FP64 for T = d_Tx (nao rows of kd contiguous
// values): split-K partial tiles by the fastest of the library tiles and
// layouts for this shape and the ordered reduction, or rocBLAS (gemm = 1).
void k_product_f64(dfks_ctx* h, int64_t kd, double* Kout, double sg, double beta, hipStream_t st) {
  const int nao = h->nao;
  const int ntn = (nao + 127) / 128;
  const int nsplit = split_for((int64_t)ntn * (ntn + 1) / 2, kd);
  const int nsx = 16;   // the split by XCD: each XCD's tiles on the same k-window of T
  if (h->opt.gemm != 1) {
    const GemmArgs g{nao, nao, kd, h->d_Tx, kd, true, h->d_Tx, kd, true};
    const int v = tuned_product(h, 3, nao, kd, nsplit, 3, 1, [&](int c) {
      HIPCHK(launch_gemm_partials(kF64, g, c == 2 ? nsx : nsplit, c == 2 ? 3 : 1, h->d_gws, c == 2 ? 1 : c, st));
    }, st);
    HIPCHK(launch_gemm_reduce(kF64, g, Kout, sg, beta, v == 2 ? nsx : nsplit, 1, h->d_gws, st));
  } else {
    RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_transpose, rocblas_operation_none, nao, nao, kd, &sg, h->d_Tx,
                           kd, h->d_Tx, kd, &beta, Kout, nao));
  }
}

// Long-range exchange of a range-separated hybrid from the factor of P alone:
// K_lr = sum_i s_i T_i^T T_i with T = B_lr^T c over the fitted tensor of the
// erf(omega r)/r operator (terms.py:604-612), into d_Klr. FP64.
void k_lr_factored(dfks_ctx* h, const double* Cf, const std::vector<double>& sgn, int nf, hipStream_t st) {
  const int nao = h->nao;
  // the products read the handle's fitted tensor and its width: point them at the long-range one
  struct Swap {
    dfks_ctx* h;
    double* B;
    int nfit;
    Swap(dfks_ctx* hh) : h(hh), B(hh->d_B), nfit(hh->nfit) {
      h->d_B = h->d_Blr;
      h->nfit = h->nfit_lr;
    }
    ~Swap() {
      h->d_B = B;
      h->nfit = nfit;
    }
  } swap(h);
  bool first = true;
  int i0 = 0;
  while (i0 < nf) {
    int i1 = i0 + 1;
    while (i1 < nf && i1 - i0 < h->nf_chunk && sgn[i1] == sgn[i0]) ++i1;
    const int nc = i1 - i0;
    (void)t_product(h, Cf + (int64_t)i0 * nao, nc, RowDot(), st);
    k_product_f64(h, (int64_t)h->nfit_lr * nc, h->d_Klr, sgn[i0], first ? 0.0 : 1.0, st);
    first = false;
    i0 = i1;
  }
  if (first) HIPCHK(hipMemsetAsync(h->d_Klr, 0, sizeof(double) * nao * nao, st));
}

// gamma, J and K of P = sum_i s_i c_i c_i^T from its factor alone:
// gamma = sum_i s_i T_i c_i, J = B gamma, K = sum_i s_i T_i^T T_i.
void jk_factored(dfks_ctx* h, const double* Cf, const std::vector<double>& sgn, int nf, hipStream_t st) {
  const int nao = h->nao, nfit = h->nfit;
  const int64_t nn = (int64_t)nao * nao;
  const bool mixed = h->mixed_now;
  // signs to the device (host copy kept in the handle until the next call)
  if (nf > (mixed ? nao : 2 * nao)) fail(1, "too many factor columns");
  h->h_sgn.assign(sgn.begin(), sgn.end());
  h->last_nf = nf;
  HIPCHK(hipMemcpyAsync(h->d_sgn, h->h_sgn.data(), sizeof(double) * nf, hipMemcpyHostToDevice, st));
  if (mixed) HIPCHK(launch_to_f32((int64_t)nao * nf, Cf, h->d_Cf32, st));
  // Coulomb + exchange over runs of equal sign, chunked to nf_chunk columns
  bool first = true;
  HIPCHK(hipMemsetAsync(h->d_gamma, 0, sizeof(double) * nfit, st));
  HIPCHK(hipMemsetAsync(h->d_K, 0, sizeof(double) * nn, st));
  int i0 = 0;
  while (i0 < nf) {
    int i1 = i0 + 1;
    while (i1 < nf && i1 - i0 < h->nf_chunk && sgn[i1] == sgn[i0]) ++i1;
    const int nc = i1 - i0;
    const double sg = sgn[i0];
    const double beta = first ? 0.0 : 1.0;
    RowDot rd;
    rd.cf = Cf + (int64_t)i0 * nao;
    rd.sg = h->d_sgn + i0;
    rd.nao = nao;
    rd.nq = nfit;
    rd.out = h->d_gpart;
    bool rows_done = false;
    // T[m,Q,i] = sum_l B[l,m,Q] Cf[l,i]
    if (h->recompute) {
      exchange_recompute(h, Cf + (int64_t)i0 * nao, mixed ? h->d_Cf32 + (int64_t)i0 * nao : nullptr, nc, st);
    } else if (mixed) {
      // split-BF16 T: the 128-row tile of k_gemm or the 8-wavefront 256-row tile with one or two k-steps
      // of operands in flight, whichever is fastest here
      const GemmArgs g{(int)((int64_t)nao * nfit), nc, nao, h->d_B32, (int64_t)nao * nfit, false,
                       h->d_Cf32 + (int64_t)i0 * nao, nao, true};
      const bool fuse = nc <= 160 && nfit >= 256;
      const int v = tuned_product(h, 4, (int64_t)nao * nfit, nc, nao, 3, 1, [&](int c) {
        if (c >= 1) HIPCHK(launch_gemm_b3w8(g, h->d_T32, fuse ? rd : RowDot(), st, c));
        else HIPCHK(launch_gemm_f32out(g, h->d_T32, st));
      }, st);
      rows_done = fuse && v >= 1;
    } else {
      rows_done = t_product(h, Cf + (int64_t)i0 * nao, nc, rd, st);
    }
    h->mark(kStExchangeGemm + 1, st);
    // gamma += sum_i s_i T_i c_i (timed with the exchange product K below); its row terms come
    // with T from the 8-wavefront kernels, otherwise from a pass over T
    if (rows_done)
      HIPCHK(launch_colsum(nao, nfit, h->d_gpart, h->d_gamma, 1, st));
    else if (mixed)
      HIPCHK(launch_gamma_f32(nao, nfit, nc, h->d_T32, Cf + (int64_t)i0 * nao, h->d_sgn + i0, h->d_gamma, h->d_gpart,
                              1, st));
    else
      HIPCHK(launch_gamma(nao, nfit, nc, h->d_Tx, Cf + (int64_t)i0 * nao, h->d_sgn + i0, h->d_gamma, h->d_gpart, 1,
                          st));
    if (h->hf != 0.0) {
      // K += s T T^T (symmetric: upper tiles, mirrored by the split reduction)
      const int64_t kd = (int64_t)nfit * nc;
      const int ntn = (nao + 127) / 128;
      const int nsplit = split_for((int64_t)ntn * (ntn + 1) / 2, kd);
      // split-K layouts: the tile-major split (nsplit by the tile count) or the split by
      // XCD (16 splits, each XCD's tiles on the same k-window of T, read once per XCD)
      const int nsx = 16;
      if (mixed) {
        GemmArgs g{nao, nao, kd, h->d_T32, kd, true, h->d_T32, kd, true};
        const int v = tuned_product(h, 6, nao, kd, nsplit, 2, 1, [&](int c) {
          HIPCHK(launch_gemm_partials(kBF16x3, g, c == 1 ? nsx : nsplit, c == 1 ? 3 : 1, h->d_gws, 0, st));
        }, st);
        HIPCHK(launch_gemm_reduce(kBF16x3, g, h->d_K, sg, beta, v == 1 ? nsx : nsplit, 1, h->d_gws, st));
      } else {
        k_product_f64(h, kd, h->d_K, sg, beta, st);
      }
    }
    h->mark(kStExchangeSyrk + 1, st);
    first = false;
    i0 = i1;
  }
  // J = B gamma
  if (h->recompute) coulomb_recompute(h, st);
  // B = (mn|P) L is symmetric in (m, n) (the three-centre kernels write both orders):
  // the upper rows only, J mirrored
  else if (mixed) HIPCHK(launch_jvec_f32(nn, nfit, h->d_B32, h->d_gamma, h->d_J, nao, st));
  else HIPCHK(launch_jvec(nn, nfit, h->d_B, h->d_gamma, h->d_J, nao, st));
  h->mark(kStCoulomb + 1, st);
}

// XC on the compact grid, bin by bin: Pg = P gathered to each block's columns,
// Z_b = Phi_b Pg_b (batched), the pointwise kernel on the compact rows, then
// prod_b = Phi_b^T W_b (batched, over Pg's storage) and the row assembly of
// Vh = sum_b prod_b scattered to the blocks' columns.
//   kXcEnergy:   k_xc, E_xc into E + DFKS_E_XC (FP64 or mixed)
//   kXcRho:      rho0 and grad rho0 per padded point into d_rho0 (FP64, no Vh)
//   kXcResponse: Z1 = Phi dP, the response kernel with the records in d_fxc (FP64)
enum XcMode { kXcEnergy = 0, kXcRho = 1, kXcResponse = 2 };

// The table of the fused mixed XC launch (once per handle: the bins are fixed at
// build): every bin up to kXfMaxN wide, its offsets and first workgroup ids; the
// workspace holds all bins' partial sums, then their W coefficients.
void xcf_prepare(dfks_ctx* h, hipStream_t st) {
  if (h->xcf_ready) return;
  h->xcf_tab.clear();
  h->xcf_fused.assign(h->xbins.size(), 0);
  int64_t part = 0, coef = 0, wz = 0, wp = 0, wv = 0;
  // the fused kernels address P with 32-bit offsets: larger systems take the per-bin path
  const bool small_p = (int64_t)h->nao * h->nao <= 0x7fffffff;
  for (size_t i = 0; i < h->xbins.size(); ++i) {
    const auto& bn = h->xbins[i];
    if (!small_p || bn.n > kXfMaxN || bn.n % 32 != 0) continue;
    int nt = 0, nks = 0;
    int64_t pd = 0, cf = 0;
    xcf_bin_geometry(bn.n, (int)bn.nblk, nt, nks, pd, cf);
    XcfBin e{bn.n, (int32_t)bn.nblk, nt, nks, bn.ao_off, bn.col_off, bn.q_off, bn.pt0, part, coef,
             (int32_t)wz, (int32_t)wp, (int32_t)wv, 0};
    h->xcf_tab.push_back(e);
    h->xcf_fused[i] = 1;
    part += pd;
    coef += cf;
    wz += bn.nblk * nt * nks;
    wp += bn.nblk;
    wv += bn.nblk * nt * nt;
  }
  if (wz > 0x7fffffff / 2 || wv > 0x7fffffff / 2) fail(1, "compact XC grid too large for the fused launch");
  h->xcf_ztot = (int)wz;
  h->xcf_ptot = (int)wp;
  h->xcf_vtot = (int)wv;
  h->xcf_part = part;
  h->d_xcf_ws = dalloc<double>((size_t)(part + (coef + 1) / 2 + 2));
  h->d_xcf_tab = dalloc<XcfBin>(std::max<size_t>(1, h->xcf_tab.size()));
  if (!h->xcf_tab.empty())
    HIPCHK(hipMemcpyAsync(h->d_xcf_tab, h->xcf_tab.data(), sizeof(XcfBin) * h->xcf_tab.size(),
                          hipMemcpyHostToDevice, st));
  h->xcf_ready = true;
}

void xc_compact(dfks_ctx* h, const double* P, double* E, int mode, hipStream_t st) {
  const int nao = h->nao;
  const bool mixed = h->mixed_now && mode == kXcEnergy;
  const double one = 1.0, zero = 0.0;
  const double thresh = h->opt.rho_thresh;
  // mixed: the bins gather from an FP32 copy of P (the same rounding as converting each gathered element)
  if (mixed) HIPCHK(launch_to_f32((int64_t)nao * nao, P, h->d_P32, st));
  // mixed, library kernels: every bin up to kXfMaxN wide in one launch of each fused kernel
  const bool fused = mixed && h->opt.gemm != 1;
  if (fused) {
    xcf_prepare(h, st);
    HIPCHK(launch_xcf_all(h->d_xcf_tab, (int)h->xcf_tab.size(), h->xcf_ztot, h->xcf_ptot, h->xcf_vtot,
                          h->d_xc_ao32, h->d_xc_cols, h->d_P32, nao, h->d_xc_w, h->xc_id, thresh, h->d_xc_exc,
                          h->d_xcf_ws, reinterpret_cast<float*>(h->d_xcf_ws + h->xcf_part),
                          reinterpret_cast<float*>(h->d_xc_q), st));
  }
  for (size_t bi = 0; bi < h->xbins.size(); ++bi) {
    if (fused && h->xcf_fused[bi]) continue;
    const auto& bn = h->xbins[bi];
    const int n = bn.n;
    const int64_t nb = bn.nblk, rows = nb * kXcG;
    const int32_t* cols = h->d_xc_cols + bn.col_off;
    if (mixed) {
      // options.gemm = rocblas (or a bin wider than kXfMaxN): FP32 operands and FP32 MFMA products
      // (rocBLAS strided batched)
      const float onef = 1.0f, zerof = 0.0f;
      float* q = reinterpret_cast<float*>(h->d_xc_q) + bn.q_off;
      const float* ao = h->d_xc_ao32 + bn.ao_off;
      float* zw = reinterpret_cast<float*>(h->d_xc_zw);
      HIPCHK((launch_gather_sym<float, float>(nb, n, cols, h->d_P32, nao, q, st)));
      RBCHK(rocblas_sgemm_strided_batched(h->rb, rocblas_operation_transpose, rocblas_operation_none, n, kXcG, n,
                                          &onef, q, n, (rocblas_stride)n * n, ao, n, (rocblas_stride)kXcG * n,
                                          &zerof, zw, n, (rocblas_stride)kXcG * n, (int)nb));
      HIPCHK(launch_xc_f32(rows, n, ao, h->d_xc_w + bn.pt0, zw, h->xc_id, thresh, h->d_xc_exc + bn.pt0, st));
      RBCHK(rocblas_sgemm_strided_batched(h->rb, rocblas_operation_none, rocblas_operation_transpose, n, n, kXcG,
                                          &onef, zw, n, (rocblas_stride)kXcG * n, ao, n, (rocblas_stride)kXcG * n,
                                          &zerof, q, n, (rocblas_stride)n * n, (int)nb));
      continue;
    }
    double* q = h->d_xc_q + bn.q_off;
    const double* ao = h->d_xc_ao + bn.ao_off;
    double* zw = h->d_xc_zw;
    HIPCHK((launch_gather_sym<double, double>(nb, n, cols, P, nao, q, st)));
    // Z_b (G x n, row-major) = Phi_b Pg_b; rocBLAS column-major: Z_b^T = Pg_b^T Phi_b^T
    tuned_product(h, 1, n, nb, 0, 2, 0, [&](int choice) {
      if (choice == 1) {
        const GemmArgs g{kXcG, n, n, ao, n, true, q, n, true};
        HIPCHK(launch_gemm_f64_batched(g, zw, (int)nb, (int64_t)kXcG * n, (int64_t)n * n, (int64_t)kXcG * n, st));
      } else {
        RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_transpose, rocblas_operation_none, n, kXcG, n,
                                            &one, q, n, (rocblas_stride)n * n, ao, n, (rocblas_stride)kXcG * n, &zero,
                                            zw, n, (rocblas_stride)kXcG * n, (int)nb));
      }
    }, st);
    if (mode == kXcEnergy)
      HIPCHK(launch_xc(rows, n, ao, h->d_xc_w + bn.pt0, zw, h->xc_id, thresh, h->d_xc_exc + bn.pt0, nullptr, st));
    else if (mode == kXcRho)
      HIPCHK(launch_xc(rows, n, ao, h->d_xc_w + bn.pt0, zw, h->xc_id, thresh, h->d_xc_exc + bn.pt0,
                       h->d_rho0 + 4 * bn.pt0, st));
    else
      HIPCHK(launch_fxc(rows, n, ao, zw, h->d_fxc + 8 * bn.pt0, st));
    if (mode == kXcRho) continue;
    // prod_b[i][j] = sum_g Phi[g][i] W[g][j]; rocBLAS column-major: prod_b^T = W_b^T Phi_b
    tuned_product(h, 2, n, nb, 0, 2, 0, [&](int choice) {
      if (choice == 1) {
        const GemmArgs g{n, n, kXcG, ao, n, false, zw, n, false};
        HIPCHK(launch_gemm_f64_batched(g, q, (int)nb, (int64_t)kXcG * n, (int64_t)kXcG * n, (int64_t)n * n, st));
      } else {
        RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_none, rocblas_operation_transpose, n, n, kXcG,
                                            &one, zw, n, (rocblas_stride)kXcG * n, ao, n, (rocblas_stride)kXcG * n,
                                            &zero, q, n, (rocblas_stride)n * n, (int)nb));
      }
    }, st);
  }
  if (mode == kXcRho) return;
  h->mark(kStXcKernel + 1, st);
  // slices of each row's entries so that about 32768 workgroups share the assembly (at most 32 x nao^2 of
  // partial rows in the split-K workspace, which holds 64 x nao^2)
  const int nslice = std::max(1, std::min(32, (32768 + nao - 1) / nao));
  if (mixed)
    HIPCHK(launch_vassemble<float>(nao, h->d_xc_vrow, h->d_xc_vprow, h->d_xc_vpcol, h->d_xc_vlen,
                                   reinterpret_cast<const float*>(h->d_xc_q), h->d_xc_cols, h->d_Vxc, h->d_gws,
                                   nslice, st));
  else
    HIPCHK(launch_vassemble<double>(nao, h->d_xc_vrow, h->d_xc_vprow, h->d_xc_vpcol, h->d_xc_vlen, h->d_xc_q,
                                    h->d_xc_cols, h->d_Vxc, h->d_gws, nslice, st));
  if (mode == kXcEnergy) HIPCHK(launch_dot(h->xc_npts, h->d_xc_exc, nullptr, h->d_part, 1.0, E + DFKS_E_XC, 0, st));
  h->mark(kStXcSyr2k + 1, st);
}

// Spin-polarised XC on the compact grid (FP64): per bin both channels' Z_s =
// Phi_b Pg_s (rocBLAS batched), the spin pointwise kernel (W_a, W_b over Z_a,
// Z_b), Phi_b^T W_s per channel, then the row assembly of each channel's half
// potential: Vh_a in d_Vxc, Vh_b in d_Vxc2; E_xc into E + DFKS_E_XC.
void xc_compact_uks(dfks_ctx* h, const double* Pa, const double* Pb, double* E, hipStream_t st) {
  const int nao = h->nao;
  const double one = 1.0, zero = 0.0;
  if (!h->d_xc_zw2) h->d_xc_zw2 = dalloc<double>(h->xc_zw);
  if (!h->d_xc_q2) h->d_xc_q2 = dalloc<double>(h->xc_qtot);
  if (!h->d_Vxc2) h->d_Vxc2 = dalloc<double>((size_t)nao * nao);
  for (const auto& bn : h->xbins) {
    const int n = bn.n;
    const int64_t nb = bn.nblk, rows = nb * kXcG;
    const int32_t* cols = h->d_xc_cols + bn.col_off;
    const double* ao = h->d_xc_ao + bn.ao_off;
    double* q[2] = {h->d_xc_q + bn.q_off, h->d_xc_q2 + bn.q_off};
    double* zw[2] = {h->d_xc_zw, h->d_xc_zw2};
    const double* Ps[2] = {Pa, Pb};
    for (int sp = 0; sp < 2; ++sp) {
      HIPCHK((launch_gather_sym<double, double>(nb, n, cols, Ps[sp], nao, q[sp], st)));
      RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_transpose, rocblas_operation_none, n, kXcG, n,
                                          &one, q[sp], n, (rocblas_stride)n * n, ao, n, (rocblas_stride)kXcG * n,
                                          &zero, zw[sp], n, (rocblas_stride)kXcG * n, (int)nb));
    }
    HIPCHK(launch_xcs(rows, n, ao, h->d_xc_w + bn.pt0, zw[0], zw[1], h->xc_id, h->opt.rho_thresh,
                      h->d_xc_exc + bn.pt0, st));
    for (int sp = 0; sp < 2; ++sp)
      RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_none, rocblas_operation_transpose, n, n, kXcG,
                                          &one, zw[sp], n, (rocblas_stride)kXcG * n, ao, n,
                                          (rocblas_stride)kXcG * n, &zero, q[sp], n, (rocblas_stride)n * n,
                                          (int)nb));
  }
  h->mark(kStXcKernel + 1, st);
  const int nslice = std::max(1, std::min(32, (32768 + nao - 1) / nao));
  HIPCHK(launch_vassemble<double>(nao, h->d_xc_vrow, h->d_xc_vprow, h->d_xc_vpcol, h->d_xc_vlen, h->d_xc_q,
                                  h->d_xc_cols, h->d_Vxc, h->d_gws, nslice, st));
  HIPCHK(launch_vassemble<double>(nao, h->d_xc_vrow, h->d_xc_vprow, h->d_xc_vpcol, h->d_xc_vlen, h->d_xc_q2,
                                  h->d_xc_cols, h->d_Vxc2, h->d_gws, nslice, st));
  HIPCHK(launch_dot(h->xc_npts, h->d_xc_exc, nullptr, h->d_part, 1.0, E + DFKS_E_XC, 0, st));
  h->mark(kStXcSyr2k + 1, st);
}

// XC energy (E + DFKS_E_XC) and the half potential Vh (d_Vxc, Vxc = Vh + Vh^T) at P.
void xc_potential(dfks_ctx* h, const double* P, double* E, hipStream_t st) {
  const int nao = h->nao;
  const int64_t nn = (int64_t)nao * nao;
  const int64_t ng = h->ng;
  const double one = 1.0, zero = 0.0;
  const bool mixed = h->mixed_now;
  const bool lib = h->opt.gemm >= 2;
  if (h->xcb) {
    xc_compact(h, P, E, kXcEnergy, st);
    return;
  }
  // XC: Z = Phi P, W (in place), Vh = Phi^T W; V = Vh + Vh^T in the Fock assembly
  if (ng > 0 && h->ao_block > 0) {
    // AO values and gradients per block of points: Z_b = Phi_b P, W_b, Vh += Phi_b^T W_b
    const ShellSet O = h->orb.view();
    const int64_t gb = std::min<int64_t>(h->ao_block, ng);
    if (mixed) HIPCHK(launch_to_f32(nn, P, h->d_P32, st));
    const int ntm = (nao + 127) / 128;
    for (int64_t g0 = 0; g0 < ng; g0 += gb) {
      const int64_t nb = std::min(gb, ng - g0);
      HIPCHK(launch_ao(nb, h->d_gxyz + 3 * g0, O, nao, h->d_aob, st));
      if (mixed) {
        HIPCHK(launch_to_f32(4 * nb * nao, h->d_aob, h->d_aob32, st));
        const GemmArgs gz{(int)nb, nao, nao, h->d_aob32, nao, true, h->d_P32, nao, true};
        HIPCHK(launch_gemm_f32out(gz, h->d_zwb32, st));
        HIPCHK(launch_xc_f32(nb, nao, h->d_aob32, h->d_gw + g0, h->d_zwb32, h->xc_id, h->opt.rho_thresh,
                             h->d_exc + g0, st));
        const GemmArgs gv{nao, nao, nb, h->d_aob32, nao, false, h->d_zwb32, nao, false};
        HIPCHK(launch_gemm(kBF16x3, gv, h->d_Vxc, 1.0, g0 == 0 ? 0.0 : 1.0, split_for((int64_t)ntm * ntm, nb), 0,
                           h->d_gws, st));
        continue;
      }
      RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_none, rocblas_operation_none, nao, nb, nao, &one, P, nao,
                             h->d_aob, nao, &zero, h->d_zwb, nao));
      HIPCHK(launch_xc(nb, nao, h->d_aob, h->d_gw + g0, h->d_zwb, h->xc_id, h->opt.rho_thresh, h->d_exc + g0, nullptr,
                       st));
      const double beta = g0 == 0 ? 0.0 : 1.0;
      RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_none, rocblas_operation_transpose, nao, nao, nb, &one, h->d_aob,
                             nao, h->d_zwb, nao, &beta, h->d_Vxc, nao));
    }
    h->mark(kStXcKernel + 1, st);
    HIPCHK(launch_dot(ng, h->d_exc, nullptr, h->d_part, 1.0, E + DFKS_E_XC, 0, st));
    h->mark(kStXcSyr2k + 1, st);
  } else if (ng > 0) {
    const int ntm = (nao + 127) / 128;
    const int nsplit = split_for((int64_t)ntm * ntm, ng);
    if (mixed) {
      HIPCHK(launch_to_f32(nn, P, h->d_P32, st));
      GemmArgs gz{(int)ng, nao, nao, h->d_ao32, nao, true, h->d_P32, nao, true};
      HIPCHK(launch_gemm_f32out(gz, h->d_zw32, st));
      h->mark(kStXcGemm + 1, st);
      HIPCHK(launch_xc_f32(ng, nao, h->d_ao32, h->d_gw, h->d_zw32, h->xc_id, h->opt.rho_thresh, h->d_exc, st));
      HIPCHK(launch_dot(ng, h->d_exc, nullptr, h->d_part, 1.0, E + DFKS_E_XC, 0, st));
      h->mark(kStXcKernel + 1, st);
      GemmArgs gv{nao, nao, ng, h->d_ao32, nao, false, h->d_zw32, nao, false};
      HIPCHK(launch_gemm(kBF16x3, gv, h->d_Vxc, 1.0, 0.0, nsplit, 0, h->d_gws, st));
    } else if (lib) {
      GemmArgs gz{(int)ng, nao, nao, h->d_ao, nao, true, P, nao, true};
      HIPCHK(launch_gemm(kF64, gz, h->d_zw, 1.0, 0.0, 1, 0, nullptr, st));
      h->mark(kStXcGemm + 1, st);
      HIPCHK(launch_xc(ng, nao, h->d_ao, h->d_gw, h->d_zw, h->xc_id, h->opt.rho_thresh, h->d_exc, nullptr, st));
      HIPCHK(launch_dot(ng, h->d_exc, nullptr, h->d_part, 1.0, E + DFKS_E_XC, 0, st));
      h->mark(kStXcKernel + 1, st);
      GemmArgs gv{nao, nao, ng, h->d_ao, nao, false, h->d_zw, nao, false};
      HIPCHK(launch_gemm(kF64, gv, h->d_Vxc, 1.0, 0.0, nsplit, 0, h->d_gws, st));
    } else {
      RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_none, rocblas_operation_none, nao, ng, nao, &one, P, nao,
                             h->d_ao, nao, &zero, h->d_zw, nao));
      h->mark(kStXcGemm + 1, st);
      HIPCHK(launch_xc(ng, nao, h->d_ao, h->d_gw, h->d_zw, h->xc_id, h->opt.rho_thresh, h->d_exc, nullptr, st));
      HIPCHK(launch_dot(ng, h->d_exc, nullptr, h->d_part, 1.0, E + DFKS_E_XC, 0, st));
      h->mark(kStXcKernel + 1, st);
      if (ntm <= 6) {
        // nao <= 768: few output tiles and K = ng; the split-K library kernel fills the chip
        // where the plain library GEMM cannot (above that size rocBLAS is the faster one)
        GemmArgs gv{nao, nao, ng, h->d_ao, nao, false, h->d_zw, nao, false};
        HIPCHK(launch_gemm(kF64, gv, h->d_Vxc, 1.0, 0.0, nsplit, 0, h->d_gws, st));
      } else {
        RBCHK(rocblas_dgemm_64(h->rb, rocblas_operation_none, rocblas_operation_transpose, nao, nao, ng, &one,
                               h->d_ao, nao, h->d_zw, nao, &zero, h->d_Vxc, nao));
      }
    }
    h->mark(kStXcSyr2k + 1, st);
  }
}

void energy_fock_factored(dfks_ctx* h, const double* P, const double* Cf, const std::vector<double>& sgn, int nf,
                          double* E, double* F, hipStream_t st) {
  if (!h->built) fail(6, "dfks_build has not been called");
  RBCHK(rocblas_set_stream(h->rb, st));
  const int nao = h->nao, nfit = h->nfit;
  const int64_t nn = (int64_t)nao * nao;
  const int64_t ng = h->ng;
  if (h->mixed_now && h->opt.precision != 1) fail(1, "mixed precision needs options.precision = 1 at create");
  h->mark(0, st);
  HIPCHK(hipMemsetAsync(E, 0, 8 * sizeof(double), st));
  // one-electron energy
  HIPCHK(launch_dot(nn, P, h->d_H, h->d_part, 1.0, E + DFKS_E_ONE, 0, st));
  jk_factored(h, Cf, sgn, nf, st);
  if (h->hf_lr != 0.0) k_lr_factored(h, Cf, sgn, nf, st);
  // E_J = 1/2 |gamma|^2
  HIPCHK(launch_dot(nfit, h->d_gamma, h->d_gamma, h->d_part, 0.5, E + DFKS_E_J, 0, st));
  // E_K = -1/4 sum P o (a_x K + a_lr K_lr) (E was cleared above)
  if (h->hf != 0.0) HIPCHK(launch_dot(nn, P, h->d_K, h->d_part, -0.25 * h->hf, E + DFKS_E_K, 1, st));
  if (h->hf_lr != 0.0) HIPCHK(launch_dot(nn, P, h->d_Klr, h->d_part, -0.25 * h->hf_lr, E + DFKS_E_K, 1, st));
  xc_potential(h, P, E, st);
  HIPCHK(launch_fock(nao, h->d_H, h->d_J, h->hf != 0.0 ? h->d_K : nullptr, -0.5 * h->hf,
                     ng > 0 ? h->d_Vxc : nullptr, F, st));
  if (h->hf_lr != 0.0) HIPCHK(launch_axpy(nn, -0.5 * h->hf_lr, h->d_Klr, F, 1, st));
  // totals: E_tot = e1 + eJ + eK + exc + enn
  HIPCHK(hipMemcpyAsync(E + DFKS_E_NN, &h->e_nn, sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(launch_dot(5, E + 1, nullptr, h->d_part, 1.0, E + DFKS_E_TOT, 0, st));
  h->mark(kStFock + 1, st);
}

// signed eigen-factorisation of a symmetric P (positive columns first); signs to sg
int factor_density(dfks_ctx* h, const double* P, std::vector<double>& sg, hipStream_t st) {
  const int nao = h->nao;
  RBCHK(rocblas_set_stream(h->rb, st));
  HIPCHK(hipMemcpyAsync(h->d_W1, P, sizeof(double) * nao * nao, hipMemcpyDeviceToDevice, st));
  RBCHK(rocsolver_dsyevd(h->rb, rocblas_evect_original, rocblas_fill_upper, nao, h->d_W1, nao, h->d_evals, h->d_offd,
                         h->d_info));
  std::vector<double> lam(nao);
  HIPCHK(hipMemcpyAsync(lam.data(), h->d_evals, sizeof(double) * nao, hipMemcpyDeviceToHost, st));
  sync(st);
  double amax = 0.0;
  for (double v : lam) amax = std::max(amax, std::fabs(v));
  std::vector<int32_t> sel;
  sg.clear();
  for (int i = nao - 1; i >= 0; --i)
    if (lam[i] > 1e-14 * amax) { sel.push_back(i); sg.push_back(1.0); }
  for (int i = 0; i < nao; ++i)
    if (lam[i] < -1e-14 * amax) { sel.push_back(i); sg.push_back(-1.0); }
  const int nf = (int)sel.size();
  HIPCHK(hipMemcpyAsync(h->d_sel, sel.data(), sizeof(int32_t) * std::max(nf, 1), hipMemcpyHostToDevice, st));
  HIPCHK(launch_factor_cols(nao, nf, h->d_W1, h->d_evals, h->d_sel, h->d_Cf, h->d_sgn, st));
  sync(st);
  return nf;
}

// Spin-polarised energy and Fock matrices of P = [P_a; P_b] (2 nao^2): the
// one-electron and Coulomb terms of P_a + P_b, exchange -1/2 a_x sum_s P_s . K(P_s)
// (terms.py:500-509), the spin-polarised XC (terms.py:838-874), and
// F_s = H + J(P_a + P_b) - a_x K(P_s) + Vxc_s into F = [F_a; F_b] (FP64).
void energy_fock_uks(dfks_ctx* h, const double* P, double* E, double* F, hipStream_t st) {
  if (!h->built) fail(6, "dfks_build has not been called");
  if (h->mixed_now) fail(1, "nspin = 2 Fock builds run in FP64: call dfks_set_precision(h, 0)");
  if (h->ng > 0 && !h->xcb)
    fail(1, "nspin = 2 needs the compact XC grid (options.xc_blocks = 1 with options.ao_block = 0)");
  RBCHK(rocblas_set_stream(h->rb, st));
  const int nao = h->nao, nfit = h->nfit;
  const int64_t nn = (int64_t)nao * nao;
  if (!h->d_uP) {
    h->d_uP = dalloc<double>(nn);
    h->d_uJ = dalloc<double>(nn);
    h->d_uK = dalloc<double>(2 * nn);
    h->d_ug = dalloc<double>(nfit);
  }
  h->mark(0, st);
  HIPCHK(hipMemsetAsync(E, 0, 8 * sizeof(double), st));
  HIPCHK(launch_axpy(nn, 1.0, P, h->d_uP, 0, st));
  HIPCHK(launch_axpy(nn, 1.0, P + nn, h->d_uP, 1, st));
  HIPCHK(launch_dot(nn, h->d_uP, h->d_H, h->d_part, 1.0, E + DFKS_E_ONE, 0, st));
  HIPCHK(hipMemsetAsync(h->d_uJ, 0, sizeof(double) * nn, st));
  HIPCHK(hipMemsetAsync(h->d_ug, 0, sizeof(double) * nfit, st));
  HIPCHK(hipMemsetAsync(h->d_uK, 0, sizeof(double) * 2 * nn, st));
  for (int sp = 0; sp < 2; ++sp) {
    const double* Ps = P + sp * nn;
    std::vector<double> sg;
    const int nf = factor_density(h, Ps, sg, st);
    if (nf == 0) continue;                       // an empty channel: no gamma, J or K
    jk_factored(h, h->d_Cf, sg, nf, st);
    HIPCHK(launch_axpy(nfit, 1.0, h->d_gamma, h->d_ug, 1, st));
    HIPCHK(launch_axpy(nn, 1.0, h->d_J, h->d_uJ, 1, st));
    if (h->hf != 0.0) {
      HIPCHK(launch_axpy(nn, 1.0, h->d_K, h->d_uK + sp * nn, 0, st));
      HIPCHK(launch_dot(nn, Ps, h->d_K, h->d_part, -0.5 * h->hf, E + DFKS_E_K, 1, st));
    }
  }
  // E_J = 1/2 |gamma_a + gamma_b|^2
  HIPCHK(launch_dot(nfit, h->d_ug, h->d_ug, h->d_part, 0.5, E + DFKS_E_J, 0, st));
  if (h->ng > 0) xc_compact_uks(h, P, P + nn, E, st);
  for (int sp = 0; sp < 2; ++sp)
    HIPCHK(launch_fock(nao, h->d_H, h->d_uJ, h->hf != 0.0 ? h->d_uK + sp * nn : nullptr, -h->hf,
                       h->ng > 0 ? (sp == 0 ? h->d_Vxc : h->d_Vxc2) : nullptr, F + sp * nn, st));
  HIPCHK(hipMemcpyAsync(E + DFKS_E_NN, &h->e_nn, sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(launch_dot(5, E + 1, nullptr, h->d_part, 1.0, E + DFKS_E_TOT, 0, st));
  h->mark(kStFock + 1, st);
}

void energy_fock_general(dfks_ctx* h, const double* P, double* E, double* F, hipStream_t st) {
  std::vector<double> sg;
  int nf = factor_density(h, P, sg, st);
  if (nf == 0) {
    sg.push_back(1.0);
    HIPCHK(hipMemsetAsync(h->d_Cf, 0, sizeof(double) * h->nao, st));
    nf = 1;
  }
  energy_fock_factored(h, P, h->d_Cf, sg, nf, E, F, st);
}

// ---------------------------------------------------------------------------
// SCF
// ---------------------------------------------------------------------------
void gemm(dfks_ctx* h, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
          const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  RBCHK(rocblas_dgemm_64(h->rb, ta ? rocblas_operation_transpose : rocblas_operation_none,
                         tb ? rocblas_operation_transpose : rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb,
                         &beta, C, ldc));
}

// ---------------------------------------------------------------------------
// Response: R = J[dP] - a_x/2 K[dP] + f_xc(P0)[dP], the derivative of
// F = dE/dP at P0 along dP (jax.jvp(jax.grad(E)) of dftax/ks/newton.py:134-136)
// ---------------------------------------------------------------------------
// rho0 and grad rho0 at the reference density P0, then the per-point records.
void response_begin(dfks_ctx* h, const double* P, hipStream_t st) {
  if (!h->built) fail(6, "dfks_build has not been called");
  if (h->mixed_now) fail(1, "the response runs in FP64: call dfks_set_precision(h, 0) first");
  RBCHK(rocblas_set_stream(h->rb, st));
  h->resp_ready = false;
  const int nao = h->nao;
  const int64_t ng = h->ng;
  const double thresh = h->opt.rho_thresh;
  if (ng > 0) {
    const int64_t npt = h->xcb ? h->xc_npts : ng;
    if (!h->d_rho0) h->d_rho0 = dalloc<double>(4 * (size_t)npt);
    if (!h->d_fxc) h->d_fxc = dalloc<double>(8 * (size_t)npt);
    if (h->xcb) {
      xc_compact(h, P, nullptr, kXcRho, st);
      HIPCHK(launch_fxc_coef(npt, h->d_rho0, h->d_xc_w, h->xc_id, thresh, h->d_fxc, st));
      h->resp_ready = true;
      return;
    }
    if (h->ao_block == 0) {
      gemm(h, false, false, nao, ng, nao, 1.0, P, nao, h->d_ao, nao, 0.0, h->d_zw, nao);
      HIPCHK(launch_xc(ng, nao, h->d_ao, h->d_gw, h->d_zw, h->xc_id, thresh, h->d_exc, h->d_rho0, st));
    } else {
      const ShellSet O = h->orb.view();
      const int64_t gb = std::min<int64_t>(h->ao_block, ng);
      for (int64_t g0 = 0; g0 < ng; g0 += gb) {
        const int64_t nb = std::min(gb, ng - g0);
        HIPCHK(launch_ao(nb, h->d_gxyz + 3 * g0, O, nao, h->d_aob, st));
        gemm(h, false, false, nao, nb, nao, 1.0, P, nao, h->d_aob, nao, 0.0, h->d_zwb, nao);
        HIPCHK(launch_xc(nb, nao, h->d_aob, h->d_gw + g0, h->d_zwb, h->xc_id, thresh, h->d_exc + g0,
                         h->d_rho0 + 4 * g0, st));
      }
    }
    HIPCHK(launch_fxc_coef(ng, h->d_rho0, h->d_gw, h->xc_id, thresh, h->d_fxc, st));
  }
  h->resp_ready = true;
}

// R for dP = U V^T + V U^T (U, V column-major nao x r). J and K come from the
// signed factor dP = 1/2 (U+V)(U+V)^T - 1/2 (U-V)(U-V)^T; the XC part from
// Z1 = Phi dP and the per-point records.
void response_fock(dfks_ctx* h, const double* U, const double* V, int r, double* R, hipStream_t st) {
  if (!h->resp_ready) fail(6, "dfks_response_begin has not been called");
  if (h->mixed_now) fail(1, "the response runs in FP64: call dfks_set_precision(h, 0) first");
  const int nao = h->nao;
  const int64_t ng = h->ng;
  if (r < 1 || r > nao) fail(1, "r must be in 1..nao");
  RBCHK(rocblas_set_stream(h->rb, st));
  if (h->rf_cap < 2 * r) {
    dfree(h->d_Rf);
    h->d_Rf = dalloc<double>((size_t)nao * 2 * r);
    h->rf_cap = 2 * r;
  }
  const double c = std::sqrt(0.5), mc = -c;
  RBCHK(rocblas_dgeam(h->rb, rocblas_operation_none, rocblas_operation_none, nao, r, &c, U, nao, &c, V, nao, h->d_Rf,
                      nao));
  RBCHK(rocblas_dgeam(h->rb, rocblas_operation_none, rocblas_operation_none, nao, r, &c, U, nao, &mc, V, nao,
                      h->d_Rf + (int64_t)nao * r, nao));
  std::vector<double> sg(2 * r, 1.0);
  for (int i = r; i < 2 * r; ++i) sg[i] = -1.0;
  jk_factored(h, h->d_Rf, sg, 2 * r, st);
  if (ng > 0) {
    // dense dP, then V1 = Phi^T W1 (the assembly adds its transpose)
    gemm(h, false, true, nao, nao, r, 1.0, U, nao, V, nao, 0.0, h->d_W1, nao);
    gemm(h, false, true, nao, nao, r, 1.0, V, nao, U, nao, 1.0, h->d_W1, nao);
    if (h->xcb) {
      xc_compact(h, h->d_W1, nullptr, kXcResponse, st);
    } else if (h->ao_block == 0) {
      gemm(h, false, false, nao, ng, nao, 1.0, h->d_W1, nao, h->d_ao, nao, 0.0, h->d_zw, nao);
      HIPCHK(launch_fxc(ng, nao, h->d_ao, h->d_zw, h->d_fxc, st));
      gemm(h, false, true, nao, nao, ng, 1.0, h->d_ao, nao, h->d_zw, nao, 0.0, h->d_Vxc, nao);
    } else {
      const ShellSet O = h->orb.view();
      const int64_t gb = std::min<int64_t>(h->ao_block, ng);
      for (int64_t g0 = 0; g0 < ng; g0 += gb) {
        const int64_t nb = std::min(gb, ng - g0);
        HIPCHK(launch_ao(nb, h->d_gxyz + 3 * g0, O, nao, h->d_aob, st));
        gemm(h, false, false, nao, nb, nao, 1.0, h->d_W1, nao, h->d_aob, nao, 0.0, h->d_zwb, nao);
        HIPCHK(launch_fxc(nb, nao, h->d_aob, h->d_zwb, h->d_fxc + 8 * g0, st));
        gemm(h, false, true, nao, nao, nb, 1.0, h->d_aob, nao, h->d_zwb, nao, g0 == 0 ? 0.0 : 1.0, h->d_Vxc, nao);
      }
    }
  }
  HIPCHK(launch_fock(nao, nullptr, h->d_J, h->hf != 0.0 ? h->d_K : nullptr, -0.5 * h->hf,
                     ng > 0 ? h->d_Vxc : nullptr, R, st));
}

// ---------------------------------------------------------------------------
// Nuclear gradient dE/dR at a frozen density (dftax/ks/forces.py:128-270): the
// energy is differentiated with the basis and auxiliary centres and the Becke
// grid following their atoms, and P = 2 Z (Z^T S Z)^-1 Z^T held at the
// converged orbitals, so the density's response to S(R) is the energy-weighted
// term -Tr(W dS), W = 1/2 P F P. Parts, each with its share of that term
// (W_k = 1/2 P F_k P): [0] nuclear repulsion, [1] one-electron (F = hcore),
// [2] Coulomb (F = J), [3] exchange (F = -a_x/2 K), [4] XC (F = Vxc).
// Density-fitting weights, with d = V+ gamma, (ij|Q) of the orbital factor Cf
// (P = Cf Cf^T) and D^P_ij = sum_Q V+_PQ (ij|Q):
//   3-centre  Gamma_P = d_P P - a_x/2 Cf D^P Cf^T,
//   2-centre  Gamma2  = -1/2 d d^T + a_x/4 sum_ij D_ij D_ij^T,
// the metric's derivative taken as d(V+) = -V+ dV V+ (terms.py:195-204).
// ---------------------------------------------------------------------------
const double kOne = 1.0, kZero = 0.0;   // rocBLAS scalars (host pointers)

__global__ __launch_bounds__(256) void k_bcast_scale(int64_t nrep, int64_t n, const double* __restrict__ d,
                                                     const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nrep * n) return;
  y[i] = d[i / n] * x[i % n];
}

// atom of each shell, by its centre
std::vector<int32_t> shell_atoms(const Basis& B, const std::vector<double>& atm4, int natom) {
  std::vector<int32_t> out(B.n(), -1);
  for (int s = 0; s < B.n(); ++s) {
    for (int a = 0; a < natom; ++a) {
      double d2 = 0.0;
      for (int x = 0; x < 3; ++x) {
        const double d = B.centre[3 * s + x] - atm4[4 * a + x];
        d2 += d * d;
      }
      if (d2 < 1e-20) { out[s] = a; break; }
    }
    if (out[s] < 0) fail(1, "a basis shell is not centred on an atom; forces need atom-centred shells");
  }
  return out;
}

void forces_run(dfks_ctx* h, const double* Co, int nocc, int64_t ng, const double* gxyz, const double* graw,
                const int32_t* gown, const double* bragg, double* out, hipStream_t st) {
  if (!h->built) fail(6, "dfks_build has not been called");
  if (h->recompute || !h->d_B) fail(1, "forces need the materialised fitted tensor (df(chunk=None))");
  if (nocc != h->nocc) fail(1, "nocc does not match the handle");
  if (nocc > h->nf_chunk) fail(1, "too many occupied orbitals for the exchange intermediate");
  RBCHK(rocblas_set_stream(h->rb, st));
  const bool mixed_was = h->mixed_now;
  h->mixed_now = false;
  struct Restore {
    dfks_ctx* h;
    bool m;
    ~Restore() { h->mixed_now = m; }
  } restore{h, mixed_was};
  const int nao = h->nao, naux = h->naux, nfit = h->nfit, natom = h->natom;
  const int64_t nn = (int64_t)nao * nao, na3 = 3 * (int64_t)natom;
  const double ax = h->hf;
  std::vector<void*> owned;
  struct Free {
    std::vector<void*>& p;
    ~Free() { for (void* q : p) dfree(q); }
  } fr{owned};
  auto alloc = [&](size_t n) { double* p = dalloc<double>(n); owned.push_back(p); return p; };
  auto alloc_i = [&](size_t n) { int32_t* p = dalloc<int32_t>(n); owned.push_back(p); return p; };
  auto up = [&](auto* dst, const auto* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(*src), hipMemcpyHostToDevice, st));
  };

  // density, Fock parts and the energy-weighted densities
  double* Cf = alloc((size_t)nao * nocc);
  HIPCHK(launch_axpy((int64_t)nao * nocc, std::sqrt(2.0), Co, Cf, 0, st));
  double* P = alloc(nn);
  gemm(h, false, true, nao, nao, nocc, 1.0, Cf, nao, Cf, nao, 0.0, P, nao);
  double* F = alloc(nn);
  double* E = alloc(8);
  energy_fock_factored(h, P, Cf, std::vector<double>(nocc, 1.0), nocc, E, F, st);
  double* Fk = alloc(nn);
  double* tmp = alloc(nn);
  double* zero = alloc(nn);
  HIPCHK(hipMemsetAsync(zero, 0, sizeof(double) * nn, st));
  double* Wk[4];
  for (int k = 0; k < 4; ++k) {
    Wk[k] = alloc(nn);
    // the term's Fock contribution, assembled as the Fock build does (K symmetrised,
    // Vxc = Vh + Vh^T from the half potential the XC pass leaves)
    if (k == 0) HIPCHK(launch_fock(nao, h->d_H, zero, nullptr, 0.0, nullptr, Fk, st));
    else if (k == 1) HIPCHK(launch_fock(nao, nullptr, h->d_J, nullptr, 0.0, nullptr, Fk, st));
    else if (k == 2) HIPCHK(launch_fock(nao, nullptr, zero, ax != 0.0 ? h->d_K : nullptr, -0.5 * ax, nullptr, Fk, st));
    else HIPCHK(launch_fock(nao, nullptr, zero, nullptr, 0.0, h->ng > 0 ? h->d_Vxc : nullptr, Fk, st));
    gemm(h, false, false, nao, nao, nao, 1.0, Fk, nao, P, nao, 0.0, tmp, nao);
    gemm(h, false, false, nao, nao, nao, 0.5, P, nao, tmp, nao, 0.0, Wk[k], nao);
  }
  double* g = alloc(4 * na3);   // parts 1..4 on the device
  HIPCHK(hipMemsetAsync(g, 0, sizeof(double) * 4 * na3, st));
  double* gp[4] = {g, g + na3, g + 2 * na3, g + 3 * na3};

  // atoms of the shells, device copies
  const std::vector<int32_t> at_o = shell_atoms(h->orb, h->atm_host, natom);
  const std::vector<int32_t> at_x = shell_atoms(h->aux, h->atm_host, natom);
  int32_t* d_at_o = alloc_i(at_o.size());
  int32_t* d_at_x = alloc_i(at_x.size());
  up(d_at_o, at_o.data(), at_o.size());
  up(d_at_x, at_x.data(), at_x.size());
  const ShellSet O = h->orb.view(), X = h->aux.view();

  // one-electron and the S-derivative shares of every part
  HIPCHK(launch_int1e_grad(O, d_at_o, natom, h->d_atm, h->d_boys, P, Wk[0], nao, gp[0], st));
  for (int k = 1; k < 4; ++k)
    HIPCHK(launch_int1e_grad(O, d_at_o, natom, h->d_atm, h->d_boys, zero, Wk[k], nao, gp[k], st));

  // density fitting: d = L gamma', (ij|Q) of Cf, D = L (ij|Q)
  double* dvec = alloc(naux);
  RBCHK(rocblas_dgemv(h->rb, rocblas_operation_transpose, nfit, naux, &kOne, h->d_Lt, nfit, h->d_gamma, 1, &kZero,
                      dvec, 1));
  const int64_t no2 = (int64_t)nocc * nocc;
  double* Aq = alloc((size_t)nfit * no2);   // [Q][i][j]
  RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_none, rocblas_operation_none, nocc, nocc, nao, &kOne,
                                      h->d_Tx, (int64_t)nfit * nocc, nocc, Cf, nao, 0, &kZero, Aq, nocc, no2, nfit));
  double* D = alloc((size_t)naux * no2);    // [P][i][j]
  gemm(h, false, false, no2, naux, nfit, 1.0, Aq, no2, h->d_Lt, nfit, 0.0, D, no2);
  // pairs (ia >= ib) of orbital shells, and the auxiliary shells with a unit partner
  std::vector<int2> pr;
  for (int a = 0; a < h->orb.n(); ++a)
    for (int b = 0; b <= a; ++b) pr.push_back(make_int2(a, b));
  std::vector<int2> apr;
  for (int c = 0; c < h->aux.n(); ++c) apr.push_back(make_int2(c, c));
  int2* d_pr = reinterpret_cast<int2*>(alloc_i(2 * pr.size()));
  int2* d_apr = reinterpret_cast<int2*>(alloc_i(2 * apr.size()));
  up(d_pr, pr.data(), pr.size());
  up(d_apr, apr.data(), apr.size());
  double* G2 = alloc((size_t)naux * naux);
  double* G3 = alloc((size_t)naux * nn);
  // Coulomb: Gamma_P = d_P P, Gamma2 = -1/2 d d^T
  {
    const int64_t tot = (int64_t)naux * nn;
    hipLaunchKernelGGL(k_bcast_scale, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, (int64_t)naux, nn, dvec,
                       P, G3);
    HIPCHK(hipGetLastError());
    gemm(h, false, true, naux, naux, 1, -0.5, dvec, naux, dvec, naux, 0.0, G2, naux);
    HIPCHK(launch_eri_grad(d_pr, (int)pr.size(), h->aux.n(), O, X, d_at_o, d_at_x, 0, h->d_boys, G3, nao, gp[1], st));
    HIPCHK(launch_eri_grad(d_apr, (int)apr.size(), h->aux.n(), X, X, d_at_x, d_at_x, 1, h->d_boys, G2, naux, gp[1],
                           st));
  }
  // exchange: Gamma_P = -a_x/2 Cf D^P Cf^T, Gamma2 = a_x/4 D D^T
  if (ax != 0.0) {
    double* Xp = alloc((size_t)naux * nao * nocc);
    RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_none, rocblas_operation_none, nao, nocc, nocc, &kOne,
                                        Cf, nao, 0, D, nocc, no2, &kZero, Xp, nao, (int64_t)nao * nocc, naux));
    const double sk = -0.5 * ax;
    RBCHK(rocblas_dgemm_strided_batched(h->rb, rocblas_operation_none, rocblas_operation_transpose, nao, nao, nocc,
                                        &sk, Xp, nao, (int64_t)nao * nocc, Cf, nao, 0, &kZero, G3, nao, nn, naux));
    gemm(h, true, false, naux, naux, no2, 0.25 * ax, D, no2, D, no2, 0.0, G2, naux);
    HIPCHK(launch_eri_grad(d_pr, (int)pr.size(), h->aux.n(), O, X, d_at_o, d_at_x, 0, h->d_boys, G3, nao, gp[2], st));
    HIPCHK(launch_eri_grad(d_apr, (int)apr.size(), h->aux.n(), X, X, d_at_x, d_at_x, 1, h->d_boys, G2, naux, gp[2],
                           st));
  }

  // XC on the caller's atom-centred grid (no weight cutoff), in chunks of points
  if (ng > 0 && h->xc_id > 0) {
    std::vector<double> invR((size_t)natom * natom), aAB((size_t)natom * natom), atm3(3 * (size_t)natom);
    for (int a = 0; a < natom; ++a) {
      for (int x = 0; x < 3; ++x) atm3[3 * a + x] = h->atm_host[4 * a + x];
      for (int b = 0; b < natom; ++b) {
        double d2 = 0.0;
        for (int x = 0; x < 3; ++x) {
          const double d = h->atm_host[4 * a + x] - h->atm_host[4 * b + x];
          d2 += d * d;
        }
        invR[(size_t)a * natom + b] = 1.0 / std::sqrt(d2 + (a == b ? 1.0 : 0.0));
        const double chi = bragg[a] / bragg[b], u = (chi - 1.0) / (chi + 1.0);
        aAB[(size_t)a * natom + b] = std::min(0.45, std::max(-0.45, u / (u * u - 1.0)));
      }
    }
    for (int64_t i = 0; i < ng; ++i)
      if (gown[i] < 0 || gown[i] >= natom) fail(1, "grid owner atom index out of range");
    double* d_pts = alloc(3 * (size_t)ng);
    double* d_raw = alloc(ng);
    double* d_w = alloc(ng);
    int32_t* d_own = alloc_i(ng);
    double* d_atm3 = alloc(atm3.size());
    double* d_invR = alloc(invR.size());
    double* d_aAB = alloc(aAB.size());
    up(d_pts, gxyz, 3 * (size_t)ng);
    up(d_raw, graw, ng);
    up(d_own, gown, ng);
    up(d_atm3, atm3.data(), atm3.size());
    up(d_invR, invR.data(), invR.size());
    up(d_aAB, aAB.data(), aAB.size());
    HIPCHK(launch_becke(ng, d_pts, d_raw, d_own, natom, d_atm3, d_invR, d_aAB, d_w, st));
    const int64_t chunk = std::max<int64_t>(256, std::min<int64_t>(ng, (int64_t)(1.5e9 / (8.0 * 16.0 * nao))));
    double* ao = alloc(10 * (size_t)chunk * nao);
    double* Z = alloc((size_t)chunk * nao);
    double* Xg = alloc((size_t)chunk * nao);
    double* q = alloc(3 * (size_t)chunk * nao);
    double* rho = alloc(chunk);
    double* grad = alloc(3 * (size_t)chunk);
    double* sigma = alloc(chunk);
    double* xco = alloc(3 * (size_t)chunk);
    double* fm = alloc(chunk);
    double* Sm = alloc(3 * (size_t)nao);
    HIPCHK(hipMemsetAsync(Sm, 0, sizeof(double) * 3 * nao, st));
    for (int64_t g0 = 0; g0 < ng; g0 += chunk) {
      const int64_t nc = std::min(chunk, ng - g0);
      HIPCHK(launch_ao_hess(nc, d_pts + 3 * g0, O, nao, ao, st));
      // Z = Phi P and X = (grad rho . grad Phi) P, row-major nc x nao
      gemm(h, false, false, nao, nc, nao, 1.0, P, nao, ao, nao, 0.0, Z, nao);
      HIPCHK(launch_xcg_rho(nc, nao, ao, Z, rho, grad, sigma, st));
      HIPCHK(launch_xc_pointwise(nc, h->xc_id, rho, sigma, xco, st));
      HIPCHK(launch_xcg_gdot(nc, nao, ao, grad, q, st));
      gemm(h, false, false, nao, nc, nao, 1.0, P, nao, q, nao, 0.0, Xg, nao);
      HIPCHK(launch_xcg_q(nc, nao, ao, Z, Xg, rho, grad, xco, d_w + g0, d_own + g0, h->opt.rho_thresh, q, gp[3], fm,
                          st));
      for (int j = 0; j < 3; ++j)
        RBCHK(rocblas_dgemv(h->rb, rocblas_operation_none, nao, nc, &kOne, q + (int64_t)j * nc * nao, nao, d_w + g0, 1,
                            &kOne, Sm + (int64_t)j * nao, 1));
      HIPCHK(launch_becke_grad(nc, d_pts + 3 * g0, d_raw + g0, d_own + g0, fm, natom, d_atm3, d_invR, d_aAB, gp[3],
                               st));
    }
    // basis part: -2 sum over the functions of each atom
    std::vector<double> hS(3 * (size_t)nao), hg(na3);
    HIPCHK(hipMemcpyAsync(hS.data(), Sm, sizeof(double) * 3 * nao, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hg.data(), gp[3], sizeof(double) * na3, hipMemcpyDeviceToHost, st));
    sync(st);
    for (int s = 0; s < h->orb.n(); ++s) {
      const int a = at_o[s];
      for (int f = 0; f < nfun(h->orb.l[s], h->orb.spherical); ++f)
        for (int j = 0; j < 3; ++j) hg[3 * a + j] -= 2.0 * hS[(size_t)j * nao + h->orb.ao0[s] + f];
    }
    up(gp[3], hg.data(), na3);
  }
  // out: [0] nuclear repulsion (host), [1..4] from the device
  std::vector<double> hall(4 * na3);
  HIPCHK(hipMemcpyAsync(hall.data(), g, sizeof(double) * 4 * na3, hipMemcpyDeviceToHost, st));
  sync(st);
  for (int a = 0; a < natom; ++a)
    for (int x = 0; x < 3; ++x) {
      double v = 0.0;
      for (int b = 0; b < natom; ++b) {
        if (b == a) continue;
        double d2 = 0.0, dx = h->atm_host[4 * a + x] - h->atm_host[4 * b + x];
        for (int y = 0; y < 3; ++y) {
          const double d = h->atm_host[4 * a + y] - h->atm_host[4 * b + y];
          d2 += d * d;
        }
        v -= h->atm_host[4 * a + 3] * h->atm_host[4 * b + 3] * dx / (d2 * std::sqrt(d2));
      }
      out[3 * a + x] = v;
    }
  for (int64_t i = 0; i < 4 * na3; ++i) out[na3 + i] = hall[i];
}

void scf_begin(dfks_ctx* h, const double* P0, double lindep, int m, double level_shift, hipStream_t st) {
  if (!h->built) fail(6, "dfks_build has not been called");
  if (m < 1 || m > 16) fail(1, "diis_space must be in 1..16");
  RBCHK(rocblas_set_stream(h->rb, st));
  const int nao = h->nao;
  const int64_t nn = (int64_t)nao * nao;
  // canonical orthonormaliser X = U_keep s_keep^{-1/2} (dftax/ks/scf.py:72-86)
  HIPCHK(hipMemcpyAsync(h->d_W1, h->d_S, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
  RBCHK(rocsolver_dsyevd(h->rb, rocblas_evect_original, rocblas_fill_upper, nao, h->d_W1, nao, h->d_evals, h->d_offd,
                         h->d_info));
  std::vector<double> s(nao);
  HIPCHK(hipMemcpyAsync(s.data(), h->d_evals, sizeof(double) * nao, hipMemcpyDeviceToHost, st));
  sync(st);
  int k0 = 0;
  while (k0 < nao && !(s[k0] > lindep)) ++k0;
  const int nmo = nao - k0;
  if (nmo < h->nocc) fail(7, "orthonormaliser keeps fewer orbitals than occupied");
  h->nmo = nmo;
  h->m = m;
  h->it = 0;
  h->diis_base = 0;
  h->restart_diis = false;
  h->level_shift = level_shift;
  auto realloc = [](double*& p, size_t n) { dfree(p); p = dalloc<double>(n); };
  realloc(h->d_X, (size_t)nao * nmo);
  HIPCHK(launch_scaled_cols(nao, k0, h->d_W1, h->d_evals, h->d_X, st));
  realloc(h->d_P, nn);
  realloc(h->d_F, nn);
  realloc(h->d_Fx, nn);
  realloc(h->d_Fs, (size_t)m * nn);
  realloc(h->d_Es, (size_t)m * nmo * nmo);
  realloc(h->d_gram, (size_t)m * m);
  realloc(h->d_coef, m);
  realloc(h->d_dots, m);
  realloc(h->d_C, (size_t)nao * nmo);
  realloc(h->d_eps, nmo);
  realloc(h->d_Ws, (size_t)nao * nao);
  realloc(h->d_lohi, 2 * (size_t)nmo);
  realloc(h->d_Fp, (size_t)nmo * nmo);
  realloc(h->d_pb, 2);
  h->pb_valid_n = -1;
  realloc(h->d_dvec, nmo);
  realloc(h->d_L, (size_t)nmo * h->nocc);
  realloc(h->d_Lw, (size_t)nmo * h->nocc);
  realloc(h->d_Lq, (size_t)nmo * h->nocc);
  realloc(h->d_Ls, (size_t)h->nocc * h->nocc);
  h->l_valid = false;
  h->p_fac = nullptr;
  {
    // load rocSOLVER's Cholesky here rather than in a timed step: a 1 x 1 factor of 1
    const double one = 1.0;
    HIPCHK(hipMemcpyAsync(h->d_Ls, &one, sizeof(double), hipMemcpyHostToDevice, st));
    RBCHK(rocsolver_dpotrf(h->rb, rocblas_fill_lower, 1, h->d_Ls, 1, h->d_info));
    sync(st);
  }
  dfree(h->d_piv);
  h->d_piv = dalloc<int32_t>(h->nocc);
  h->sp2_guess = 24;
  h->c_stale = false;
  h->n_eigh_steps = 0;
  h->n_pchol_steps = 0;
  HIPCHK(hipMemsetAsync(h->d_gram, 0, sizeof(double) * m * m, st));
  HIPCHK(hipMemsetAsync(h->d_Fs, 0, sizeof(double) * m * nn, st));
  HIPCHK(hipMemsetAsync(h->d_Es, 0, sizeof(double) * m * nmo * nmo, st));
  HIPCHK(hipMemsetAsync(h->d_coef, 0, sizeof(double) * m, st));
  HIPCHK(hipMemcpyAsync(h->d_P, P0, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
  energy_fock_general(h, h->d_P, h->d_Edev, h->d_F, st);
  double e[8];
  HIPCHK(hipMemcpyAsync(e, h->d_Edev, sizeof(e), hipMemcpyDeviceToHost, st));
  sync(st);
  h->e_prev = e[0];
}

// Split-K factor of the purification product X X (n x n, upper 128 x 128 tiles, one
// 512-thread workgroup per CU): the s in 1..8 (K chunks of at least 64) that minimises
// rounds x (chunk + 128), rounds = ceil(tiles s / ncu) and 128 the fixed cost of a
// workgroup (pipeline fill, partial-tile epilogue) in units of K. n = 1376 (66 tiles,
// 256 CUs): 3, one round of 198 workgroups (2 used half the CUs, 4 needs a second
// round for 8 of them).
int sp2_split(int n) {
  static const int ncu = [] {
    int dev = 0, c = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1)
      c = 256;
    return c;
  }();
  const int64_t ntn = (n + 127) / 128, tiles = ntn * (ntn + 1) / 2;
  int best = 1;
  double best_t = 1e300;
  for (int sp = 1; sp <= 8 && (sp == 1 || n / sp >= 64); ++sp) {
    const double t = (double)((tiles * sp + ncu - 1) / ncu) * ((n + sp - 1) / sp + 128);
    if (t < best_t - 1e-12) {
      best_t = t;
      best = sp;
    }
  }
  return best;
}

// kSp2Block purification iterations, launched eagerly. Up to kSp2Groups upper 64 x 64
// tiles (n <= 2015): one fused kernel per iteration (product, update, trace sums) that
// alternates X between d_W1 and the split-K workspace, so the block (an even number of
// iterations) leaves X in d_W1. Larger n: the split-K product's partial tiles, then the
// update kernel.
hipError_t sp2_block(dfks_ctx* h, hipStream_t st) {
  const int n = h->nmo;
  static_assert(kSp2Block % 2 == 0, "the fused iterations end in d_W1");
  if (sp2_step_tiles(n) <= kSp2Groups) {
    for (int c = 0; c < kSp2Block; ++c) {
      const double* X = c % 2 == 0 ? h->d_W1 : h->d_gws;
      double* Xn = c % 2 == 0 ? h->d_gws : h->d_W1;
      const hipError_t e = launch_sp2_step(n, X, Xn, h->nocc, h->d_sp2, st);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  const int nsplit = sp2_split(n);
  const GemmArgs g{n, n, n, h->d_W1, n, true, h->d_W1, n, true};
  for (int c = 0; c < kSp2Block; ++c) {
    hipError_t e = launch_gemm_partials(kF64, g, nsplit, 1, h->d_gws, h->sp2_variant, st);
    if (e == hipSuccess) e = launch_sp2_update(n, nsplit, h->d_gws, h->d_W1, h->nocc, h->d_sp2, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// The same block built once per handle and nmo as a HIP graph of kernel
// nodes (fixed buffers, no memset or copy nodes); one graph launch replaces
// 4 kSp2Block kernel launches. The graph is recorded node by node (launch_k
// with a GraphRec installed), not by stream capture: a stream in capture mode
// would make other threads' legacy-stream work fail while it lasts. Falls
// back to eager launches if the graph cannot be built.
void sp2_run_blocks(dfks_ctx* h, int nblocks, hipStream_t st) {
  if (h->sp2_exec_n != h->nmo) {
    if (h->sp2_exec) (void)hipGraphExecDestroy(h->sp2_exec);
    h->sp2_exec = nullptr;
    h->sp2_exec_n = h->nmo;
    // the products' tile, timed eagerly on the current X (the partials are recomputed by the graph)
    const int n = h->nmo;
    if (sp2_step_tiles(n) > kSp2Groups) {
      const int nsplit = sp2_split(n);
      const GemmArgs g{n, n, n, h->d_W1, n, true, h->d_W1, n, true};
      h->sp2_variant = tuned_product(h, 5, n, n, nsplit, 2, 1, [&](int c) {
        HIPCHK(launch_gemm_partials(kF64, g, nsplit, 1, h->d_gws, c, st));
      }, st);
    }
    hipGraph_t gr = nullptr;
    if (hipGraphCreate(&gr, 0) == hipSuccess && gr) {
      GraphRec rec;
      rec.g = gr;
      g_rec = &rec;
      const hipError_t e1 = sp2_block(h, st);
      g_rec = nullptr;
      if (e1 == hipSuccess && rec.err == hipSuccess &&
          hipGraphInstantiate(&h->sp2_exec, gr, nullptr, nullptr, 0) != hipSuccess)
        h->sp2_exec = nullptr;
      (void)hipGraphDestroy(gr);
      (void)hipGetLastError();
    }
  }
  for (int b = 0; b < nblocks; ++b) {
    if (h->sp2_exec) HIPCHK(hipGraphLaunch(h->sp2_exec, st));
    else HIPCHK(sp2_block(h, st));
  }
}

// Lanczos spectral bounds of F' = d_Ws into the purification state (after the
// Gershgorin bounds), replayed as one graph per handle and nmo (49 small kernels).
void sp2_bounds(dfks_ctx* h, hipStream_t st) {
  const int n = h->nmo;
  if (n > 4 * kLzParts) return;   // beyond the partial-sum buffer: Gershgorin bounds only
  if (h->lz_exec_n != n) {
    if (h->lz_exec) (void)hipGraphExecDestroy(h->lz_exec);
    h->lz_exec = nullptr;
    h->lz_exec_n = n;
    hipGraph_t gr = nullptr;
    if (hipGraphCreate(&gr, 0) == hipSuccess && gr) {
      GraphRec rec;
      rec.g = gr;
      g_rec = &rec;
      const hipError_t e1 = launch_sp2_lanczos(n, h->d_Ws, h->d_lz, h->d_sp2, st);
      g_rec = nullptr;
      if (e1 == hipSuccess && rec.err == hipSuccess &&
          hipGraphInstantiate(&h->lz_exec, gr, nullptr, nullptr, 0) != hipSuccess)
        h->lz_exec = nullptr;
      (void)hipGraphDestroy(gr);
      (void)hipGetLastError();
    }
  }
  if (h->lz_exec) HIPCHK(hipGraphLaunch(h->lz_exec, st));
  else HIPCHK(launch_sp2_lanczos(n, h->d_Ws, h->d_lz, h->d_sp2, st));
}

// The occupied factor of the purified X from the previous step's factor L_p
// (orthonormal columns: L_p L_p^T was a rank-nocc projector): W = X L_p,
// S = W^T W = R R^T (Cholesky), L = W R^{-T}. Then L L^T = W S^{-1} W^T is the
// orthogonal projector onto range(X L_p), which is range(X) while X L_p has
// full rank, i.e. L L^T = X, and L^T L = I. Rounding in L grows with the
// condition of S (whose eigenvalues are the squared cosines of the angles
// between the two occupied spaces), and so does the deviation of the computed
// L^T L from I: the factor is taken when max |L^T L - I| <= 1e-10 (S
// conditioned to ~1e6 or better), the pivoted Cholesky runs otherwise. Two
// GEMMs of nmo^2 nocc and nmo nocc^2 and a Cholesky of nocc^2 instead of the
// pivoted Cholesky's nocc sequential pivots.
bool project_occupied_factor_chol(dfks_ctx* h, const double* X, hipStream_t st) {
  const int n = h->nmo, nocc = h->nocc;
  gemm(h, false, false, n, nocc, n, 1.0, X, n, h->d_L, n, 0.0, h->d_Lw, n);
  gemm(h, true, false, nocc, nocc, n, 1.0, h->d_Lw, n, h->d_Lw, n, 0.0, h->d_Ls, nocc);
  RBCHK(rocsolver_dpotrf(h->rb, rocblas_fill_lower, nocc, h->d_Ls, nocc, h->d_info));
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
  sync(st);
  if (info != 0) return false;
  const double one = 1.0;
  RBCHK(rocblas_dtrsm(h->rb, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                      rocblas_diagonal_non_unit, n, nocc, &one, h->d_Ls, nocc, h->d_Lw, n));
  // L^T L into d_Ls (the Cholesky factor is no longer needed)
  gemm(h, true, false, nocc, nocc, n, 1.0, h->d_Lw, n, h->d_Lw, n, 0.0, h->d_Ls, nocc);
  std::vector<double> G((size_t)nocc * nocc);
  HIPCHK(hipMemcpyAsync(G.data(), h->d_Ls, sizeof(double) * G.size(), hipMemcpyDeviceToHost, st));
  sync(st);
  double dev = 0.0;
  for (int j = 0; j < nocc; ++j)
    for (int i = 0; i < nocc; ++i) dev = std::max(dev, std::fabs(G[(size_t)j * nocc + i] - (i == j ? 1.0 : 0.0)));
  if (!(dev <= 1e-10)) return false;
  HIPCHK(hipMemcpyAsync(h->d_L, h->d_Lw, sizeof(double) * n * nocc, hipMemcpyDeviceToDevice, st));
  // the step's factor check reads the residual trace and smallest pivot of the pivoted
  // Cholesky: this factor reproduces X as a converged pivoted one would
  const double chk[2] = {0.0, 1.0};
  HIPCHK(hipMemcpyAsync(h->d_sp2 + kSpResid, chk, sizeof(chk), hipMemcpyHostToDevice, st));
  sync(st);
  return true;
}

// The same factor as the orthogonal polar factor of W = X L_p, L = W (W^T W)^{-1/2},
// by the Newton-Schulz iteration W <- W (3 I - W^T W) / 2, two thin products per
// step. With the singular values of W at sqrt(1 - e), one step takes e to
// 3/4 e^2 (1 + e/3): two steps cover e up to ~3e-4 (the occupied space of a settling
// SCF turns less), then one read of the last Gram matrix gives the check
// (max |L^T L - I| <= 1e-10) and ||W^T W - I||_F, an upper bound on every |e|, from
// which further steps are counted (at most six in all; none when the bound is 0.5
// or more). L = W Z with Z a polynomial in W^T W, so L L^T is the projector onto
// range(W) = range(X). No Cholesky or triangular solve (rocSOLVER's dpotrf and
// rocBLAS's dtrsm at nocc = 160 took ~0.5 ms together). Returns false, with L_p
// untouched, when the occupied space turned too far or the check fails.
bool project_occupied_factor(dfks_ctx* h, const double* X, hipStream_t st) {
  const int n = h->nmo, nocc = h->nocc;
  double* W = h->d_Lw;
  double* Wn = h->d_Lq;
  gemm(h, false, false, n, nocc, n, 1.0, X, n, h->d_L, n, 0.0, W, n);
  gemm(h, true, false, nocc, nocc, n, 1.0, W, n, W, n, 0.0, h->d_Ls, nocc);
  auto polar_steps = [&](int k) {
    for (int i = 0; i < k; ++i) {
      HIPCHK(hipMemcpyAsync(Wn, W, sizeof(double) * n * nocc, hipMemcpyDeviceToDevice, st));
      gemm(h, false, false, n, nocc, nocc, -0.5, W, n, h->d_Ls, nocc, 1.5, Wn, n);
      std::swap(W, Wn);
      gemm(h, true, false, nocc, nocc, n, 1.0, W, n, W, n, 0.0, h->d_Ls, nocc);
    }
  };
  auto read_dev = [&](double* dev) {
    HIPCHK(launch_eye_dev(nocc, h->d_Ls, h->d_part, st));
    HIPCHK(hipMemcpyAsync(dev, h->d_part, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    sync(st);
  };
  int done = 2;
  polar_steps(done);
  double dev[2] = {0.0, 0.0};
  read_dev(dev);
  if (!(dev[0] <= 1e-10)) {
    int more = 0;
    for (double e = dev[1]; e > 1e-14 && done + more < 6; e = 0.75 * e * e * (1.0 + e)) ++more;
    if (!(dev[1] < 0.5) || more == 0) return project_occupied_factor_chol(h, X, st);
    polar_steps(more);
    read_dev(dev);
    if (!(dev[0] <= 1e-10)) return project_occupied_factor_chol(h, X, st);
  }
  HIPCHK(hipMemcpyAsync(h->d_L, W, sizeof(double) * n * nocc, hipMemcpyDeviceToDevice, st));
  // the step's factor check reads the residual trace and smallest pivot of the pivoted
  // Cholesky: this factor reproduces X as a converged pivoted one would
  HIPCHK(launch_store2(h->d_sp2 + kSpResid, 0.0, 1.0, st));
  return true;
}

// Occupied projector of F' = d_Ws (nmo x nmo) by SP2 purification, then its
// factor L (projected from the previous step's factor, else by pivoted Cholesky)
// and Cf = sqrt(2) X L. Returns false when the purification does not reach an
// idempotent projector of trace nocc (no gap at the Fermi level); the caller then
// takes the eigensolver path.
bool purify_density(dfks_ctx* h, hipStream_t st) {
  const int n = h->nmo, nocc = h->nocc, nao = h->nao;
  double* X = h->d_W1;
  HIPCHK(launch_gersh(n, h->d_Ws, h->d_lohi, h->d_sp2, st));
  // spectral bounds: Lanczos on an SCF's first purification (or after one that failed),
  // afterwards the previous bounds widened by ||F' - F'_prev||_F (Weyl)
  const bool weyl = h->pb_valid_n == n;
  if (weyl) {
    HIPCHK(launch_sp2_weyl(n, h->d_Ws, h->d_Fp, h->d_pb, h->d_sp2, st));
  } else {
    sp2_bounds(h, st);
    HIPCHK(hipMemcpyAsync(h->d_Fp, h->d_Ws, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(launch_save_bounds(h->d_sp2, h->d_pb, st));
  h->pb_valid_n = -1;   // valid again once this purification succeeds
  HIPCHK(launch_sp2_init(n, h->d_Ws, h->d_sp2, X, st));
  HIPCHK(launch_sp2_begin(n, X, nocc, h->d_sp2, st));
  int done = 0;
  int chunk = h->sp2_guess + 1;
  double stv[kSpHist];
  bool frozen = false;
  while (done < kSp2MaxIter) {
    int nblocks = (chunk + kSp2Block - 1) / kSp2Block;
    nblocks = std::min(nblocks, (kSp2MaxIter - done) / kSp2Block);
    sp2_run_blocks(h, nblocks, st);
    done += nblocks * kSp2Block;
    HIPCHK(hipMemcpyAsync(stv, h->d_sp2, sizeof(stv), hipMemcpyDeviceToHost, st));
    sync(st);
    if (stv[kSpFlag] < 0.0) {
      frozen = true;
      break;
    }
    chunk = kSp2Block;
  }
  if (!frozen) return false;
  h->sp2_guess = std::max(8, (int)stv[kSpIter]);
  if (!(std::fabs(stv[kSpErr]) < 1e-10 * std::max(1, nocc) && std::fabs(stv[kSpTrace] - nocc) < 1e-8 * std::max(1, nocc)))
    return false;
  // Weyl bounds only widen from step to step: once they are 2 % wider than the last
  // Lanczos bounds (about 0.03 more SP2 iterations), the next step runs Lanczos again
  const double range = stv[kSpEmax] - stv[kSpEmin];
  if (!weyl) h->lz_range = range;
  h->pb_valid_n = range <= 1.02 * h->lz_range ? n : -1;
  // X = L L^T: projected from the previous step's factor, or by pivoted Cholesky, panel by
  // panel (X is overwritten by the residual)
  if (!(h->l_valid && project_occupied_factor(h, X, st))) {
    HIPCHK(launch_diag(n, X, h->d_dvec, st));
    for (int k0 = 0; k0 < nocc; k0 += kPanel) {
      const int b = std::min(kPanel, nocc - k0);
      const int last = k0 + b >= nocc;
      HIPCHK(launch_pchol_panel(n, k0, b, X, h->d_dvec, h->d_L, h->d_piv, h->d_sp2, last, st));
      if (!last) {
        const GemmArgs u{n, n, b, h->d_L + (int64_t)k0 * n, n, false, h->d_L + (int64_t)k0 * n, n, false};
        HIPCHK(launch_gemm(kF64, u, X, -1.0, 1.0, 1, 1, h->d_gws, st));
      }
    }
    h->n_pchol_steps += 1;
  }
  h->l_valid = true;
  gemm(h, false, false, nao, nocc, n, std::sqrt(2.0), h->d_X, nao, h->d_L, n, 0.0, h->d_Cf, nao);
  return true;
}

// One iteration of dftax/ks/scf.py:343-379 on the device.
void scf_step(dfks_ctx* h, double* out3, hipStream_t st) {
  if (!h->d_X) fail(6, "dfks_scf_begin has not been called");
  RBCHK(rocblas_set_stream(h->rb, st));
  const int nao = h->nao, nmo = h->nmo, m = h->m;
  const int64_t nn = (int64_t)nao * nao, mm = (int64_t)nmo * nmo;
  const int slot = (h->it - h->diis_base) % m;
  const int count = std::min(h->it - h->diis_base + 1, m);
  double* err = h->d_Es + slot * mm;
  if (h->timing) {
    if (!h->scf_ev[0]) for (auto& e : h->scf_ev) (void)hipEventCreate(&e);
    (void)hipEventRecord(h->scf_ev[0], st);
  }
  const int nocc = h->nocc;
  // F' = X^T F X, kept as this step's DIIS history entry: the extrapolation runs in the
  // orthogonal basis, Fx' = sum_i c_i F'_i = X^T (sum_i c_i F_i) X
  double* Fo = h->d_Fs + slot * mm;
  gemm(h, true, false, nmo, nao, nao, 1.0, h->d_X, nao, h->d_F, nao, 0.0, h->d_W2, nmo);
  gemm(h, false, false, nmo, nmo, nao, 1.0, h->d_W2, nmo, h->d_X, nao, 0.0, Fo, nmo);
  if (h->p_fac) {
    // err = X^T (F P S - S P F) X = 2 (F' D D^T - D D^T F') when P = 2 X D D^T X^T and
    // X^T S X = I: two thin products instead of four of nao^3
    const double* D = h->p_fac;
    gemm(h, false, false, nmo, nocc, nmo, 1.0, Fo, nmo, D, nmo, 0.0, h->d_Lw, nmo);
    gemm(h, false, true, nmo, nmo, nocc, 2.0, h->d_Lw, nmo, D, nmo, 0.0, err, nmo);
    gemm(h, false, true, nmo, nmo, nocc, -2.0, D, nmo, h->d_Lw, nmo, 1.0, err, nmo);
  } else {
    // err = X^T (F P S - S P F) X ; S P F = (F P S)^T for symmetric F, P, S
    gemm(h, false, false, nao, nao, nao, 1.0, h->d_F, nao, h->d_P, nao, 0.0, h->d_W1, nao);
    gemm(h, false, false, nao, nao, nao, 1.0, h->d_W1, nao, h->d_S, nao, 0.0, h->d_W2, nao);
    HIPCHK(launch_antisym(nao, h->d_W2, h->d_W1, st));
    gemm(h, true, false, nmo, nao, nao, 1.0, h->d_X, nao, h->d_W1, nao, 0.0, h->d_W2, nmo);
    gemm(h, false, false, nmo, nmo, nao, 1.0, h->d_W2, nmo, h->d_X, nao, 0.0, err, nmo);
  }
  // DIIS history
  for (int i = 0; i < count; ++i)
    HIPCHK(launch_dot(mm, err, h->d_Es + i * mm, h->d_part, 1.0, h->d_dots + i, 0, st));
  HIPCHK(launch_gram_update(m, slot, count, h->d_dots, h->d_gram, st));
  HIPCHK(launch_diis_solve(m, count, h->d_gram, h->d_coef, st));
  HIPCHK(launch_diis_extrap(m, mm, h->d_coef, h->d_Fs, h->d_Ws, st));
  if (h->level_shift != 0.0) {
    // F' += b X^T (S - 1/2 S P S) X (the AO term in d_Fx)
    gemm(h, false, false, nao, nao, nao, 1.0, h->d_S, nao, h->d_P, nao, 0.0, h->d_W1, nao);
    HIPCHK(hipMemcpyAsync(h->d_Fx, h->d_S, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
    gemm(h, false, false, nao, nao, nao, -0.5, h->d_W1, nao, h->d_S, nao, 1.0, h->d_Fx, nao);
    gemm(h, true, false, nmo, nao, nao, 1.0, h->d_X, nao, h->d_Fx, nao, 0.0, h->d_W2, nmo);
    gemm(h, false, false, nmo, nmo, nao, h->level_shift, h->d_W2, nmo, h->d_X, nao, 1.0, h->d_Ws, nmo);
  }
  if (h->timing) (void)hipEventRecord(h->scf_ev[1], st);
  const bool purified = h->opt.density == 0 && purify_density(h, st);
  if (!purified) {
    // FP64 eigensolve (also in mixed steps: an FP32 solve's eigenvector error,
    // ~eps_32 |F| / gap, stalls the orbital gradient above the switch threshold)
    RBCHK(rocsolver_dsyevd(h->rb, rocblas_evect_original, rocblas_fill_upper, nmo, h->d_Ws, nmo, h->d_eps,
                           h->d_offd, h->d_info));
    gemm(h, false, false, nao, nmo, nmo, 1.0, h->d_X, nao, h->d_Ws, nmo, 0.0, h->d_C, nao);
    HIPCHK(launch_axpy((int64_t)nao * nocc, std::sqrt(2.0), h->d_C, h->d_Cf, 0, st));
    h->c_stale = false;
    h->n_eigh_steps += 1;
    h->l_valid = false;
  } else {
    h->c_stale = true;
  }
  gemm(h, false, true, nao, nao, nocc, 1.0, h->d_Cf, nao, h->d_Cf, nao, 0.0, h->d_P, nao);
  h->p_fac = purified ? h->d_L : h->d_Ws;
  if (h->timing) (void)hipEventRecord(h->scf_ev[2], st);
  std::vector<double> sg(nocc, 1.0);
  energy_fock_factored(h, h->d_P, h->d_Cf, sg, nocc, h->d_Edev, h->d_F, st);
  // derr = |err|_F
  HIPCHK(launch_dot(mm, err, err, h->d_part, 1.0, h->d_Edev + 7, 0, st));
  double e[8];
  HIPCHK(hipMemcpyAsync(e, h->d_Edev, sizeof(e), hipMemcpyDeviceToHost, st));
  int info = 0;
  double chol[2] = {0.0, 1.0};
  if (purified) HIPCHK(hipMemcpyAsync(chol, h->d_sp2 + kSpResid, sizeof(chol), hipMemcpyDeviceToHost, st));
  else HIPCHK(hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
  sync(st);
  if (info != 0) fail(5, "Fock eigensolver failed, info=" + std::to_string(info));
  // a converged projector has residual trace ~1e-13 after nocc pivots, and pivots >= 1/nmo
  if (purified && !(std::fabs(chol[0]) < 1e-8 && chol[1] > 0.1 / nmo))
    fail(5, "occupied factor of the purified density failed: residual trace " + std::to_string(chol[0]) +
                ", smallest pivot " + std::to_string(chol[1]));
  out3[0] = e[0];
  out3[1] = e[0] - h->e_prev;
  out3[2] = std::sqrt(e[7]);
  h->e_prev = e[0];
  if (h->restart_diis) {
    // the history so far holds mixed-precision Fock matrices; from the next
    // step on it holds only FP64 ones, starting with the F just built
    h->diis_base = h->it + 1;
    h->restart_diis = false;
  }
  h->it += 1;
}

}  // namespace

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Allocation-free Becke weights: the pair tables of the partition live in a
// file-scope device array instead of buffers the call allocates.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxAtoms = 1024;
// 1/R_AB and the size-adjustment a_AB, natom^2 each (16 MiB at the maximum)
constexpr size_t kScratchDoubles = 2 * (size_t)kNoAllocMaxAtoms * kNoAllocMaxAtoms;
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_addr;  // device -> address of that device's copy of scratch

// The device address of the current device's scratch array, looked up once per
// device with that device current (each device holds its own copy).
double* scratch_address() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(dev);
  if (it != g_scratch_addr.end()) return it->second;
  double* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) fail(2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_addr[dev] = ptr;
  return ptr;
}

// Pair tables of the Becke partition (becke.py:88-95), one thread per (A, B):
// invR = 1 / R_AB (1 on the diagonal), aAB = clip(u / (u^2 - 1), -0.45, 0.45)
// with u = (chi - 1) / (chi + 1), chi = R_bragg(A) / R_bragg(B).
// Contraction is off so every product and sum rounds as in dfks_becke_weights'
// host loop (x86-64 without FMA): the tables, and so the weights, come out bit
// for bit the same. A fused d*d + d2 moves 1/R_AB by an ulp for some pairs, and
// where a cell function s = (1 - f3) / 2 is tiny its cancellation turns that ulp
// into a relative change of ~1e-8 in the weight.
__global__ __launch_bounds__(256) void k_becke_pairs(int natm, const double* __restrict__ atm,
                                                     const double* __restrict__ bragg, double* __restrict__ invR,
                                                     double* __restrict__ aAB) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)natm * natm) return;
  const int a = (int)(i / natm), b = (int)(i % natm);
  double d2 = 0.0;
  for (int x = 0; x < 3; ++x) {
    const double d = atm[3 * a + x] - atm[3 * b + x];
    d2 += d * d;
  }
  invR[i] = 1.0 / sqrt(d2 + (a == b ? 1.0 : 0.0));
  const double chi = bragg[a] / bragg[b];
  const double u = (chi - 1.0) / (chi + 1.0);
  aAB[i] = fmin(0.45, fmax(-0.45, u / (u * u - 1.0)));
}

// Entry points of the restricted closed-shell path: an nspin = 2 handle gets
// status 7 (its SCF loop runs on the host over dfks_energy_fock).
void no_range_sep(const dfks_ctx* h, const char* what) {
  if (h && h->hf_lr != 0.0)
    fail(7, std::string(what) + " is not implemented for range-separated functionals; their handles support "
                                "dfks_energy_fock, the device SCF and the matrix and setup calls");
}

void closed_shell_only(const dfks_ctx* h, const char* what) {
  if (h && h->nspin == 2)
    fail(7, std::string(what) + " is implemented for nspin = 1; an nspin = 2 handle supports dfks_energy_fock "
                                "(FP64) and the matrix and setup calls");
}

extern "C" {

const char* dfks_last_error(void) { return g_err.c_str(); }

int dfks_create(const dfks_system_desc* sys, const dfks_options* opt, dfks_handle* out) {
  return guard([&] {
    if (!sys || !out) fail(1, "null argument");
    // checked before any device work, so a host without a GPU gets the same answer
    if (sys->nspin != 1 && sys->nspin != 2) fail(1, "nspin must be 1 (restricted) or 2 (unrestricted)");
    if (sys->nspin == 1 && sys->spin != 0) fail(1, "nspin = 1 needs spin = 0 (n_alpha = n_beta)");
    if (sys->nspin == 2 && (sys->spin < 0 || sys->nelec < sys->spin || (sys->nelec + sys->spin) % 2 != 0))
      fail(1, "nspin = 2 needs 0 <= spin <= nelec with nelec + spin even (n_alpha - n_beta = spin)");
    std::lock_guard<std::mutex> lk(g_mu);
    static bool tables = false;
    if (!tables) {
      upload_c2s_tables();
      tables = true;
    }
    auto* h = new dfks_ctx();
    try {
      if (opt) h->opt = *opt;
      else h->opt = dfks_options{0, 1e-7, 1e-10, 0, 0, 0, 0, 0.0, 0, 1};
      if (h->opt.precision != 0 && h->opt.precision != 1) fail(1, "precision must be 0 (FP64) or 1 (mixed)");
      if (h->opt.gemm < 0 || h->opt.gemm > 4)
        fail(1, "gemm must be 0 (auto), 1 (rocBLAS), 2 (library MFMA), 3 (library 8-wave exchange tile) or 4 "
                "(3 with the further variants)");
      if (h->opt.density != 0 && h->opt.density != 1) fail(1, "density must be 0 (purification) or 1 (eigensolver)");
      if (h->opt.aux_chunk < 0 || h->opt.ao_block < 0 || h->opt.screen < 0.0 || h->opt.screen >= 1.0)
        fail(1, "aux_chunk and ao_block must be >= 0 and screen in [0, 1)");
      if (h->opt.xc_blocks != 0 && h->opt.xc_blocks != 1) fail(1, "xc_blocks must be 0 (dense rows) or 1 (compact blocks)");
      h->recompute = h->opt.aux_chunk > 0;
      h->ao_block = h->opt.ao_block;

      h->xc_id = sys->xc_id;
      h->hf = hf_coeff(sys->xc_id);
      h->hf_lr = hf_lr_coeff(sys->xc_id);
      h->omega = omega_of(sys->xc_id);
      h->nspin = sys->nspin;
      if (h->hf_lr != 0.0) {
        if (sys->nspin != 1) fail(7, "range-separated functionals are implemented for nspin = 1");
        if (h->opt.precision != 0) fail(7, "range-separated functionals run in FP64 (options.precision = 0)");
        if (h->recompute) fail(7, "range-separated functionals need the materialised fitted tensors (aux_chunk = 0)");
      }
      if (h->nspin == 1) {
        if (sys->nelec <= 0 || sys->nelec % 2) fail(1, "closed shell needs an even, positive electron count");
        h->nocc = sys->nelec / 2;
      } else {
        if (sys->nelec <= 0) fail(1, "needs a positive electron count");
        h->spin = sys->spin;
        h->nocc = (sys->nelec + sys->spin) / 2;      // alpha (the larger channel)
        h->nocc_b = (sys->nelec - sys->spin) / 2;
      }
      h->orb.load(sys->orbital);
      h->aux.load(sys->aux);
      if (sys->ngrid < 0) fail(1, "negative grid size");
      if (h->orb.maxl() > 3) fail(1, "orbital shells above f (l > 3) are not supported");
      if (h->aux.maxl() > 4) fail(1, "auxiliary shells above g (l > 4) are not supported");
      h->nao = h->orb.nfunc;
      h->naux = h->aux.nfunc;
      if (h->nocc > h->nao) fail(1, "more occupied orbitals than basis functions");
      h->natom = sys->natom;
      h->atm_host.resize(4 * std::max(1, sys->natom));
      double enn = 0.0;
      for (int a = 0; a < sys->natom; ++a) {
        for (int x = 0; x < 3; ++x) h->atm_host[4 * a + x] = sys->atom_xyz[3 * a + x];
        h->atm_host[4 * a + 3] = sys->atom_z[a];
        for (int b = 0; b < a; ++b) {
          double d2 = 0;
          for (int x = 0; x < 3; ++x) {
            double d = sys->atom_xyz[3 * a + x] - sys->atom_xyz[3 * b + x];
            d2 += d * d;
          }
          enn += sys->atom_z[a] * sys->atom_z[b] / std::sqrt(d2);
        }
      }
      h->e_nn = enn;
      h->boys_host = boys_table();
      h->ng = sys->ngrid;
      if (h->ng > 0) {
        if (!sys->grid_xyz || !sys->grid_w) fail(1, "null grid arrays");
        h->gxyz_host.assign(sys->grid_xyz, sys->grid_xyz + 3 * h->ng);
        h->gw_host.assign(sys->grid_w, sys->grid_w + h->ng);
      }
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int dfks_set_precision(dfks_handle h, int32_t mixed) {
  return guard([&] {
    if (mixed) closed_shell_only(h, "dfks_set_precision(h, 1)");
    if (!h) fail(1, "null handle");
    if (mixed && h->opt.precision != 1) fail(1, "mixed steps need options.precision = 1 at create");
    const bool was = h->mixed_now;
    h->mixed_now = mixed != 0;
    // leaving mixed steps inside an SCF: the next step extrapolates over the
    // mixed history and builds F in FP64; the DIIS history restarts after it
    if (was && !h->mixed_now && h->d_X) h->restart_diis = true;
  });
}

int dfks_set_timing(dfks_handle h, int32_t on) {
  return guard([&] {
    if (!h) fail(1, "null handle");
    h->timing = on != 0;
  });
}

int dfks_stage_times(dfks_handle h, double* ms_out, int32_t n) {
  return guard([&] {
    if (!h || !ms_out) fail(1, "null argument");
    for (int i = 0; i < n; ++i) ms_out[i] = -1.0;
    if (!h->timing) fail(1, "timing is off (dfks_set_timing)");
    for (int i = 0; i < kNumStages && i < n; ++i) {
      if (i < kStScfDiis) {
        int a = i, b = i + 1;
        // stages whose start event was not recorded begin where the previous one ended
        while (a > 0 && !h->ev_used[a]) --a;
        if (!h->ev_used[a] || !h->ev_used[b]) { ms_out[i] = 0.0; continue; }
        HIPCHK(hipEventSynchronize(h->ev[b]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->ev[a], h->ev[b]));
        ms_out[i] = ms;
      } else if (h->scf_ev[0]) {
        int a = i - kStScfDiis, b = a + 1;
        HIPCHK(hipEventSynchronize(h->scf_ev[b]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->scf_ev[a], h->scf_ev[b]));
        ms_out[i] = ms;
      }
    }
    for (int i = 0; i < 10 && kNumStages + i < n; ++i) ms_out[kNumStages + i] = h->build_ms[i];
  });
}

int dfks_destroy(dfks_handle h) {
  return guard([&] {
    if (!h) return;
    StreamScope ss(h->st);
    delete h;
  });
}

int dfks_build(dfks_handle h, dfks_stream stream) {
  return guard([&] {
    if (!h) fail(1, "null handle");
    if (h->built) fail(1, "handle already built");
    StreamScope ss(h->st = S_(stream));
    build_all(h, S_(stream));
  });
}

int dfks_sizes(dfks_handle h, int64_t* out5) {
  return guard([&] {
    if (!h || !out5) fail(1, "null argument");
    out5[0] = h->nao;
    out5[1] = h->naux;
    out5[2] = h->nfit;
    out5[3] = h->nocc;
    out5[4] = h->ng;
  });
}

int dfks_stats(dfks_handle h, int64_t* out, int32_t n) {
  return guard([&] {
    if (!h || !out) fail(1, "null argument");
    int64_t width = 0;
    for (const auto& bn : h->xbins) width += bn.nblk * bn.n;
    const int64_t v[11] = {h->pairs_kept, h->pairs_total, (int64_t)h->slab_np.size(), h->max_np, h->ao_block,
                           h->n_eigh_steps, h->sp2_guess, h->xc_nblk, (int64_t)h->xbins.size(), width,
                           h->n_pchol_steps};
    for (int i = 0; i < n; ++i) out[i] = i < 11 ? v[i] : 0;
  });
}

int dfks_get_matrix(dfks_handle h, int32_t which, double* dev_out, dfks_stream stream) {
  return guard([&] {
    if (!h || !dev_out) fail(1, "null argument");
    if (!h->built) fail(6, "dfks_build has not been called");
    // the handle's buffers are last used on this stream: dfks_destroy orders
    // their release after it
    StreamScope ss(h->st = S_(stream));
    const size_t nn = (size_t)h->nao * h->nao;
    const double* src = nullptr;
    size_t n = 0;
    switch (which) {
      case DFKS_S: src = h->d_S; n = nn; break;
      case DFKS_HCORE: src = h->d_H; n = nn; break;
      case DFKS_KINETIC: src = h->d_T; n = nn; break;
      case DFKS_NUCLEAR: src = h->d_V; n = nn; break;
      case DFKS_INT2C: src = h->d_int2c; n = (size_t)h->naux * h->naux; break;
      case DFKS_INT3C:
        if (!h->d_int3c) fail(1, "int3c was not kept (options.keep_int3c = 0)");
        src = h->d_int3c; n = nn * h->naux; break;
      case DFKS_BFIT: src = h->d_B; n = nn * h->nfit; break;
      case DFKS_AO:
        if (!h->d_ao && h->xcb) {
          // compact grid: evaluate the dense rows on demand in the caller's buffer
          HIPCHK(launch_ao(h->ng, h->d_gxyz, h->orb.view(), h->nao, dev_out, S_(stream)));
          return;
        }
        src = h->d_ao; n = 4 * (size_t)h->ng * h->nao; break;
      case DFKS_GRID_W: src = h->d_gw; n = (size_t)h->ng; break;
      default: fail(1, "unknown matrix selector");
    }
    if (!src) fail(1, "this matrix is not held by the handle (recomputed slabs or a streamed AO grid)");
    HIPCHK(hipMemcpyAsync(dev_out, src, n * sizeof(double), hipMemcpyDeviceToDevice, S_(stream)));
  });
}

int dfks_energy_fock(dfks_handle h, const double* P, int32_t nspin, double* E, double* F, dfks_stream stream) {
  return guard([&] {
    if (!h || !P || !E || !F) fail(1, "null argument");
    if (nspin != h->nspin)
      fail(1, "P has " + std::to_string(nspin) + " spin blocks but the handle was created with nspin = " +
                  std::to_string(h->nspin));
    StreamScope ss(h->st = S_(stream));
    if (h->nspin == 2) energy_fock_uks(h, P, E, F, S_(stream));
    else energy_fock_general(h, P, E, F, S_(stream));
  });
}

int dfks_energy_fock_factored(dfks_handle h, const double* P, const double* Cf, const double* signs_host, int32_t nf,
                              double* E, double* F, dfks_stream stream) {
  return guard([&] {
    closed_shell_only(h, "dfks_energy_fock_factored");
    if (!h || !P || !Cf || !signs_host || !E || !F) fail(1, "null argument");
    if (nf <= 0) fail(1, "nf must be positive");
    std::vector<double> sg(signs_host, signs_host + nf);
    for (double s : sg)
      if (s != 1.0 && s != -1.0) fail(1, "signs must be +1 or -1");
    StreamScope ss(h->st = S_(stream));
    energy_fock_factored(h, P, Cf, sg, nf, E, F, S_(stream));
  });
}

int dfks_xc_eval(int32_t xc_id, int64_t n, const double* rho, const double* sigma, double* out3, dfks_stream stream) {
  return guard([&] {
    hf_coeff(xc_id);
    HIPCHK(launch_xc_pointwise(n, xc_id, rho, sigma, out3, S_(stream)));
  });
}

int dfks_xc_eval_pol(int32_t xc_id, int64_t n, const double* in5, double* out6, dfks_stream stream) {
  return guard([&] {
    hf_coeff(xc_id);
    if (hf_lr_coeff(xc_id) != 0.0) fail(7, "spin-polarised range-separated functionals are not implemented");
    if (n < 0 || (n > 0 && (!in5 || !out6))) fail(1, "bad arguments");
    HIPCHK(launch_xc_pointwise_pol(n, xc_id, in5, out6, S_(stream)));
  });
}

int dfks_xc_eval2(int32_t xc_id, int64_t n, const double* rho, const double* sigma, double* out6,
                  dfks_stream stream) {
  return guard([&] {
    hf_coeff(xc_id);
    HIPCHK(launch_xc_pointwise2(n, xc_id, rho, sigma, out6, S_(stream)));
  });
}

int dfks_response_begin(dfks_handle h, const double* P, dfks_stream stream) {
  return guard([&] {
    closed_shell_only(h, "dfks_response_begin");
    no_range_sep(h, "dfks_response_begin");
    if (!h || !P) fail(1, "null argument");
    StreamScope ss(h->st = S_(stream));
    response_begin(h, P, S_(stream));
  });
}

int dfks_response_fock(dfks_handle h, const double* U, const double* V, int32_t r, double* R, dfks_stream stream) {
  return guard([&] {
    closed_shell_only(h, "dfks_response_fock");
    no_range_sep(h, "dfks_response_fock");
    if (!h || !U || !V || !R) fail(1, "null argument");
    StreamScope ss(h->st = S_(stream));
    response_fock(h, U, V, r, R, S_(stream));
  });
}

int dfks_scf_begin(dfks_handle h, const double* P0, double lindep, int32_t diis_space, double level_shift,
                   dfks_stream stream) {
  return guard([&] {
    closed_shell_only(h, "dfks_scf_begin");
    if (!h || !P0) fail(1, "null argument");
    StreamScope ss(h->st = S_(stream));
    scf_begin(h, P0, lindep, diis_space, level_shift, S_(stream));
  });
}

int dfks_scf_step(dfks_handle h, double* out3, dfks_stream stream) {
  return guard([&] {
    closed_shell_only(h, "dfks_scf_step");
    if (!h || !out3) fail(1, "null argument");
    StreamScope ss(h->st = S_(stream));
    scf_step(h, out3, S_(stream));
  });
}

int dfks_scf_result(dfks_handle h, double* P, double* C, double* eps, int32_t* nmo, dfks_stream stream) {
  return guard([&] {
    closed_shell_only(h, "dfks_scf_result");
    if (!h) fail(1, "null handle");
    if (!h->d_X) fail(6, "dfks_scf_begin has not been called");
    hipStream_t st = S_(stream);
    StreamScope ss(h->st = st);
    const int nao = h->nao, nm = h->nmo;
    if (h->c_stale && (C || eps)) {
      // orbitals of the last step's F' (its density came from purification)
      RBCHK(rocblas_set_stream(h->rb, st));
      HIPCHK(hipMemcpyAsync(h->d_W2, h->d_Ws, sizeof(double) * nm * nm, hipMemcpyDeviceToDevice, st));
      RBCHK(rocsolver_dsyevd(h->rb, rocblas_evect_original, rocblas_fill_upper, nm, h->d_W2, nm, h->d_eps,
                             h->d_offd, h->d_info));
      gemm(h, false, false, nao, nm, nm, 1.0, h->d_X, nao, h->d_W2, nm, 0.0, h->d_C, nao);
      int info = 0;
      HIPCHK(hipMemcpyAsync(&info, h->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
      sync(st);
      if (info != 0) fail(5, "orbital eigensolver failed, info=" + std::to_string(info));
      h->c_stale = false;
    }
    if (P) HIPCHK(hipMemcpyAsync(P, h->d_P, sizeof(double) * h->nao * h->nao, hipMemcpyDeviceToDevice, st));
    if (C) HIPCHK(hipMemcpyAsync(C, h->d_C, sizeof(double) * h->nao * h->nmo, hipMemcpyDeviceToDevice, st));
    if (eps) HIPCHK(hipMemcpyAsync(eps, h->d_eps, sizeof(double) * h->nmo, hipMemcpyDeviceToDevice, st));
    if (nmo) *nmo = h->nmo;
  });
}

int dfks_overlap(const dfks_basis_desc* basis, double* host_out, dfks_stream stream) {
  return guard([&] {
    if (!basis || !host_out) fail(1, "null argument");
    hipStream_t st = S_(stream);
    StreamScope ss(st);
    {
      std::lock_guard<std::mutex> lk(g_mu);
      static bool tables = false;
      if (!tables) {
        upload_c2s_tables();
        tables = true;
      }
    }
    Basis b;
    b.load(*basis);
    b.upload();
    struct Rel {
      Basis* b;
      ~Rel() { b->release(); }
    } rel{&b};
    const int64_t n = b.nfunc;
    double *dS = dalloc<double>(n * n), *dT = dalloc<double>(n * n), *dV = dalloc<double>(n * n);
    double* dummy = dalloc<double>(kBoysNg * kBoysN);
    struct Free {
      std::vector<void*> p;
      ~Free() { for (void* q : p) dfree(q); }
    } fr{{dS, dT, dV, dummy}};
    HIPCHK(launch_int1e(b.view(), 0, dummy, dummy, dS, dT, dV, n, st));
    HIPCHK(hipMemcpyAsync(host_out, dS, n * n * sizeof(double), hipMemcpyDeviceToHost, st));
    sync(st);
  });
}

int dfks_release_memory(dfks_stream stream) {
  return guard([&] {
    hipStream_t st = S_(stream);
    StreamScope ss(st);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    cache_release(dev);
    HIPCHK(hipStreamSynchronize(st));
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) HIPCHK(hipMemPoolTrimTo(pool, 0));
  });
}

int dfks_device_free(int64_t* out) {
  return guard([&] {
    if (!out) fail(1, "null argument");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    // memory the default pool reserves but has not handed out
    uint64_t held = 0;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t res = 0, used = 0;
      if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &res) == hipSuccess &&
          hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess && res > used)
        held = res - used;
    }
    size_t kept = 0;
    {
      std::lock_guard<std::mutex> lk(g_cache_mu);
      for (const auto& f : g_big_free)
        if (f.second.dev == dev) kept += f.second.bytes;
    }
    *out = (int64_t)(fr + held + kept);
  });
}

int dfks_becke_weights(int64_t n, const double* pts, const double* raw_w, const int32_t* owner, int32_t natom,
                       const double* atom_xyz, const double* bragg, double* w_out, dfks_stream stream) {
  return guard([&] {
    hipStream_t st = S_(stream);
    StreamScope ss(st);
    if (n < 0 || natom <= 0) fail(1, "bad sizes");
    if (n == 0) return;
    for (int64_t i = 0; i < n; ++i)
      if (owner[i] < 0 || owner[i] >= natom) fail(1, "owner atom index out of range");
    std::vector<double> invR((size_t)natom * natom, 0.0), aAB((size_t)natom * natom, 0.0);
    for (int a = 0; a < natom; ++a)
      for (int b = 0; b < natom; ++b) {
        double d2 = 0;
        for (int x = 0; x < 3; ++x) {
          double d = atom_xyz[3 * a + x] - atom_xyz[3 * b + x];
          d2 += d * d;
        }
        double R = std::sqrt(d2 + (a == b ? 1.0 : 0.0));
        invR[(size_t)a * natom + b] = 1.0 / R;
        double chi = bragg[a] / bragg[b];
        double u = (chi - 1.0) / (chi + 1.0);
        double av = u / (u * u - 1.0);
        aAB[(size_t)a * natom + b] = std::min(0.45, std::max(-0.45, av));
      }
    double *d_pts = dalloc<double>(3 * n), *d_raw = dalloc<double>(n), *d_w = dalloc<double>(n);
    int32_t* d_own = dalloc<int32_t>(n);
    double *d_atm = dalloc<double>(3 * natom), *d_invR = dalloc<double>(invR.size()), *d_a = dalloc<double>(aAB.size());
    struct Free {
      std::vector<void*> p;
      ~Free() { for (void* q : p) dfree(q); }
    } fr{{d_pts, d_raw, d_w, d_own, d_atm, d_invR, d_a}};
    HIPCHK(hipMemcpyAsync(d_pts, pts, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_raw, raw_w, n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_own, owner, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_atm, atom_xyz, 3 * natom * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_invR, invR.data(), invR.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_a, aAB.data(), aAB.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(launch_becke(n, d_pts, d_raw, d_own, natom, d_atm, d_invR, d_a, d_w, st));
    HIPCHK(hipMemcpyAsync(w_out, d_w, n * sizeof(double), hipMemcpyDeviceToHost, st));
    sync(st);
  });
}

int dfks_becke_weights_noalloc(int64_t n, const double* pts, const double* raw_w, const int32_t* owner,
                               int32_t natom, const double* atom_xyz, const double* bragg, double* w_out,
                               dfks_stream stream) {
  return guard([&] {
    if (n < 0 || natom <= 0) fail(1, "bad sizes");
    if (natom > kNoAllocMaxAtoms)
      fail(1, "dfks_becke_weights_noalloc takes at most " + std::to_string(kNoAllocMaxAtoms) + " atoms (got " +
                  std::to_string(natom) + "); use dfks_becke_weights");
    if (n == 0) return;
    if (!pts || !raw_w || !owner || !atom_xyz || !bragg || !w_out) fail(1, "null argument");
    hipStream_t st = S_(stream);
    double* invR = scratch_address();
    double* aAB = invR + (size_t)natom * natom;
    const int64_t np = (int64_t)natom * natom;
    hipLaunchKernelGGL(k_becke_pairs, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, natom, atom_xyz, bragg,
                       invR, aAB);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_becke(n, pts, raw_w, owner, natom, atom_xyz, invR, aAB, w_out, st));
  });
}

int dfks_forces(dfks_handle h, const double* C_occ, int32_t nocc, int64_t ngrid, const double* grid_xyz,
                const double* grid_raw_w, const int32_t* grid_owner, const double* bragg, double* grad_parts,
                dfks_stream stream) {
  return guard([&] {
    closed_shell_only(h, "dfks_forces");
    no_range_sep(h, "dfks_forces");
    if (!h || !C_occ || !grad_parts || !bragg) fail(1, "null argument");
    if (ngrid < 0 || (ngrid > 0 && (!grid_xyz || !grid_raw_w || !grid_owner))) fail(1, "bad grid arguments");
    StreamScope ss(h->st = S_(stream));
    forces_run(h, C_occ, nocc, ngrid, grid_xyz, grid_raw_w, grid_owner, bragg, grad_parts, S_(stream));
  });
}

}  // extern "C"
