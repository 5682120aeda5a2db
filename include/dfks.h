/*
 * dfks.h -- C ABI of the MI355X density-fitted Kohn-Sham library (libdfks.so).
 *
 * One handle holds one molecule's device state: shells, grid, S/hcore, the
 * fitted three-centre tensor B = (mn|P) L (V+ = L L^T), the AO values and
 * gradients on the grid, and the SCF working set (DIIS history, orbitals).
 * All matrices are row-major nao x nao FP64 device buffers owned by the
 * caller unless stated otherwise; calls enqueue on the given HIP stream and
 * do not synchronise it unless documented. Every call returns 0 on success
 * and a non-zero status otherwise; dfks_last_error() gives the message
 * (thread-local).
 *
 * Reference interfaces replaced (dftax, Python/JAX):
 *   dfks_create + dfks_build     KS.__init__ / _build_integrals (dftax/ks/energy.py:439-676, 368-436)
 *                                with coulomb=df(...) resolved by _make_coulomb (dftax/ks/terms.py:951-988)
 *   dfks_energy_fock             KS.electronic(P) + jax.grad -> sym(dE/dP)
 *                                (dftax/ks/energy.py:704-712, dftax/ks/scf.py:42-46);
 *                                its terms DFCoulomb.energy (terms.py:592-613) and
 *                                GridXC.energy (terms.py:810-828)
 *   dfks_scf_begin/step/result   the body of _scf_solve (dftax/ks/scf.py:279-391)
 *   dfks_becke_weights           becke_partition + weight assembly (dftax/grid/becke.py:68-108,
 *                                dftax/grid/grid.py:240-248)
 *   dfks_get_matrix              KS.S / KS.hcore and the integral tensors (energy.py:369-436)
 *   dfks_overlap                 overlap_matrix / cross_overlap_matrix as used by the minao
 *                                guess (dftax/ks/guess.py:319-337)
 *   dfks_forces                  forces(mol, xc, result) = -jax.grad of the frozen-density
 *                                energy (dftax/ks/forces.py:128-270)
 */
#ifndef DFKS_H
#define DFKS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfks_ctx* dfks_handle;
typedef void* dfks_stream; /* a hipStream_t; NULL = the default stream */

/* Shell-major contracted Gaussian basis (host arrays, copied at create). */
typedef struct {
  int32_t nshell;
  const int32_t* l;        /* [nshell] angular momentum, <= 4 */
  const int32_t* nprim;    /* [nshell] */
  const int32_t* prim0;    /* [nshell] offset into exps/coefs */
  const int32_t* ao0;      /* [nshell] first function in the final span */
  const double* centre;    /* [3*nshell] Bohr */
  const double* exps;      /* [nprim_total] */
  const double* coefs;     /* [nprim_total] normalised contraction coefficients */
  int32_t nprim_total;
  int32_t spherical;       /* 1: 2l+1 functions for l >= 2 */
  int32_t nfunc;           /* functions in the final span */
} dfks_basis_desc;

typedef struct {
  dfks_basis_desc orbital;
  dfks_basis_desc aux;     /* density-fitting set */
  int32_t natom;
  const double* atom_xyz;  /* [3*natom] Bohr */
  const double* atom_z;    /* [natom] nuclear charges */
  int64_t ngrid;
  const double* grid_xyz;  /* [3*ngrid] Bohr */
  const double* grid_w;    /* [ngrid] quadrature weights */
  int32_t xc_id;           /* 1 PBE, 2 PBE0, 3 B3LYP, 4 CAM-B3LYP (range-separated, dftax/energy/xc.py:
                              711-750: exact exchange 0.19 K + 0.46 K_lr with K_lr from a second fitted
                              tensor of the erf(0.33 r)/r operator, terms.py:578-612; nspin = 1, FP64
                              and materialised tensors only, dfks_create returns status 7 otherwise;
                              response and forces return status 7; E[DFKS_E_K] holds both exchange
                              terms) */
  int32_t nelec;           /* electrons; nspin = 1: nocc = nelec / 2; nspin = 2: n_alpha, n_beta =
                              (nelec +- spin) / 2 */
  int32_t nspin;           /* spin blocks of P and F (the leading axis of dftax's P[nspin, nao, nao],
                              dftax/ks/terms.py:520-527): 1 restricted closed shell; 2 spin-unrestricted
                              (alpha, beta): dfks_energy_fock in FP64 on the compact XC grid; the
                              closed-shell-only calls (device SCF, response, forces, factored Fock,
                              mixed precision) return status 7 for such a handle */
  int32_t spin;            /* 2S = n_alpha - n_beta (dftax Molecule.spin); must be 0 when nspin = 1 */
} dfks_system_desc;

typedef struct {
  int32_t precision;       /* 0: FP64 only; 1: also keep FP32 operands for mixed steps */
  double metric_cutoff;    /* relative eigenvalue cutoff of the metric pseudo-inverse (1e-7) */
  double rho_thresh;       /* density mask of the XC quadrature (1e-10) */
  int32_t keep_int3c;      /* 1: keep the unfitted (mn|P) tensor for inspection */
  int32_t gemm;            /* FP64 products: 0 library symmetric exchange kernel, and for the exchange
                              product T and the compact-grid XC products the fastest of rocBLAS and the
                              library kernels, timed once per shape and device (default); 1 rocBLAS
                              for all; 2 library kernels for all; 3 as 2 with the 8-wavefront tile for
                              T; 4 as 3 with each product's further library variant where it has one
                              (split-BF16 T with two k-steps of operands in flight, FP64 K split by
                              XCD). Mixed steps always use the library's split-BF16 kernels. */
  int32_t density;         /* SCF density from the orthonormal-basis Fock matrix: 0 spectral projection
                              (SP2 purification + pivoted Cholesky factor; the eigensolver runs only
                              when purification does not converge, and once for dfks_scf_result's
                              orbitals); 1 eigensolver in every step (dftax/ks/scf.py:363-367). */
  int32_t aux_chunk;       /* 0: materialise B = (mn|P) L (nao^2 x nfit words). > 0: recompute (mn|P) in
                              auxiliary slabs of about this many functions in every Fock build
                              (dftax/ks/terms.py:321-493, df(chunk=<int>)): one pass for the exchange
                              intermediate and gamma, one for J. */
  double screen;           /* relative Schwarz threshold on bra shell pairs for the recomputed slabs:
                              pairs with max sqrt((mn|mn)) < screen * max over pairs are skipped
                              (eri3c_bucketed.py:140-169); 0 keeps every pair. */
  int64_t ao_block;        /* 0: keep AO values and gradients on the grid resident; > 0: recompute them
                              per block of this many points in every Fock build (energy.py:158-212). */
  int32_t xc_blocks;       /* 1: XC on blocks of 128 spatially sorted points, each with only the AOs
                              whose value or gradient bound reaches 1e-14 there (compact rows, batched
                              per-block products); 0: dense AO rows (grid.py:10-44 as written). */
} dfks_options;

/* Matrix selectors for dfks_get_matrix. */
enum {
  DFKS_S = 0,        /* overlap, nao x nao */
  DFKS_HCORE = 1,    /* T + V */
  DFKS_KINETIC = 2,
  DFKS_NUCLEAR = 3,
  DFKS_INT2C = 4,    /* (P|Q), naux x naux */
  DFKS_INT3C = 5,    /* (mn|P), nao x nao x naux (needs keep_int3c) */
  DFKS_BFIT = 6,     /* B = (mn|P) L, nao x nao x nfit */
  DFKS_AO = 7,       /* AO values and gradients, 4 x ngrid x nao */
  DFKS_GRID_W = 8    /* grid weights, ngrid */
};

/* Energy components written by dfks_energy_fock (device double[8]). */
enum {
  DFKS_E_TOT = 0, DFKS_E_ONE = 1, DFKS_E_J = 2, DFKS_E_K = 3, DFKS_E_XC = 4, DFKS_E_NN = 5
};

const char* dfks_last_error(void);

int dfks_create(const dfks_system_desc* sys, const dfks_options* opt, dfks_handle* out);
int dfks_destroy(dfks_handle h);

/* Integrals, metric factor, fitted B and AO grid. Synchronises the stream. */
int dfks_build(dfks_handle h, dfks_stream stream);

/* Sizes: [nao, naux, nfit, nocc] and ngrid. */
int dfks_sizes(dfks_handle h, int64_t* out5);

/* Build and solver counters: [0] bra shell pairs kept, [1] bra shell pairs in total,
 * [2] auxiliary slabs (recompute mode, else 0), [3] largest slab in functions,
 * [4] AO block (0: resident grid), [5] SCF steps that fell back to the eigensolver,
 * [6] SP2 iterations of the last purified step, [7] compact XC blocks (0: dense
 * rows), [8] their bins of equal width, [9] the sum of their padded widths,
 * [10] purified SCF steps whose occupied factor came from the pivoted Cholesky
 * (the others project the previous step's factor). Fills n values (zeros past 11). */
int dfks_stats(dfks_handle h, int64_t* out, int32_t n);

int dfks_get_matrix(dfks_handle h, int32_t which, double* dev_out, dfks_stream stream);

/* E(P) and F = sym(dE/dP) = hcore + J - a_x/2 K + Vxc for a symmetric density P
 * (any P: K uses a signed eigen-factorisation of P). P and F hold nspin
 * consecutive nao x nao blocks (dftax's P[nspin, nao, nao]); nspin must equal
 * the handle's (status 1 otherwise). E: device double[8].
 * nspin = 2 (replaces CoulombTerm.energy + GridXC.energy of terms.py:520-527,
 * 838-874 for P[2, nao, nao]): J of P_a + P_b, exchange -1/2 a_x sum_s P_s . K(P_s)
 * (terms.py:500-509), the spin-polarised functional, and F_s = sym(dE/dP_s) =
 * hcore + J - a_x K(P_s) + Vxc_s. */
int dfks_energy_fock(dfks_handle h, const double* P, int32_t nspin, double* E, double* F, dfks_stream stream);

/* Same for P = sum_i s_i c_i c_i^T given its factor: Cf column-major
 * [nf][nao] (column i contiguous), signs s_i in {+1,-1} (host array). */
int dfks_energy_fock_factored(dfks_handle h, const double* P, const double* Cf, const double* signs_host,
                              int32_t nf, double* E, double* F, dfks_stream stream);

/* Pointwise functional f = eps_xc * rho and its derivatives f_rho, f_sigma. */
int dfks_xc_eval(int32_t xc_id, int64_t n, const double* rho, const double* sigma, double* out3,
                 dfks_stream stream);

/* Spin-polarised pointwise functional (dftax/energy/xc.py spin branches):
 * in5[5 i ..] = {rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb}, out6[6 i ..] =
 * {f, f_rho_a, f_rho_b, f_sigma_aa, f_sigma_ab, f_sigma_bb} of
 * f = eps_xc(rho_a, rho_b, ...) (rho_a + rho_b), unmasked. Device pointers. */
int dfks_xc_eval_pol(int32_t xc_id, int64_t n, const double* in5, double* out6, dfks_stream stream);

/* Pointwise second derivatives: out6[6 i ..] = {f, f_rho, f_sigma, f_rho rho,
 * f_rho sigma, f_sigma sigma} (the XC kernel of the response below). */
int dfks_xc_eval2(int32_t xc_id, int64_t n, const double* rho, const double* sigma, double* out6,
                  dfks_stream stream);

/* Response Fock matrix, the derivative of F = dE/dP at P0 along dP:
 *   R = J[dP] - a_x/2 K[dP] + f_xc(P0)[dP],
 * which replaces jax.jvp(jax.grad(E), (P0,), (dP,)) in the Hessian-vector
 * products of dftax/ks/newton.py:134-136. dfks_response_begin fixes P0 (device
 * nao x nao) and caches the functional's second derivatives on the grid;
 * dfks_response_fock then takes dP = U V^T + V U^T (U, V device, column-major
 * nao x r, 1 <= r <= nao; the orbital-rotation tangent of P = 2 C_o C_o^T has
 * U = C_o, V = 2 dC_o) and writes R (device nao x nao). FP64 builds only. */
int dfks_response_begin(dfks_handle h, const double* P0, dfks_stream stream);
int dfks_response_fock(dfks_handle h, const double* U, const double* V, int32_t r, double* R, dfks_stream stream);

/* Device SCF (dftax/ks/scf.py:279-391): canonical orthonormaliser of S with
 * eigenvalue cutoff lindep, Pulay DIIS of depth m, level shift. P0: device
 * nao x nao initial density. Synchronises the stream. */
int dfks_scf_begin(dfks_handle h, const double* P0, double lindep, int32_t diis_space, double level_shift,
                   dfks_stream stream);
/* One iteration; host out3 = {E_new, E_new - E_prev, |err|}. Synchronises. */
int dfks_scf_step(dfks_handle h, double* out3, dfks_stream stream);
/* Current density, orbitals (nao x nmo, column-major), orbital energies of the last
 * step's extrapolated Fock matrix. Any may be NULL. */
int dfks_scf_result(dfks_handle h, double* P, double* C, double* eps, int32_t* nmo, dfks_stream stream);

/* Arithmetic of the following Fock builds: 0 FP64; 1 mixed (FP32 operands,
 * three-term split-BF16 MFMA products with FP32 accumulation for the
 * exchange and XC GEMMs; FP64 reductions, functional, J sum, DIIS, eigh).
 * Needs options.precision = 1 at create. Switching from mixed to FP64 inside
 * an SCF: the next step still extrapolates over the mixed DIIS history, then
 * builds its Fock matrix in FP64; the DIIS history restarts after that step,
 * so it holds FP64 Fock matrices only. */
int dfks_set_precision(dfks_handle h, int32_t mixed);

/* Stage timers (HIP events on the work stream). Stages of the last Fock
 * build / SCF step, in ms: 0 exchange GEMM T = B^T Cf, 1 gamma from T and the
 * exchange SYRK K = T^T T, 2 Coulomb J = B gamma, 3 XC GEMM Z = Phi P, 4 XC kernel,
 * 5 XC SYR2K, 6 Fock assembly + energies, 7 commutator + DIIS, 8 eigensolve +
 * density; then build wall times 9 one-electron, 10 (P|Q), 11 (mn|P),
 * 12 metric factor + B, 13 AO grid, and the compact grid's parts of 13:
 * 14 Morton sort, 15 block shell scans, 16 host tables, 17 allocations and
 * uploads, 18 AO evaluation to the final wait. On the compact XC grid stage 3 is empty,
 * 4 holds the per-bin gathers, products and pointwise kernel, and 5 the Vxc
 * row assembly. */
int dfks_set_timing(dfks_handle h, int32_t on);
int dfks_stage_times(dfks_handle h, double* ms_out, int32_t n);

/* Device memory this library can obtain on the current device, in bytes: what
 * hipMemGetInfo reports free, plus what the default memory pool reserves but has
 * not handed out, plus the buffers of 8 MiB and more that the library keeps after
 * they are freed for reuse by later handles. The memory plan sizes the DF chunks
 * from it, where the reference reads the accelerator's memory statistics
 * (dftax/ks/energy.py:197-212). */
int dfks_device_free(int64_t* out);

/* Returns the buffers kept for reuse on the current device (after their last
 * users have finished) and the default pool's unused reserve to the device, for
 * callers that need the memory outside this library (the device then reports it
 * free). Frees on the stream and synchronises it (only it). No counterpart in
 * the reference, whose arrays are owned by the JAX runtime. */
int dfks_release_memory(dfks_stream stream);

/* Overlap matrix of one basis (host out, nfunc x nfunc); used by the minao guess.
 * Runs on the stream and synchronises it (only it). */
int dfks_overlap(const dfks_basis_desc* basis, double* host_out, dfks_stream stream);

/* Becke partition weights for an atom-centred grid (host arrays in and out).
 * Runs on the stream and synchronises it (only it). */
int dfks_becke_weights(int64_t n, const double* pts, const double* raw_w, const int32_t* owner, int32_t natom,
                       const double* atom_xyz, const double* bragg, double* w_out, dfks_stream stream);

/* Nuclear gradient dE/dR (Ha/Bohr) at a frozen converged density, the quantity
 * dftax's forces() differentiates (dftax/ks/forces.py:128-270: the basis and
 * auxiliary centres and the Becke grid follow their atoms, P = 2 Z (Z^T S Z)^-1 Z^T
 * with Z the occupied orbitals held fixed). C_occ: device, column-major
 * nao x nocc occupied orbitals of the converged SCF (nocc = the handle's).
 * The grid is the unpruned-by-weight atom-centred grid of the energy's
 * Becke spec (host arrays: points [3 ngrid] Bohr, raw radial x angular
 * weights [ngrid], owner atom [ngrid]) and the atoms' Bragg radii [natom];
 * the library forms the Becke weights and their derivatives itself.
 * grad_parts: host [5][natom][3], the terms of dE/dR: nuclear repulsion,
 * one-electron, Coulomb, exchange, exchange-correlation, each with its share
 * of the density's response to the overlap (-Tr(W_k dS), W_k = 1/2 P F_k P);
 * their sum is dE/dR and forces are its negative. Needs the materialised
 * fitted tensor (options.aux_chunk = 0). Synchronises the stream. */
int dfks_forces(dfks_handle h, const double* C_occ, int32_t nocc, int64_t ngrid, const double* grid_xyz,
                const double* grid_raw_w, const int32_t* grid_owner, const double* bragg, double* grad_parts,
                dfks_stream stream);

/* Becke partition weights of device-resident points, with no device memory
 * allocated by the call. No counterpart in the reference: this is
 * dfks_becke_weights for callers that cannot allocate (code under HIP graph
 * capture, serving loops that set up grids on the device). All pointers are
 * device memory: pts [3n], raw_w [n], owner [n] (an owner outside 0..natom-1
 * gets weight 0), atom_xyz [3 natom], bragg [natom], w_out [n]. The
 * partition's pair tables go to a fixed device array of the library, 16 MiB
 * per device, so natom is at most 1024 (status 1 above). Enqueues on the
 * stream and returns without waiting. Calls that overlap on one device share
 * that array and race: order them on one stream, or synchronise in between. */
int dfks_becke_weights_noalloc(int64_t n, const double* pts, const double* raw_w, const int32_t* owner, int32_t natom,
                               const double* atom_xyz, const double* bragg, double* w_out, dfks_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* DFKS_H */
