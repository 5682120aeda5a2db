"""GPU parity: the HIP path through the C ABI against the CPU FP64 oracle.

Tolerances: 1e-10 max-abs on (ij|P) and on the
Fock matrix, 1e-8 Ha on the total energy.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_INT = 1e-10
TOL_FOCK = 1e-10
TOL_E = 1e-8


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _oracle(mol, xc, grid=None):
    from oracle import ref
    return ref.OracleKS(mol, xc, grid=grid)


def _hip(mol, xc, grid=None, keep=True):
    from dftax_amd import KS, df, hip
    return KS(mol, xc, grid=grid, coulomb=df(engine=hip(keep_int3c=keep)))


@pytest.fixture(scope="module")
def water_pair(torch_cuda, water_svp):
    _say("case: water def2-SVP PBE0, default grid")
    return _oracle(water_svp, "PBE0"), _hip(water_svp, "PBE0")


@pytest.fixture(scope="module")
def dimer_pair(torch_cuda, water_dimer_tzvp):
    from dftax_amd import becke
    _say("case: water dimer def2-TZVP B3LYP, becke(35, 110)")
    g = becke(35, 110)
    return _oracle(water_dimer_tzvp, "B3LYP", grid=g), _hip(water_dimer_tzvp, "B3LYP", grid=g)


def _mat(ks, which, shape):
    return ks.coulomb.matrix(which, shape).cpu().numpy()


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_one_electron(case, request):
    from dftax_amd import _native as N
    o, k = request.getfixturevalue(case)
    n = k.nao
    assert np.abs(_mat(k, N.DFKS_S, (n, n)) - o.S).max() < TOL_INT
    assert np.abs(_mat(k, N.DFKS_HCORE, (n, n)) - o.hcore).max() < TOL_INT


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_int2c(case, request):
    from dftax_amd import _native as N
    o, k = request.getfixturevalue(case)
    m = k.coulomb.naux
    assert np.abs(_mat(k, N.DFKS_INT2C, (m, m)) - o.int2c).max() < TOL_INT


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_int3c(case, request):
    from dftax_amd import _native as N
    o, k = request.getfixturevalue(case)
    n, m = k.nao, k.coulomb.naux
    g = _mat(k, N.DFKS_INT3C, (n, n, m))
    assert np.abs(g - o.int3c).max() < TOL_INT
    assert np.abs(g - g.transpose(1, 0, 2)).max() < 1e-13


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_fitted_tensor(case, request):
    """B B^T reproduces the fitted 4-index (mn|V+|ls) of the oracle (basis-independent of L)."""
    from dftax_amd import _native as N
    o, k = request.getfixturevalue(case)
    n = k.nao
    B = _mat(k, N.DFKS_BFIT, (n * n, k.coulomb.nfit))
    I = o.int3c.reshape(n * n, -1)
    ref = I @ o.Vp @ I.T
    got = B @ B.T
    assert k.coulomb.nfit == o.L.shape[1]
    assert np.abs(got - ref).max() < 1e-9 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_grid_and_ao(case, request):
    from dftax_amd import _native as N
    o, k = request.getfixturevalue(case)
    assert k.grid_w.shape == o.grid_w.shape
    assert np.abs(k.grid_xyz - o.grid_xyz).max() == 0.0
    assert np.abs(k.grid_w - o.grid_w).max() < 1e-13 * max(1.0, np.abs(o.grid_w).max())
    ng, n = k.grid_w.shape[0], k.nao
    ao = _mat(k, N.DFKS_AO, (4, ng, n))
    assert np.abs(ao[0] - o.ao).max() < 1e-12
    assert np.abs(np.moveaxis(ao[1:], 0, -1) - o.dao).max() < 1e-11


@pytest.mark.parametrize("xc", ["PBE", "PBE0", "B3LYP"])
def test_xc_pointwise(torch_cuda, xc):
    import ctypes
    from dftax_amd import _native as N
    from dftax_amd.xc import resolve
    from oracle import ref
    torch = torch_cuda
    rng = np.random.default_rng(7)
    rho = 10.0 ** rng.uniform(-9.5, 3, size=4000)
    sigma = (rho ** (8 / 3)) * 10.0 ** rng.uniform(-4, 2, size=4000)
    sigma[:10] = 0.0
    r = torch.tensor(rho, device="cuda")
    s = torch.tensor(sigma, device="cuda")
    out = torch.empty((4000, 3), dtype=torch.float64, device="cuda")
    N.check(N.lib().dfks_xc_eval(resolve(xc).xc_id, 4000, ctypes.c_void_p(r.data_ptr()),
                                 ctypes.c_void_p(s.data_ptr()), ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    # oracle: the torch transcription with a gradient vector along x
    rt = torch.tensor(rho, requires_grad=True)
    gt = torch.zeros((4000, 3), dtype=torch.float64)
    gt[:, 0] = torch.tensor(np.sqrt(sigma))
    gt.requires_grad_(True)
    f = ref.eps_xc(xc, rt, gt) * rt
    fr, fg = torch.autograd.grad(f.sum(), (rt, gt))
    f = f.detach().numpy()
    fr = fr.numpy()
    # d f / d grad_x = 2 f_sigma grad_x
    gx = np.sqrt(sigma)
    mask = gx > 0
    fs = np.where(mask, fg.numpy()[:, 0] / np.where(mask, 2 * gx, 1.0), got[:, 2])
    # value to 1e-12 relative; derivatives to 1e-9: autograd through |grad| and forward-mode
    # dual numbers round differently at near-zero gradients (5e-10 measured for PBE at
    # rho = 1.6e-8, sigma = 3e-25), far below what reaches the Fock matrix (tested at 1e-10)
    for col, (a, b, tol) in enumerate(((got[:, 0], f, 1e-12), (got[:, 1], fr, 1e-9), (got[:, 2], fs, 1e-9))):
        err = np.abs(a - b) / np.maximum(1.0, np.abs(b))
        i = int(np.argmax(err))
        assert err[i] <= tol, (col, i, rho[i], sigma[i], a[i], b[i], err[i])


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_energy_fock_guess_density(case, request, torch_cuda):
    """Non-idempotent minao P0: general signed-factor path, all energy terms and F."""
    from oracle import ref
    o, k = request.getfixturevalue(case)
    torch = torch_cuda
    P0 = ref.minao_density(o.mol, o.basis)
    P0 *= o.mol.nelectron / np.sum(P0 * o.S)
    e_ref, F_ref, parts = o.energy_fock(P0)
    E, F = k.coulomb.energy_fock(torch.tensor(P0, device="cuda"))
    E = E.cpu().numpy()
    F = F.cpu().numpy()
    _say(f"E gpu {E[0]:.12f} oracle {e_ref:.12f}")
    assert abs(E[1] - parts["e1"]) < 1e-9
    assert abs(E[2] - parts["eJ"]) < 1e-9
    assert abs(E[3] - parts["eK"]) < 1e-9
    assert abs(E[4] - parts["exc"]) < 1e-9
    assert abs(E[0] - e_ref) < TOL_E
    assert np.abs(F - F_ref).max() < TOL_FOCK


def test_scf_water_pbe0(water_pair):
    """Full SCF: device DIIS + eigensolve against the oracle's restated loop."""
    from dftax_amd import scf
    from oracle import ref
    o, k = water_pair
    r_ref = ref.scf(o, e_tol=1e-9)
    r = scf(k, e_tol=1e-9)
    _say(f"SCF gpu {r.e_tot:.12f} ({r.n_iter} it) oracle {r_ref['e_tot']:.12f} ({r_ref['n_iter']} it)")
    assert r.converged and r_ref["converged"]
    assert abs(r.e_tot - r_ref["e_tot"]) < TOL_E
    P = r.P[0].cpu().numpy()
    # idempotency P S P = 2 P and Tr(P S) = N
    S = o.S
    assert np.abs(P @ S @ P - 2 * P).max() < 1e-8
    assert abs(np.sum(P * S) - o.mol.nelectron) < 1e-10


# ---------------------------------------------------------------------------
# MFMA GEMM backends and the mixed-precision path
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def dimer_backends(torch_cuda, water_dimer_tzvp):
    from dftax_amd import KS, becke, df, hip
    _say("case: water dimer def2-TZVP PBE0, becke(35,110): rocBLAS vs MFMA vs mixed vs eigensolver density")
    g = becke(35, 110)
    out = {}
    for name, eng in (("rocblas", hip(gemm="rocblas")), ("mfma", hip(precision="mixed", gemm="mfma")),
                      ("eigh", hip(gemm="rocblas", density="eigh"))):
        out[name] = KS(water_dimer_tzvp, "PBE0", grid=g, coulomb=df(engine=eng))
    return out


def _p0(ks):
    from oracle import ref
    P0 = ref.minao_density(ks.mol, ks.basis)
    S = ks.S.cpu().numpy()
    return P0 * (ks.mol.nelectron / np.sum(P0 * S))


def test_mfma_fp64_matches_rocblas(dimer_backends, torch_cuda):
    torch = torch_cuda
    a, b = dimer_backends["rocblas"], dimer_backends["mfma"]
    P0 = torch.tensor(_p0(a), device="cuda")
    Ea, Fa = a.coulomb.energy_fock(P0)
    b.coulomb.set_precision(False)
    Eb, Fb = b.coulomb.energy_fock(P0)
    assert np.abs((Fa - Fb).cpu().numpy()).max() < 1e-11
    assert np.abs((Ea - Eb).cpu().numpy()).max() < 1e-10


def test_mixed_fock_error_is_bounded(dimer_backends, torch_cuda):
    """Split-BF16 products (FP32 accumulate) against FP64: relative Fock error ~1e-5."""
    torch = torch_cuda
    k = dimer_backends["mfma"]
    P0 = torch.tensor(_p0(k), device="cuda")
    k.coulomb.set_precision(False)
    E64, F64 = k.coulomb.energy_fock(P0)
    k.coulomb.set_precision(True)
    Em, Fm = k.coulomb.energy_fock(P0)
    k.coulomb.set_precision(False)
    dF = np.abs((Fm - F64).cpu().numpy()).max()
    dE = abs(float(Em[0] - E64[0]))
    _say(f"mixed vs fp64: max|dF| = {dF:.3e}, |dE| = {dE:.3e}")
    assert dF < 1e-3 and dE < 1e-3
    assert dF > 0.0


def test_mixed_scf_refines_to_fp64(dimer_backends):
    from dftax_amd import scf
    k = dimer_backends["mfma"]
    r64 = scf(k, precision="fp64", e_tol=1e-9)
    rmx = scf(k, precision="mixed", e_tol=1e-9)
    nmix = sum(1 for h in rmx.history if h[3] == "mixed")
    _say(f"fp64 {r64.e_tot:.12f} ({r64.n_iter} it); mixed {rmx.e_tot:.12f} ({rmx.n_iter} it, {nmix} mixed)")
    assert r64.converged and rmx.converged and nmix >= 1
    assert abs(rmx.e_tot - r64.e_tot) < 1e-8


def test_purification_graph_survives_scf_restart(dimer_backends):
    """A second SCF on the same handle (a new dfks_scf_begin) purifies every
    step again and repeats the first SCF: the recorded purification graph only
    refers to buffers that live as long as the handle. (Repeated builds agree
    to rounding, not bitwise: the Vxc row assembly sums in LDS with atomics.)"""
    from dftax_amd import scf
    k = dimer_backends["rocblas"]
    r1 = scf(k, precision="fp64", e_tol=1e-9)
    r2 = scf(k, precision="fp64", e_tol=1e-9)
    st = k.coulomb.stats()
    _say(f"SCF twice on one handle: {r1.e_tot:.12f} ({r1.n_iter} it), {r2.e_tot:.12f} ({r2.n_iter} it), "
         f"eigensolver steps {st['eigh_steps']}")
    assert r1.converged and r2.converged and st["eigh_steps"] == 0
    assert r2.n_iter == r1.n_iter and abs(r2.e_tot - r1.e_tot) < 1e-12


def test_purified_density_matches_eigensolver(dimer_backends, torch_cuda):
    """SP2 purification + occupied factor (pivoted Cholesky, then projections of the
    previous step's factor) against an FP64 eigensolve in every step.

    Both runs start from the same guess; with a HOMO-LUMO gap the purified
    projector is the aufbau density, so the two SCF trajectories agree step by
    step and the lazily computed orbitals of the last step match.
    """
    from dftax_amd import scf
    a, e = dimer_backends["rocblas"], dimer_backends["eigh"]
    ra = scf(a, e_tol=1e-10, d_tol=1e-7)
    npc = a.coulomb.stats()["pchol_steps"]
    re = scf(e, e_tol=1e-10, d_tol=1e-7)
    _say(f"purified {ra.e_tot:.12f} ({ra.n_iter} it, {npc} pivoted-Cholesky factors, the rest projected); "
         f"eigensolver {re.e_tot:.12f} ({re.n_iter} it)")
    assert ra.converged and re.converged
    assert ra.n_iter == re.n_iter
    # the first purified step factors X by pivoted Cholesky; later steps project the previous factor
    assert 1 <= npc < ra.n_iter - 2
    dEs = [abs(x[0] - y[0]) for x, y in zip(ra.history, re.history)]
    assert max(dEs) < 1e-9, dEs
    assert abs(ra.e_tot - re.e_tot) < 1e-10
    assert np.abs((ra.P - re.P).cpu().numpy()).max() < 1e-8
    assert np.abs((ra.mo_energy - re.mo_energy).cpu().numpy()).max() < 1e-8
    nocc = a.nocc[0]
    Ca, Ce = ra.mo_coeff[0][:, :nocc], re.mo_coeff[0][:, :nocc]
    assert np.abs((Ca @ Ca.T - Ce @ Ce.T).cpu().numpy()).max() < 1e-8


# ---------------------------------------------------------------------------
# Recomputed auxiliary slabs, streamed AO grid, Schwarz screening
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def dimer_recompute(torch_cuda, water_dimer_tzvp):
    from dftax_amd import KS, becke, df, hip
    _say("case: water dimer def2-TZVP PBE0 becke(35,110): recomputed slabs / streamed grid / screening vs materialised")
    g = becke(35, 110)
    eng = dict(gemm="rocblas")
    specs = {
        "mat": df(chunk=None, engine=hip(ao_block=0, xc_grid="dense", **eng)),
        "rec": df(chunk=48, engine=hip(ao_block=1000, xc_grid="dense", **eng)),
        "scr": df(chunk=48, screen=1e-12, engine=hip(**eng)),
        "scr4": df(chunk=None, screen=1e-4, engine=hip(**eng)),
        "recmx": df(chunk=48, screen=1e-12, engine=hip(precision="mixed", ao_block=1000, xc_grid="dense", **eng)),
        "blk": df(chunk=None, engine=hip(ao_block=0, **eng)),
        "blkmx": df(chunk=48, screen=1e-12, engine=hip(precision="mixed", **eng)),
    }
    return {k: KS(water_dimer_tzvp, "PBE0", grid=g, coulomb=v) for k, v in specs.items()}


def test_recompute_matches_materialised(dimer_recompute, torch_cuda):
    """Slab-by-slab (mn|P) with the metric applied to the exchange intermediate,
    J from a second slab pass, AO grid per 1000-point block: same E and F."""
    torch = torch_cuda
    m, r = dimer_recompute["mat"], dimer_recompute["rec"]
    st = r.coulomb.stats()
    _say(f"recompute stats {st}")
    assert st["aux_slabs"] > 4 and st["max_slab"] <= 48 + 15 and st["ao_block"] == 1000
    P0 = torch.tensor(_p0(m), device="cuda")
    Em, Fm = m.coulomb.energy_fock(P0)
    Er, Fr = r.coulomb.energy_fock(P0)
    dF = np.abs((Fm - Fr).cpu().numpy()).max()
    dE = np.abs((Em - Er).cpu().numpy()).max()
    _say(f"recompute vs materialised: max|dF| = {dF:.2e}, max|dE| = {dE:.2e}")
    assert dF < 1e-10 and dE < 1e-9


def test_schwarz_screening(dimer_recompute, torch_cuda):
    """screen=1e-12 changes F by less than 1e-9; a loose 1e-4 drops pairs and
    moves F by a bounded amount (the dropped pairs really are skipped)."""
    torch = torch_cuda
    m = dimer_recompute["mat"]
    P0 = torch.tensor(_p0(m), device="cuda")
    Em, Fm = m.coulomb.energy_fock(P0)
    s12, s4 = dimer_recompute["scr"], dimer_recompute["scr4"]
    E12, F12 = s12.coulomb.energy_fock(P0)
    E4, F4 = s4.coulomb.energy_fock(P0)
    st12, st4 = s12.coulomb.stats(), s4.coulomb.stats()
    d12 = np.abs((F12 - Fm).cpu().numpy()).max()
    d4 = np.abs((F4 - Fm).cpu().numpy()).max()
    _say(f"screen 1e-12: kept {st12['pairs_kept']}/{st12['pairs_total']} max|dF| {d12:.2e}; "
         f"screen 1e-4: kept {st4['pairs_kept']}/{st4['pairs_total']} max|dF| {d4:.2e}")
    assert d12 < 1e-9
    assert st4["pairs_kept"] < st4["pairs_total"]
    assert 0.0 < d4 < 1e-2


def test_recompute_scf(dimer_recompute):
    from dftax_amd import scf
    m, r = dimer_recompute["mat"], dimer_recompute["rec"]
    rm = scf(m, e_tol=1e-9)
    rr = scf(r, e_tol=1e-9)
    _say(f"materialised {rm.e_tot:.12f} ({rm.n_iter} it); recomputed {rr.e_tot:.12f} ({rr.n_iter} it)")
    assert rm.converged and rr.converged
    assert abs(rm.e_tot - rr.e_tot) < 1e-9


def test_recompute_mixed(dimer_recompute, torch_cuda):
    """Mixed steps on the recompute + streamed-grid path: split-BF16 slab and
    metric products (batched), FP32 streamed XC; bounded Fock error, and the
    SCF refined in FP64 lands on the materialised FP64 energy."""
    from dftax_amd import scf
    torch = torch_cuda
    m, r = dimer_recompute["mat"], dimer_recompute["recmx"]
    P0 = torch.tensor(_p0(m), device="cuda")
    Em, Fm = m.coulomb.energy_fock(P0)
    r.coulomb.set_precision(False)
    E64, F64 = r.coulomb.energy_fock(P0)
    r.coulomb.set_precision(True)
    Emx, Fmx = r.coulomb.energy_fock(P0)
    r.coulomb.set_precision(False)
    d64 = np.abs((F64 - Fm).cpu().numpy()).max()
    dmx = np.abs((Fmx - Fm).cpu().numpy()).max()
    _say(f"recompute mixed: FP64 max|dF| {d64:.2e}, mixed max|dF| {dmx:.2e}")
    assert d64 < 1e-9
    assert 0.0 < dmx < 1e-3 and abs(float(Emx[0] - Em[0])) < 1e-3
    rm = scf(m, e_tol=1e-9)
    rx = scf(r, precision="mixed", e_tol=1e-9)
    nmix = sum(1 for h in rx.history if h[3] == "mixed")
    _say(f"materialised {rm.e_tot:.12f}; recompute mixed {rx.e_tot:.12f} ({rx.n_iter} it, {nmix} mixed)")
    assert rx.converged and nmix >= 1
    assert abs(rx.e_tot - rm.e_tot) < 1e-8
    # mixed to the switch, then FP64 steps until the energy settles: within the 1e-6 Ha budget
    rr = scf(r, precision="mixed", refine=3)
    nfp = sum(1 for h in rr.history if h[3] == "fp64")
    _say(f"mixed + {nfp} FP64 steps {rr.e_tot:.12f} ({rr.n_iter} it), |dE| {abs(rr.e_tot - rm.e_tot):.1e}")
    assert rr.energy_converged and 2 <= nfp <= 3 and abs(rr.e_tot - rm.e_tot) < 1e-6


# ---------------------------------------------------------------------------
# Response Fock matrix and the Newton solver (dftax/ks/newton.py)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("xc", ["PBE", "PBE0", "B3LYP"])
def test_xc_second_derivatives(torch_cuda, xc):
    """f_rr, f_rs, f_ss (dual of duals) against double autograd of the oracle's
    functional written on sigma."""
    import ctypes
    from dftax_amd import _native as N
    from dftax_amd.xc import resolve
    from oracle import ref
    torch = torch_cuda
    rng = np.random.default_rng(11)
    n = 3000
    rho = 10.0 ** rng.uniform(-6, 2, size=n)
    sigma = (rho ** (8 / 3)) * 10.0 ** rng.uniform(-3, 1, size=n)
    r = torch.tensor(rho, device="cuda")
    s = torch.tensor(sigma, device="cuda")
    out = torch.empty((n, 6), dtype=torch.float64, device="cuda")
    N.check(N.lib().dfks_xc_eval2(resolve(xc).xc_id, n, ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(s.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    rt = torch.tensor(rho, requires_grad=True)
    st = torch.tensor(sigma, requires_grad=True)
    grad = torch.stack([torch.sqrt(st), torch.zeros_like(st), torch.zeros_like(st)], dim=1)
    f = ref.eps_xc(xc, rt, grad) * rt
    fr, fs = torch.autograd.grad(f.sum(), (rt, st), create_graph=True)
    frr, frs = torch.autograd.grad(fr.sum(), (rt, st), retain_graph=True)
    fsr, fss = torch.autograd.grad(fs.sum(), (rt, st))
    want = [f.detach().numpy(), fr.detach().numpy(), fs.detach().numpy(), frr.numpy(), frs.numpy(), fss.numpy()]
    assert np.abs(frs.numpy() - fsr.numpy()).max() <= 1e-9 * np.abs(frs.numpy()).max()
    # value to 1e-12; derivatives to 1e-7 / 1e-6 relative: in PBE the exchange and
    # correlation gradient terms cancel at small reduced gradient (mu = beta pi^2 / 3),
    # so the two FP64 evaluation orders differ by ~2e-8 of f_sigma there
    for col, (b, tol) in enumerate(zip(want, (1e-12, 1e-7, 1e-7, 1e-6, 1e-6, 1e-6))):
        a = got[:, col]
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * np.abs(b).max())
        i = int(np.argmax(err))
        _say(f"{xc} col {col}: max rel err {err[i]:.2e}")
        assert err[i] <= tol, (col, i, rho[i], sigma[i], a[i], b[i])


def _uv(n, r, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, r)), rng.standard_normal((n, r))


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_response_fock_matches_oracle(case, request, torch_cuda):
    """R = J[dP] - a_x/2 K[dP] + f_xc(P)[dP] against the oracle's forward-over-reverse
    derivative of F = sym(dE/dP), at the minao density, dP = U V^T + V U^T."""
    torch = torch_cuda
    o, k = request.getfixturevalue(case)
    P = _p0(k)
    U, V = _uv(k.nao, 4, 5)
    R_ref = o.response_fock(P, U @ V.T + V @ U.T)
    k.coulomb.response_begin(torch.tensor(P, device="cuda"))
    R = k.coulomb.response_fock(torch.tensor(U, device="cuda"), torch.tensor(V, device="cuda")).cpu().numpy()
    err = np.abs(R - R_ref).max()
    _say(f"{case}: response max|dR| {err:.2e} (max|R| {np.abs(R_ref).max():.2e})")
    assert err < 1e-9 * max(1.0, np.abs(R_ref).max())


def test_response_fock_recompute(dimer_recompute, torch_cuda):
    """Recomputed slabs + streamed grid blocks give the materialised response."""
    torch = torch_cuda
    m, r = dimer_recompute["mat"], dimer_recompute["rec"]
    P = torch.tensor(_p0(m), device="cuda")
    U, V = (torch.tensor(a, device="cuda") for a in _uv(m.nao, 7, 6))
    out = []
    for ks in (m, r):
        ks.coulomb.set_precision(False)
        ks.coulomb.response_begin(P)
        out.append(ks.coulomb.response_fock(U, V).cpu().numpy())
    err = np.abs(out[0] - out[1]).max()
    _say(f"recompute response max|dR| {err:.2e}")
    assert err < 1e-10 * max(1.0, np.abs(out[0]).max())


def test_newton_hessian_quadratic_form(water_pair, torch_cuda):
    """<v, H v> from the response against the second difference of E(C exp(K(v)));
    H symmetric: <w, H v> = <v, H w>."""
    import math
    from dftax_amd import scf
    from dftax_amd.newton import _rotate
    torch = torch_cuda
    _, k = water_pair
    res = scf(k, e_tol=1e-10)
    C = res.mo_coeff[0]
    no = k.nocc[0]
    c = k.coulomb
    E0, F, P = c.energy_fock_factored(math.sqrt(2.0) * C[:, :no])
    Ft = C.T @ F @ C
    Co, Cv = C[:, :no], C[:, no:]
    c.response_begin(P)

    def hvp(v):
        R = c.response_fock(Co, -2.0 * (Cv @ v.T))
        return 4.0 * (v @ Ft[no:, no:] - Ft[:no, :no] @ v) - 4.0 * (Co.T @ R @ Cv)

    rng = np.random.default_rng(2)
    v = torch.tensor(rng.standard_normal((no, C.shape[1] - no)), device="cuda")
    v = v / torch.linalg.norm(v)
    w = torch.tensor(rng.standard_normal((no, C.shape[1] - no)), device="cuda")
    Hv, Hw = hvp(v), hvp(w)
    a, b = float(torch.sum(w * Hv)), float(torch.sum(v * Hw))
    assert abs(a - b) < 1e-9 * max(1.0, abs(a))

    def energy(kappa):
        Cr = _rotate(C, kappa, no)
        E, _, _ = c.energy_fock_factored(math.sqrt(2.0) * Cr[:, :no])
        return float(E[0])

    eps = 1e-3
    quad = (energy(eps * v) - 2.0 * float(E0[0]) + energy(-eps * v)) / eps ** 2
    vHv = float(torch.sum(v * Hv))
    _say(f"<v,Hv> = {vHv:.8f}, second difference {quad:.8f}")
    # |eps v| = 1e-3: quartic terms ~1e-6 relative, rounding of E ~1e-7 relative
    assert abs(quad - vHv) < 1e-4 * abs(vHv)


@pytest.mark.parametrize("case", ["water_pair", "dimer_pair"])
def test_newton_matches_scf(case, request, torch_cuda):
    """Trust-region Newton lands on the DIIS SCF minimum and on the oracle's energy."""
    from dftax_amd import scf
    from dftax_amd.newton import newton
    from oracle import ref
    o, k = request.getfixturevalue(case)
    rs = scf(k, e_tol=1e-10)
    rn = newton(k)
    ro = ref.scf(o, e_tol=1e-10, d_tol=1e-7)
    nh = sum(h[5] for h in rn.history)
    _say(f"{case}: newton {rn.e_tot:.12f} ({rn.n_iter} it, {nh} Hessian products); scf {rs.e_tot:.12f}; "
          f"oracle {ro['e_tot']:.12f}")
    assert rn.converged
    assert abs(rn.e_tot - rs.e_tot) < 1e-8
    assert abs(rn.e_tot - ro["e_tot"]) < TOL_E
    Pn = rn.P[0].cpu().numpy()
    assert np.abs(Pn - rs.P[0].cpu().numpy()).max() < 1e-5


# ---------------------------------------------------------------------------
# Compact XC grid (blocks of 128 sorted points, significant AOs only)
# ---------------------------------------------------------------------------

def test_xc_blocks_match_dense(dimer_recompute, torch_cuda):
    """Compact-block XC against dense AO rows: energy, Fock and response; the
    dropped AO values are bounded by 1e-14."""
    torch = torch_cuda
    d, b = dimer_recompute["mat"], dimer_recompute["blk"]
    st = b.coulomb.stats()
    _say(f"compact XC: {st['xc_blocks']} blocks in {st['xc_bins']} bins, mean width "
         f"{st['xc_width_sum'] / max(st['xc_blocks'], 1):.0f} of nao {b.nao}")
    assert st["xc_blocks"] > 0 and d.coulomb.stats()["xc_blocks"] == 0
    P = torch.tensor(_p0(d), device="cuda")
    Ed, Fd = d.coulomb.energy_fock(P)
    Eb, Fb = b.coulomb.energy_fock(P)
    dF = np.abs((Fd - Fb).cpu().numpy()).max()
    dE = abs(float(Ed[4] - Eb[4]))
    _say(f"compact vs dense: max|dF| {dF:.2e}, |dE_xc| {dE:.2e}")
    assert dF < 1e-10 and dE < 1e-10
    U, V = (torch.tensor(a, device="cuda") for a in _uv(d.nao, 5, 8))
    out = []
    for ks in (d, b):
        ks.coulomb.response_begin(P)
        out.append(ks.coulomb.response_fock(U, V).cpu().numpy())
    dR = np.abs(out[0] - out[1]).max()
    _say(f"compact vs dense response: max|dR| {dR:.2e}")
    assert dR < 1e-10 * max(1.0, np.abs(out[0]).max())


def test_xc_blocks_mixed_scf(dimer_recompute, torch_cuda):
    """Mixed precision on the compact grid with recomputed, screened slabs: the
    refined SCF lands on the dense materialised FP64 energy."""
    from dftax_amd import scf
    torch = torch_cuda
    d, bm = dimer_recompute["mat"], dimer_recompute["blkmx"]
    P = torch.tensor(_p0(d), device="cuda")
    _, Fd = d.coulomb.energy_fock(P)
    bm.coulomb.set_precision(True)
    _, Fm = bm.coulomb.energy_fock(P)
    bm.coulomb.set_precision(False)
    dmx = np.abs((Fm - Fd).cpu().numpy()).max()
    _say(f"compact mixed Fock max|dF| {dmx:.2e}")
    assert 0.0 < dmx < 1e-3
    rd = scf(d, e_tol=1e-9)
    rb = scf(bm, precision="mixed", e_tol=1e-9)
    _say(f"dense FP64 {rd.e_tot:.12f}; compact mixed + FP64 {rb.e_tot:.12f} ({rb.n_iter} it)")
    assert rb.converged and abs(rb.e_tot - rd.e_tot) < 1e-8


def test_exchange_t160_matches_rocblas(torch_cuda):
    """The library kernels of the exchange product T = B^T C against rocBLAS:
    the 160-wide FP64 tile, the 8-wavefront 128 x 160 tile and the autotuned
    choice give the same Fock matrix. (H2O)_28 / def2-SVP has 140 occupied
    orbitals (one partial N tile, nao = 672 a multiple of the k-step);
    (H2O)_33 has 165 (two N tiles, the second partial) and nao = 792 (a
    partial last k-step)."""
    from dftax_amd import KS, becke, df, hip, water_cluster
    from dftax_amd.scf import initial_density
    torch = torch_cuda
    for shape, nocc in (((4, 7, 1), 140), ((3, 11, 1), 165)):
        mol = water_cluster(shape, seed=0, basis="def2-svp")
        F = {}
        for g in ("rocblas", "mfma", "mfma8", "auto"):
            ks = KS(mol, "PBE0", grid=becke(30, 110), coulomb=df(engine=hip(gemm=g)))
            assert ks.nocc[0] == nocc
            E, F[g] = ks.coulomb.energy_fock(initial_density(ks, None))
            F[g] = F[g].cpu().numpy()
            ks.coulomb.close()
        d = {g: float(np.abs(F[g] - F["rocblas"]).max()) for g in ("mfma", "mfma8", "auto")}
        _say(f"(H2O)_{len(mol.symbols) // 3} SVP, nocc {nocc}: T vs rocBLAS max|dF| 160-wide {d['mfma']:.2e}, "
             f"8-wave {d['mfma8']:.2e}, autotuned {d['auto']:.2e}")
        assert max(d.values()) < 1e-11
        # mixed steps: the split-BF16 T of k_gemm and of the 8-wavefront 256-row tile
        Fm = {}
        for g in ("mfma", "mfma8", "mfma8d"):
            ks = KS(mol, "PBE0", grid=becke(30, 110), coulomb=df(engine=hip(precision="mixed", gemm=g)))
            ks.coulomb.set_precision(True)
            _, Fm[g] = ks.coulomb.energy_fock(initial_density(ks, None))
            Fm[g] = Fm[g].cpu().numpy()
            ks.coulomb.close()
        dm = {g: float(np.abs(Fm[g] - F["rocblas"]).max()) for g in Fm}
        dmm = float(np.abs(Fm["mfma8"] - Fm["mfma"]).max())
        dmd = float(np.abs(Fm["mfma8d"] - Fm["mfma8"]).max())
        _say(f"  mixed vs FP64 max|dF| k_gemm {dm['mfma']:.2e}, 8-wave {dm['mfma8']:.2e}; between them {dmm:.2e}; "
             f"8-wave with two k-steps in flight vs one {dmd:.2e}")
        assert max(dm.values()) < 1e-3 and dmm < 1e-5
        # the same products in the same order (only the load schedule differs): equal up to the
        # run-to-run spread of the Fock build (the row assembly's LDS sums)
        assert dmd <= 1e-12
