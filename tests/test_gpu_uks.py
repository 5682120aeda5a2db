"""Spin-unrestricted (nspin = 2) Fock builds and SCF on the GPU against the oracle.

dftax's open-shell path (dftax/ks/terms.py:500-527, 838-874; the spin branches
of dftax/energy/xc.py; dftax/ks/scf.py:279-391 with two channels): the library
builds E and F_s = sym(dE/dP_s) of a spin-stacked P[2, nao, nao] through
dfks_energy_fock, the host loop runs DIIS over both channels. Cases follow the
reference's UKS tests (tests/unit/test_uks.py): the CH3 doublet and the O2
triplet, and the closed-shell limit UKS = RKS.

Tolerances: the pointwise functional to 1e-12 relative (values) and 1e-9
relative (derivatives), as the restricted functionals; Fock matrices 1e-10
max-abs and energy components 1e-9 Ha, the restricted path's bar; SCF
energies 1e-8 Ha.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CH3 = "C 0 0 0; H 0 1.079 0; H 0.934 -0.539 0; H -0.934 -0.539 0"
O2 = "O 0 0 0; O 0 0 1.21"
WATER = "O 0 0 0; H 0.757 0.587 0; H -0.757 0.587 0"


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _oracle_p0(oks, mol):
    from oracle import ref
    P = ref.minao_density(mol, oks.basis)
    P = P * (mol.nelectron / np.sum(P * oks.S))
    na, nb = oks.nocc_ab
    return np.stack([na / mol.nelectron * P, nb / mol.nelectron * P])


@pytest.mark.parametrize("xc", ["PBE", "PBE0", "B3LYP"])
def test_spin_pointwise_functional(xc, torch_cuda):
    """dfks_xc_eval_pol against autograd of the oracle's spin-polarised forms:
    f and df/drho_s directly, df/dgrad rho_a = 2 f_saa grad rho_a + f_sab grad rho_b
    (and the beta analogue) through the chain rule."""
    torch = torch_cuda
    from dftax_amd import _native as N
    from oracle import ref
    rng = np.random.default_rng(3)
    n = 64
    ra = 10.0 ** rng.uniform(-3, 1.5, n)
    rb = 10.0 ** rng.uniform(-3, 1.5, n)
    ga = rng.normal(size=(n, 3)) * ra[:, None] ** (4 / 3)
    gb = rng.normal(size=(n, 3)) * rb[:, None] ** (4 / 3)
    in5 = np.stack([ra, rb, np.sum(ga * ga, 1), np.sum(ga * gb, 1), np.sum(gb * gb, 1)], 1)
    din = torch.tensor(in5, device="cuda")
    out = torch.empty((n, 6), dtype=torch.float64, device="cuda")
    N.check(N.lib().dfks_xc_eval_pol({"PBE": 1, "PBE0": 2, "B3LYP": 3}[xc], n, ctypes.c_void_p(din.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), N.current_stream()))
    got = out.cpu().numpy()
    t = [torch.tensor(x, requires_grad=True) for x in (ra, rb, ga, gb)]
    f = ref.eps_xc_pol(xc, *t) * (t[0] + t[1])
    d = torch.autograd.grad(f.sum(), t)
    f = f.detach().numpy()
    d_ra, d_rb, d_ga, d_gb = (x.numpy() for x in d)
    ch_ga = 2 * got[:, 3:4] * ga + got[:, 4:5] * gb
    ch_gb = 2 * got[:, 5:6] * gb + got[:, 4:5] * ga
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    _say(f"{xc}: f {rel(got[:, 0], f):.1e}, d/drho_a {rel(got[:, 1], d_ra):.1e}, d/drho_b {rel(got[:, 2], d_rb):.1e}, "
         f"d/dgrad_a {rel(ch_ga, d_ga):.1e}, d/dgrad_b {rel(ch_gb, d_gb):.1e}")
    assert rel(got[:, 0], f) < 1e-12
    assert rel(got[:, 1], d_ra) < 1e-9 and rel(got[:, 2], d_rb) < 1e-9
    assert rel(ch_ga, d_ga) < 1e-9 and rel(ch_gb, d_gb) < 1e-9


CASES = {"ch3_svp_pbe0": (CH3, "def2-svp", 1, "PBE0"), "ch3_svp_b3lyp": (CH3, "def2-svp", 1, "B3LYP"),
         "o2_sto3g_pbe": (O2, "sto-3g", 2, "PBE")}


@pytest.fixture(scope="module")
def uks_cases(torch_cuda):
    from dftax_amd import KS, becke, df, hip
    from dftax_amd.system import Molecule
    from oracle import ref
    out = {}
    for name, (geom, basis, spin, xc) in CASES.items():
        _say(f"case: {name} (spin {spin}), becke(30, 110)")
        mol = Molecule.from_xyz(geom, basis, spin=spin, spherical=True)
        g = becke(30, 110)
        oks = ref.OracleKS(mol, xc, grid=g)
        ks = KS(mol, xc, grid=g, coulomb=df(engine=hip()))
        out[name] = (mol, oks, ks)
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_uks_fock_matches_oracle(case, uks_cases, torch_cuda):
    torch = torch_cuda
    mol, oks, ks = uks_cases[case]
    assert ks.nspin == 2 and ks.nocc == oks.nocc_ab
    P0 = _oracle_p0(oks, mol)
    e_ref, F_ref, parts = oks.energy_fock_uks(P0)
    E, F = ks.coulomb.energy_fock(torch.tensor(P0, device="cuda"))
    E = E.cpu().numpy()
    dF = np.abs(F.cpu().numpy() - F_ref).max()
    dE = [abs(E[0] - e_ref)] + [abs(E[k] - parts[n]) for k, n in ((1, "e1"), (2, "eJ"), (3, "eK"), (4, "exc"))]
    _say(f"{case}: max|dF| alpha/beta {dF:.2e}, |dE| total/1e/J/K/XC {' '.join(f'{x:.1e}' for x in dE)}")
    assert dF < 1e-10
    assert max(dE) < 1e-9


def test_uks_scf_matches_oracle(uks_cases, torch_cuda):
    from dftax_amd import scf
    from oracle import ref
    for case in ("ch3_svp_pbe0", "o2_sto3g_pbe"):
        mol, oks, ks = uks_cases[case]
        r = scf(ks, e_tol=1e-10, d_tol=1e-7)
        o = ref.scf_uks(oks, e_tol=1e-10, d_tol=1e-7)
        P = r.P.cpu().numpy()
        S = oks.S
        na, nb = oks.nocc_ab
        _say(f"{case}: GPU UKS {r.e_tot:.10f} ({r.n_iter} it), oracle {o['e_tot']:.10f} ({o['n_iter']} it); "
             f"Tr(P_a S) {np.sum(P[0] * S):.8f} Tr(P_b S) {np.sum(P[1] * S):.8f}")
        assert r.converged and o["converged"]
        assert abs(r.e_tot - o["e_tot"]) < 1e-8
        assert abs(np.sum(P[0] * S) - na) < 1e-8 and abs(np.sum(P[1] * S) - nb) < 1e-8
        assert r.mo_coeff.shape[0] == 2 and r.nocc == (na, nb)


def test_uks_closed_shell_limit(torch_cuda):
    """KS(mol, spin=0) builds two channels; at P_a = P_b = P/2 its energy and
    Fock matrices differ from the restricted handle's only where the spin path
    masks differently (terms.py:838-874 clamps each channel at 1e-10, so points
    with 1e-10 < rho <= 2e-10 enter as rho_s = 1e-10 with zero gradient): that
    difference (~1e-9 Ha, ~1e-8 in F for water/def2-SVP) must equal the oracle's
    to 1e-10, and the SCF lands on the restricted energy (test_uks.py: UKS
    reduces to RKS, there at 1e-7)."""
    torch = torch_cuda
    from dftax_amd import KS, becke, df, hip, scf
    from dftax_amd.system import Molecule
    from oracle import ref
    mol = Molecule.from_xyz(WATER, "def2-svp", spherical=True)
    g = becke(30, 110)
    kr = KS(mol, "B3LYP", grid=g, coulomb=df(engine=hip()))
    ku = KS(mol, "B3LYP", grid=g, coulomb=df(engine=hip()), spin=0)
    assert kr.nspin == 1 and ku.nspin == 2
    P = ref.minao_density(mol, kr.basis)
    P = torch.tensor(P * (mol.nelectron / np.sum(P * kr.S.cpu().numpy())), device="cuda")
    Er, Fr = kr.coulomb.energy_fock(P)
    Eu, Fu = ku.coulomb.energy_fock(torch.stack([P / 2, P / 2]))
    dE = float(Eu[0] - Er[0])
    dF = (Fu - Fr[None]).cpu().numpy()
    oks = ref.OracleKS(mol, "B3LYP", grid=g)
    Pn = P.cpu().numpy()
    e_r, F_r, _ = oks.energy_fock_autograd(Pn)
    e_u, F_u, _ = oks.energy_fock_uks(np.stack([Pn / 2, Pn / 2]))
    rr, ru = scf(kr, e_tol=1e-10), scf(ku, e_tol=1e-10)
    _say(f"closed-shell limit: UKS - RKS dE {dE:.3e} (oracle {e_u - e_r:.3e}), max|dF| {np.abs(dF).max():.2e} "
         f"(vs the oracle's difference {np.abs(dF - (F_u - F_r[None])).max():.1e}); SCF RKS {rr.e_tot:.10f} "
         f"UKS {ru.e_tot:.10f}")
    assert abs(dE - (e_u - e_r)) < 1e-10
    assert np.abs(dF - (F_u - F_r[None])).max() < 1e-10
    assert abs(rr.e_tot - ru.e_tot) < 1e-8


def test_uks_handle_refuses_closed_shell_calls(uks_cases, torch_cuda):
    """The device SCF, response, forces, factored Fock and mixed precision are
    restricted-only: status 7 on an nspin = 2 handle."""
    from dftax_amd import _native as N
    _, _, ks = uks_cases["o2_sto3g_pbe"]
    L, h = N.lib(), ks.coulomb.handle
    assert L.dfks_set_precision(h, 1) == 7
    assert L.dfks_scf_step(h, None, None) == 7
    assert L.dfks_response_begin(h, None, None) == 7
    assert b"nspin = 1" in L.dfks_last_error()
