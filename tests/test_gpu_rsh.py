"""Range-separated hybrids (CAM-B3LYP) on the GPU against the oracle.

dftax's range-separated path (dftax/ks/terms.py:578-612, energy.py:419-433,
xc.py:604-750): a second fitted tensor B_lr from the erf(omega r)/r integrals
(attenuated Boys values, eri3c_bucketed.py:340-346) and the attenuated
metric's pseudo-inverse factor, K_lr = sum_i s_i T_i^T T_i over it, and the
CAM-B3LYP semilocal part 0.35 B88 + 0.46 SR-B88 (ITYH) + 0.81 LYP + 0.19 VWN5.
E_K = -1/4 sum P o (0.19 K + 0.46 K_lr), F = H + J - 1/2 (0.19 K + 0.46 K_lr)
+ Vxc.

Tolerances as the global hybrids': the pointwise functional to 1e-12
relative (value) and 1e-9 (derivatives), Fock matrices 1e-10 max-abs and
energy components 1e-9 Ha at the minao density, SCF energies 1e-8 Ha.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAM_ID = 4


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_cam_b3lyp_pointwise(torch_cuda):
    torch = torch_cuda
    from dftax_amd import _native as N
    from oracle import ref
    rng = np.random.default_rng(11)
    n = 4000
    rho = 10.0 ** rng.uniform(-9.5, 3, size=n)
    sigma = (rho ** (8 / 3)) * 10.0 ** rng.uniform(-4, 2, size=n)
    sigma[:10] = 0.0
    r = torch.tensor(rho, device="cuda")
    s = torch.tensor(sigma, device="cuda")
    out = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    N.check(N.lib().dfks_xc_eval(CAM_ID, n, ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(s.data_ptr()),
                                 ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    rt = torch.tensor(rho, requires_grad=True)
    gt = torch.zeros((n, 3), dtype=torch.float64)
    gt[:, 0] = torch.tensor(np.sqrt(sigma))
    gt.requires_grad_(True)
    f = ref.eps_xc("CAM-B3LYP", rt, gt) * rt
    fr, fg = torch.autograd.grad(f.sum(), (rt, gt))
    f, fr = f.detach().numpy(), fr.numpy()
    gx = np.sqrt(sigma)
    mask = gx > 0
    fs = np.where(mask, fg.numpy()[:, 0] / np.where(mask, 2 * gx, 1.0), got[:, 2])
    errs = []
    for col, (a, b, tol) in enumerate(((got[:, 0], f, 1e-12), (got[:, 1], fr, 1e-9), (got[:, 2], fs, 1e-9))):
        err = np.abs(a - b) / np.maximum(1.0, np.abs(b))
        i = int(np.argmax(err))
        errs.append(err[i])
        assert err[i] <= tol, (col, i, rho[i], sigma[i], a[i], b[i], err[i])
    _say(f"CAM-B3LYP pointwise: f {errs[0]:.1e}, f_rho {errs[1]:.1e}, f_sigma {errs[2]:.1e}")


CASES = {"water_svp": ("O 0 0 0; H 0.757 0.587 0; H -0.757 0.587 0", "def2-svp", True),
         "water_sto3g_cart": ("O 0 0 0; H 0.7586 0 0.5043; H 0.7586 0 -0.5043", "sto-3g", False)}


@pytest.fixture(scope="module")
def cam_cases(torch_cuda):
    from dftax_amd import KS, becke, df, hip
    from dftax_amd.system import Molecule
    from oracle import ref
    out = {}
    for name, (geom, basis, sph) in CASES.items():
        _say(f"case: {name} CAM-B3LYP, becke(30, 110)")
        mol = Molecule.from_xyz(geom, basis, spherical=sph)
        g = becke(30, 110)
        out[name] = (mol, ref.OracleKS(mol, "CAM-B3LYP", grid=g), KS(mol, "CAM-B3LYP", grid=g, coulomb=df(engine=hip())))
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_cam_fock_matches_oracle(case, cam_cases, torch_cuda):
    torch = torch_cuda
    from oracle import ref
    mol, oks, ks = cam_cases[case]
    P0 = ref.minao_density(mol, oks.basis)
    P0 = P0 * (mol.nelectron / np.sum(P0 * oks.S))
    e_ref, F_ref, parts = oks.energy_fock(P0)
    E, F = ks.coulomb.energy_fock(torch.tensor(P0, device="cuda"))
    E = E.cpu().numpy()
    dF = np.abs(F.cpu().numpy() - F_ref).max()
    dE = [abs(E[0] - e_ref)] + [abs(E[k] - parts[n]) for k, n in ((1, "e1"), (2, "eJ"), (3, "eK"), (4, "exc"))]
    _say(f"{case}: max|dF| {dF:.2e}, |dE| total/1e/J/K(+K_lr)/XC {' '.join(f'{x:.1e}' for x in dE)}; "
         f"E_K {parts['eK']:.8f} of which long-range {-0.25 * 0.46 * np.sum(P0 * parts['K_lr']):.8f}")
    assert dF < 1e-10
    assert max(dE) < 1e-9


def test_cam_scf_matches_oracle(cam_cases, torch_cuda):
    from dftax_amd import scf
    from oracle import ref
    mol, oks, ks = cam_cases["water_svp"]
    r = scf(ks, e_tol=1e-10, d_tol=1e-7)
    o = ref.scf(oks, e_tol=1e-10, d_tol=1e-7)
    _say(f"water def2-SVP CAM-B3LYP: GPU {r.e_tot:.10f} ({r.n_iter} it), oracle {o['e_tot']:.10f} ({o['n_iter']} it)")
    assert r.converged and o["converged"]
    assert abs(r.e_tot - o["e_tot"]) < 1e-8


def test_cam_handle_refuses_unbuilt_paths(cam_cases, torch_cuda):
    """Response, forces, mixed precision and the spin-polarised form are not
    built for range-separated functionals: status 7, nothing runs."""
    from dftax_amd import KS, becke, df, hip
    from dftax_amd import _native as N
    mol, _, ks = cam_cases["water_sto3g_cart"]
    L, h = N.lib(), ks.coulomb.handle
    assert L.dfks_response_begin(h, None, None) == 7
    assert b"range-separated" in L.dfks_last_error()
    assert L.dfks_set_precision(h, 1) != 0
    with pytest.raises(N.DfksError, match="range-separated"):
        KS(mol, "CAM-B3LYP", grid=becke(20, 50), coulomb=df(engine=hip(precision="mixed")))
    assert L.dfks_xc_eval_pol(CAM_ID, 0, None, None, None) == 7
