"""Pins of the CPU FP64 oracle (no GPU): closed forms, derivative relations,
invariants and golden data, chosen so that a dropped term, a wrong sign or
index, or a transposed operand in the restatement fails one of them.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from dftax_amd.basis import (BasisSet, build_basis, cart2sph_block, contracted_norm, gto_norm,
                             cart_components, ncart, nsph)
from dftax_amd.system import WATER, Molecule
from oracle import ref


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------

def make_basis(shells, spherical=False):
    """shells: list of (l, centre(3), exps, raw_coefs), normalised like build_basis."""
    ls, centre, prim0, nprim, ao0, cart0, exps, coefs = [], [], [], [], [], [], [], []
    p = ao = cart = 0
    for l, c, e, r in shells:
        e = np.atleast_1d(np.asarray(e, float))
        r = np.atleast_1d(np.asarray(r, float))
        cp = r * np.array([gto_norm(l, x) for x in e])
        if l <= 1:
            cp = cp * contracted_norm(e, cp, cart_components(l)[0])
        ls.append(l); centre.append(np.asarray(c, float)); prim0.append(p); nprim.append(len(e))
        ao0.append(ao); cart0.append(cart); exps.extend(e); coefs.extend(cp)
        p += len(e)
        ao += nsph(l) if (spherical and l >= 2) else ncart(l)
        cart += ncart(l)
    return BasisSet(l=np.array(ls, np.int32), atom=np.zeros(len(ls), np.int32), centre=np.array(centre),
                    prim0=np.array(prim0, np.int32), nprim=np.array(nprim, np.int32), ao0=np.array(ao0, np.int32),
                    cart0=np.array(cart0, np.int32), exps=np.array(exps), coefs=np.array(coefs),
                    spherical=spherical, nao=ao, nao_cart=cart)


def boys_quad(n, t):
    return integrate.quad(lambda u: u ** (2 * n) * math.exp(-t * u * u), 0.0, 1.0, epsabs=1e-15, epsrel=1e-14)[0]


# ---------------------------------------------------------------------------
# Boys function (dftax/energy/boys.py; test strategy of test_boys_function.py)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 3, 6, 10, 14])
def test_boys_quadrature_and_limits(n):
    for t in [0.0, 1e-4, 0.05, 0.33, 1.0, 2.71, 7.9, 15.5, 29.96, 39.9, 40.0, 63.0, 200.0]:
        f = ref.boys(n, t)
        exact = 1.0 / (2 * n + 1) if t == 0.0 else boys_quad(n, t)
        assert abs(f - exact) <= 2e-12 * max(exact, 1e-300) + 1e-16, (n, t)


def test_boys_closed_form_and_recursions():
    for t in np.linspace(1e-3, 45.0, 97):
        f0 = math.sqrt(math.pi) / (2 * math.sqrt(t)) * math.erf(math.sqrt(t))
        assert abs(ref.boys(0, t) - f0) < 1e-13
        for n in range(1, 12):
            # downward recursion F_{n-1} = (2t F_n + e^{-t}) / (2n - 1)
            lhs = ref.boys(n - 1, t)
            rhs = (2 * t * ref.boys(n, t) + math.exp(-t)) / (2 * n - 1)
            # beyond t = 40 the large-t asymptote drops the e^{-t}-scale terms (boys.py:170-176),
            # an absolute error of order e^{-40}
            if t < 40:
                assert abs(lhs - rhs) < 1e-12 * lhs
            else:
                assert abs(lhs - rhs) < 1e-17 + 1e-15 * lhs
            # dF_n/dt = -F_{n+1}
            h = 1e-4
            d = (ref.boys(n, t + h) - ref.boys(n, t - h)) / (2 * h) if t > h else None
            if d is not None:
                assert abs(d + ref.boys(n + 1, t)) < 1e-7 * ref.boys(n + 1, t) + 1e-13


def test_boys_incomplete_gamma():
    for n in (0, 2, 5, 9):
        for t in (0.7, 5.0, 22.0, 38.0):
            a = n + 0.5
            exact = special.gamma(a) * special.gammainc(a, t) / (2 * t ** a)
            assert abs(ref.boys(n, t) - exact) < 1e-12 * exact


# ---------------------------------------------------------------------------
# Integrals: closed forms for s functions
# ---------------------------------------------------------------------------

def _norm_s(a):
    return (2 * a / math.pi) ** 0.75


def test_s_one_electron_closed_forms():
    A, B, C = np.array([0.1, -0.2, 0.3]), np.array([0.9, 0.4, -0.5]), np.array([-0.3, 0.7, 0.2])
    a, b = 1.3, 0.45
    bs = make_basis([(0, A, a, 1.0), (0, B, b, 1.0)])
    S, T, V = ref.int1e(bs, C[None], np.array([2.0]))
    p = a + b
    ab2 = float(np.sum((A - B) ** 2))
    K = math.exp(-a * b / p * ab2)
    s_ab = _norm_s(a) * _norm_s(b) * (math.pi / p) ** 1.5 * K
    t_ab = a * b / p * (3 - 2 * a * b / p * ab2) * s_ab
    P = (a * A + b * B) / p
    v_ab = -2.0 * 2 * math.pi / p * K * _norm_s(a) * _norm_s(b) * ref.boys(0, p * float(np.sum((P - C) ** 2)))
    assert abs(S[0, 1] - s_ab) < 1e-14
    assert abs(T[0, 1] - t_ab) < 1e-14
    assert abs(V[0, 1] - v_ab) < 1e-14
    assert abs(S[0, 0] - 1.0) < 1e-14


def test_s_two_and_three_centre_closed_forms():
    A, B, C = np.array([0.0, 0.0, 0.0]), np.array([0.7, -0.1, 0.4]), np.array([1.9, 0.8, -1.1])
    a, b, g = 0.8, 2.1, 0.37
    orb = make_basis([(0, A, a, 1.0), (0, B, b, 1.0)])
    aux = make_basis([(0, C, g, 1.0), (0, A, 1.7, 1.0)])
    I3 = ref.int3c(orb, aux)
    p = a + b
    P = (a * A + b * B) / p
    K = math.exp(-a * b / p * float(np.sum((A - B) ** 2)))
    rho = p * g / (p + g)
    val = (2 * math.pi ** 2.5 / (p * g * math.sqrt(p + g)) * K * ref.boys(0, rho * float(np.sum((P - C) ** 2)))
           * _norm_s(a) * _norm_s(b) * _norm_s(g))
    assert abs(I3[0, 1, 0] - val) < 1e-13 * abs(val)
    assert abs(I3[1, 0, 0] - val) < 1e-13 * abs(val)
    V2 = ref.int2c(aux)
    q = 1.7
    v2 = (2 * math.pi ** 2.5 / (g * q * math.sqrt(g + q)) * ref.boys(0, g * q / (g + q) * float(np.sum((C - A) ** 2)))
          * _norm_s(g) * _norm_s(q))
    assert abs(V2[0, 1] - v2) < 1e-13 * abs(v2)


# ---------------------------------------------------------------------------
# Integrals: higher l from centre derivatives of s integrals
#   d/dA_x [N_s e^{-a|r-A|^2}] = 2 a (x - A_x) N_s e^{...} = (2 a N_s / N_p) p_x
# ---------------------------------------------------------------------------

def _coef(l, a):
    bs = make_basis([(l, [0, 0, 0], a, 1.0)])
    return bs.coefs[0]


def test_p_and_d_from_centre_derivatives():
    A, B, C = np.array([0.2, -0.1, 0.05]), np.array([-0.6, 0.5, 0.9]), np.array([0.4, 1.1, -0.7])
    a, b, g = 0.9, 0.6, 0.45
    h = 1e-4

    def ss_s(Ax):
        orb = make_basis([(0, Ax, a, 1.0), (0, B, b, 1.0)])
        aux = make_basis([(0, C, g, 1.0)])
        return ref.int3c(orb, aux)[0, 1, 0]

    # (p_x s | s)
    orb = make_basis([(1, A, a, 1.0), (0, B, b, 1.0)])
    aux = make_basis([(0, C, g, 1.0)])
    I = ref.int3c(orb, aux)
    fd = (ss_s(A + [h, 0, 0]) - ss_s(A - [h, 0, 0])) / (2 * h)
    scale = 2 * a * _coef(0, a) / _coef(1, a)
    assert abs(I[0, 3, 0] * scale - fd) < 1e-8 * abs(fd)
    fdy = (ss_s(A + [0, h, 0]) - ss_s(A - [0, h, 0])) / (2 * h)
    assert abs(I[1, 3, 0] * scale - fdy) < 1e-8 * abs(fdy)
    # (d_xy s | s), cartesian: d^2/dA_x dA_y = 4 a^2 N_s/N_d d_xy
    orb = make_basis([(2, A, a, 1.0), (0, B, b, 1.0)])
    I = ref.int3c(orb, aux)
    e = np.eye(3)
    fd2 = (ss_s(A + h * (e[0] + e[1])) - ss_s(A + h * (e[0] - e[1])) - ss_s(A - h * (e[0] - e[1]))
           + ss_s(A - h * (e[0] + e[1]))) / (4 * h * h)
    scale2 = 4 * a * a * _coef(0, a) / _coef(2, a)
    assert abs(I[1, 6, 0] * scale2 - fd2) < 1e-6 * abs(fd2)
    # (s s | p_z) on the auxiliary side: d/dC_z
    def ss_sC(Cx):
        orb2 = make_basis([(0, A, a, 1.0), (0, B, b, 1.0)])
        return ref.int3c(orb2, make_basis([(0, Cx, g, 1.0)]))[0, 1, 0]
    orb2 = make_basis([(0, A, a, 1.0), (0, B, b, 1.0)])
    I = ref.int3c(orb2, make_basis([(1, C, g, 1.0)]))
    fdz = (ss_sC(C + [0, 0, h]) - ss_sC(C - [0, 0, h])) / (2 * h)
    assert abs(I[0, 1, 2] * 2 * g * _coef(0, g) / _coef(1, g) - fdz) < 1e-8 * abs(fdz)
    # translation invariance: d/dA + d/dB + d/dC = 0 on (s s | s)
    def f3(dA, dB, dC):
        return ref.int3c(make_basis([(0, A + dA, a, 1.0), (0, B + dB, b, 1.0)]),
                         make_basis([(0, C + dC, g, 1.0)]))[0, 1, 0]
    z = np.zeros(3)
    dx = np.array([h, 0, 0])
    tot = ((f3(dx, z, z) - f3(-dx, z, z)) + (f3(z, dx, z) - f3(z, -dx, z)) + (f3(z, z, dx) - f3(z, z, -dx)))
    assert abs(tot / (2 * h)) < 1e-9


def _rotation(seed=3):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_rotation_invariants_high_l():
    """Spherical shells rotate orthogonally: spectra and norms are invariant (catches
    cart->sph and recursion index errors up to (ff|g))."""
    geom = "O 0 0 0; H 0.757 0.587 0; H -0.757 0.587 0.1"
    m1 = Molecule.from_xyz(geom, "def2-tzvp", spherical=True)
    R = _rotation()
    m2 = Molecule(m1.symbols, m1.coords @ R.T + np.array([0.3, -0.2, 0.5]), "def2-tzvp", spherical=True)
    out = []
    for m in (m1, m2):
        b = build_basis(m.symbols, m.coords, m.basis, spherical=True)
        x = build_basis(m.symbols, m.coords, "even-tempered-jk", spherical=True)
        S, T, V = ref.int1e(b, m.coords, m.atom_charges())
        out.append((np.linalg.eigvalsh(S), np.linalg.eigvalsh(T + V), np.linalg.eigvalsh(ref.int2c(x)),
                    np.linalg.norm(ref.int3c(b, x))))
    for a, c in zip(out[0], out[1]):
        assert np.allclose(a, c, rtol=1e-10, atol=1e-11)


def test_overlap_diagonal_and_counts():
    m = Molecule.from_xyz(WATER, "def2-tzvp", spherical=True)
    for name, n in (("sto-3g", 7), ("def2-svp", 24), ("def2-tzvp", 43), ("even-tempered-jk", 115)):
        b = build_basis(m.symbols, m.coords, name, spherical=True)
        assert b.nao == n
        S, _, _ = ref.int1e(b, m.coords, m.atom_charges())
        assert np.abs(np.diag(S) - 1.0).max() < 1e-13


def test_metric_pseudo_inverse_identities():
    m = Molecule.from_xyz(WATER, "def2-svp", spherical=True)
    x = build_basis(m.symbols, m.coords, "even-tempered-jk", spherical=True)
    V = ref.int2c(x)
    Vp, L = ref.metric_pinv(V)
    assert np.allclose(Vp, Vp.T, atol=1e-12)
    assert np.allclose(L @ L.T, Vp, rtol=1e-12, atol=1e-9)
    assert np.abs(V @ Vp @ V - V).max() < 1e-9 * np.abs(V).max()
    # positive definite here: cutoff keeps everything and V+ = V^-1
    w = np.linalg.eigvalsh(V)
    assert w.min() > 1e-7 * w.max()
    assert np.allclose(Vp @ V, np.eye(V.shape[0]), atol=1e-8)


def test_cart2sph_matches_reference_table():
    with np.load("tests/golden/cart2sph_ref.npz") as z:
        for l in range(7):
            assert np.abs(cart2sph_block(l) - z[f"l{l}"]).max() < 1e-13


# ---------------------------------------------------------------------------
# Grid
# ---------------------------------------------------------------------------

def test_lebedev_moments():
    from dftax_amd.grid import LEBEDEV_LADDER, lebedev_grid
    for n in LEBEDEV_LADDER:
        p, w = lebedev_grid(n)
        assert abs(w.sum() - 1.0) < 1e-13
        assert np.allclose(np.linalg.norm(p, axis=1), 1.0, atol=1e-13)
        assert abs(np.sum(w * p[:, 0] ** 2) - 1.0 / 3.0) < 1e-12
        if n >= 14:
            assert abs(np.sum(w * p[:, 2] ** 4) - 1.0 / 5.0) < 1e-12


def test_becke_grid_integrates_gaussians():
    from dftax_amd.grid import becke
    m = Molecule.from_xyz(WATER, "def2-svp", spherical=True)
    pts, w = ref.becke_grid(m.symbols, m.coords, becke())
    for a in range(3):
        for al in (0.3, 2.0, 40.0):
            r2 = np.sum((pts - m.coords[a]) ** 2, axis=1)
            val = np.sum(w * np.exp(-al * r2)) / (math.pi / al) ** 1.5
            assert abs(val - 1.0) < 2e-6, (a, al, val)


# ---------------------------------------------------------------------------
# Functionals
# ---------------------------------------------------------------------------

def test_functional_limits():
    import torch
    rho = torch.tensor(10.0 ** np.linspace(-8, 3, 40), dtype=torch.float64)
    zero = torch.zeros((40, 3), dtype=torch.float64)
    lda = -(3.0 / 4.0) * (3.0 / math.pi) ** (1.0 / 3.0) * rho ** (1.0 / 3.0)
    assert torch.allclose(ref.lda_x(rho), lda, rtol=1e-12)
    assert torch.allclose(ref.pbe_x(rho, zero), lda, rtol=1e-10)
    assert torch.allclose(ref.b88_x(rho, zero), lda, rtol=1e-10)
    # Lieb-Oxford style bound of the PBE enhancement and its s -> inf limit 1 + kappa
    g = torch.zeros((40, 3), dtype=torch.float64)
    g[:, 0] = 1e6 * rho ** (4.0 / 3.0)
    fx = ref.pbe_x(rho, g) / lda
    assert torch.all(fx <= 1.804 + 1e-12) and torch.all(fx > 1.80)
    # PBE correlation: gradient correction H >= 0 and eps_c -> 0 as t -> inf
    c0 = ref.pbe_c(rho, zero)
    assert torch.all(c0 < 0)
    c1 = ref.pbe_c(rho, g * 0.01)
    assert torch.all(c1 >= c0)
    cinf = ref.pbe_c(rho, g * 1e6)
    assert torch.all(cinf.abs() < 1e-6 * c0.abs() + 1e-300)
    # LYP closed shell vanishes as rho -> 0 faster than rho^(1/3); B3LYP composition at sigma=0
    b3 = ref.eps_xc("B3LYP", rho, zero)
    comp = 0.8 * lda + 0.19 * ref.vwn_rpa_c(rho) + 0.81 * ref.lyp_c(rho, zero)
    assert torch.allclose(b3, comp, rtol=1e-12)


# ---------------------------------------------------------------------------
# DF energy, Fock and SCF
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def water_oks():
    from dftax_amd.grid import becke
    m = Molecule.from_xyz(WATER, "def2-svp", spherical=True)
    return ref.OracleKS(m, "PBE0", grid=becke(40, 110))


def _guess(oks):
    P0 = ref.minao_density(oks.mol, oks.basis)
    return P0 * (oks.mol.nelectron / np.sum(P0 * oks.S))


def test_fock_is_energy_gradient(water_oks):
    """F = sym(dE/dP): autograd, analytic J/K/Vxc and a finite difference agree."""
    import torch
    P = _guess(water_oks)
    e, F, parts = water_oks.energy_fock(P)
    e2, F2, _ = water_oks.energy_fock_autograd(P)
    assert abs(e - e2) < 1e-10
    assert np.abs(F - F2).max() < 1e-10
    rng = np.random.default_rng(1)
    D = rng.standard_normal(P.shape)
    D = 0.5 * (D + D.T)
    h = 1e-5
    ep = float(water_oks.energy_t(torch.tensor(P + h * D))[0])
    em = float(water_oks.energy_t(torch.tensor(P - h * D))[0])
    assert abs((ep - em) / (2 * h) - np.sum(F * D)) < 1e-6 * max(1.0, abs(np.sum(F * D)))
    assert parts["eJ"] > 0
    assert np.abs(parts["J"] - parts["J"].T).max() < 1e-12
    assert np.abs(parts["K"] - parts["K"].T).max() < 1e-12


def test_exchange_factor_form_equals_full_density(water_oks):
    """K(P) from the signed eigen-factor of P equals the full-density contraction."""
    P = _guess(water_oks)
    B = np.einsum("mnP,PQ->mnQ", water_oks.int3c, water_oks.L)
    K_full = sum(B[:, :, q] @ P @ B[:, :, q] for q in range(B.shape[2]))
    lam, U = np.linalg.eigh(P)
    Cf = U * np.sqrt(np.abs(lam))
    T = np.einsum("lmq,li->mqi", B, Cf)
    K_fac = np.einsum("mqi,nqi,i->mn", T, T, np.sign(lam))
    assert np.abs(K_full - K_fac).max() < 1e-11 * np.abs(K_full).max()


def test_scf_invariants(water_oks):
    r = ref.scf(water_oks, e_tol=1e-10, d_tol=1e-7)
    assert r["converged"]
    P, S = r["P"], water_oks.S
    assert np.abs(P @ S @ P - 2 * P).max() < 1e-10
    assert abs(np.sum(P * S) - water_oks.mol.nelectron) < 1e-10
    rho = np.einsum("gm,mn,gn->g", water_oks.ao, P, water_oks.ao)
    assert abs(np.sum(water_oks.grid_w * rho) - 10.0) < 1e-3
    # stationarity: an occupied-virtual rotation raises the energy to second order
    e0, F, _ = water_oks.energy_fock(P)
    C = r["C"]
    nocc = water_oks.nocc
    kappa = np.zeros((C.shape[1], C.shape[1]))
    kappa[nocc, 0] = 1e-3
    kappa = kappa - kappa.T
    from scipy.linalg import expm
    C2 = C @ expm(kappa)
    P2 = 2 * C2[:, :nocc] @ C2[:, :nocc].T
    e2, _, _ = water_oks.energy_fock(P2)
    assert e2 > e0 - 1e-10


def test_response_fock_pins(water_oks):
    """The oracle's response Fock R = dF/dP . dP: its J and K parts equal J(dP) and
    K(dP) (both linear), and the whole matches a central difference of the
    analytic Fock matrix (the XC second derivative is the only nonlinear part)."""
    rng = np.random.default_rng(3)
    P = _guess(water_oks)
    n = P.shape[0]
    U, V = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    dP = U @ V.T + V @ U.T
    R = water_oks.response_fock(P, dP)
    assert np.abs(R - R.T).max() < 1e-12
    # the difference error falls as h^2 (1.8e-4, 1.8e-6, 1.8e-8 at h = 1e-3, 1e-4, 1e-5)
    h = 1e-5
    _, Fp, _ = water_oks.energy_fock(P + h * dP)
    _, Fm, _ = water_oks.energy_fock(P - h * dP)
    fd = (Fp - Fm) / (2 * h)
    assert np.abs(R - fd).max() < 1e-8 * max(1.0, np.abs(fd).max())
    # linear parts: J(dP) - a_x/2 K(dP) is what remains after removing the XC response
    _, _, parts = water_oks.energy_fock(dP)
    _, _, p_p = water_oks.energy_fock(P + h * dP)
    _, _, p_m = water_oks.energy_fock(P - h * dP)
    vxc_fd = (p_p["Vxc"] - p_m["Vxc"]) / (2 * h)
    lin = parts["J"] - 0.5 * water_oks.hf * parts["K"]
    assert np.abs(R - vxc_fd - lin).max() < 1e-8 * max(1.0, np.abs(fd).max())
    assert np.abs(lin).max() > 1e-2 and np.abs(vxc_fd).max() > 1e-2
