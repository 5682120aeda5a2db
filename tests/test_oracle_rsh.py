"""Pins of the oracle's range-separated hybrid pieces (no GPU).

The reference checks these against PySCF's range-separated integrals and
libxc (tests/unit/test_rsh.py), neither of which is installed here; the pins
below are what the mathematics fixes:
- the erf(omega r)/r integrals of s Gaussians in closed form (a Gaussian charge
  meets erf(omega r)/r as a Coulomb interaction with an extra Gaussian
  smearing of exponent omega^2), higher l from centre derivatives, and the
  omega -> infinity limit (the Coulomb operator);
- the short-range attenuation F(a) against a direct quadrature of the
  erfc-screened exchange hole of the uniform gas;
- the ITYH short-range B88 at zero gradient reduces to short-range LDA
  exchange with a = omega / (2 k_F);
- CAM-B3LYP's exchange fractions sum to one in both limits (omega -> 0: B88
  and K; omega -> infinity: B88 and K + K_lr), which fixes the four mixing
  constants given the published (alpha, beta) = (0.19, 0.46);
- density-fitted K_lr against the exact four-centre K_lr, and F = dE/dP of the
  CAM-B3LYP energy (analytic against autograd).
Each check fails when its constant or index is perturbed.
"""
import math

import numpy as np
import pytest
import torch
from scipy import integrate
from scipy.special import erfc

from oracle import ref
from test_oracle_pins import _coef, _norm_s, make_basis

OMEGA = 0.33


def _erf_pair(p, g, R, omega):
    """Interaction of unit-charge Gaussians of exponents p, g at distance R
    through erf(omega r)/r (omega = inf: the Coulomb operator)."""
    L = math.sqrt(1.0 / p + 1.0 / g + (0.0 if math.isinf(omega) else 1.0 / omega ** 2))
    return 2.0 / (math.sqrt(math.pi) * L) if R < 1e-14 else math.erf(R / L) / R


def test_attenuated_s_closed_forms():
    A, B, C = np.array([0.0, 0.0, 0.0]), np.array([0.7, -0.1, 0.4]), np.array([1.9, 0.8, -1.1])
    a, b, g, q = 0.8, 2.1, 0.37, 1.7
    orb = make_basis([(0, A, a, 1.0), (0, B, b, 1.0)])
    aux = make_basis([(0, C, g, 1.0), (0, A, q, 1.0)])
    for omega in (OMEGA, 1.7):
        I3 = ref.int3c(orb, aux, omega)
        p = a + b
        P = (a * A + b * B) / p
        K = math.exp(-a * b / p * float(np.sum((A - B) ** 2)))
        charge = _norm_s(a) * _norm_s(b) * K * (math.pi / p) ** 1.5 * _norm_s(g) * (math.pi / g) ** 1.5
        val = charge * _erf_pair(p, g, float(np.linalg.norm(P - C)), omega)
        assert abs(I3[0, 1, 0] - val) < 1e-13 * abs(val)
        V2 = ref.int2c(aux, omega)
        cq = _norm_s(g) * (math.pi / g) ** 1.5 * _norm_s(q) * (math.pi / q) ** 1.5
        assert abs(V2[0, 1] - cq * _erf_pair(g, q, float(np.linalg.norm(C - A)), omega)) < 1e-13 * abs(V2[0, 1])
        assert abs(V2[1, 1] - _norm_s(q) ** 2 * (math.pi / q) ** 3 * _erf_pair(q, q, 0.0, omega)) < 1e-13 * V2[1, 1]
    # the attenuated operator is weaker than 1/r, and omega -> infinity is the Coulomb operator
    V_lr, V_full, V_inf = ref.int2c(aux, OMEGA), ref.int2c(aux), ref.int2c(aux, 1e9)
    assert np.all(np.abs(V_lr) < np.abs(V_full))
    assert np.abs(V_inf - V_full).max() < 1e-12 * np.abs(V_full).max()
    assert abs(V_full[0, 1] - cq * _erf_pair(g, q, float(np.linalg.norm(C - A)), math.inf)) < 1e-13 * V_full[0, 1]


def test_attenuated_higher_l_from_centre_derivatives():
    A, B, C = np.array([0.2, -0.1, 0.05]), np.array([-0.6, 0.5, 0.9]), np.array([0.4, 1.1, -0.7])
    a, b, g = 0.9, 0.6, 0.45
    h = 1e-4

    def ss_s(Ax, Cx=C):
        return ref.int3c(make_basis([(0, Ax, a, 1.0), (0, B, b, 1.0)]), make_basis([(0, Cx, g, 1.0)]), OMEGA)[0, 1, 0]

    I = ref.int3c(make_basis([(1, A, a, 1.0), (0, B, b, 1.0)]), make_basis([(0, C, g, 1.0)]), OMEGA)
    scale = 2 * a * _coef(0, a) / _coef(1, a)
    for x in range(3):
        e = np.eye(3)[x] * h
        fd = (ss_s(A + e) - ss_s(A - e)) / (2 * h)
        assert abs(I[x, 3, 0] * scale - fd) < 1e-8 * abs(fd)
    # (d_xy s | s), cartesian
    I = ref.int3c(make_basis([(2, A, a, 1.0), (0, B, b, 1.0)]), make_basis([(0, C, g, 1.0)]), OMEGA)
    e = np.eye(3)
    fd2 = (ss_s(A + h * (e[0] + e[1])) - ss_s(A + h * (e[0] - e[1])) - ss_s(A - h * (e[0] - e[1]))
           + ss_s(A - h * (e[0] + e[1]))) / (4 * h * h)
    assert abs(I[1, 6, 0] * 4 * a * a * _coef(0, a) / _coef(2, a) - fd2) < 1e-6 * abs(fd2)
    # (s s | p_z) on the auxiliary side
    I = ref.int3c(make_basis([(0, A, a, 1.0), (0, B, b, 1.0)]), make_basis([(1, C, g, 1.0)]), OMEGA)
    dz = np.array([0, 0, h])
    fdz = (ss_s(A, C + dz) - ss_s(A, C - dz)) / (2 * h)
    assert abs(I[0, 1, 2] * 2 * g * _coef(0, g) / _coef(1, g) - fdz) < 1e-8 * abs(fdz)


def _f_quadrature(a):
    """eps_x^SR / eps_x^LDA of the uniform gas (k_F = 1) from its exchange hole
    n_x(r) = -(9n/2) [j1(k_F r)/(k_F r)]^2 and the erfc(omega r)/r interaction,
    omega = 2 a k_F: eps = 2 pi int n_x(r) erfc(omega r) r dr."""
    n, om = 1.0 / (3 * math.pi ** 2), 2.0 * a

    def f(r):
        j = 1.0 / 3.0 - r * r / 30.0 if r < 1e-3 else (math.sin(r) - r * math.cos(r)) / r ** 3
        return -(9 * n / 2) * j * j * erfc(om * r) * r
    edges = np.linspace(0.0, 400.0, 2001)
    val = sum(integrate.quad(f, x0, x1, limit=200, epsabs=1e-15, epsrel=1e-13)[0]
              for x0, x1 in zip(edges[:-1], edges[1:]))
    return 2 * math.pi * val / (-3.0 / (4 * math.pi))


@pytest.mark.parametrize("a,tol", [(0.02, 1e-13), (0.1, 1e-13), (0.3, 1e-13), (1.0, 1e-13), (3.0, 1e-12),
                                   (7.5, 1e-11)])
def test_erf_attenuation_against_exchange_hole(a, tol):
    F = float(ref.erf_attenuation(torch.tensor(a, dtype=torch.float64)))
    assert abs(F - _f_quadrature(a)) < tol


def test_erf_attenuation_asymptote():
    # from a = 8 on the reference switches to 1/(36 a^2); the exact F(a) is
    # 1/(36 a^2) - 1/(960 a^4) + ..., so the switch's relative error is 3/(80 a^2)
    for a in (8.5, 20.0):
        F = float(ref.erf_attenuation(torch.tensor(a, dtype=torch.float64)))
        exact = _f_quadrature(a)
        assert abs((F - exact) / exact / (3.0 / (80.0 * a * a)) - 1.0) < 0.05


def test_ityh_zero_gradient_is_short_range_lda():
    rho = torch.tensor([1e-3, 0.05, 0.7, 12.0, 300.0], dtype=torch.float64)
    grad = torch.zeros(len(rho), 3, dtype=torch.float64)
    kF = (3 * math.pi ** 2 * rho) ** (1 / 3)
    want = ref.lda_x(rho) * ref.erf_attenuation(OMEGA / (2 * kF))
    got = ref.ityh_b88_x(rho, grad)
    assert torch.max(torch.abs(got - want) / torch.abs(want)) < 1e-12
    # the attenuation cuts the exchange hole's long-range part only: |SR| < |full| everywhere
    g = torch.tensor([[0.3, -0.2, 0.1]] * len(rho), dtype=torch.float64) * rho[:, None] ** (4 / 3)
    assert torch.all(torch.abs(ref.ityh_b88_x(rho, g)) < torch.abs(ref.b88_x(rho, g)))


def test_cam_b3lyp_exchange_fractions():
    """omega -> 0: K_lr -> 0 and SR-B88 -> B88; omega -> infinity: K_lr -> K and
    SR-B88 -> 0. Both limits are a global hybrid whose exchange fractions (DFT +
    exact) sum to one; the correlation fractions sum to one."""
    b88, sr, lyp, vwn = ref.MIX["CAM-B3LYP"]
    a_x = ref.HF_COEFF["CAM-B3LYP"]
    a_lr, omega = ref.HF_LR["CAM-B3LYP"]
    assert abs((b88 + sr) + a_x - 1.0) < 1e-15        # omega -> 0
    assert abs(b88 + (a_x + a_lr) - 1.0) < 1e-15      # omega -> infinity
    assert abs(lyp + vwn - 1.0) < 1e-15
    assert (a_x, a_lr, omega) == (0.19, 0.46, 0.33)   # Yanai-Tew-Handy alpha, beta, mu
    rho = torch.tensor([0.02, 0.4, 5.0], dtype=torch.float64)
    g = torch.tensor([[0.1, 0.2, -0.3]] * 3, dtype=torch.float64) * rho[:, None] ** (4 / 3)
    assert torch.max(torch.abs(ref.ityh_b88_x(rho, g, omega=1e-9) / ref.b88_x(rho, g) - 1.0)) < 1e-8
    assert torch.max(torch.abs(ref.ityh_b88_x(rho, g, omega=1e6) / ref.b88_x(rho, g))) < 1e-8


@pytest.fixture(scope="module")
def water_cam():
    from dftax_amd.grid import becke
    from dftax_amd.system import Molecule
    mol = Molecule.from_xyz("O 0 0 0; H 0.7586 0 0.5043; H 0.7586 0 -0.5043", "sto-3g", spherical=False)
    df = ref.OracleKS(mol, "CAM-B3LYP", grid=becke(20, 50))
    ex = ref.OracleKS(mol, "CAM-B3LYP", grid=becke(20, 50), exact=True)
    return mol, df, ex


def test_lr_exchange_density_fitted_against_exact(water_cam):
    mol, df, ex = water_cam
    P = ref.minao_density(mol, df.basis)
    P = P * (mol.nelectron / np.sum(P * df.S))
    t = torch.tensor(P)
    K_df = df._exchange_lr_t(t).numpy()
    K_ex = ex._exchange_lr_t(t).numpy()
    rel = np.abs(K_df - K_ex).max() / np.abs(K_ex).max()
    # the fitting error of the attenuated operator (same order as the full-range K's)
    K_df_full = df._exchange_matrix_t(t).numpy()
    K_ex_full = ex._exchange_matrix_t(t).numpy()
    rel_full = np.abs(K_df_full - K_ex_full).max() / np.abs(K_ex_full).max()
    # measured 1.0e-6 (the smooth long-range operator fits better than 1/r: 3.8e-5); a
    # transposed index (the Coulomb-like contraction) is off by 2.9
    assert rel < 1e-4 and rel < rel_full
    # the long-range part is the smaller one, and symmetric
    assert np.abs(K_ex).max() < np.abs(K_ex_full).max()
    assert np.abs(K_df - K_df.T).max() < 1e-11 * np.abs(K_df).max()
    # one doubly occupied orbital: -1/4 sum P K_lr = -1/2 E_J,lr (prefactor of the exchange energy)
    c = np.linalg.eigh(df.S)[1][:, -1]
    P1 = 2.0 * np.outer(c, c)
    eK = -0.25 * float(np.sum(P1 * ex._exchange_lr_t(torch.tensor(P1)).numpy()))
    eJ = 0.5 * float(np.einsum("mn,mnls,ls->", P1, ex.eri_lr, P1))
    assert abs(eK + 0.5 * eJ) < 1e-12 * abs(eJ)


def test_cam_fock_is_energy_gradient(water_cam):
    mol, df, _ = water_cam
    P = ref.minao_density(mol, df.basis)
    P = P * (mol.nelectron / np.sum(P * df.S))
    e_a, F_a, parts = df.energy_fock(P)
    e_g, F_g, _ = df.energy_fock_autograd(P)
    assert abs(e_a - e_g) < 1e-10
    assert np.abs(F_a - F_g).max() < 1e-10
    # the long-range exchange is in both: removing it changes E by -1/4 a_lr sum P K_lr
    assert abs(parts["eK"] - (-0.25 * df.hf * np.sum(P * parts["K"]) - 0.25 * df.hf_lr * np.sum(P * parts["K_lr"]))) < 1e-12
