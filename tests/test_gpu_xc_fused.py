"""The fused mixed-precision compact-grid XC pass against its unfused form and FP64.

In a mixed-precision Fock build the compact XC grid runs, per bin of blocks,
two split-BF16 MFMA kernels (`k_xcf_rho`: Z = Phi P gathered and reduced to
rho and grad rho in LDS, then the functional; `k_xcf_v`: Phi^T W with W formed
while staging). `gemm="rocblas"` keeps the unfused form (P gathered to HBM,
FP32 rocBLAS batched products, the pointwise kernel), so the two handles differ
only in the XC products. PBE (no exact exchange) on the water dimer,
def2-TZVP (f functions, so blocks of up to a few hundred functions): both
mixed Fock matrices against the FP64 one, at the minao density.

Tolerances. Split-BF16 operands carry ~2^-18 relative representation error
against FP32's 2^-24, and rho = sum_mn phi_m P_mn phi_n cancels heavily, so
the fused pass's E_xc error is ~40x the unfused pass's (measured 1.9e-5 against
4.7e-7 Ha of E_xc = -18.6 Ha): |dE_xc| <= 5e-5 Ha fused, 2e-6 unfused. The
Fock matrix is as accurate as the unfused pass's (measured 1.1e-5 against
1.3e-5): max |dF| <= 2e-5 for both, the level the mixed SCF needs (the exchange
products of a hybrid alone give ~1e-4).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DFKS_E_XC = 4      # index of E_xc in the energy vector (include/dfks.h)


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _p0(ks):
    from oracle import ref
    P0 = ref.minao_density(ks.mol, ks.basis)
    S = ks.S.cpu().numpy()
    return P0 * (ks.mol.nelectron / np.sum(P0 * S))


def test_fused_mixed_xc(torch_cuda, water_dimer_tzvp):
    from dftax_amd import KS, becke, df, hip
    torch = torch_cuda
    g = becke(40, 110)
    _say("case: water dimer def2-TZVP PBE, becke(40,110): fused vs unfused mixed XC vs FP64")
    fused = KS(water_dimer_tzvp, "PBE", grid=g, coulomb=df(engine=hip(precision="mixed")))
    plain = KS(water_dimer_tzvp, "PBE", grid=g, coulomb=df(engine=hip(precision="mixed", gemm="rocblas")))
    P0 = torch.tensor(_p0(fused), device="cuda")
    fused.coulomb.set_precision(False)
    E64, F64 = fused.coulomb.energy_fock(P0)
    fused.coulomb.set_precision(True)
    Ef, Ff = fused.coulomb.energy_fock(P0)
    Ef2, Ff2 = fused.coulomb.energy_fock(P0)
    fused.coulomb.set_precision(False)
    plain.coulomb.set_precision(True)
    Ep, Fp = plain.coulomb.energy_fock(P0)
    plain.coulomb.set_precision(False)
    F64, Ff, Ff2, Fp = (x.cpu().numpy() for x in (F64, Ff, Ff2, Fp))
    dff = np.abs(Ff - F64).max()
    dfp = np.abs(Fp - F64).max()
    dexf = abs(float(Ef[DFKS_E_XC] - E64[DFKS_E_XC]))
    dexp = abs(float(Ep[DFKS_E_XC] - E64[DFKS_E_XC]))
    _say(f"  fused vs FP64: max|dF| {dff:.2e}, |dE_xc| {dexf:.2e}; unfused vs FP64: max|dF| {dfp:.2e}, "
         f"|dE_xc| {dexp:.2e}; fused vs unfused max|dF| {np.abs(Ff - Fp).max():.2e}; E_xc {float(E64[DFKS_E_XC]):.6f}")
    assert np.abs(Ff - Ff2).max() <= 1e-12       # repeatable (the row assembly's LDS sums aside)
    assert np.abs(Ff - Ff.T).max() < 1e-12
    assert 0.0 < dff <= 2e-5
    assert dexf <= 5e-5
    assert dfp <= 2e-5 and dexp <= 2e-6
