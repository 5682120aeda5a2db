"""The workload configurations at their real sizes, on the GPU, through the C ABI.

- (H2O)_8 def2-TZVP PBE0, default grid: (ij|P) samples, the Fock matrix and
  energy parts at the minao density, and the SCF energy against the CPU
  oracle's golden vectors (tests/golden/make_h2o8.py), at 1e-10 (integrals,
  Fock) and 1e-8 Ha (energies).
- (H2O)_32 def2-TZVP PBE0, B resident: the fitted tensor B B^T on sampled bra
  pairs against the oracle's (mn|P) V+ (P|kl), the metric against the oracle,
  and the FP64 / mixed SCFs: idempotency, electron count, commutator below
  d_tol, mixed refined within 1e-6 Ha of FP64.
- (H2O)_64: tests/test_gpu_configs_h2o64.py (a module of its own, so the
  handles here are gone before it allocates).
- The serving stream (def2-SVP B3LYP molecules of 2-24 heavy atoms, default
  grid): every energy the same alone and in the mix.
Every system is built from the same seeded generators as the benchmark.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "h2o8_pbe0_oracle.npz"


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _release(ks):
    """Destroy the handle and return the library's kept buffers and torch's cache,
    so the next system of the session starts from an empty device."""
    import torch
    from dftax_amd import _native as N
    ks.coulomb.close()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    N.check(N.lib().dfks_release_memory(N.current_stream()))


def _p0(ks):
    """Renormalised minao density of the oracle, as the golden vectors used."""
    from oracle import ref
    P0 = ref.minao_density(ks.mol, ks.basis)
    S = ks.S.cpu().numpy()
    return P0 * (ks.mol.nelectron / np.sum(P0 * S))


def _scf_properties(ks, r, label):
    """PSP = 2P, Tr(PS) = N, commutator below d_tol at the converged SCF."""
    P = r.P[0]
    S = ks.S
    idem = float((P @ S @ P - 2.0 * P).abs().max())
    ntr = float((P * S).sum())
    derr = r.history[-1][2]
    _say(f"{label}: E {r.e_tot:.10f} ({r.n_iter} it) |PSP-2P| {idem:.1e} Tr(PS) {ntr:.10f} |[F,P]| {derr:.1e}")
    assert r.converged
    assert idem < 1e-8
    assert abs(ntr - ks.nelec) < 1e-8
    assert derr < 1e-6


# ---------------------------------------------------------------------------
# (H2O)_8 def2-TZVP PBE0 against the oracle's golden vectors
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def h2o8(torch_cuda):
    from dftax_amd import KS, df, hip
    from dftax_amd.system import water_cluster
    g = np.load(GOLDEN)
    mol = water_cluster((2, 2, 2), seed=0, basis="def2-tzvp")
    _say("case: (H2O)_8 def2-TZVP PBE0, default grid, against the oracle's golden vectors")
    assert np.abs(mol.coords - g["coords"]).max() < 1e-12   # the generator is the one the vectors came from
    ks = KS(mol, "PBE0", coulomb=df(engine=hip(keep_int3c=True)))
    yield ks, g
    _release(ks)


def test_h2o8_integrals(h2o8, torch_cuda):
    from dftax_amd import _native as N
    ks, g = h2o8
    nao, naux = ks.nao, ks.coulomb.naux
    I = ks.coulomb.matrix(N.DFKS_INT3C, (nao, nao, naux))
    idx = torch_cuda.tensor(g["int3c_idx"].astype(np.int64), device=I.device)
    got = I[idx[:, 0], idx[:, 1], idx[:, 2]].cpu().numpy()
    del I
    err = np.abs(got - g["int3c_val"]).max()
    _say(f"(ij|P) on {len(got)} samples: max |d| {err:.2e}")
    assert err < 1e-10


def test_h2o8_fock_at_guess(h2o8, torch_cuda):
    ks, g = h2o8
    P0 = _p0(ks)
    assert abs(P0.sum() - float(g["p0_sum"])) < 1e-10 and np.abs(P0 @ g["vecs"][0] - g["p0_v"]).max() < 1e-10
    E, F = ks.coulomb.energy_fock(torch_cuda.tensor(P0, device="cuda"))
    E, F = E.cpu().numpy(), F.cpu().numpy()
    dE = np.abs(E[[0, 1, 2, 3, 4]] - g["e_parts"])
    dF = max(np.abs(F[g["fock_row_idx"]] - g["fock_rows"]).max(), np.abs(g["vecs"] @ F.T - g["fock_v"]).max())
    _say(f"F(P0) rows and F v: max |d| {dF:.2e}; E parts |d| {dE.max():.2e}")
    assert dF < 1e-10
    assert dE[0] < 1e-8 and np.all(dE[1:] < 1e-8)


def test_h2o8_scf(h2o8):
    from dftax_amd import scf
    ks, g = h2o8
    r = scf(ks, e_tol=1e-9, d_tol=1e-6)
    _scf_properties(ks, r, "(H2O)_8 FP64 SCF")
    _say(f"(H2O)_8 SCF {r.e_tot:.10f} oracle {float(g['e_scf'][0]):.10f}")
    assert abs(r.e_tot - float(g["e_scf"][0])) < 1e-8


# ---------------------------------------------------------------------------
# (H2O)_32 def2-TZVP PBE0, B resident
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def h2o32(torch_cuda):
    from dftax_amd import KS, df, hip
    from dftax_amd.system import water_cluster
    _say("case: (H2O)_32 def2-TZVP PBE0, B resident, FP64 and mixed")
    mol = water_cluster((4, 4, 2), seed=0, basis="def2-tzvp")
    ks = KS(mol, "PBE0", coulomb=df(chunk=None, engine=hip(precision="mixed")))
    yield ks, {}
    _release(ks)


def _shell_of(basis):
    out = np.empty(basis.nao, dtype=np.int64)
    for s in range(basis.nshell):
        a0 = int(basis.ao0[s])
        a1 = int(basis.ao0[s + 1]) if s + 1 < basis.nshell else basis.nao
        out[a0:a1] = s
    return out


def test_h2o32_fitted_tensor_and_metric(h2o32, torch_cuda):
    """B B^T on 40 sampled bra function pairs equals (mn|P) V+ (P|kl) built by
    the oracle's MD engine from the same shells; (P|Q) equals the oracle's."""
    from dftax_amd import _native as N
    from oracle import ref
    ks, _ = h2o32
    nao, naux, nfit = ks.nao, ks.coulomb.naux, ks.coulomb.nfit
    V = ks.coulomb.matrix(N.DFKS_INT2C, (naux, naux)).cpu().numpy()
    V_ref = ref.int2c(ks.aux_basis)
    dV = np.abs(V - V_ref).max()
    rng = np.random.default_rng(32)
    m = rng.integers(0, nao, 40)
    n = rng.integers(0, nao, 40)
    # the copy of B (51 GB) comes from torch's allocator: hand back the buffers the library keeps
    # for reuse (earlier handles of this process) first
    N.release_memory(ks.coulomb.device)
    B = ks.coulomb.matrix(N.DFKS_BFIT, (nao * nao, nfit))
    Bs = B[torch_cuda.tensor(m * nao + n, device=B.device)].cpu().numpy()
    del B
    torch_cuda.cuda.empty_cache()
    sh, shx = _shell_of(ks.basis), ks.aux_basis
    trip = [(sh[a], sh[b], c) for a, b in zip(m, n) for c in range(shx.nshell)]
    blocks = ref.int3c_triples(ks.basis, shx, trip)
    Is = np.zeros((40, naux))
    k = 0
    for i, (a, b) in enumerate(zip(m, n)):
        for c in range(shx.nshell):
            blk = blocks[k]
            k += 1
            p0 = int(shx.ao0[c])
            Is[i, p0:p0 + blk.shape[2]] = blk[a - int(ks.basis.ao0[sh[a]]), b - int(ks.basis.ao0[sh[b]])]
    Vp, _ = ref.metric_pinv(V_ref)
    G_ref = Is @ Vp @ Is.T
    G = Bs @ Bs.T
    dG = np.abs(G - G_ref).max() / np.abs(G_ref).max()
    _say(f"(H2O)_32: (P|Q) max |d| {dV:.2e}; sampled B B^T vs (mn|P) V+ (P|kl): rel {dG:.2e}")
    assert dV < 1e-10
    assert dG < 1e-9


def test_h2o32_fp64_scf(h2o32):
    from dftax_amd import scf
    ks, res = h2o32
    r = scf(ks, precision="fp64")
    _scf_properties(ks, r, "(H2O)_32 FP64 SCF")
    res["fp64"] = r.e_tot


def test_h2o32_mixed_scf(h2o32):
    """Mixed to the switch, then FP64: to d_tol within 1e-8 Ha of the FP64 SCF,
    and with energy refinement (refine=3) within the 1e-6 Ha budget."""
    from dftax_amd import scf
    ks, res = h2o32
    e64 = res.get("fp64")
    if e64 is None:
        pytest.skip("needs the FP64 SCF of the test before")
    rm = scf(ks, precision="mixed")
    _scf_properties(ks, rm, "(H2O)_32 mixed SCF to d_tol")
    assert sum(1 for h in rm.history if h[3] == "mixed") >= 3
    assert abs(rm.e_tot - e64) < 1e-8
    rr = scf(ks, precision="mixed", refine=3)
    _say(f"(H2O)_32 mixed + FP64 energy refinement: {rr.e_tot:.10f} ({rr.n_iter} it), |dE| {abs(rr.e_tot - e64):.1e}")
    assert rr.energy_converged and abs(rr.e_tot - e64) < 1e-6


# ---------------------------------------------------------------------------
# Serving stream: the same energy alone and in the mix
# ---------------------------------------------------------------------------

def test_serving_stream_alone_equals_mix(torch_cuda):
    """8 molecules of the serving stream (seed 0: 2-24 heavy atoms, def2-SVP,
    B3LYP, default grid, mixed precision with FP64 energy refinement), solved
    one at a time and then arriving together on 8 slots."""
    import time
    from dftax_amd.serve import Request, Server, latency_stats
    from dftax_amd.system import serving_stream
    mols = serving_stream(8, seed=0)
    srv = Server("B3LYP", precision="mixed", slots=8)
    alone = []
    for i, m in enumerate(mols):
        t = time.perf_counter()
        alone.append(srv.solve(Request(rid=i, mol=m)))
        _say(f"alone {i}: {len(m.symbols)} atoms, {alone[-1].extra.get('nao')} functions, "
             f"{time.perf_counter() - t:.2f} s")
        assert not alone[-1].error, alone[-1].error
    reqs = [Request(rid=i, mol=m, arrival=0.0) for i, m in enumerate(mols)]
    srv.run(reqs)
    st = latency_stats(reqs)
    _say(f"serving {len(mols)} molecules ({min(len(m.symbols) for m in mols)}-{max(len(m.symbols) for m in mols)} atoms): "
         f"{st['solves_per_min']:.0f} solves/min in the mix")
    for a, r in zip(alone, reqs):
        assert not r.error, r.error
        assert a.energy_converged and r.energy_converged
        assert abs(a.e_tot - r.e_tot) < 1e-8, (r.rid, a.e_tot, r.e_tot)
