"""Host logic (no GPU): basis assembly, grid layout, workload generators, specs."""
import math

import numpy as np
import pytest

from dftax_amd.basis import build_basis, gto_norm
from dftax_amd.grid import atom_centred_points, becke, becke_radial, nwchem_orders, BRAGG_BOHR
from dftax_amd.system import ANGSTROM_TO_BOHR, WATER, Molecule, serving_stream, tree_molecule, water_cluster


def test_gto_norm_s_and_p():
    for a in (0.1, 1.0, 37.5):
        # s: (2a/pi)^(3/4) times the 1/sqrt(4 pi) angular factor folded in
        assert abs(gto_norm(0, a) - (2 * a / math.pi) ** 0.75 * math.sqrt(4 * math.pi)) < 1e-12 * gto_norm(0, a)


def test_basis_sizes():
    m = water_cluster((4, 4, 2), seed=0)
    b = build_basis(m.symbols, m.coords, "def2-tzvp", spherical=True)
    x = build_basis(m.symbols, m.coords, "even-tempered-jk", spherical=True)
    assert (b.nao, x.nao) == (1376, 3680)
    bc = build_basis(m.symbols, m.coords, "def2-tzvp", spherical=False)
    assert bc.nao == 1536
    w = Molecule.from_xyz(WATER, "def2-svp", spherical=True)
    assert build_basis(w.symbols, w.coords, "def2-svp", spherical=True).nao == 24
    assert build_basis(w.symbols, w.coords, "even-tempered-jk", spherical=False).nao == 137


def test_grid_layout_and_pruning():
    r, w = becke_radial(75, 0.5)
    assert r.shape == (75,) and np.all(np.diff(r) < 0) and np.all(w > 0)
    # H: 50/86/266/302/266 regions; O (row 2) the same ladder with other bounds
    orders = nwchem_orders(1, np.array([0.1, 0.3, 0.6, 2.0, 10.0]) * BRAGG_BOHR[1], 302)
    assert list(orders) == [50, 86, 266, 302, 266]
    orders = nwchem_orders(8, np.array([0.1, 0.3, 0.6, 2.0, 10.0]) * BRAGG_BOHR[8], 302)
    assert list(orders) == [50, 86, 266, 302, 266]
    assert list(nwchem_orders(8, np.array([0.1, 5.0]), 38)) == [38, 38]
    m = Molecule.from_xyz(WATER, "def2-svp", spherical=True)
    pts, raw, own = atom_centred_points(m.symbols, m.coords, becke())
    assert pts.shape[0] == raw.shape[0] == own.shape[0]
    assert set(np.unique(own)) == {0, 1, 2}
    # each atom's points sit on spheres around it
    for a in range(3):
        d = np.linalg.norm(pts[own == a] - m.coords[a], axis=1)
        assert d.max() <= 45.0 + 1e-12


def test_water_cluster_generator():
    m1 = water_cluster((2, 2, 2), seed=0)
    m2 = water_cluster((2, 2, 2), seed=0)
    assert np.array_equal(m1.coords, m2.coords)
    assert not np.array_equal(m1.coords, water_cluster((2, 2, 2), seed=1).coords)
    xyz = m1.coords / ANGSTROM_TO_BOHR
    for k in range(8):
        o, h1, h2 = xyz[3 * k: 3 * k + 3]
        assert abs(np.linalg.norm(h1 - o) - 0.957) < 1e-12
        cosang = np.dot(h1 - o, h2 - o) / (np.linalg.norm(h1 - o) * np.linalg.norm(h2 - o))
        assert abs(math.degrees(math.acos(cosang)) - 104.5) < 1e-9
    mol_id = np.repeat(np.arange(8), 3)
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    inter = mol_id[:, None] != mol_id[None, :]
    assert d[inter].min() >= 1.7


def test_tree_molecules():
    rng = np.random.default_rng(5)
    for n in (2, 7, 24):
        m = tree_molecule(n, rng)
        heavy = [s for s in m.symbols if s != "H"]
        assert len(heavy) == n
        assert m.nelectron % 2 == 0
        xyz = m.coords / ANGSTROM_TO_BOHR
        d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1) + 10 * np.eye(len(xyz))
        assert d.min() > 0.9
    stream = serving_stream(6, seed=1)
    assert len(stream) == 6 and all(2 <= sum(s != "H" for s in m.symbols) <= 24 for m in stream)


def test_df_spec_validation():
    from dftax_amd import df, hip
    with pytest.raises(NotImplementedError):
        df(spherical=False)
    for bad in (0, -3, 1.5, "big", True):
        with pytest.raises(ValueError):
            df(chunk=bad)
    with pytest.raises(ValueError):
        df(screen=1.0)
    with pytest.raises(ValueError):
        hip(precision="fp16")
    with pytest.raises(ValueError):
        hip(ao_block=-1)
    spec = df()
    assert spec.auxbasis == "even-tempered-jk" and spec.engine.precision == "fp64"
    assert df(chunk=64, screen=1e-12).chunk == 64


def test_memory_plan():
    """Materialise when B and the grid fit; else slabs of ~1/8 of the budget and
    (if still needed) a streamed grid; explicit values pass through."""
    from dftax_amd.ks import memory_plan
    gb = 1 << 30
    small = memory_plan(24, 115, 36000, 5, "fp64", "auto", None, 200 * gb)
    assert small.aux_chunk == 0 and small.ao_block == 0
    # (H2O)_32-like: B ~ 56 GB fits a 245 GB budget
    c3 = memory_plan(1376, 3680, 1_033_616, 160, "mixed", "auto", None, 245 * gb)
    assert c3.aux_chunk == 0 and c3.ao_block == 0 and c3.bytes_est < 245 * gb
    # (H2O)_64-like: B ~ 450 GB does not fit: recompute, grid streamed, estimate within budget
    c5 = memory_plan(2752, 7360, 2_070_000, 320, "fp64", "auto", None, 245 * gb)
    assert 32 <= c5.aux_chunk <= 2048 and 4096 <= c5.ao_block < 2_070_000 and c5.bytes_est < 245 * gb
    forced = memory_plan(24, 115, 36000, 5, "fp64", 16, 100, 200 * gb)
    assert forced.aux_chunk == 16 and forced.ao_block == 100
    assert memory_plan(2752, 7360, 2_070_000, 320, "fp64", None, 0, 245 * gb).aux_chunk == 0


def test_steihaug_cg_host_logic():
    """The Newton step solver (newton.py:44-94 semantics) on explicit quadratics:
    interior solution H p = -g for a positive definite H and a large radius, a
    boundary step of norm radius when the radius binds, and the boundary along
    negative curvature for an indefinite H."""
    import torch
    from dftax_amd.newton import _steihaug_cg
    torch.manual_seed(0)
    n = 12
    Q, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64))
    H = Q @ torch.diag(torch.linspace(0.5, 4.0, n, dtype=torch.float64)) @ Q.T
    g = torch.randn(n, dtype=torch.float64)
    hvp = lambda v: H @ v
    p, nh = _steihaug_cg(hvp, g, 1e3, 64)
    assert torch.allclose(p, torch.linalg.solve(H, -g), atol=1e-10)
    p, _ = _steihaug_cg(hvp, g, 0.05, 64)
    assert abs(float(torch.linalg.norm(p)) - 0.05) < 1e-12 and float(g @ p) < 0.0
    Hn = Q @ torch.diag(torch.linspace(-1.0, 4.0, n, dtype=torch.float64)) @ Q.T
    p, _ = _steihaug_cg(lambda v: Hn @ v, g, 0.7, 64)
    assert abs(float(torch.linalg.norm(p)) - 0.7) < 1e-12
    assert float(g @ p + 0.5 * p @ (Hn @ p)) < 0.0


def test_rotate_is_orthogonal_and_density_preserving():
    """C exp(K) keeps C^T S C = 1 for S-orthonormal C (newton.py:97-106)."""
    import torch
    from dftax_amd.newton import _rotate
    torch.manual_seed(1)
    n, no = 9, 3
    C = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64))[0]
    k = 0.3 * torch.randn(no, n - no, dtype=torch.float64)
    C2 = _rotate(C, k, no)
    assert torch.allclose(C2.T @ C2, torch.eye(n, dtype=torch.float64), atol=1e-12)
    assert torch.allclose(_rotate(C, torch.zeros_like(k), no), C, atol=1e-14)


def test_serving_latency_slowdowns():
    """Per-request slowdown against each molecule's own isolated time."""
    from dftax_amd.serve import Request, latency_stats
    reqs = []
    for i, (arr, start, end) in enumerate([(0.0, 0.0, 1.0), (0.5, 0.5, 2.5), (1.0, 1.2, 1.6)]):
        r = Request(rid=i, mol=None, arrival=arr)
        r.start, r.end, r.converged = start, end, True
        reqs.append(r)
    st = latency_stats(reqs, {0: 1.0, 1: 1.0, 2: 0.2})
    assert st["ok"] == 3 and abs(st["p50_s"] - 1.0) < 1e-12
    assert abs(st["slowdown_p50"] - 2.0) < 1e-12 and st["slowdown_p99"] > 2.9
    assert abs(st["mean_queue_wait_s"] - 0.2 / 3) < 1e-12


def test_mixed_phase_switch_rule():
    """The mixed phase of scf ends at the gradient threshold, at the step cap, or
    on a stall at the split-BF16 floor -- never on an early, large-gradient stall."""
    from dftax_amd.scf import mixed_phase_done
    steady = [1e-1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1.2e-4, 9e-5]
    assert [mixed_phase_done(steady[:k]) for k in range(1, len(steady) + 1)] == [False] * 7 + [True]
    # a plateau just above the threshold: switch once three steps fail to improve on the best by 30 %
    plateau = [1e-1, 1e-2, 1e-3, 3e-4, 2.0e-4, 1.9e-4, 2.2e-4, 1.8e-4]
    done = [mixed_phase_done(plateau[:k]) for k in range(1, len(plateau) + 1)]
    assert done == [False] * 7 + [True]
    # early oscillation far above the floor does not count as a stall
    early = [1e-1, 2e-1, 1.5e-1, 1.2e-1, 1.1e-1, 9e-2]
    assert not any(mixed_phase_done(early[:k]) for k in range(1, len(early) + 1))
    # the cap
    assert mixed_phase_done([1e-2] * 30) and not mixed_phase_done([1e-2] * 29)
    # stall detection off
    assert not mixed_phase_done(plateau, stall=0)
