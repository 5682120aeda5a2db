"""Serving loop: admission by HBM budget and FIFO order (CPU, fake solves), and
the same energy alone and in the mix (GPU)."""
import threading
import time

import numpy as np
import pytest

from dftax_amd.serve import Request, Server, latency_stats


class FakeServer(Server):
    """Solves sleep; tracks the budget in use to check admission."""

    def __init__(self, budget, slots, durations):
        super().__init__("B3LYP", slots=slots, budget_bytes=budget)
        self.durations = durations
        self.in_use = 0
        self.max_in_use = 0
        self.max_active = 0
        self.active = 0
        self.order = []
        self._lk = threading.Lock()

    def estimate(self, req):
        return req.extra["fp"]

    def stream_for(self, slot):
        return None

    def solve(self, req, stream=None):
        with self._lk:
            self.in_use += req.footprint
            self.active += 1
            self.max_in_use = max(self.max_in_use, self.in_use)
            self.max_active = max(self.max_active, self.active)
            self.order.append(req.rid)
        time.sleep(self.durations[req.rid])
        req.e_tot, req.converged, req.n_iter = -1.0 * req.rid, True, 3
        with self._lk:
            self.in_use -= req.footprint
            self.active -= 1
        return req


def test_admission_respects_budget_and_slots():
    fps = [40, 30, 50, 20, 60, 10, 45, 35]
    reqs = [Request(rid=i, mol=None, arrival=0.01 * i, extra={"fp": fp}) for i, fp in enumerate(fps)]
    reqs.append(Request(rid=8, mol=None, arrival=0.0, extra={"fp": 500}))   # never fits
    srv = FakeServer(budget=100, slots=3, durations={i: 0.05 for i in range(9)})
    srv.run(reqs)
    assert srv.max_in_use <= 100
    assert srv.max_active <= 3
    ok = [r for r in reqs if not r.error]
    assert len(ok) == 8 and all(r.converged for r in ok)
    assert "exceeds the HBM budget" in reqs[8].error
    # FIFO: admission order equals arrival order
    assert srv.order == sorted(srv.order)
    assert all(r.start >= r.arrival - 1e-9 and r.end > r.start for r in ok)
    st = latency_stats(reqs)
    assert st["ok"] == 8 and st["p99_s"] >= st["p50_s"] > 0


def test_requests_overlap_when_memory_allows():
    reqs = [Request(rid=i, mol=None, arrival=0.0, extra={"fp": 10}) for i in range(6)]
    srv = FakeServer(budget=1000, slots=6, durations={i: 0.2 for i in range(6)})
    t = time.perf_counter()
    srv.run(reqs)
    assert time.perf_counter() - t < 0.6
    assert srv.max_active == 6


@pytest.mark.gpu
def test_same_energy_alone_and_in_the_mix():
    import torch
    from dftax_amd import becke
    from dftax_amd.system import serving_stream
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mols = serving_stream(4, seed=3, max_heavy=4)
    srv = Server("B3LYP", grid=becke(40, 110), precision="mixed", slots=4)
    alone = [srv.solve(Request(rid=i, mol=m)).e_tot for i, m in enumerate(mols)]
    reqs = [Request(rid=i, mol=m, arrival=0.0) for i, m in enumerate(mols)]
    srv.run(reqs)
    for a, r in zip(alone, reqs):
        assert not r.error and r.energy_converged
        print(f"alone {a:.10f} mix {r.e_tot:.10f}", flush=True)
        assert abs(a - r.e_tot) < 1e-8
