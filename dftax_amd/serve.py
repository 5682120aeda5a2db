"""Continuous-batching SCF server on one GPU.

Requests (molecules) arrive over time; each one needs a full DF-KS solve.
The server keeps a FIFO of arrived requests and admits the head as soon as
a worker slot is free and its HBM footprint fits the remaining budget. The
footprint counts B (nao^2 * nfit words, plus its FP32 copy and the unfitted
tensor during the build), the AO grid and the SCF workspaces. Each admitted
request runs on its own HIP stream from its own worker thread: its native
handle is built, the SCF runs (mixed precision with FP64 refinement by
default) and the handle is freed, returning its budget. Kernels of
different molecules therefore overlap on the device, and a request that
arrives while others run starts as soon as memory and a slot free up.

There is no counterpart in dftax; the nearest is ``scf_batched``
(dftax/ks/batched.py:82-304), which vmaps one molecule over geometries.

Command line (one JSON line on stdout):
    python -m dftax_amd.serve --n 64 --rate 1.0 --seed 0
where ``--rate`` is the arrival rate in multiples of the isolated-solve
throughput measured first on the same requests.
"""

from __future__ import annotations

import argparse
import json
import os
import sys

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from dftax_amd.basis import build_basis
from dftax_amd import _native
from dftax_amd.grid import Becke, atom_centred_points, becke


@dataclass
class Request:
    rid: int
    mol: object
    arrival: float = 0.0
    footprint: int = 0
    start: float = 0.0
    end: float = 0.0
    e_tot: float = float("nan")
    n_iter: int = 0
    converged: bool = False
    energy_converged: bool = False
    error: str = ""
    extra: dict = field(default_factory=dict)


def footprint_bytes(mol, auxbasis: str, grid: Becke, precision: str) -> int:
    """HBM bytes one request holds at its peak (build and SCF)."""
    b = build_basis(mol.symbols, mol.coords, mol.basis, spherical=mol.spherical)
    x = build_basis(mol.symbols, mol.coords, auxbasis, spherical=True)
    nao, naux = b.nao, x.nao
    ng = atom_centred_points(mol.symbols, mol.coords, grid)[0].shape[0]
    nocc = mol.nelectron // 2
    w3 = nao * nao * naux
    f32 = precision == "mixed"
    build_peak = 2 * 8 * w3                                   # (mn|P) and B together
    steady = 8 * w3 + (4 * w3 if f32 else 0)                  # B (+ FP32 copy)
    steady += ng * nao * (8 * 5 + (4 * 5 if f32 else 0))      # AO planes + Z/W (+ FP32 copies)
    steady += nao * naux * max(nocc, 1) * (8 + (4 if f32 else 0))
    steady += 8 * nao * nao * (64 + 40)                       # split-K workspace, DIIS, matrices
    return int(max(build_peak, steady) * 1.1) + (64 << 20)


class Server:
    """Admission by HBM budget, one worker thread and HIP stream per slot."""

    def __init__(self, xc: str = "B3LYP", *, auxbasis: str = "even-tempered-jk", grid: Becke | None = None,
                 precision: str = "mixed", slots: int = 16, budget_bytes: int | None = None, e_tol: float = 1e-8,
                 d_tol: float = 1e-6, device: int = 0, refine: int | None = 3):
        self.xc = xc
        self.auxbasis = auxbasis
        self.grid = grid if grid is not None else becke()
        self.precision = precision
        self.slots = slots
        self.e_tol, self.d_tol = e_tol, d_tol
        # mixed solves end once the FP64 energy has converged (at most `refine` FP64
        # steps: time to converged energy); None runs the FP64 steps to d_tol as well
        self.refine = refine if precision == "mixed" else None
        self.device_index = device
        if budget_bytes is None:
            budget_bytes = int(0.85 * _native.device_free(device))
        self.budget = budget_bytes
        self._free = budget_bytes
        self._cv = threading.Condition()
        self._streams = {}

    @property
    def device(self):
        import torch
        return torch.device("cuda", self.device_index)

    def stream_for(self, slot: int):
        """The HIP stream of a worker slot (created on first use)."""
        import torch
        if slot not in self._streams:
            self._streams[slot] = torch.cuda.Stream(device=self.device)
        return self._streams[slot]

    def estimate(self, req: Request) -> int:
        return footprint_bytes(req.mol, self.auxbasis, self.grid, self.precision)

    # -- one solve -----------------------------------------------------------
    def solve(self, req: Request, stream=None) -> Request:
        import torch
        from dftax_amd import _native
        from dftax_amd.ks import KS, df, hip
        from dftax_amd.scf import scf
        stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.device(self.device), torch.cuda.stream(stream):
            ks = None
            try:
                t0 = time.perf_counter()
                ks = KS(req.mol, self.xc, grid=self.grid,
                        coulomb=df(self.auxbasis, engine=hip(precision=self.precision)), device=self.device)
                t1 = time.perf_counter()
                r = scf(ks, e_tol=self.e_tol, d_tol=self.d_tol, precision=self.precision, refine=self.refine)
                req.e_tot, req.n_iter, req.converged = r.e_tot, r.n_iter, r.converged
                req.energy_converged = r.energy_converged
                # the library's build phases (host timers, always kept; read with timing on for the call)
                hnd = ks.coulomb.handle
                _native.check(ks.coulomb.L.dfks_set_timing(hnd, 1))
                lib_build = {k[6:]: round(v, 1) for k, v in _native.stage_times(hnd).items() if k.startswith("build")}
                _native.check(ks.coulomb.L.dfks_set_timing(hnd, 0))
                req.extra = {"nao": ks.nao, "ngrid": int(ks.grid_w.shape[0]),
                             "n_mixed": sum(1 for h in r.history if h[3] == "mixed"),
                             "build_s": t1 - t0, "scf_s": time.perf_counter() - t1,
                             "build_lib_ms": lib_build}
            except Exception as exc:  # reported per request, the server keeps going
                req.error = f"{type(exc).__name__}: {exc}"
            finally:
                if ks is not None:
                    ks.coulomb.close()
                stream.synchronize()
        return req

    # -- the loop ------------------------------------------------------------
    def run(self, requests: list[Request], progress=None) -> list[Request]:
        """Serve requests at their ``arrival`` offsets (seconds from now);
        ``progress(msg)`` is called every 64 completed requests."""
        for r in requests:
            r.footprint = self.estimate(r)
            if r.footprint > self.budget:
                r.error = f"footprint {r.footprint} exceeds the HBM budget {self.budget}"
        pending = sorted([r for r in requests if not r.error], key=lambda r: r.arrival)
        t0 = time.perf_counter()
        waiting: list[Request] = []
        free_slots = list(range(self.slots))
        threads = []
        done: list[Request] = []
        lock = threading.Lock()

        def worker(req: Request, slot: int):
            req.start = time.perf_counter() - t0
            self.solve(req, self.stream_for(slot))
            req.end = time.perf_counter() - t0
            with self._cv:
                self._free += req.footprint
                free_slots.append(slot)
                with lock:
                    done.append(req)
                    if progress is not None and len(done) % 64 == 0:
                        progress(f"served {len(done)} of {len(requests)} requests")
                self._cv.notify_all()

        i = 0
        n = len(pending)
        while len(done) < n:
            now = time.perf_counter() - t0
            while i < n and pending[i].arrival <= now:
                waiting.append(pending[i])
                i += 1
            with self._cv:
                # admit FIFO while the head fits
                while waiting and free_slots and waiting[0].footprint <= self._free:
                    req = waiting.pop(0)
                    self._free -= req.footprint
                    slot = free_slots.pop(0)
                    th = threading.Thread(target=worker, args=(req, slot), daemon=True)
                    threads.append(th)
                    th.start()
                next_arrival = pending[i].arrival - now if i < n else None
                timeout = 0.05 if next_arrival is None else max(0.0, min(0.05, next_arrival))
                self._cv.wait(timeout=timeout)
        for th in threads:
            th.join()
        return requests


def latency_stats(reqs: list[Request], alone: dict | None = None) -> dict:
    """Latency percentiles; with ``alone`` (request id -> isolated solve seconds) also the
    per-request slowdown latency / isolated time of the same molecule."""
    ok = [r for r in reqs if not r.error]
    lat = np.array([r.end - r.arrival for r in ok])
    span = max(r.end for r in ok) - min(r.arrival for r in ok) if ok else float("nan")
    build = np.array([r.extra.get("build_s", np.nan) for r in ok])
    scf_t = np.array([r.extra.get("scf_s", np.nan) for r in ok])
    wait = np.array([r.start - r.arrival for r in ok])
    out = {
        "n": len(reqs), "ok": len(ok), "converged": sum(r.converged for r in ok),
        "energy_converged": sum(r.energy_converged for r in ok),
        "mean_build_s": float(np.nanmean(build)) if ok else None,
        "mean_scf_s": float(np.nanmean(scf_t)) if ok else None,
        "mean_queue_wait_s": float(wait.mean()) if ok else None,
        "solves_per_min": 60.0 * len(ok) / span if ok else 0.0,
        "p50_s": float(np.percentile(lat, 50)) if ok else None,
        "p99_s": float(np.percentile(lat, 99)) if ok else None,
        "mean_s": float(lat.mean()) if ok else None,
        "span_s": span,
    }
    if alone and ok:
        slow = np.array([(r.end - r.arrival) / alone[r.rid] for r in ok if r.rid in alone])
        if slow.size:
            out["slowdown_p50"] = float(np.percentile(slow, 50))
            out["slowdown_p99"] = float(np.percentile(slow, 99))
    return out


def run_stream(n: int = 64, rates=(1.0, 2.0), *, seed: int = 0, slots: int = 16, xc: str = "B3LYP",
               precision: str = "mixed", isolated: int = 0, refine: int | None = 3, progress=None) -> dict:
    """The serving workload end to end: ``n`` molecules of the seeded stream,
    the isolated-solve time of the first ``isolated`` of them (0: all; after an
    untimed pass over the same requests, so the autotuner has seen their
    shapes), then Poisson arrivals at each rate x the isolated throughput.
    Returns the isolated times and, per rate, ``latency_stats`` with the
    per-request slowdown over the isolated sample. ``progress(msg)`` is called
    as the phases advance."""
    import torch
    from dftax_amd.system import serving_stream
    say = progress or (lambda msg: None)
    mols = serving_stream(n, seed=seed)
    srv = Server(xc, precision=precision, slots=slots, refine=refine or None)
    n_iso = len(mols) if isolated <= 0 else min(isolated, len(mols))
    warm = []
    for k in range(n_iso):
        t = time.perf_counter()
        srv.solve(Request(rid=-1, mol=mols[k]))
        warm.append(time.perf_counter() - t)
    say(f"untimed first pass over {n_iso} requests: {sum(warm):.1f} s")
    alone, iso_req = {}, []
    for k in range(n_iso):
        t = time.perf_counter()
        r = srv.solve(Request(rid=k, mol=mols[k]))
        alone[k] = time.perf_counter() - t
        iso_req.append(r)
        if r.error:
            raise RuntimeError(r.error)
    iso = list(alone.values())
    tau = float(np.mean(iso))
    say(f"isolated solves: {n_iso}, mean {tau:.3f} s")
    out = {"isolated_solve_s_mean": tau, "isolated_solve_s_p99": float(np.percentile(iso, 99)),
           "isolated_build_s_mean": float(np.mean([r.extra["build_s"] for r in iso_req])),
           "isolated_scf_s_mean": float(np.mean([r.extra["scf_s"] for r in iso_req])),
           "isolated_requests": n_iso,
           # per request, in request order: the untimed first solve, the isolated solve and its build part
           "isolated_first_pass_s": [round(x, 4) for x in warm],
           "isolated_s": [round(alone[k], 4) for k in range(n_iso)],
           "isolated_build_s": [round(r.extra["build_s"], 4) for r in iso_req],
           "isolated_build_lib_ms": [r.extra["build_lib_ms"] for r in iso_req]}
    rng = np.random.default_rng(seed + 1)
    for rate in rates:
        lam = rate / tau
        gaps = rng.exponential(1.0 / lam, size=len(mols))
        arr = np.cumsum(gaps) - gaps[0]
        reqs = [Request(rid=k, mol=m, arrival=float(a)) for k, (m, a) in enumerate(zip(mols, arr))]
        srv.run(reqs, progress=progress)
        st = latency_stats(reqs, alone)
        st["p99_over_isolated"] = st["p99_s"] / tau if st["p99_s"] else None
        st["p99_over_isolated_p99"] = st["p99_s"] / out["isolated_solve_s_p99"] if st["p99_s"] else None
        out[f"rate_{rate:g}x"] = st
        say(f"rate {rate:g}x: {st['solves_per_min']:.0f} solves/min, p99 {st['p99_s']:.2f} s")
    out.update({"n_requests": len(mols), "xc": xc, "basis": "def2-svp / even-tempered-jk",
                "precision": precision, "refine_fp64_steps": srv.refine, "slots": slots,
                "budget_GB": srv.budget / 1e9, "device": torch.cuda.get_device_name(0)})
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--rate", type=float, nargs="+", default=[1.0, 2.0])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--xc", default="B3LYP")
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--isolated", type=int, default=0,
                    help="requests timed alone to set the rate (0: all, which also gives per-request slowdowns)")
    ap.add_argument("--refine", type=int, default=3,
                    help="at most this many FP64 steps after the mixed phase, stopping when the energy "
                         "has converged (0: run them to the density tolerance as well)")
    args = ap.parse_args()
    # one hardware queue per worker stream (HIP's default is 4), set before the runtime starts
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(min(max(4, args.slots), 16)))
    out = run_stream(args.n, args.rate, seed=args.seed, slots=args.slots, xc=args.xc, precision=args.precision,
                     isolated=args.isolated, refine=args.refine,
                     progress=lambda msg: print(f"serve: {msg}", file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
