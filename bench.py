#!/usr/bin/env python3
"""SCF iterations per second of the DF-KS engine on one MI355X.

A step is one SCF iteration on the device (commutator, DIIS, eigensolve,
density, and the fused energy + Fock build: exchange GEMMs, Coulomb J over
the resident B, XC on the resident AO grid). The default workload is the
(H2O)_32 cluster (4x4x2 lattice, seed 0), def2-TZVP, PBE0, FP64, with B and
the AO grid resident in HBM before the timed region starts.

Prints one JSON line (see the fields below); roofline figures come from HIP
events recorded inside the library on the work stream.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# three bf16 MFMA products per multiply-add: a third of the 2.5 PFLOP/s dense bf16 peak
PEAK_BF16X3_TFLOPS = 2500.0 / 3.0

WORKLOADS = {
    "h2o": dict(shape=None, basis="def2-svp", screen=None),
    "h2o8": dict(shape=(2, 2, 2), basis="def2-tzvp", screen=None),
    "h2o32": dict(shape=(4, 4, 2), basis="def2-tzvp", screen=None),
    # B (~450 GB FP64) exceeds HBM: auxiliary slabs recomputed per Fock build, Schwarz 1e-12.
    # At this size the split-BF16 fixed point sits ~6e-3 (orbital gradient) from the FP64 one,
    # so the mixed phase ends at 1e-3 and the energy needs up to 8 FP64 steps.
    "h2o64": dict(shape=(4, 4, 4), basis="def2-tzvp", screen=1e-12, refine=8, mixed_d_tol=1e-3),
}


def measured_traffic(workload: str, stage: str):
    """HBM bytes per launch of a stage's FP64 kernel, from the committed counter summary
    (profiles/<workload>_pbe0_hbm_traffic.txt: FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_pbe0_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == f"{stage} (FP64)":
            return round((float(cells[2]) + float(cells[3])) * 1e9)
    return None


def note(msg: str) -> None:
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    """The host CPU's model name (Linux /proc/cpuinfo), for the CPU baseline's record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="h2o32", choices=sorted(WORKLOADS))
    ap.add_argument("--refine", type=int, default=None,
                    help="FP64 steps at most after the mixed phase in the energy-refined mixed SCF "
                         "(default 3; 8 for h2o64)")
    ap.add_argument("--mixed-d-tol", type=float, default=None,
                    help="orbital gradient that ends the mixed phase of the mixed SCFs (default 1e-4; 1e-3 for h2o64)")
    ap.add_argument("--xc", default="PBE0")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scf", action="store_true", help="skip the two full SCF solves (FP64 and mixed)")
    ap.add_argument("--serve-n", type=int, default=512,
                    help="requests of the serving leg (the def2-SVP B3LYP stream at 1x the isolated throughput); "
                         "0 skips it")
    ap.add_argument("--serve-isolated", type=int, default=64,
                    help="requests of the serving stream timed alone (the sample the rate and the per-request "
                         "slowdowns are taken from)")
    ap.add_argument("--counters", action="store_true",
                    help="short run for hardware-counter passes: a plain 1 GiB device copy (the byte-counter "
                         "calibration), then --steps timed steps per precision, no full SCF solves")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding of the aux axis is a design note only)", file=sys.stderr)
        return 2

    if args.serve_n > 0 and not args.counters:
        # one hardware queue per serving slot (HIP's default is 4), set before the runtime starts
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
    import ctypes
    import numpy as np
    import torch
    from dftax_amd import KS, Molecule, WATER, df, hip, scf, water_cluster
    from dftax_amd import _native as N
    from dftax_amd.scf import initial_density

    if args.counters:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)                    # reads 1 GiB and writes 1 GiB
        torch.cuda.synchronize()
        del src, dst

    wl = WORKLOADS[args.workload]
    if args.refine is None:
        args.refine = wl.get("refine", 3)
    if args.mixed_d_tol is None:
        args.mixed_d_tol = wl.get("mixed_d_tol", 1e-4)
    if wl["shape"] is None:
        mol = Molecule.from_xyz(WATER, wl["basis"], spherical=True)
    else:
        mol = water_cluster(wl["shape"], seed=0, basis=wl["basis"])
    t0 = time.perf_counter()
    ks = KS(mol, args.xc, coulomb=df(screen=wl["screen"], engine=hip(precision="mixed")))
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    note(f"built {args.workload} in {t_build:.1f} s")
    c = ks.coulomb
    L, h = c.L, c.handle
    stream = c._stream()
    N.check(L.dfks_set_timing(h, 1))
    P0 = initial_density(ks, None)
    out = (ctypes.c_double * 3)()

    def timed(mixed: bool):
        c.set_precision(mixed)
        N.check(L.dfks_scf_begin(h, ctypes.c_void_p(P0.data_ptr()), 1e-7, 8, 0.0, stream))
        for _ in range(args.warmup):
            N.check(L.dfks_scf_step(h, out, stream))
        acc = {}
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            N.check(L.dfks_scf_step(h, out, stream))
            for k, v in N.stage_times(h).items():
                if not k.startswith("build"):
                    acc[k] = acc.get(k, 0.0) + v
        torch.cuda.synchronize()
        return time.perf_counter() - t, {k: v / args.steps for k, v in acc.items()}

    dt, stage_avg = timed(False)          # FP64: the headline value
    note(f"FP64 steps: {1e3 * dt / args.steps:.1f} ms each")
    dt_mx, stage_mx = timed(True)         # mixed: split-BF16 MFMA products
    note(f"mixed steps: {1e3 * dt_mx / args.steps:.1f} ms each")
    build = {k: v for k, v in N.stage_times(h).items() if k.startswith("build")}
    if args.counters:
        print(json.dumps({"counters_run": args.workload, "steps": args.steps, "copy_bytes": 1 << 30,
                          "fp64_ms": round(1e3 * dt / args.steps, 3), "mixed_ms": round(1e3 * dt_mx / args.steps, 3)}),
              flush=True)
        return 0

    # both SCFs to convergence (e_tol 1e-8, d_tol 1e-6): final-energy agreement and time to solution
    r64 = rmx = rrf = None
    t64 = tmx = trf = float("nan")
    if not args.no_scf:
        torch.cuda.synchronize()
        t = time.perf_counter()
        r64 = scf(ks, precision="fp64")
        t64 = time.perf_counter() - t
        note(f"FP64 SCF: {r64.n_iter} iterations, {t64:.1f} s")
        t = time.perf_counter()
        rmx = scf(ks, precision="mixed", mixed_d_tol=args.mixed_d_tol)
        tmx = time.perf_counter() - t
        note(f"mixed SCF: {rmx.n_iter} iterations, {tmx:.1f} s")
        t = time.perf_counter()
        rrf = scf(ks, precision="mixed", refine=args.refine, mixed_d_tol=args.mixed_d_tol)
        trf = time.perf_counter() - t
        note(f"mixed SCF + FP64 energy refinement: {rrf.n_iter} iterations, {trf:.1f} s")

    # response Fock matrix of the Newton Hessian-vector product at the converged density:
    # dP = U V^T + V U^T with U = C_occ and V = -2 C_virt v^T (one CG direction); materialised
    # workloads only (with recomputed slabs one product costs two slab passes per factor sign)
    resp = None
    if r64 is not None and c.plan.aux_chunk == 0:
        c.set_precision(False)
        C = r64.mo_coeff[0]
        no = ks.nocc[0]
        g = torch.Generator(device=C.device).manual_seed(0)
        v = torch.randn((no, C.shape[1] - no), dtype=torch.float64, device=C.device, generator=g)
        U, V = C[:, :no].contiguous(), (-2.0 * (C[:, no:] @ v.T)).contiguous()
        torch.cuda.synchronize()
        t = time.perf_counter()
        c.response_begin(r64.P[0])
        torch.cuda.synchronize()
        t_begin = time.perf_counter() - t
        t = time.perf_counter()
        for _ in range(3):
            c.response_fock(U, V)
        torch.cuda.synchronize()
        resp = {"begin_ms": round(1e3 * t_begin, 3), "response_fock_ms": round(1e3 * (time.perf_counter() - t) / 3, 3),
                "rank": int(no)}
        note(f"response Fock: {resp['response_fock_ms']:.1f} ms per Hessian-vector product")

    nao, nfit, nocc, ng = c.nao, c.nfit, c.nocc, c.ngrid
    recompute = c.plan.aux_chunk > 0 or c.stats()["ao_block"] > 0
    flops = {
        "exchange_gemm": 2.0 * nao * nao * nfit * nocc,
        # the symmetric product computes the upper tiles only: nao (nao+1) / 2 dot products
        "exchange_syrk": 1.0 * nao * (nao + 1) * nfit * nocc,
        "xc_gemm": 2.0 * ng * nao * nao,
        "xc_syr2k": 2.0 * ng * nao * nao,
    }
    compact = c.stats().get("xc_blocks", 0) > 0
    if compact:
        # compact XC grid: Z, the pointwise kernel and the per-block products share one stage
        # (xc_kernel) and the row assembly is xc_syr2k; no dense GEMM counts apply
        flops.pop("xc_gemm")
        flops.pop("xc_syr2k")
    if recompute:
        # stages hold integral recomputation: only K keeps a flop count
        flops = {"exchange_syrk": flops["exchange_syrk"]}

    def stage_table(avg, elem_bytes, peak_tf):
        # J = B gamma reads the rows m >= l of B (symmetric in its two orbital indices)
        bytes_ = {} if recompute else {"coulomb_j": elem_bytes * nao * (nao + 1) / 2 * nfit}
        if not recompute and not compact:
            bytes_["xc_kernel"] = elem_bytes * ng * nao * 6
        tab = {}
        for k, ms in avg.items():
            ent = {"ms": round(ms, 4)}
            if flops.get(k) and ms > 0:
                a = flops[k] / (ms * 1e-3) / 1e12
                ent.update(bound="mfma", achieved_tflops=round(a, 2), frac=round(a / peak_tf, 3))
            if k in bytes_ and ms > 0:
                a = bytes_[k] / (ms * 1e-3) / 1e9
                ent.update(bound="hbm", achieved_gbs=round(a, 1), frac=round(a / PEAK_HBM_GBS, 3))
            tab[k] = ent
        return tab

    stages = stage_table(stage_avg, 8.0, PEAK_FP64_TFLOPS)
    stages_mx = stage_table(stage_mx, 4.0, PEAK_BF16X3_TFLOPS)
    dom = max((k for k in stage_avg if k in flops and stage_avg[k] > 0), key=lambda k: stage_avg[k])
    ach = flops[dom] / (stage_avg[dom] * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": measured_traffic(args.workload, dom),
            "kernel": dom, "per_launch": f"{flops[dom]:.4g} flop, {stage_avg[dom]:.3f} ms"}

    cpu = None
    if not args.no_cpu_baseline and not recompute:
        from oracle.ref import cpu_iteration_estimate
        threads = torch.get_num_threads()
        sec, parts = cpu_iteration_estimate(nao, nfit, nocc, ng)
        cpu = {"value": round(1.0 / sec, 6), "unit": "SCF iter/s", "cores": threads, "kind": "port",
               "cpu_model": _cpu_model(),
               "sample": ("CPU FP64 restatement (numpy/BLAS, torch for the functional) of one iteration on "
                          f"synthetic operands of the workload's shapes: exchange/J on 96 of {nfit} fitted aux "
                          f"functions; XC on 40000 of {ng} grid points (Z = Phi P, rho and grad rho, PBE0 "
                          "with its first derivatives, W, V = Phi^T W); each slice scaled linearly to the "
                          f"full size; full-size DIIS/transform GEMMs and {nao}x{nao} eigh"),
               "stages_s": {k: round(v, 3) for k, v in parts.items()}}

    mixed = {
        "iter_s": round(args.steps / dt_mx, 4),
        "ms_per_step": round(1e3 * dt_mx / args.steps, 3),
        "speedup_per_iter_vs_fp64": round(dt / dt_mx, 3),
        "dtype": "fp32 operands, 3-term split-bf16 MFMA products, fp32 accumulate; fp64 J sum, functional, DIIS",
        "stages": stages_mx,
    }
    if r64 is not None:
        mixed.update({
            "scf_fp64": {"e_tot": r64.e_tot, "n_iter": r64.n_iter, "converged": r64.converged,
                         "wall_s": round(t64, 3)},
            "scf_mixed": {"e_tot": rmx.e_tot, "n_iter": rmx.n_iter, "converged": rmx.converged,
                          "wall_s": round(tmx, 3), "n_mixed_steps": sum(1 for x in rmx.history if x[3] == "mixed")},
            "abs_dE_mixed_vs_fp64": abs(rmx.e_tot - r64.e_tot),
            "time_to_solution_speedup": round(t64 / tmx, 3),
            # converge in mixed precision, then FP64 Fock builds until the energy settles (at most `refine`)
            "scf_mixed_refine": {"e_tot": rrf.e_tot, "n_iter": rrf.n_iter, "energy_converged": rrf.energy_converged,
                                  "refine": args.refine, "mixed_d_tol": args.mixed_d_tol,
                                  "wall_s": round(trf, 3),
                                  "n_mixed_steps": sum(1 for x in rrf.history if x[3] == "mixed"),
                                  "n_fp64_steps": sum(1 for x in rrf.history if x[3] == "fp64"),
                                  "last_step_orbital_gradient": rrf.history[-1][2],
                                  "abs_dE_vs_fp64": abs(rrf.e_tot - r64.e_tot),
                                  "time_to_solution_speedup": round(t64 / trf, 3)},
        })
    # serving leg: the (H2O)_32 handle goes first, then the request stream runs on the whole device
    # three-centre integral build: unique (mn|P) values (bra shell pairs a >= b, every auxiliary
    # function) per build, and the bytes of the (mn|P) tensor written (both bra triangles, FP64)
    eri = None
    if not recompute and build.get("build_3c", 0) > 0:
        b = ks.basis
        nf = np.where((b.l >= 2) & bool(b.spherical), 2 * b.l + 1, (b.l + 1) * (b.l + 2) // 2)
        pair_funcs = (np.sum(nf) ** 2 + np.sum(nf ** 2)) / 2            # sum over a >= b of nf_a nf_b
        n_int = float(pair_funcs) * c.naux
        sec = build["build_3c"] * 1e-3
        wbytes = 8.0 * nao * nao * c.naux
        eri = {"integrals": n_int, "ms": round(build["build_3c"], 1),
               "integrals_per_s": round(n_int / sec, 1), "bytes_written": wbytes,
               "achieved_gbs_written": round(wbytes / sec / 1e9, 1),
               "hbm_frac": round(wbytes / sec / 1e9 / PEAK_HBM_GBS, 4),
               "note": "bound by FP64 VALU work (Boys function and Hermite recursions), not by HBM writes"}
    serving = None
    mem_plan = {"aux_chunk": c.plan.aux_chunk, "ao_block": c.plan.ao_block, **c.stats()}
    if args.serve_n > 0:
        c.close()
        del ks, c
        N.check(N.lib().dfks_release_memory(N.current_stream()))
        torch.cuda.empty_cache()
        from dftax_amd.serve import run_stream
        t = time.perf_counter()
        sv = run_stream(args.serve_n, (1.0,), seed=0, slots=16, xc="B3LYP", precision="mixed",
                        isolated=args.serve_isolated, refine=3, progress=note)
        r1 = sv["rate_1x"]
        serving = {
            "workload": f"{args.serve_n} molecules of the seeded stream (2-24 heavy atoms C:N:O 6:1:1), def2-svp / "
                        "even-tempered-jk, B3LYP, default grid, mixed precision + FP64 energy refinement (<= 3 steps)",
            "arrivals": "Poisson at 1x the isolated-solve throughput",
            "isolated_sample": f"the first {sv['isolated_requests']} requests, each solved alone after an untimed "
                               "first pass over them",
            "slots": sv["slots"], "solves_per_min": round(r1["solves_per_min"], 1),
            "p50_s": round(r1["p50_s"], 4), "p99_s": round(r1["p99_s"], 4),
            "isolated_solve_s_mean": round(sv["isolated_solve_s_mean"], 4),
            "isolated_solve_s_p99": round(sv["isolated_solve_s_p99"], 4),
            "p99_over_isolated_mean": round(r1["p99_over_isolated"], 3),
            "p99_over_isolated_p99": round(r1["p99_over_isolated_p99"], 3),
            "slowdown_p50": round(r1.get("slowdown_p50", float("nan")), 3),
            "slowdown_p99": round(r1.get("slowdown_p99", float("nan")), 3),
            "ok": r1["ok"], "energy_converged": r1["energy_converged"], "wall_s": round(time.perf_counter() - t, 1),
        }

    rec = {
        "metric": "SCF iter/s FP64 vs mixed (|dE|<1e-6 Ha); solves/min + p99 latency",
        "value": round(args.steps / dt, 4),
        "unit": "SCF iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded water-cluster geometry)",
        "config": {"workload": f"{args.workload}: {mol.natm} atoms, {wl['basis']} / even-tempered-jk, "
                               f"{args.xc}, becke(75,302) grid", "nao": nao, "naux_fit": nfit, "nocc": nocc,
                   "ngrid": ng, "precision": "fp64 (value); mixed reported under 'mixed'", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages": stages,
        "mixed": mixed,
        "memory_plan": mem_plan,
        "newton_response": resp,
        "eri_build": eri,
        "serving": serving,
        "build_s": round(t_build, 3),
        "build_stages_ms": {k: round(v, 1) for k, v in build.items()},
    }
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
