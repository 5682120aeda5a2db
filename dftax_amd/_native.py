"""ctypes binding of libdfks.so (declarations in include/dfks.h).

The library is loaded from ``dftax_amd/_lib/libdfks.so``; there is no
fallback: if it is missing, every entry point raises. Device buffers cross
the boundary as raw pointers (torch tensors supply them here).
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libdfks.so"

c_int32, c_int64, c_double, c_void_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_f64 = ctypes.POINTER(ctypes.c_double)


class BasisDesc(ctypes.Structure):
    _fields_ = [("nshell", c_int32), ("l", P_i32), ("nprim", P_i32), ("prim0", P_i32), ("ao0", P_i32),
                ("centre", P_f64), ("exps", P_f64), ("coefs", P_f64), ("nprim_total", c_int32),
                ("spherical", c_int32), ("nfunc", c_int32)]


class SystemDesc(ctypes.Structure):
    _fields_ = [("orbital", BasisDesc), ("aux", BasisDesc), ("natom", c_int32), ("atom_xyz", P_f64),
                ("atom_z", P_f64), ("ngrid", c_int64), ("grid_xyz", P_f64), ("grid_w", P_f64),
                ("xc_id", c_int32), ("nelec", c_int32), ("nspin", c_int32), ("spin", c_int32)]


class Options(ctypes.Structure):
    _fields_ = [("precision", c_int32), ("metric_cutoff", c_double), ("rho_thresh", c_double),
                ("keep_int3c", c_int32), ("gemm", c_int32), ("density", c_int32), ("aux_chunk", c_int32),
                ("screen", c_double), ("ao_block", c_int64), ("xc_blocks", c_int32)]


DFKS_S, DFKS_HCORE, DFKS_KINETIC, DFKS_NUCLEAR, DFKS_INT2C, DFKS_INT3C, DFKS_BFIT, DFKS_AO, DFKS_GRID_W = range(9)

EXPORTS = [
    "dfks_last_error", "dfks_create", "dfks_destroy", "dfks_build", "dfks_sizes", "dfks_get_matrix",
    "dfks_energy_fock", "dfks_energy_fock_factored", "dfks_xc_eval", "dfks_scf_begin", "dfks_scf_step",
    "dfks_scf_result", "dfks_becke_weights", "dfks_overlap", "dfks_set_timing", "dfks_stage_times",
    "dfks_set_precision", "dfks_stats", "dfks_xc_eval2", "dfks_response_begin", "dfks_response_fock",
    "dfks_device_free", "dfks_release_memory", "dfks_becke_weights_noalloc", "dfks_forces",
    "dfks_xc_eval_pol",
]

STAGES = ["exchange_gemm", "exchange_syrk", "coulomb_j", "xc_gemm", "xc_kernel", "xc_syr2k", "fock_assembly",
          "scf_diis", "scf_density", "build_1e", "build_2c", "build_3c", "build_metric_B", "build_ao",
          "build_ao_sort", "build_ao_scan", "build_ao_tables", "build_ao_upload", "build_ao_eval"]

_lib = None


class DfksError(RuntimeError):
    pass


def lib():
    """Load libdfks.so; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise DfksError(f"{LIB_PATH} is missing: build it with `python -m dftax_amd.build` "
                        "(the DF-KS path has no CPU fallback)")
    L = ctypes.CDLL(str(LIB_PATH))
    L.dfks_last_error.restype = ctypes.c_char_p
    for name in EXPORTS[1:]:
        getattr(L, name).restype = c_int32
    L.dfks_create.argtypes = [ctypes.POINTER(SystemDesc), ctypes.POINTER(Options), ctypes.POINTER(c_void_p)]
    L.dfks_destroy.argtypes = [c_void_p]
    L.dfks_build.argtypes = [c_void_p, c_void_p]
    L.dfks_sizes.argtypes = [c_void_p, ctypes.POINTER(c_int64)]
    L.dfks_get_matrix.argtypes = [c_void_p, c_int32, c_void_p, c_void_p]
    L.dfks_energy_fock.argtypes = [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]
    L.dfks_energy_fock_factored.argtypes = [c_void_p, c_void_p, c_void_p, P_f64, c_int32, c_void_p, c_void_p,
                                            c_void_p]
    L.dfks_xc_eval.argtypes = [c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]
    L.dfks_xc_eval2.argtypes = [c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]
    L.dfks_response_begin.argtypes = [c_void_p, c_void_p, c_void_p]
    L.dfks_response_fock.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]
    L.dfks_scf_begin.argtypes = [c_void_p, c_void_p, c_double, c_int32, c_double, c_void_p]
    L.dfks_scf_step.argtypes = [c_void_p, P_f64, c_void_p]
    L.dfks_scf_result.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, P_i32, c_void_p]
    L.dfks_becke_weights.argtypes = [c_int64, P_f64, P_f64, P_i32, c_int32, P_f64, P_f64, P_f64, c_void_p]
    L.dfks_overlap.argtypes = [ctypes.POINTER(BasisDesc), P_f64, c_void_p]
    L.dfks_set_timing.argtypes = [c_void_p, c_int32]
    L.dfks_set_precision.argtypes = [c_void_p, c_int32]
    L.dfks_stage_times.argtypes = [c_void_p, P_f64, c_int32]
    L.dfks_stats.argtypes = [c_void_p, ctypes.POINTER(c_int64), c_int32]
    L.dfks_device_free.argtypes = [ctypes.POINTER(c_int64)]
    L.dfks_release_memory.argtypes = [c_void_p]
    L.dfks_xc_eval_pol.argtypes = [c_int32, c_int64, c_void_p, c_void_p, c_void_p]
    L.dfks_forces.argtypes = [c_void_p, c_void_p, c_int32, c_int64, P_f64, P_f64, P_i32, P_f64, P_f64, c_void_p]
    L.dfks_becke_weights_noalloc.argtypes = [c_int64, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p,
                                             c_void_p, c_void_p]
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        msg = lib().dfks_last_error().decode(errors="replace")
        raise DfksError(f"dfks error {status}: {msg}")


def f64p(a: np.ndarray):
    return a.ctypes.data_as(P_f64)


def i32p(a: np.ndarray):
    return a.ctypes.data_as(P_i32)


class BasisArrays:
    """Keeps the numpy arrays of a BasisDesc alive."""

    def __init__(self, basis):
        self.l = np.ascontiguousarray(basis.l, np.int32)
        self.nprim = np.ascontiguousarray(basis.nprim, np.int32)
        self.prim0 = np.ascontiguousarray(basis.prim0, np.int32)
        self.ao0 = np.ascontiguousarray(basis.ao0, np.int32)
        self.centre = np.ascontiguousarray(basis.centre, np.float64)
        self.exps = np.ascontiguousarray(basis.exps, np.float64)
        self.coefs = np.ascontiguousarray(basis.coefs, np.float64)
        self.desc = BasisDesc(len(self.l), i32p(self.l), i32p(self.nprim), i32p(self.prim0), i32p(self.ao0),
                              f64p(self.centre), f64p(self.exps), f64p(self.coefs), len(self.exps),
                              int(basis.spherical), int(basis.nao))


def current_stream() -> c_void_p:
    """The calling thread's current HIP stream (torch's), as a ctypes pointer."""
    import torch
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def becke_weights(pts: np.ndarray, raw: np.ndarray, owner: np.ndarray, atom_xyz: np.ndarray,
                  bragg: np.ndarray) -> np.ndarray:
    pts = np.ascontiguousarray(pts, np.float64)
    raw = np.ascontiguousarray(raw, np.float64)
    owner = np.ascontiguousarray(owner, np.int32)
    xyz = np.ascontiguousarray(atom_xyz, np.float64)
    br = np.ascontiguousarray(bragg, np.float64)
    out = np.empty(raw.shape[0])
    check(lib().dfks_becke_weights(raw.shape[0], f64p(pts), f64p(raw), i32p(owner), xyz.shape[0], f64p(xyz),
                                   f64p(br), f64p(out), current_stream()))
    return out


def device_free(device) -> int:
    """Bytes the library can obtain on ``device``: free device memory, plus the
    default pool's unused reserve and the large buffers the library keeps for
    reuse (both of which the device counts as in use)."""
    import torch
    out = c_int64()
    with torch.cuda.device(device):
        check(lib().dfks_device_free(ctypes.byref(out)))
    return int(out.value)


def release_memory(device) -> None:
    """Give the large buffers the library keeps for reuse, and the pool's unused
    reserve, back to ``device`` (waits for their last users)."""
    import torch
    with torch.cuda.device(device):
        check(lib().dfks_release_memory(current_stream()))


def overlap(basis) -> np.ndarray:
    """Overlap matrix of a basis, computed on the GPU (used by the minao guess)."""
    arrs = BasisArrays(basis)
    out = np.empty((basis.nao, basis.nao))
    check(lib().dfks_overlap(ctypes.byref(arrs.desc), f64p(out), current_stream()))
    return out


def stage_times(handle) -> dict:
    out = (ctypes.c_double * len(STAGES))()
    check(lib().dfks_stage_times(handle, out, len(STAGES)))
    return {k: float(v) for k, v in zip(STAGES, out)}


STATS = ["pairs_kept", "pairs_total", "aux_slabs", "max_slab", "ao_block", "eigh_steps", "sp2_iterations",
         "xc_blocks", "xc_bins", "xc_width_sum", "pchol_steps"]


def stats(handle) -> dict:
    out = (c_int64 * len(STATS))()
    check(lib().dfks_stats(handle, out, len(STATS)))
    return {k: int(v) for k, v in zip(STATS, out)}
