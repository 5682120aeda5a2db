"""SCF driver on the device (dftax/ks/scf.py:394-495 surface).

Each iteration runs entirely on the GPU (libdfks ``dfks_scf_step``): the
commutator error X^T(FPS - SPF)X, Pulay DIIS (ridge 1e-10, identity-masked
empty slots), the optional level shift, eigh(X^T F X), the aufbau density and
the fused energy + Fock build at the new density. The host reads back three
scalars per iteration for the convergence test |dE| < e_tol and |err| < d_tol.
"""

from __future__ import annotations

import ctypes
import time
import warnings
from dataclasses import dataclass, field

import numpy as np

from dftax_amd import _native
from dftax_amd.guess import CoreSpec, MinAOSpec, minao_density


@dataclass
class KSResult:
    """Outcome of a solve (fields of dftax's KSResult, scf.py:49-69, plus timings)."""

    e_tot: float
    e_elec: float
    converged: bool
    n_iter: int
    nocc: tuple
    mo_energy: object
    mo_coeff: object
    P: object
    ts: float = 0.0
    # the energy test alone (|dE| < e_tol) held at the last step; with
    # scf(refine=n) the solve stops on it while ``converged`` keeps the
    # reference's meaning (|dE| < e_tol and |err| < d_tol, scf.py:372)
    energy_converged: bool = False
    iter_times: list = field(default_factory=list)
    history: list = field(default_factory=list)


def initial_density(ks, guess):
    """Device P0 for a guess spec or an explicit array (guess.py:602-683): nao x nao
    for a restricted KS, 2 x nao x nao (alpha, beta) for an open-shell one."""
    if getattr(ks, "nspin", 1) == 2:
        return _initial_density_uks(ks, guess)
    torch = __import__("torch")
    n = ks.nao
    if guess is None or isinstance(guess, MinAOSpec):
        name = guess.basis if isinstance(guess, MinAOSpec) else "sto-3g"
        P = minao_density(ks.symbols, ks.atom_coords, ks.basis, name)
        Pt = torch.tensor(P, dtype=torch.float64, device=ks.device)
        Pt = Pt * (ks.nelec / float(torch.sum(Pt * ks.S)))
        return Pt
    if isinstance(guess, CoreSpec):
        return _core_density(ks)
    P0 = torch.as_tensor(guess, dtype=torch.float64, device=ks.device)
    if P0.dim() == 3:
        P0 = P0[0]
    if tuple(P0.shape) != (n, n):
        raise ValueError(f"guess density has shape {tuple(P0.shape)}, expected {(n, n)}")
    return P0.contiguous()


def _orthonormalizer(ks, thresh=1e-7):
    torch = __import__("torch")
    s, U = torch.linalg.eigh(ks.S)
    keep = s > thresh
    return U[:, keep] / torch.sqrt(s[keep])[None, :]


def _initial_density_uks(ks, guess):
    """Spin-stacked P0 (guess.py:603-626): the minao density renormalised to the
    electron count and split by the channel occupations, the aufbau fill of the
    hcore eigenvectors per channel, or an explicit (2, nao, nao) array."""
    torch = __import__("torch")
    n = ks.nao
    na, nb = ks.nocc
    if guess is None or isinstance(guess, MinAOSpec):
        name = guess.basis if isinstance(guess, MinAOSpec) else "sto-3g"
        P = torch.tensor(minao_density(ks.symbols, ks.atom_coords, ks.basis, name), dtype=torch.float64,
                         device=ks.device)
        P = P * (ks.nelec / float(torch.sum(P * ks.S)))
        return torch.stack([(na / ks.nelec) * P, (nb / ks.nelec) * P]).contiguous()
    if isinstance(guess, CoreSpec):
        X = _orthonormalizer(ks)
        _, Cp = torch.linalg.eigh(X.T @ ks.hcore @ X)
        C = X @ Cp
        return torch.stack([C[:, :na] @ C[:, :na].T, C[:, :nb] @ C[:, :nb].T]).contiguous()
    P0 = torch.as_tensor(guess, dtype=torch.float64, device=ks.device)
    if tuple(P0.shape) != (2, n, n):
        raise ValueError(f"an open-shell guess density has shape (2, {n}, {n}), got {tuple(P0.shape)}")
    return P0.contiguous()


def _diis_coefficients(errs, count: int, m: int):
    """Pulay coefficients of the first ``count`` of m stored error vectors (rows of
    ``errs``), the restated system of scf.py:252-276 (a 1e-10 ridge on the valid
    diagonal, identity rows for the empty slots); solved on the host."""
    B = (errs @ errs.T).cpu().numpy()
    valid = np.arange(m) < count
    B = np.where(valid[:, None] & valid[None, :], B, np.eye(m)) + 1e-10 * np.diag(valid.astype(float))
    A = np.zeros((m + 1, m + 1))
    A[:m, :m] = B
    A[:m, m] = A[m, :m] = np.where(valid, -1.0, 0.0)
    rhs = np.zeros(m + 1)
    rhs[m] = -1.0
    return np.linalg.solve(A, rhs)[:m]


def _scf_uks(ks, *, max_iter, e_tol, d_tol, diis_space, lindep_thresh, level_shift, guess, verbose):
    """scf.py:279-391 with two spin channels, on the device: the channel-stacked
    Fock and error matrices share one DIIS, each channel takes the aufbau fill of
    its own nocc[s] lowest orbitals of X^T F_s X; the Fock builds are the
    library's (dfks_energy_fock with nspin = 2)."""
    torch = __import__("torch")
    S = ks.S
    X = _orthonormalizer(ks, lindep_thresh)
    n, nmo = X.shape
    nocc = ks.nocc
    m = int(diis_space)
    P = _initial_density_uks(ks, guess)
    e_prev, F = ks.energy_fock(P)
    dF = torch.zeros((m, 2 * n, n), dtype=torch.float64, device=ks.device)
    dE = torch.zeros((m, 2 * nmo, nmo), dtype=torch.float64, device=ks.device)
    converged = False
    it = 0
    times, hist = [], []
    eps = C = None
    while it < max_iter and not converged:
        t0 = time.perf_counter()
        FPS = F @ P @ S
        err = X.T @ (FPS - FPS.transpose(1, 2)) @ X            # S P F = (F P S)^T
        derr = float(torch.linalg.norm(err))
        slot = it % m
        dF[slot] = F.reshape(2 * n, n)
        dE[slot] = err.reshape(2 * nmo, nmo)
        c = torch.tensor(_diis_coefficients(dE.reshape(m, -1), min(it + 1, m), m), dtype=torch.float64,
                         device=ks.device)
        Fx = torch.einsum("i,iab->ab", c, dF).reshape(2, n, n)
        if level_shift != 0.0:
            Fx = Fx + level_shift * (S[None] - S[None] @ P @ S[None])
        eps, Cp = torch.linalg.eigh(X.T @ Fx @ X)
        C = X @ Cp                                              # (2, nao, nmo)
        P = torch.stack([C[s, :, :nocc[s]] @ C[s, :, :nocc[s]].T for s in range(2)]).contiguous()
        e, F = ks.energy_fock(P)
        de = e - e_prev
        converged = abs(de) < e_tol and derr < d_tol
        times.append(time.perf_counter() - t0)
        hist.append((e, de, derr, "fp64"))
        if verbose:
            print(f"  uks {it}: E={e:.10f} dE={de:+.2e} |[F,P]|={derr:.2e}")
        e_prev = e
        it += 1
    res = KSResult(e_tot=float(e_prev), e_elec=float(e_prev) - ks.e_nn, converged=converged, n_iter=it,
                   nocc=tuple(nocc), mo_energy=eps, mo_coeff=C, P=P, energy_converged=converged,
                   iter_times=times, history=hist)
    if not converged:
        warnings.warn(f"SCF did NOT converge in {it} iterations (e_tol={e_tol}, d_tol={d_tol}); the returned "
                      f"energy is unreliable. Increase max_iter or try level_shift>0.", stacklevel=3)
    return res


def _core_density(ks):
    """Aufbau fill of the hcore eigenvectors in the orthonormal basis."""
    torch = __import__("torch")
    s, U = torch.linalg.eigh(ks.S)
    keep = s > 1e-7
    X = U[:, keep] / torch.sqrt(s[keep])[None, :]
    _, Cp = torch.linalg.eigh(X.T @ ks.hcore @ X)
    C = X @ Cp
    Co = C[:, : ks.nocc[0]]
    return (2.0 * Co @ Co.T).contiguous()


def mixed_phase_done(derrs, mixed_d_tol: float = 1e-4, mixed_max_iter: int = 30, stall: int = 3) -> bool:
    """Whether the mixed phase ends after the steps whose orbital gradients are
    ``derrs`` (oldest first): the gradient is below ``mixed_d_tol``, the phase
    has run ``mixed_max_iter`` steps, or it has stalled at the split-BF16 floor
    -- the last ``stall`` steps all stay above 0.7 x the best gradient before
    them while that best is already below 10 x ``mixed_d_tol`` (the floor of
    the split-BF16 Fock error grows with the system: ~1e-4 for (H2O)_64)."""
    n = len(derrs)
    if n == 0:
        return False
    if derrs[-1] < mixed_d_tol or n >= mixed_max_iter:
        return True
    if stall > 0 and n > stall:
        best = min(derrs[:-stall])
        if best < 10.0 * mixed_d_tol and min(derrs[-stall:]) > 0.7 * best:
            return True
    return False


def scf(ks, *, max_iter: int = 128, e_tol: float = 1e-8, d_tol: float = 1e-6, diis_space: int = 8,
        lindep_thresh: float = 1e-7, level_shift: float = 0.0, guess=None, verbose: bool = False,
        precision: str | None = None, mixed_d_tol: float = 1e-4, mixed_max_iter: int = 30,
        mixed_stall: int = 3, refine: int | None = None) -> KSResult:
    """Run the KS SCF to self-consistency (closed shell).

    ``precision="mixed"`` (needs a KS built with ``hip(precision="mixed")``)
    iterates with split-BF16 Fock builds until the orbital gradient
    ``|err| < mixed_d_tol``, ``mixed_max_iter`` mixed steps, or a stall of
    ``mixed_stall`` steps at the split-BF16 floor (``mixed_phase_done``), then
    switches to FP64: the next step extrapolates over the mixed DIIS history and builds
    its Fock matrix in FP64, the DIIS history restarts after it, and the FP64
    steps run to the usual ``e_tol``/``d_tol``.
    The switch keys on the gradient only: the split-BF16 energy carries noise
    of order 1e-5 relative, so |dE| is not a usable test in that phase. The
    default is the precision the KS was built for, FP64 unless it was built
    mixed.

    ``refine=n`` (mixed only, n >= 2) converges the energy only: the FP64
    steps stop at the first one (from the second on) whose energy change is
    below ``e_tol`` -- ``energy_converged`` -- or after n of them. The energy
    is second order in the density error, so it lands within ~1e-9 Ha of the
    fully converged FP64 energy while the orbital gradient is left at the
    size the split-BF16 Fock error imposes (~1e-4), so ``converged`` (which
    also needs ``|err| < d_tol``) is normally False; use the default for
    densities and orbitals converged to ``d_tol``.
    """
    if getattr(ks, "nspin", 1) == 2:
        if precision not in (None, "fp64") or refine is not None:
            raise ValueError("an open-shell (nspin = 2) SCF runs in FP64")
        return _scf_uks(ks, max_iter=max_iter, e_tol=e_tol, d_tol=d_tol, diis_space=diis_space,
                        lindep_thresh=lindep_thresh, level_shift=level_shift, guess=guess, verbose=verbose)
    torch = __import__("torch")
    L = ks.coulomb.L
    h = ks.coulomb.handle
    stream = ks.coulomb._stream()
    if precision is None:
        precision = getattr(ks.coulomb, "precision", "fp64")
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
    mixed = precision == "mixed"
    if refine is not None and (not mixed or int(refine) < 2):
        raise ValueError("refine needs precision='mixed' and at least 2 FP64 steps")
    P0 = initial_density(ks, guess)
    ks.coulomb.set_precision(mixed)
    _native.check(L.dfks_scf_begin(h, ctypes.c_void_p(P0.data_ptr()), float(lindep_thresh), int(diis_space),
                                   float(level_shift), stream))
    out = (ctypes.c_double * 3)()
    converged = e_conv = False
    stop = False
    it = 0
    e = float("nan")
    times, hist = [], []
    while it < max_iter and not stop:
        t0 = time.perf_counter()
        _native.check(L.dfks_scf_step(h, out, stream))
        times.append(time.perf_counter() - t0)
        e, de, derr = out[0], out[1], out[2]
        hist.append((e, de, derr, "mixed" if mixed else "fp64"))
        if verbose:
            print(f"  scf {it} [{'mixed' if mixed else 'fp64'}]: E={e:.10f} dE={de:+.2e} |[F,P]|={derr:.2e}")
        if mixed:
            if mixed_phase_done([h[2] for h in hist if h[3] == "mixed"], mixed_d_tol, mixed_max_iter, mixed_stall):
                ks.coulomb.set_precision(False)
                mixed = False
        else:
            n_fp64 = sum(1 for h in hist if h[3] == "fp64")
            converged = abs(de) < e_tol and derr < d_tol
            if refine is not None:
                e_conv = n_fp64 >= 2 and abs(de) < e_tol
                stop = converged or e_conv or n_fp64 >= int(refine)
            else:
                e_conv = abs(de) < e_tol
                stop = converged
        it += 1
    n = ks.nao
    P = torch.empty((n, n), dtype=torch.float64, device=ks.device)
    nmo = ctypes.c_int32()
    _native.check(L.dfks_scf_result(h, None, None, None, ctypes.byref(nmo), stream))
    C = torch.empty((nmo.value, n), dtype=torch.float64, device=ks.device)   # column-major nao x nmo
    eps = torch.empty(nmo.value, dtype=torch.float64, device=ks.device)
    _native.check(L.dfks_scf_result(h, ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(C.data_ptr()),
                                    ctypes.c_void_p(eps.data_ptr()), ctypes.byref(nmo), stream))
    torch.cuda.current_stream(ks.device).synchronize()
    res = KSResult(e_tot=float(e), e_elec=float(e) - ks.e_nn, converged=converged, n_iter=it, nocc=ks.nocc,
                   mo_energy=eps[None], mo_coeff=C.T[None], P=P[None], energy_converged=e_conv or converged,
                   iter_times=times, history=hist)
    if refine is not None and not res.energy_converged:
        warnings.warn(f"SCF energy did NOT converge in {it} iterations ({refine} FP64 refinement steps, "
                      f"e_tol={e_tol}); the returned energy is unreliable.", stacklevel=2)
    elif refine is None and not converged:
        warnings.warn(f"SCF did NOT converge in {it} iterations (e_tol={e_tol}, d_tol={d_tol}); the returned "
                      f"energy is unreliable. Increase max_iter or try level_shift>0.", stacklevel=2)
    return res
