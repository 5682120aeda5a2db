"""Second-order SCF: trust-region Newton on occupied-virtual rotations
(dftax/ks/newton.py:311-379 surface, closed shell).

The energy as a function of an orbital rotation ``C(kappa) = C exp(K)``, K
antisymmetric with the occupied-virtual block kappa (no x nv), is expanded at
kappa = 0 with ``P = 2 C_o C_o^T`` and ``F = dE/dP`` (the Fock matrix):

* gradient ``g = -4 (C_o^T F C_v)`` -- jax.grad of the energy in the reference;
* Hessian-vector product (the reference's ``jax.jvp(jax.grad(E))``,
  newton.py:134-136) ``H v = 4 (v F_vv - F_oo v) - 4 C_o^T R C_v``, where
  ``F_oo = C_o^T F C_o``, ``F_vv = C_v^T F C_v`` and ``R`` is the response
  Fock matrix ``J[dP] - a_x/2 K[dP] + f_xc(P)[dP]`` along
  ``dP = -2 (C_o v C_v^T + C_v v^T C_o^T)``, computed on the GPU by
  ``dfks_response_fock`` with ``U = C_o`` and ``V = -2 C_v v^T``.

The outer loop is the reference's: Steihaug-Toint truncated CG for the step
(newton.py:44-94), acceptance on energy decrease, trust radius doubled (capped
at 1) on success and halved otherwise, convergence on ``|g| < g_tol`` and
``|dE| < e_tol``. The CG runs ``cg_iters`` iterations and stops early only at
the trust boundary, as the reference's frozen fixed-length loop does.
The Fock matrix and energy of an accepted point are kept for the next
iteration instead of being rebuilt (same values: the build is deterministic).
"""

from __future__ import annotations

import math
import time
import warnings

import numpy as np

from dftax_amd.scf import KSResult, initial_density


def _steihaug_cg(hvp, g, radius: float, maxiter: int):
    """Truncated CG for min <g,p> + 1/2 <p,Hp> over ||p|| <= radius
    (dftax/ks/newton.py:44-94): follows negative curvature or a step that
    would cross the boundary to the boundary, and then stops."""
    torch = __import__("torch")

    def dot(a, b):
        return float(torch.sum(a * b))

    def to_boundary(z, d):
        dd, zd, zz = dot(d, d), dot(z, d), dot(z, z)
        disc = math.sqrt(max(zd * zd + dd * (radius * radius - zz), 0.0))
        tau = (-zd + disc) / max(dd, 1e-30)
        return z + tau * d

    z = torch.zeros_like(g)
    r = g.clone()
    d = -g
    rr = dot(g, g)
    p = z
    n_hvp = 0
    for _ in range(maxiter):
        Hd = hvp(d)
        n_hvp += 1
        dHd = dot(d, Hd)
        alpha = rr / (1.0 if dHd == 0.0 else dHd)
        z_next = z + alpha * d
        hit = dHd <= 0.0 or dot(z_next, z_next) >= radius * radius
        if hit:
            return to_boundary(z, d), n_hvp
        p = z_next
        r = r + alpha * Hd
        rr_next = dot(r, r)
        beta = rr_next / (1.0 if rr == 0.0 else rr)
        d = -r + beta * d
        z, rr = z_next, rr_next
    return p, n_hvp


def _rotate(C, kappa, nocc: int):
    """C exp(K) with K[:no, no:] = kappa, K[no:, :no] = -kappa^T (newton.py:97-106)."""
    torch = __import__("torch")
    nmo = C.shape[1]
    K = torch.zeros((nmo, nmo), dtype=C.dtype, device=C.device)
    K[:nocc, nocc:] = kappa
    K[nocc:, :nocc] = -kappa.T
    return C @ torch.linalg.matrix_exp(K)


def newton(ks, *, max_iter: int = 64, g_tol: float = 1e-6, e_tol: float = 1e-10, cg_iters: int = 32,
           trust: float = 0.4, guess=None, verbose: bool = False, lindep_thresh: float = 1e-7) -> KSResult:
    """Second-order SCF (trust-region Newton on orbital rotations), FP64.

    Same arguments as dftax's ``newton`` (newton.py:311-379): the reference
    orbitals are the aufbau eigenvectors of the Fock matrix at the guess
    density, and each Hessian-vector product costs one response-Fock build
    (J and K of a rank-2 no factor plus the XC second-derivative kernel).
    """
    torch = __import__("torch")
    c = ks.coulomb
    if getattr(c, "precision", "fp64") == "mixed":
        c.set_precision(False)
    no = ks.nocc[0]
    S = ks.S
    s, U = torch.linalg.eigh(S)
    keep = s > lindep_thresh
    X = U[:, keep] / torch.sqrt(s[keep])[None, :]          # canonical orthonormaliser (scf.py:72-86)
    P0 = initial_density(ks, guess)
    _, F0 = ks.energy_fock(P0)
    _, Cp = torch.linalg.eigh(X.T @ F0 @ X)
    C = X @ Cp
    nmo = C.shape[1]
    if no >= nmo:
        raise ValueError(f"no virtual orbitals to rotate into (nocc={no}, nmo={nmo})")

    def state_at(Cm):
        E, F, P = c.energy_fock_factored(math.sqrt(2.0) * Cm[:, :no])
        return float(E[0]), F, P

    e_cur, F_cur, P_cur = state_at(C)
    radius = float(trust)
    e_prev = math.inf
    converged = False
    it = 0
    hist, times = [], []
    while it < max_iter and not converged:
        t0 = time.perf_counter()
        e0 = e_cur
        Co, Cv = C[:, :no], C[:, no:]
        Ft = C.T @ F_cur @ C
        Foo, Fvv = Ft[:no, :no], Ft[no:, no:]
        g = -4.0 * Ft[:no, no:]
        gnorm = float(torch.linalg.norm(g))
        c.response_begin(P_cur)

        def hvp(v):
            R = c.response_fock(Co, -2.0 * (Cv @ v.T))
            return 4.0 * (v @ Fvv - Foo @ v) - 4.0 * (Co.T @ R @ Cv)

        delta, n_hvp = _steihaug_cg(hvp, g, radius, cg_iters)
        C_new = _rotate(C, delta, no)
        e_new, F_new, P_new = state_at(C_new)
        accept = e_new < e0
        if accept:
            C, e_cur, F_cur, P_cur = C_new, e_new, F_new, P_new
        radius = min(2.0 * radius, 1.0) if accept else 0.5 * radius
        e = e_new if accept else e0
        converged = gnorm < g_tol and abs(e - e_prev) < e_tol
        times.append(time.perf_counter() - t0)
        hist.append((e, e - e_prev, gnorm, radius, bool(accept), n_hvp))
        if verbose:
            print(f"  newton {it}: E={e:.10f} |g|={gnorm:.2e} r={radius:.2e} ok={accept} hvp={n_hvp}")
        e_prev = e
        it += 1
    # canonical orbital energies at the final density (report only)
    eps, Cp = torch.linalg.eigh(X.T @ F_cur @ X)
    torch.cuda.current_stream(ks.device).synchronize()
    res = KSResult(e_tot=float(e_cur), e_elec=float(e_cur) - ks.e_nn - ks.e_disp, converged=converged, n_iter=it,
                   nocc=ks.nocc, mo_energy=eps[None], mo_coeff=(X @ Cp)[None], P=P_cur[None], iter_times=times,
                   history=hist)
    if not converged:
        warnings.warn(f"Newton SCF did NOT converge in {it} iterations (g_tol={g_tol}, e_tol={e_tol}); the "
                      f"returned energy is unreliable. Increase max_iter or loosen the trust radius.", stacklevel=2)
    return res
