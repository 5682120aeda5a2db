"""The KS energy functional with the MI355X density-fitting engine.

Mirrors dftax's builder surface (dftax/ks/energy.py:439-723,
dftax/ks/terms.py:63-148): ``KS(mol, xc, grid=becke(...), coulomb=df(aux, ...))``
owns ``S``, ``hcore``, ``e_nn`` and the Coulomb + XC terms, and exposes
``electronic(P)`` / ``total(P)``. The one addition is the engine value on the
DF spec, ``df(aux, engine=hip(...))`` (the default here), which resolves the
Coulomb term to :class:`HipDFKS`: one native handle that evaluates E(P) and
F = sym(dE/dP) (J, -a_x/2 K and Vxc together) on the GPU, with the same
``energy(P, S, nocc)`` contract as dftax's ``CoulombTerm`` (terms.py:512-527).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from dftax_amd import _native
from dftax_amd.basis import BasisSet, build_basis
from dftax_amd.grid import Becke, atom_bragg, atom_centred_points, becke
from dftax_amd.system import Molecule
from dftax_amd.xc import XCFunctional, resolve


def _torch():
    import torch
    return torch


@dataclass(frozen=True)
class HipSpec:
    """Engine value: run the DF-KS term through libdfks on the current GPU.

    ``precision="fp64"`` computes everything in FP64; ``"mixed"`` also keeps
    FP32 copies of B and the AO grid so the SCF can run its early iterations
    with split-BF16 MFMA GEMMs (FP32 accumulate) and refine in FP64.
    ``gemm`` picks the FP64 GEMM kernels: ``"auto"`` (library symmetric
    exchange kernel; for T = B^T C and the compact-grid XC products the
    fastest of rocBLAS and the library kernels, timed once per shape),
    ``"rocblas"``, ``"mfma"`` (library kernels for all), ``"mfma8"`` (as
    ``"mfma"``, with the 8-wavefront 128 x 160 tile for T) or ``"mfma8d"`` (as
    ``"mfma8"``, with each product's further variant where it has one: the
    split-BF16 T with two k-steps of operands in flight, the FP64 K split by
    XCD). Mixed steps use the library's
    split-BF16 kernels for the exchange products and FP32 rocBLAS products on
    the compact XC grid. ``metric_cutoff`` is the relative eigenvalue cutoff of the
    metric pseudo-inverse (terms.py:190-192). ``density`` is how each SCF step
    gets the aufbau density from the orthonormal-basis Fock matrix:
    ``"purify"`` (SP2 spectral projection with FP64 MFMA products and a
    pivoted-Cholesky occupied factor; the eigensolver is used only when the
    projection does not converge, i.e. without a gap at the Fermi level) or
    ``"eigh"`` (an FP64 eigensolve every step, as scf.py:363-367 does).
    ``ao_block``: ``None`` keeps the AO grid resident when it fits the HBM
    budget and otherwise recomputes it per block of points in every Fock
    build; ``0`` forces resident, an int streams with that block.
    ``xc_grid``: ``"blocks"`` evaluates the XC term on blocks of 128
    spatially sorted grid points, each holding only the AOs whose value or
    gradient bound reaches 1e-14 there (resident compact rows, batched
    per-block products; ``ao_block`` is then unused); ``"dense"`` keeps full
    AO rows per point.
    """

    precision: str = "fp64"
    metric_cutoff: float = 1e-7
    rho_thresh: float = 1e-10
    keep_int3c: bool = False
    gemm: str = "auto"
    density: str = "purify"
    ao_block: int | None = None
    xc_grid: str = "blocks"


def hip(*, precision: str = "fp64", metric_cutoff: float = 1e-7, rho_thresh: float = 1e-10,
        keep_int3c: bool = False, gemm: str = "auto", density: str = "purify",
        ao_block: int | None = None, xc_grid: str = "blocks") -> HipSpec:
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
    if gemm not in ("auto", "rocblas", "mfma", "mfma8", "mfma8d"):
        raise ValueError(f"gemm must be 'auto', 'rocblas', 'mfma', 'mfma8' or 'mfma8d', got {gemm!r}")
    if density not in ("purify", "eigh"):
        raise ValueError(f"density must be 'purify' or 'eigh', got {density!r}")
    if ao_block is not None and (not isinstance(ao_block, int) or ao_block < 0):
        raise ValueError(f"ao_block must be None or a non-negative int, got {ao_block!r}")
    if xc_grid not in ("blocks", "dense"):
        raise ValueError(f"xc_grid must be 'blocks' or 'dense', got {xc_grid!r}")
    return HipSpec(precision=precision, metric_cutoff=metric_cutoff, rho_thresh=rho_thresh, keep_int3c=keep_int3c,
                   gemm=gemm, density=density, ao_block=ao_block, xc_grid=xc_grid)


@dataclass(frozen=True)
class DFSpec:
    """Density-fitting spec (dftax/ks/terms.py:63-76) plus the engine value."""

    auxbasis: str = "even-tempered-jk"
    chunk: object = "auto"
    screen: float | None = None
    spherical: bool | None = None
    engine: HipSpec = field(default_factory=HipSpec)


def df(auxbasis: str = "even-tempered-jk", *, chunk="auto", screen: float | None = None,
       spherical: bool | None = None, engine: HipSpec | None = None) -> DFSpec:
    """Density-fitted Coulomb + exchange (terms.py:96-148).

    ``chunk``: ``"auto"`` materialises B = (mn|P) L when it fits the HBM
    budget and otherwise recomputes (mn|P) in auxiliary slabs in every Fock
    build; ``None`` forces materialised; an int recomputes with slabs of about
    that many auxiliary functions (energy.py:180-212). ``screen``: relative
    Schwarz threshold on bra shell pairs, on both paths (pairs with
    max sqrt((mn|mn)) below screen times the largest are skipped and stay
    zero). ``spherical`` False is rejected: the auxiliary span is spherical
    on every path of this engine.
    """
    if spherical is False:
        raise NotImplementedError("the HIP engine fits in the spherical auxiliary span")
    if chunk is not None and chunk != "auto" and (not isinstance(chunk, int) or isinstance(chunk, bool) or chunk <= 0):
        raise ValueError(f"chunk must be 'auto', None or a positive int, got {chunk!r}")
    if screen is not None and not (0.0 <= float(screen) < 1.0):
        raise ValueError(f"screen must be in [0, 1), got {screen!r}")
    return DFSpec(auxbasis=auxbasis, chunk=chunk, screen=screen, spherical=spherical,
                  engine=engine if engine is not None else HipSpec())


@dataclass(frozen=True)
class MemoryPlan:
    """How one handle lays out its data in HBM."""

    aux_chunk: int      # 0: B materialised; > 0: auxiliary slab size of the recompute path
    ao_block: int       # 0: AO grid resident; > 0: points per recomputed block
    bytes_est: int      # estimated device bytes at the peak


def memory_plan(nao: int, naux: int, ngrid: int, nocc: int, precision: str, chunk, ao_block,
                budget: int, tensors: int = 1) -> MemoryPlan:
    """Pick materialised vs recomputed DF and a resident vs streamed AO grid.

    Word counts (FP64, the mixed path adds FP32 copies): B is nao^2 naux and
    the build holds (mn|P) and B together (a range-separated hybrid,
    ``tensors = 2``, keeps its long-range B beside it and is materialised
    only); the recompute path holds a slab
    of nao^2 x chunk, the raw and fitted exchange intermediates (2 nao naux
    nf) and the metric factor; the AO grid is 4 ng nao values plus ng nao for
    Z/W. ``budget`` is the device bytes the handle may use.
    """
    w = 8
    f32 = 1.5 if precision == "mixed" else 1.0
    nf = max(1, min(nao, 512, max(nocc, 1)))
    exch = nao * naux * nf * w * f32 + 64 * nao * nao * w
    mat = max((1 + tensors) * nao * nao * naux * w, tensors * nao * nao * naux * w * f32 + exch)
    if chunk is None or tensors > 1:
        aux_chunk = 0
    elif isinstance(chunk, int):
        aux_chunk = chunk
    else:
        aux_chunk = 0 if mat + 5 * ngrid * nao * w * f32 <= budget else -1
    if aux_chunk == -1 or aux_chunk > 0:
        if aux_chunk == -1:
            # slab of about 1/8 of the budget, 32 .. 2048 functions
            aux_chunk = int(min(2048, max(32, budget // 8 // (nao * nao * w))))
        df_bytes = nao * nao * aux_chunk * w + 2 * nao * naux * nf * w + naux * naux * w + 64 * nao * nao * w
        if precision == "mixed":
            df_bytes += nao * nao * aux_chunk * 4 + naux * naux * 4   # FP32 slab and metric factor
    else:
        df_bytes = mat
    grid = 5 * ngrid * nao * w * f32
    if ao_block is None:
        # resident if it fits; else blocks sized to half the memory the DF part leaves
        spare = max(budget - df_bytes, 0) // 2
        ao_block = 0 if df_bytes + grid <= budget else int(max(4096, min(ngrid, spare // (5 * nao * w * f32))))
    grid_bytes = grid if ao_block == 0 else 5 * min(ao_block, ngrid) * nao * w * f32
    return MemoryPlan(aux_chunk=aux_chunk, ao_block=ao_block, bytes_est=int(df_bytes + grid_bytes))


def build_grid(symbols, coords, spec: Becke):
    """Atom-centred points, GPU Becke weights, then the weight cutoff."""
    pts, raw, owner = atom_centred_points(symbols, coords, spec)
    w = _native.becke_weights(pts, raw, owner, coords, atom_bragg(symbols))
    if spec.cutoff is not None:
        keep = np.abs(w) > spec.cutoff
        pts, w = pts[keep], w[keep]
    return np.ascontiguousarray(pts), np.ascontiguousarray(w)


class HipDFKS:
    """Native DF Coulomb/exchange + grid XC term (one libdfks handle)."""

    def __init__(self, basis: BasisSet, aux: BasisSet, mol: Molecule, grid_xyz, grid_w, xc: XCFunctional,
                 nelec: int, spec: HipSpec, device=None, chunk="auto", screen: float | None = None,
                 nspin: int = 1, spin: int = 0):
        torch = _torch()
        self.L = _native.lib()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self._arrays = (_native.BasisArrays(basis), _native.BasisArrays(aux))
        self._xyz = np.ascontiguousarray(mol.coords, np.float64)
        self._z = np.ascontiguousarray(mol.atom_charges(), np.float64)
        self._gxyz = np.ascontiguousarray(grid_xyz, np.float64)
        self._gw = np.ascontiguousarray(grid_w, np.float64)
        desc = _native.SystemDesc(self._arrays[0].desc, self._arrays[1].desc, len(self._z), _native.f64p(self._xyz),
                                  _native.f64p(self._z), self._gw.shape[0], _native.f64p(self._gxyz),
                                  _native.f64p(self._gw), xc.xc_id, int(nelec), int(nspin), int(spin))
        self.nspin = int(nspin)
        free = _native.device_free(self.device)
        self.plan = memory_plan(basis.nao, aux.nao, self._gw.shape[0], nelec // 2, spec.precision, chunk,
                                spec.ao_block, int(0.85 * free), 2 if xc.hf_coeff_lr else 1)
        opts = _native.Options(1 if spec.precision == "mixed" else 0, float(spec.metric_cutoff),
                               float(spec.rho_thresh), int(spec.keep_int3c),
                               {"auto": 0, "rocblas": 1, "mfma": 2, "mfma8": 3, "mfma8d": 4}[spec.gemm],
                               {"purify": 0, "eigh": 1}[spec.density], int(self.plan.aux_chunk),
                               float(screen or 0.0), int(self.plan.ao_block), int(spec.xc_grid == "blocks"))
        self.precision = spec.precision
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _native.check(self.L.dfks_create(ctypes.byref(desc), ctypes.byref(opts), ctypes.byref(h)))
            self.handle = h
            _native.check(self.L.dfks_build(self.handle, self._stream()))
        sizes = (ctypes.c_int64 * 5)()
        _native.check(self.L.dfks_sizes(self.handle, sizes))
        self.nao, self.naux, self.nfit, self.nocc, self.ngrid = (int(v) for v in sizes)
        self.hf_coeff = xc.hf_coeff
        self._E = torch.zeros(8, dtype=torch.float64, device=self.device)

    def _stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    def close(self) -> None:
        if getattr(self, "handle", None) is not None:
            self.L.dfks_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def matrix(self, which: int, shape):
        torch = _torch()
        out = torch.empty(shape, dtype=torch.float64, device=self.device)
        _native.check(self.L.dfks_get_matrix(self.handle, which, ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def stats(self) -> dict:
        """Build and solver counters (kept bra pairs, slabs, AO block, eigensolver fallbacks)."""
        return _native.stats(self.handle)

    def set_precision(self, mixed: bool) -> None:
        _native.check(self.L.dfks_set_precision(self.handle, int(bool(mixed))))

    def energy_fock(self, P):
        """(E components [8], F) for a device density P: nao x nao (nspin = 1), or
        2 x nao x nao (alpha, beta) for an nspin = 2 handle, F of the same shape."""
        torch = _torch()
        P = P.to(device=self.device, dtype=torch.float64).contiguous()
        nsp = int(P.shape[0]) if P.dim() == 3 else 1
        F = torch.empty_like(P)
        _native.check(self.L.dfks_energy_fock(self.handle, ctypes.c_void_p(P.data_ptr()), nsp,
                                              ctypes.c_void_p(self._E.data_ptr()), ctypes.c_void_p(F.data_ptr()),
                                              self._stream()))
        return self._E.clone(), F

    def energy_fock_factored(self, Cf):
        """(E components [8], F, P) at P = Cf Cf^T for a device factor Cf (nao x nf): the
        exchange runs over the nf factor columns, with no eigen-factorisation of P."""
        torch = _torch()
        Cf = Cf.to(device=self.device, dtype=torch.float64)
        P = (Cf @ Cf.T).contiguous()
        cm = Cf.T.contiguous()                       # column-major nao x nf
        nf = int(cm.shape[0])
        signs = (ctypes.c_double * nf)(*([1.0] * nf))
        F = torch.empty_like(P)
        _native.check(self.L.dfks_energy_fock_factored(self.handle, ctypes.c_void_p(P.data_ptr()),
                                                       ctypes.c_void_p(cm.data_ptr()), signs, nf,
                                                       ctypes.c_void_p(self._E.data_ptr()),
                                                       ctypes.c_void_p(F.data_ptr()), self._stream()))
        return self._E.clone(), F, P

    def response_begin(self, P) -> None:
        """Fix the reference density P0 of :meth:`response_fock` (device nao x nao)."""
        torch = _torch()
        self._P0 = P.to(device=self.device, dtype=torch.float64).contiguous()
        _native.check(self.L.dfks_response_begin(self.handle, ctypes.c_void_p(self._P0.data_ptr()),
                                                 self._stream()))

    def response_fock(self, U, V):
        """R = J[dP] - a_x/2 K[dP] + f_xc(P0)[dP] for dP = U V^T + V U^T (U, V: nao x r),
        the derivative of F = dE/dP at P0 along dP."""
        torch = _torch()
        Uc = U.to(device=self.device, dtype=torch.float64).T.contiguous()   # column-major nao x r
        Vc = V.to(device=self.device, dtype=torch.float64).T.contiguous()
        if Uc.shape != Vc.shape or Uc.shape[1] != self.nao:
            raise ValueError(f"U and V must both be ({self.nao}, r); got {tuple(U.shape)} and {tuple(V.shape)}")
        R = torch.empty((self.nao, self.nao), dtype=torch.float64, device=self.device)
        _native.check(self.L.dfks_response_fock(self.handle, ctypes.c_void_p(Uc.data_ptr()),
                                                ctypes.c_void_p(Vc.data_ptr()), int(Uc.shape[0]),
                                                ctypes.c_void_p(R.data_ptr()), self._stream()))
        return R

    def energy(self, P, S=None, nocc=None):
        """dftax's CoulombTerm contract: the J + K + XC energy of a spin-stacked P."""
        E, _ = self.energy_fock(P[0] if P.dim() == 3 and self.nspin == 1 else P)
        return float(E[2] + E[3] + E[4])


class KS:
    """KS energy functional on the MI355X engine (energy.py:439-723).

    ``spin=None`` follows the molecule: restricted (one doubly occupied channel)
    when ``system.spin == 0``, else spin-unrestricted with two channels;
    ``spin=<int>`` always builds the two channels (n_alpha - n_beta = spin), as
    dftax's KS does (energy.py, the spin dispatch of test_uks.py)."""

    def __init__(self, system: Molecule, xc, *, grid: Becke | None = None, coulomb: DFSpec | None = None,
                 spin: int | None = None, device=None):
        torch = _torch()
        nelec = system.nelectron
        nspin = 1 if spin is None and system.spin == 0 else 2
        spin_v = system.spin if spin is None else int(spin)
        if nspin == 1 and nelec % 2:
            raise ValueError(f"nelec={nelec} is odd but the system has spin=0")
        if nspin == 2 and (spin_v < 0 or spin_v > nelec or (nelec + spin_v) % 2):
            raise ValueError(f"inconsistent (nelec={nelec}, spin={spin_v})")
        self.xc = resolve(xc)
        spec = coulomb if coulomb is not None else df()
        if not isinstance(spec, DFSpec):
            raise TypeError("the HIP engine is a df(...) Coulomb backend")
        self.mol = system
        self.symbols = tuple(system.symbols)
        self.atom_coords = np.asarray(system.coords)
        self.nelec = nelec
        self.nspin = nspin
        self.spin = spin_v
        self.nocc = (nelec // 2,) if nspin == 1 else ((nelec + spin_v) // 2, (nelec - spin_v) // 2)
        self.basis = build_basis(system.symbols, system.coords, system.basis, spherical=system.spherical)
        self.aux_basis = build_basis(system.symbols, system.coords, spec.auxbasis, spherical=True)
        gspec = grid if grid is not None else becke()
        self.grid_spec = gspec
        self.grid_xyz, self.grid_w = build_grid(system.symbols, system.coords, gspec)
        self.coulomb = HipDFKS(self.basis, self.aux_basis, system, self.grid_xyz, self.grid_w, self.xc, nelec,
                               spec.engine, device=device, chunk=spec.chunk, screen=spec.screen, nspin=nspin,
                               spin=spin_v)
        self.xc_term = None  # XC is evaluated inside the native term
        n = self.coulomb.nao
        self.S = self.coulomb.matrix(_native.DFKS_S, (n, n))
        self.hcore = self.coulomb.matrix(_native.DFKS_HCORE, (n, n))
        self.e_nn = float(_nuclear_repulsion(system.coords, system.atom_charges()))
        self.e_disp = 0.0
        self.device = self.coulomb.device

    @property
    def nao(self) -> int:
        return self.coulomb.nao

    def _stacked(self, P):
        if self.nspin == 2:
            if P.dim() != 3 or P.shape[0] != 2:
                raise ValueError(f"an open-shell KS takes P of shape (2, nao, nao), got {tuple(P.shape)}")
            return P
        return P[0] if P.dim() == 3 else P

    def energy_fock(self, P):
        """(E_total, F) at a density P: nao x nao or 1 x nao x nao (restricted),
        2 x nao x nao (alpha, beta; F likewise per channel)."""
        E, F = self.coulomb.energy_fock(self._stacked(P))
        return float(E[0]), F

    def electronic(self, P) -> float:
        E, _ = self.coulomb.energy_fock(self._stacked(P))
        return float(E[1] + E[2] + E[3] + E[4])

    def total(self, P) -> float:
        return self.electronic(P) + self.e_nn + self.e_disp


def _nuclear_repulsion(xyz, z) -> float:
    xyz = np.asarray(xyz)
    z = np.asarray(z)
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    iu = np.triu_indices(len(z), 1)
    return float(np.sum(z[iu[0]] * z[iu[1]] / d[iu]))
