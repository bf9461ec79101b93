"""Python handle over one libmpol context (device memory through torch-ROCm).

torch provides device allocation and the stream; every computation of the
operator and the solve runs in the HIP kernels of libmpol.so.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

FP64 = _lib.MPOL_FP64
FP32 = _lib.MPOL_FP32


def require_device() -> torch.device:
    """The GPU the library runs on; raises if there is none (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("mutual-polarization operator needs a ROCm GPU (torch.cuda is unavailable); "
                           "there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def real_space_table(r, xi):
    """The six real-space coupling tables (self term first) from the library's
    FP64 host routine; same order as the reference's real_space_table."""
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    out = np.zeros((6, r.size + 1), dtype=np.float64)
    lib = _lib.load()
    _lib.check(lib.mpol_real_space_table(float(xi), _lib.dptr(r), r.size, _lib.dptr(out)))
    return tuple(out[k] for k in range(6))


def real_space_table_field(r, xi):
    """The six couplings of real_space_table without the overlap terms of
    r < 2 (the field at a point from a sphere's dipole; field-map tables)."""
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    out = np.zeros((6, r.size + 1), dtype=np.float64)
    lib = _lib.load()
    _lib.check(lib.mpol_real_space_table_field(float(xi), _lib.dptr(r), r.size, _lib.dptr(out)))
    return tuple(out[k] for k in range(6))


def real_space_table_quad(r, xi):
    """The eight quadrupole couplings (self term first) in the order
    pot_quad, field_quad_1..3, grad_quad_1..4, from the library's FP64 host
    routine."""
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    out = np.zeros((8, r.size + 1), dtype=np.float64)
    lib = _lib.load()
    _lib.check(lib.mpol_real_space_table_quad(float(xi), _lib.dptr(r), r.size, _lib.dptr(out)))
    return tuple(out[k] for k in range(8))


def real_space_table_quad_field(r, xi):
    """The eight quadrupole couplings of real_space_table_quad without the
    overlap terms of r < 2 (field-map tables), self term first."""
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    out = np.zeros((8, r.size + 1), dtype=np.float64)
    lib = _lib.load()
    _lib.check(lib.mpol_real_space_table_quad_field(float(xi), _lib.dptr(r), r.size, _lib.dptr(out)))
    return tuple(out[k] for k in range(8))


def real_space_table_sizes_field(r, radii, xi):
    """(perp, para), each (len(r), M(M+1)/2): the pair couplings of
    real_space_table_sizes without the overlap terms and without an r -> 0
    row (every r > 0), for field maps."""
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    a = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).reshape(-1))
    m = a.size * (a.size + 1) // 2
    perp = np.zeros((m, r.size), dtype=np.float64)
    para = np.zeros((m, r.size), dtype=np.float64)
    lib = _lib.load()
    _lib.check(lib.mpol_real_space_table_sizes_field(float(xi), _lib.dptr(r), r.size, _lib.dptr(a), a.size,
                                                     _lib.dptr(perp), _lib.dptr(para)))
    return perp.T.copy(), para.T.copy()


def real_space_table_sizes(r, radii, xi):
    """(perp, para), each (len(r), M(M+1)/2) for the M radii: the I - rr and
    rr pair couplings of every radius pair (i <= j), row 0 the r -> 0 value."""
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    a = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).reshape(-1))
    m = a.size * (a.size + 1) // 2
    perp = np.zeros((m, r.size), dtype=np.float64)
    para = np.zeros((m, r.size), dtype=np.float64)
    lib = _lib.load()
    _lib.check(lib.mpol_real_space_table_sizes(float(xi), _lib.dptr(r), r.size, _lib.dptr(a), a.size,
                                               _lib.dptr(perp), _lib.dptr(para)))
    return perp.T.copy(), para.T.copy()


_TABLES: dict = {}


def default_tables(xi):
    """field_dip_1, field_dip_2 and r on [0, 0.001:0.001:10]
    ("CPU - Ewald test/capacitance.py":469-477), cached per xi."""
    key = float(xi)
    if key not in _TABLES:
        r = np.arange(0.001, 10.0 + 1e-12, 0.001, dtype=np.float64)
        tabs = real_space_table(r, xi)
        _TABLES[key] = (tabs[2], tabs[3], np.concatenate(([0.0], r)))
    return _TABLES[key]


_FTABLES: dict = {}


def default_force_tables(xi):
    """field_dip_force_1, field_dip_force_2 on the rvals of default_tables
    (the pair-force couplings dA/dr and d(B - A)/dr, self entries 0)."""
    key = float(xi)
    if key not in _FTABLES:
        r = np.arange(0.001, 10.0 + 1e-12, 0.001, dtype=np.float64)
        tabs = real_space_table(r, xi)
        _FTABLES[key] = (tabs[4], tabs[5])
    return _FTABLES[key]


@dataclass
class SolveResult:
    x: torch.Tensor
    matvecs_fp64: int
    matvecs_fp32: int
    converged: int
    refinements: int
    max_relres: float
    matvecs_inner: int = 0  # of matvecs_fp32, on the inner operator (set_inner_operator)

    @property
    def flag(self) -> int:
        return 0 if self.converged == self.x.shape[0] else 1


class Context:
    """One geometry (positions, grid, tables) on one GPU."""

    def __init__(self, device: torch.device | None = None):
        self.device = device if device is not None else require_device()
        self._lib = _lib.load()
        self._ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _lib.check(self._lib.mpol_create(ctypes.byref(self._ctx), self.device.index,
                                             ctypes.c_void_p(stream)))
        self.n = 0
        self.nsets = 0
        self._ntab = 0
        self.complex = False
        self._keep = []
        self._inner = None

    # -- lifetime -------------------------------------------------------
    def close(self):
        if self._ctx:
            self._lib.mpol_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sync_stream(self):
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.mpol_set_stream(self._ctx, ctypes.c_void_p(stream)))

    def _dev(self, a, dtype=torch.float64):
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.device, dtype=dtype)
        else:
            t = torch.as_tensor(np.asarray(a), dtype=dtype, device=self.device)
        return t.contiguous()

    # -- setup ----------------------------------------------------------
    def set_tables(self, perp, para, rvals):
        perp = np.ascontiguousarray(np.real(np.asarray(perp)), dtype=np.float64).reshape(-1)
        para = np.ascontiguousarray(np.real(np.asarray(para)), dtype=np.float64).reshape(-1)
        rvals = np.ascontiguousarray(np.asarray(rvals), dtype=np.float64).reshape(-1)
        if not (perp.size == para.size == rvals.size):
            raise ValueError("perp, para and rvals must have the same length")
        self._sync_stream()
        _lib.check(self._lib.mpol_set_tables(self._ctx, _lib.dptr(perp), _lib.dptr(para), _lib.dptr(rvals),
                                             rvals.size))
        self._ntab = rvals.size

    def set_geometry(self, x, box, ngrid, P, xi, eta, rc):
        xt = self._dev(x).reshape(-1, 3)
        box = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(3))
        ng = np.ascontiguousarray(np.asarray(ngrid, dtype=np.int64).reshape(3).astype(np.int32))
        eta = np.ascontiguousarray(np.asarray(eta, dtype=np.float64).reshape(3))
        self._sync_stream()
        _lib.check(self._lib.mpol_set_geometry(self._ctx, ctypes.c_void_p(xt.data_ptr()), xt.shape[0],
                                               _lib.dptr(box), _lib.i32ptr(ng), int(P), float(xi),
                                               _lib.dptr(eta), float(rc)))
        self.n = xt.shape[0]
        self.nsets = 0

    def set_permittivity(self, eps):
        """Permittivities (nsets, N) or (N,).  Complex values switch the
        context to complex right-hand sides (every operator input and output
        is then complex)."""
        if isinstance(eps, torch.Tensor):
            cplx = eps.is_complex()
        else:
            cplx = np.iscomplexobj(np.asarray(eps))
        if cplx:
            ec = eps.to(self.device) if isinstance(eps, torch.Tensor) else torch.as_tensor(
                np.asarray(eps, dtype=np.complex128), device=self.device)
            ec = ec.to(torch.complex128)
            if ec.dim() == 1:
                ec = ec.reshape(1, -1)
            if ec.shape[1] != self.n:
                raise ValueError("permittivity must have N entries per set")
            er = ec.real.contiguous()
            ei = ec.imag.contiguous()
            self._sync_stream()
            _lib.check(self._lib.mpol_set_permittivity_complex(self._ctx, ctypes.c_void_p(er.data_ptr()),
                                                               ctypes.c_void_p(ei.data_ptr()), er.shape[0]))
            self.nsets = er.shape[0]
            self.complex = True
            self._keep = [er, ei]
            return
        et = self._dev(eps)
        if et.dim() == 1:
            et = et.reshape(1, -1)
        if et.shape[1] != self.n:
            raise ValueError("permittivity must have N entries per set")
        self._sync_stream()
        _lib.check(self._lib.mpol_set_permittivity(self._ctx, ctypes.c_void_p(et.data_ptr()), et.shape[0]))
        self.nsets = et.shape[0]
        self.complex = False
        self._keep = [et]

    def set_quad_tables(self, qtabs):
        """The seven quadrupole tables (7, ntab) on the rvals of set_tables:
        field_quad_1..3, grad_quad_1..4, entry 0 the self term."""
        q = np.ascontiguousarray(np.real(np.asarray(qtabs)), dtype=np.float64)
        if q.ndim != 2 or q.shape[0] != 7:
            raise ValueError("quadrupole tables must have shape (7, ntab)")
        _lib.check(self._lib.mpol_set_quad_tables(self._ctx, _lib.dptr(q), q.shape[1]))

    TUNING_KNOBS = ("tile_ty", "tile_tz", "chunk", "y_cols", "z_cols", "z_threads", "engine", "rhs_group", "z_rank1",
                    "cplx_nyquist", "fft_order")

    def set_tuning(self, precision=None, **knobs):
        """Override launch shapes (0 = automatic); see TUNING_KNOBS.  precision
        (FP32 or FP64) restricts them to the operator in that precision."""
        unknown = set(knobs) - set(self.TUNING_KNOBS)
        if unknown:
            raise ValueError(f"unknown tuning knobs {sorted(unknown)}")
        vals = np.array([int(knobs.get(k, 0)) for k in self.TUNING_KNOBS], dtype=np.int32)
        prec = -1 if precision is None else int(precision)
        _lib.check(self._lib.mpol_set_tuning_for(self._ctx, prec, _lib.i32ptr(vals), vals.size))

    def geometry_info(self) -> dict:
        s = _lib.GeometrySummary()
        _lib.check(self._lib.mpol_geometry_info(self._ctx, ctypes.byref(s)))
        return {"n": s.n, "npairs": s.npairs, "ngrid": tuple(s.ngrid), "P": s.P, "ncell": tuple(s.ncell),
                "tile_rows": s.tile_rows, "fft_e": tuple(s.fft_e)}

    # -- operator -------------------------------------------------------
    def _as_rhs(self, m, nc=3):
        """Device FP64 right-hand sides (nrhs, N, nc); with complex
        permittivities packed as (nrhs, 2, N, nc) = (real part, imaginary part)."""
        if self.complex:
            if isinstance(m, torch.Tensor):
                mt = m.to(self.device)
                mt = mt.to(torch.complex128) if mt.is_complex() else mt.to(torch.float64).to(torch.complex128)
            else:
                mt = torch.as_tensor(np.asarray(m, dtype=np.complex128), device=self.device)
            if mt.dim() == 2:
                mt = mt.reshape(1, -1, nc)
            if mt.shape[1] != self.n or mt.shape[2] != nc:
                raise ValueError(f"expected moments of shape (nrhs, {self.n}, {nc})")
            return torch.stack((mt.real, mt.imag), dim=1).contiguous()
        mt = self._dev(m)
        if mt.dim() == 2:
            mt = mt.reshape(1, -1, nc)
        if mt.shape[1] != self.n or mt.shape[2] != nc:
            raise ValueError(f"expected moments of shape (nrhs, {self.n}, {nc})")
        return mt.contiguous()

    def _from_rhs(self, t):
        """Inverse of _as_rhs for an output buffer."""
        if self.complex:
            return torch.complex(t[:, 0], t[:, 1])
        return t

    def matvec(self, m, precision=FP64, slab_factor=1.0, out=None):
        mt = self._as_rhs(m)
        if out is None or self.complex:
            out = torch.empty_like(mt)
        self._sync_stream()
        _lib.check(self._lib.mpol_matvec(self._ctx, ctypes.c_void_p(mt.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                         mt.shape[0], int(precision), float(slab_factor)))
        return self._from_rhs(out)

    # -- dipole + quadrupole operator -------------------------------------
    def quad_matvec(self, m, precision=FP64):
        """(nrhs, N, 8) moments (S, Q) -> (nrhs, N, 8) fields and gradients (E, G)."""
        mt = self._as_rhs(m, 8)
        out = torch.empty_like(mt)
        self._sync_stream()
        _lib.check(self._lib.mpol_quad_matvec(self._ctx, ctypes.c_void_p(mt.data_ptr()),
                                              ctypes.c_void_p(out.data_ptr()), mt.shape[0], int(precision)))
        return self._from_rhs(out)

    def quad_components(self, m, precision=FP64):
        """Wave-space and real-space parts of quad_matvec."""
        mt = self._as_rhs(m, 8)
        hk = torch.empty_like(mt)
        hr = torch.empty_like(mt)
        self._sync_stream()
        _lib.check(self._lib.mpol_quad_matvec_components(self._ctx, ctypes.c_void_p(mt.data_ptr()),
                                                         ctypes.c_void_p(hk.data_ptr()),
                                                         ctypes.c_void_p(hr.data_ptr()), mt.shape[0],
                                                         int(precision)))
        return self._from_rhs(hk), self._from_rhs(hr)

    def quad_solve(self, b, x0, tol, restart=10, maxit=100, mixed=False, max_refine=8):
        return self._solve(self._lib.mpol_quad_solve, 8, b, x0, tol, restart, maxit, mixed, max_refine, 0.0)

    # -- polydisperse operator ---------------------------------------------
    def set_sizes(self, radius, types=None, perp=None, para=None):
        """Radii (N,).  With types (N,) in [0, M) and pair tables perp, para
        (ntab, M(M+1)/2) on the rvals of set_tables the pair couplings are
        interpolated; without them they are evaluated in closed form."""
        rt = self._dev(radius).reshape(-1)
        if rt.shape[0] != self.n:
            raise ValueError("radius must have N entries")
        if types is None:
            _lib.check(self._lib.mpol_set_sizes(self._ctx, ctypes.c_void_p(rt.data_ptr()), None, 0, None, None))
            self._keep_sizes = [rt]
            return
        tt = torch.as_tensor(np.asarray(types, dtype=np.int32).reshape(-1), device=self.device).contiguous()
        M = int(tt.max().item()) + 1
        pq = np.ascontiguousarray(np.asarray(perp, dtype=np.float64).T)
        pa = np.ascontiguousarray(np.asarray(para, dtype=np.float64).T)
        if pq.shape[0] != M * (M + 1) // 2 or pa.shape != pq.shape:
            raise ValueError("tables must have M(M+1)/2 columns for M radius types")
        self._sync_stream()
        _lib.check(self._lib.mpol_set_sizes(self._ctx, ctypes.c_void_p(rt.data_ptr()), ctypes.c_void_p(tt.data_ptr()),
                                            M, _lib.dptr(pq), _lib.dptr(pa)))
        self._keep_sizes = [rt, tt]

    def sizes_matvec(self, p, precision=FP64):
        pt = self._as_rhs(p)
        out = torch.empty_like(pt)
        self._sync_stream()
        _lib.check(self._lib.mpol_sizes_matvec(self._ctx, ctypes.c_void_p(pt.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr()), pt.shape[0], int(precision)))
        return self._from_rhs(out)

    def sizes_components(self, p, precision=FP64):
        pt = self._as_rhs(p)
        hk = torch.empty_like(pt)
        hr = torch.empty_like(pt)
        self._sync_stream()
        _lib.check(self._lib.mpol_sizes_matvec_components(self._ctx, ctypes.c_void_p(pt.data_ptr()),
                                                          ctypes.c_void_p(hk.data_ptr()),
                                                          ctypes.c_void_p(hr.data_ptr()), pt.shape[0],
                                                          int(precision)))
        return self._from_rhs(hk), self._from_rhs(hr)

    def sizes_solve(self, b, x0, tol, restart=10, maxit=100, mixed=False, max_refine=8):
        return self._solve(self._lib.mpol_sizes_solve, 3, b, x0, tol, restart, maxit, mixed, max_refine, 0.0)

    STAGES = ("slab_sum", "real_space", "spread_xfft", "fft_y_fwd", "fft_z_scale", "fft_y_inv", "xifft_contract",
              "contract_reduce")

    def matvec_timed(self, m, precision=FP64, slab_factor=1.0, repeats=10):
        """Mean milliseconds of each operator stage (HIP events on the launch stream)."""
        mt = self._as_rhs(m)
        out = torch.empty_like(mt)
        ms = np.zeros(len(self.STAGES), dtype=np.float64)
        self._sync_stream()
        _lib.check(self._lib.mpol_matvec_timed(self._ctx, ctypes.c_void_p(mt.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr()), mt.shape[0], int(precision),
                                               float(slab_factor), int(repeats), _lib.dptr(ms)))
        return dict(zip(self.STAGES, ms.tolist())), self._from_rhs(out)

    def debug_spectrum(self, m, stage, precision=FP64):
        """Grid spectrum after an operator stage, shape (nrhs, 3, nz, ny, nx//2+1)."""
        mt = self._as_rhs(m)
        g = self.geometry_info()["ngrid"]
        dt = torch.complex128 if precision == FP64 else torch.complex64
        spec = torch.empty((mt.shape[0], 3, g[2], g[1], g[0] // 2 + 1), dtype=dt, device=self.device)
        self._sync_stream()
        _lib.check(self._lib.mpol_debug_spectrum(self._ctx, ctypes.c_void_p(mt.data_ptr()), mt.shape[0],
                                                 int(precision), int(stage), ctypes.c_void_p(spec.data_ptr())))
        return spec.cpu().numpy()

    def components(self, m, precision=FP64, slab_factor=1.0):
        mt = self._as_rhs(m)
        hk = torch.empty_like(mt)
        hr = torch.empty_like(mt)
        hs = torch.empty_like(mt)
        self._sync_stream()
        _lib.check(self._lib.mpol_matvec_components(
            self._ctx, ctypes.c_void_p(mt.data_ptr()), ctypes.c_void_p(hk.data_ptr()),
            ctypes.c_void_p(hr.data_ptr()), ctypes.c_void_p(hs.data_ptr()), mt.shape[0], int(precision),
            float(slab_factor)))
        return self._from_rhs(hk), self._from_rhs(hr), self._from_rhs(hs)

    def _solve(self, fn, nc, b, x0, tol, restart, maxit, mixed, max_refine, slab_factor):
        bt = self._as_rhs(b, nc)
        x0t = self._as_rhs(x0, nc)
        if x0t.shape != bt.shape:
            raise ValueError("b and x0 must have the same shape")
        x = torch.empty_like(bt)
        opts = _lib.SolveOpts(float(tol), int(restart), int(maxit), 1 if mixed else 0, int(max_refine),
                              float(slab_factor))
        info = _lib.SolveInfo()
        self._sync_stream()
        _lib.check(fn(self._ctx, ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(x0t.data_ptr()),
                      ctypes.c_void_p(x.data_ptr()), bt.shape[0], ctypes.byref(opts), ctypes.byref(info)))
        return SolveResult(self._from_rhs(x), info.matvecs_fp64, info.matvecs_fp32, info.converged,
                           info.refinements, info.max_relres, info.matvecs_inner)

    def solve(self, b, x0, tol, restart=10, maxit=100, mixed=False, max_refine=8, slab_factor=1.0):
        return self._solve(self._lib.mpol_solve, 3, b, x0, tol, restart, maxit, mixed, max_refine, slab_factor)

    def set_inner_operator(self, inner, floor=1e-3):
        """Operator of the mixed solves' correction steps: `inner`, a Context
        on the same particles and permittivity sets (coarse_operator builds
        one on a coarser grid), takes the first residual and the FP32
        correction solves, which stop at `floor` (relative) at the lowest;
        every residual that decides convergence stays on this context's FP64
        operator.  Dipole and dipole + quadrupole solves.  None removes it."""
        self._sync_stream()
        _lib.check(self._lib.mpol_set_inner_operator(self._ctx, None if inner is None else inner._ctx,
                                                     float(floor)))
        self._inner = inner  # the C context keeps a pointer to it

    # -- forces of the Ewald-summed polarisation energy --------------------
    def ewald_forces(self, p, f1, f2):
        """F = -dU/dx (N, 3) device tensor at solved dipoles p (nrhs, N, 3):
        real space from the pair-force tables f1, f2 (field_dip_force_1/2 on
        the rvals of set_tables), wave space from the Gaussian derivative."""
        pt = self._as_rhs(p)
        f1 = np.ascontiguousarray(np.asarray(f1, dtype=np.float64).reshape(-1))
        f2 = np.ascontiguousarray(np.asarray(f2, dtype=np.float64).reshape(-1))
        if f1.size != self._ntab or f2.size != self._ntab:
            raise ValueError("force tables must have one entry per rvals entry of set_tables")
        out = torch.empty((self.n, 3), dtype=torch.float64, device=self.device)
        self._sync_stream()
        _lib.check(self._lib.mpol_ewald_forces(self._ctx, ctypes.c_void_p(pt.data_ptr()), pt.shape[0], _lib.dptr(f1),
                                               _lib.dptr(f2), ctypes.c_void_p(out.data_ptr())))
        return out

    # -- fixed-dipole forces ----------------------------------------------
    def compute_forces(self, pos, dip, box, r_cut, prefactor=1.0, periodic_z=True, use_cell_list=True):
        """Truncated dipole-dipole forces (N, 3) and per-particle energies (N,)
        as FP64 device tensors.  Needs no tables or geometry."""
        pt = self._dev(pos)
        dt = self._dev(dip)
        box = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(-1))
        n = pt.shape[0] if pt.dim() == 2 else -1
        if pt.dim() != 2 or pt.shape[1] != 3 or dt.shape != pt.shape or box.size != 3:
            raise _lib.MpolError("Invalid shapes: pos(N,3), dip(N,3), box(3) required.")
        f = torch.empty((n, 3), dtype=torch.float64, device=self.device)
        pe = torch.empty((n,), dtype=torch.float64, device=self.device)
        self._sync_stream()
        _lib.check(self._lib.mpol_compute_forces(self._ctx, ctypes.c_void_p(pt.data_ptr()),
                                                 ctypes.c_void_p(dt.data_ptr()), n, _lib.dptr(box), float(r_cut),
                                                 float(prefactor), 1 if periodic_z else 0,
                                                 1 if use_cell_list else 0, ctypes.c_void_p(f.data_ptr()),
                                                 ctypes.c_void_p(pe.data_ptr())))
        return f, pe
