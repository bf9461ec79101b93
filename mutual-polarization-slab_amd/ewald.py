"""ewald.py -- drop-in for the reference's spectral-Ewald module, on MI355X.

Same entry points and argument meaning as "CPU - Ewald test/ewald.py"; the
operator (magnetic_field, magnetic_field_components) runs in the HIP kernels
of libmpol.so: a fused Gaussian spread + forward x-FFT, y- and z-FFTs with the
k-space scale fused into the z pass, an inverse x-FFT fused with the
contraction, and a CSR real-space sum, all on the GPU.

Differences a caller can see, all at rounding level:
  * Grids are real (half spectrum along x).  For real dipoles the result is
    the real part of the reference's complex result; the reference's imaginary
    part for real input comes only from its fftshift-ordered Nyquist planes
    and is below 1e-9 of the wave-space field.
  * The precomputed tuples (real_precomp, grid_precomp, k_precomp) and the
    candidate pairs p1/p2 are accepted for signature compatibility; the GPU
    rebuilds the same quantities from x (cell list at rc, min-image d < rc,
    same table interpolation, same Gaussian weights) once per geometry and
    caches them.
Host helpers (cell_list, neighbor_list, pre_calculations, matlab_round,
besselj_3_over_2) are small NumPy index utilities kept for API parity.
"""
from __future__ import annotations

import hashlib
from collections import OrderedDict
from typing import Tuple

import numpy as np

from mpol_hip import context as _ctxmod

# ---------------------------------------------------------------------------
# host helpers ("CPU - Ewald test/ewald.py":40-223)
# ---------------------------------------------------------------------------


def erfc(x):
    """Complementary error function (scipy.special.erfc)."""
    from scipy.special import erfc as _erfc
    return _erfc(np.asarray(x))


def besselj_3_over_2(x):
    """J_{3/2}(x) = sqrt(2/(pi x)) (sin x/x - cos x), J(0) = 0 ([E]ewald.py:51-65)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    nz = x != 0
    v = x[nz]
    out[nz] = np.sqrt(2.0 / (np.pi * v)) * (np.sin(v) / v - np.cos(v))
    return out


def matlab_round(x):
    """Round half away from zero ([E]ewald.py:71-78)."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x >= 0.0, np.floor(x + 0.5), -np.floor(-x + 0.5))


def cell_list(x, box, rmax) -> Tuple[np.ndarray, np.ndarray]:
    """0-based cell coordinates, max(3, floor(L/rmax)) cells per axis ([E]ewald.py:84-111)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    xw = np.mod(np.asarray(x, dtype=np.float64), box)
    ncell = np.maximum(np.floor(box / rmax).astype(int), 3)
    cell = np.minimum(np.floor((xw / box) * ncell).astype(int), ncell - 1)
    return cell, ncell


def neighbor_list(cell, Ncell) -> Tuple[np.ndarray, np.ndarray]:
    """Directed pairs (p1, p2), p1 != p2, in periodically adjacent cells ([E]ewald.py:114-199)."""
    cell = np.asarray(cell, dtype=np.int64)
    nc = np.asarray(Ncell, dtype=np.int64).reshape(3)
    n = cell.shape[0]
    if n == 0:
        e = np.zeros(0, dtype=int)
        return e, e
    cid = cell[:, 0] + nc[0] * (cell[:, 1] + nc[1] * cell[:, 2])
    order = np.argsort(cid, kind="stable")
    counts = np.bincount(cid, minlength=int(np.prod(nc)))
    start = np.cumsum(counts) - counts
    p1s, p2s = [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nb = np.mod(cell + np.array([dx, dy, dz]), nc)
                nid = nb[:, 0] + nc[0] * (nb[:, 1] + nc[1] * nb[:, 2])
                cnt = counts[nid]
                tot = int(cnt.sum())
                if tot == 0:
                    continue
                rep_i = np.repeat(np.arange(n), cnt)
                within = np.arange(tot) - np.repeat(np.cumsum(cnt) - cnt, cnt)
                rep_j = order[np.repeat(start[nid], cnt) + within]
                p1s.append(rep_i)
                p2s.append(rep_j)
    p1 = np.concatenate(p1s)
    p2 = np.concatenate(p2s)
    keep = p1 != p2
    return p1[keep], p2[keep]


def pre_calculations(P, h) -> Tuple[np.ndarray, np.ndarray]:
    """P^3 stencil offsets (x fastest) and their physical displacements ([E]ewald.py:205-223)."""
    P = int(P)
    lin = np.arange(P ** 3)
    c = (P - 1) // 2
    off = np.stack((lin % P - c, (lin // P) % P - c, lin // (P * P) - c), axis=1).astype(int)
    return off, off.astype(np.float64) * np.asarray(h, dtype=np.float64).reshape(1, 3)


def real_space_table(r, xi):
    """The six real-space couplings with the self term prepended, computed by
    the library's FP64 routine ([E]ewald.py:604-931)."""
    return _ctxmod.real_space_table(r, xi)


# ---------------------------------------------------------------------------
# GPU contexts, cached per geometry
# ---------------------------------------------------------------------------
_CACHE: "OrderedDict[tuple, _ctxmod.Context]" = OrderedDict()
_CACHE_MAX = 4


def _digest(*arrays) -> str:
    h = hashlib.sha1()
    for a in arrays:
        a = np.ascontiguousarray(np.asarray(a))
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def geometry_context(x, box, Ngrid, P, xi, eta, rc, Hperp, Hpara, rvals):
    """A Context holding the precompute for this geometry (built on first use)."""
    x = np.asarray(x, dtype=np.float64)
    key = (_digest(x), _digest(box), _digest(np.asarray(Ngrid, dtype=np.int64)), int(P), float(xi), _digest(eta),
           float(rc), _digest(np.real(np.asarray(Hperp)), np.real(np.asarray(Hpara)), rvals))
    ctx = _CACHE.get(key)
    if ctx is None:
        ctx = _ctxmod.Context()
        ctx.set_tables(Hperp, Hpara, rvals)
        ctx.set_geometry(x, box, Ngrid, P, xi, eta, rc)
        ctx._eps_key = None
        _CACHE[key] = ctx
        while len(_CACHE) > _CACHE_MAX:
            _CACHE.popitem(last=False)[1].close()
    else:
        _CACHE.move_to_end(key)
    return ctx


def _with_permittivity(ctx, lambdap):
    lam = np.ascontiguousarray(np.real(np.asarray(lambdap, dtype=np.complex128)).reshape(-1))
    k = _digest(lam)
    if getattr(ctx, "_eps_key", None) != k:
        ctx.set_permittivity(lam)
        ctx._eps_key = k
    return ctx


def _split_complex(m):
    """(nrhs, N, 3) real stack of the real (and, if present, imaginary) parts."""
    m = np.asarray(m)
    if np.iscomplexobj(m) and np.any(m.imag != 0):
        return np.stack((m.real, m.imag)), True
    return np.real(m)[None, ...].astype(np.float64), False


def _join_complex(t, had_imag):
    a = t.detach().cpu().numpy()
    if had_imag:
        return a[0] + 1j * a[1]
    return a[0].astype(np.complex128)


# ---------------------------------------------------------------------------
# the operator ("CPU - Ewald test/ewald.py":1146-1327)
# ---------------------------------------------------------------------------
def magnetic_field(x, m, lambdap, box, p1, p2, Ngrid, h, P, xi, eta, rc, offset, offsetxyz, Hperp, Hpara, rvals,
                   real_precomp=None, grid_precomp=None, k_precomp=None, slab_factor: float = 1.0,
                   precision: str = "fp64") -> np.ndarray:
    """A m = Hk + Hr - H_slab at the particles, (N,3) complex128.

    precision: "fp64" (default) or "fp32" (FP32 grids, FFT and real-space sum
    with FP64 geometry decisions)."""
    ctx = geometry_context(x, box, Ngrid, P, xi, eta, rc, Hperp, Hpara, rvals)
    _with_permittivity(ctx, lambdap)
    mm, had_imag = _split_complex(m)
    out = ctx.matvec(mm, precision=_prec(precision), slab_factor=float(slab_factor))
    return _join_complex(out, had_imag)


def magnetic_field_components(x, m, lambdap, box, p1, p2, Ngrid, h, P, xi, eta, rc, offset, offsetxyz, Hperp,
                              Hpara, rvals, real_precomp=None, grid_precomp=None, k_precomp=None,
                              slab_factor: float = 1.0, precision: str = "fp64"):
    """(Hk, Hr, H_slab), each (N,3) complex128, with A m = Hk + Hr - H_slab."""
    ctx = geometry_context(x, box, Ngrid, P, xi, eta, rc, Hperp, Hpara, rvals)
    _with_permittivity(ctx, lambdap)
    mm, had_imag = _split_complex(m)
    hk, hr, hs = ctx.components(mm, precision=_prec(precision), slab_factor=float(slab_factor))
    return _join_complex(hk, had_imag), _join_complex(hr, had_imag), _join_complex(hs, had_imag)


def _prec(precision) -> int:
    p = str(precision).lower()
    if p in ("fp64", "double", "float64"):
        return _ctxmod.FP64
    if p in ("fp32", "float", "float32", "mixed"):
        return _ctxmod.FP32
    raise ValueError(f"unknown precision {precision!r}")
