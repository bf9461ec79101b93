"""Autotuner of the operator's launch shapes and of the Ewald grid.

tune_launch(ctx, precision, nrhs=1)
    Coordinate search over the launch knobs of one context and precision,
    scored by the per-stage HIP-event times of the dipole operator
    (mpol_matvec_timed): engine, tile rows and particle chunk by the spread +
    contraction time, the y-pass column width by the two y passes, the z-pass
    column width and threads by the z pass.  The best knobs are set on the
    context for that precision (mpol_set_tuning_for) and returned; results
    are cached per (particles, grid, P, precision, right-hand sides).

ewald_candidates(box, tol)
    Ewald parameter sets by the reference's rule ("CPU - Ewald test/
    capacitance.py":357-362: rc = sqrt(-ln tol)/xi, kcut = 2 xi sqrt(-ln tol),
    Ngrid = ceil(1 + L kcut/pi), P = ceil(-2 ln tol/pi), eta = P (h xi)^2/pi),
    one per compile-time FFT size N: the largest xi whose grid on the longest
    box axis is N (shorter axes get the smallest compile-time size that holds
    their derived grid).  A larger N moves work from the real-space sum to
    the grid.  The rule's tol is not the operator's error: at fixed tol the
    real-space truncation error grows with xi (about xi^3), so candidates
    differ in accuracy.

tune_ewald(x, box, eps, target)
    Candidates over grid sizes and rule tolerances (tol, tol/10, ...); each
    one's operator error is measured on these particles against a tight
    (tol 1e-11) FP64 operator, and the fastest candidate whose error is at
    most target is returned.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import FP32, FP64

CT_SIZES = (32, 64, 96, 128, 160, 192, 256, 320, 512)

TILE_OPTIONS = (
    {"engine": 0},
    {"engine": 1, "tile_ty": 4, "tile_tz": 2, "chunk": 128},
    {"engine": 1, "tile_ty": 4, "tile_tz": 2, "chunk": 256},
    {"engine": 1, "tile_ty": 4, "tile_tz": 2, "chunk": 64},
    {"engine": 1, "tile_ty": 2, "tile_tz": 2, "chunk": 128},
    {"engine": 1, "tile_ty": 4, "tile_tz": 4, "chunk": 128},
    {"engine": 1, "tile_ty": 4, "tile_tz": 4, "chunk": 64},
    {"engine": 2},
)
Y_OPTIONS = ({"y_cols": 0}, {"y_cols": 8}, {"y_cols": 16}, {"y_cols": 32})
Z_OPTIONS = ({"z_cols": 0, "z_threads": 0}, {"z_cols": 4}, {"z_cols": 8}, {"z_cols": 16},
             {"z_cols": 8, "z_threads": 1024}, {"z_cols": 16, "z_threads": 1024})

_CACHE: dict = {}


def _stage_times(ctx, m, precision, repeats):
    stages, _ = ctx.matvec_timed(m, precision=precision, slab_factor=0.0, repeats=repeats)
    return stages


def tune_launch(ctx, precision=FP32, nrhs=1, repeats=5, seed=0, force=False):
    """Best launch knobs for ctx's current geometry in one precision (set on
    the context and returned as a dict of knob -> value)."""
    import torch
    info = ctx.geometry_info()
    key = (info["n"], tuple(info["ngrid"]), info["P"], int(precision), int(nrhs))
    if key in _CACHE and not force:
        knobs = dict(_CACHE[key])
        ctx.set_tuning(precision=precision, **knobs)
        return knobs
    g = torch.Generator().manual_seed(seed)
    m = torch.randn((nrhs, info["n"], 3), generator=g, dtype=torch.float64).to(ctx.device)
    best = {}

    def score(opts, stages_of):
        nonlocal best
        results = []
        for o in opts:
            trial = {**best, **o}
            try:
                ctx.set_tuning(precision=precision, **trial)
                ctx.matvec(m, precision=precision, slab_factor=0.0)  # warm-up (builds tile images)
                st = _stage_times(ctx, m, precision, repeats)
            except ValueError:
                continue  # knob set not applicable to this grid
            results.append((sum(st[k] for k in stages_of), trial))
        results.sort(key=lambda r: r[0])
        best = results[0][1]

    score(TILE_OPTIONS, ("spread_xfft", "xifft_contract", "contract_reduce"))
    score(Y_OPTIONS, ("fft_y_fwd", "fft_y_inv"))
    score(Z_OPTIONS, ("fft_z_scale",))
    ctx.set_tuning(precision=precision, **best)
    _CACHE[key] = dict(best)
    return dict(best)


def _ct_at_least(n):
    for s in CT_SIZES:
        if s >= n:
            return s
    return None


def ewald_candidates(box, tol, sizes=CT_SIZES):
    """[(xi, ngrid (3,), P, eta (3,), rc)] by the reference's parameter rule
    at tol, one per compile-time grid size of the longest axis."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    s = math.sqrt(-math.log(tol))
    P = int(math.ceil(-2.0 * math.log(tol) / math.pi))
    lmax = float(np.max(box))
    out = []
    for N in sizes:
        xi = (N - 1) * math.pi / (2.0 * lmax * s)
        kcut = 2.0 * xi * s
        ng = []
        for d in range(3):
            need = int(math.ceil(1.0 + box[d] * kcut / math.pi))
            ct = _ct_at_least(need)
            if ct is None:
                break
            ng.append(ct)
        if len(ng) != 3:
            continue
        ng = np.array(ng, dtype=np.int64)
        rc = s / xi
        if rc > float(np.min(box)) / 2.0 or P + 16 > int(np.min(ng)):
            continue
        eta = P * (box / ng * xi) ** 2 / math.pi
        out.append((xi, ng, P, eta, rc))
    return out


def _tight_reference(box):
    """A tight parameter set (rule tol 1e-11) with rc <= L/2."""
    tol = 1e-11
    s = math.sqrt(-math.log(tol))
    box = np.asarray(box, dtype=np.float64)
    xi = max(1.0, 2.0 * s / float(np.min(box)) * 1.001)
    kcut = 2.0 * xi * s
    ng = np.array([_ct_at_least(int(math.ceil(1.0 + b * kcut / math.pi))) or 0 for b in box], dtype=np.int64)
    if np.any(ng == 0):
        ng = np.ceil(1.0 + box * kcut / math.pi).astype(np.int64)
    P = int(math.ceil(-2.0 * math.log(tol) / math.pi))
    eta = P * (box / ng * xi) ** 2 / math.pi
    return xi, ng, P, eta, s / xi


def tune_ewald(x, box, eps, target, precision=FP32, tols=None, sizes=CT_SIZES, repeats=3, tables=None,
               device=None):
    """Fastest Ewald parameter set whose operator error on these particles is
    at most target: returns (xi, ngrid, P, eta, rc, ms_per_matvec, report),
    report = [(ms, err, xi, ngrid, P, rc)] for every candidate tried."""
    import torch

    from .context import Context, default_tables
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    tols = (1e-6, 1e-7, 1e-8, 1e-9) if tols is None else tols
    cands = [c for t in tols for c in ewald_candidates(box, t, sizes)]
    if not cands:
        raise ValueError("no Ewald grid with rc <= L/2 for these tolerances")
    tab = default_tables if tables is None else tables
    ctx = Context(device)
    g = torch.Generator().manual_seed(1)
    m = torch.randn((1, n, 3), generator=g, dtype=torch.float64).to(ctx.device)

    def setup(xi, ng, P, eta, rc):
        perp, para, rv = tab(xi)
        ctx.set_tables(perp, para, rv)
        ctx.set_geometry(x, box, ng, P, xi, eta, rc)
        ctx.set_permittivity(eps)

    setup(*_tight_reference(box))
    ref = ctx.matvec(m, slab_factor=0.0)
    scale = float(ref.abs().max())
    report = []
    for (xi, ng, P, eta, rc) in cands:
        setup(xi, ng, P, eta, rc)
        err = float((ctx.matvec(m, slab_factor=0.0) - ref).abs().max()) / scale
        ctx.matvec(m, precision=precision, slab_factor=0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            ctx.matvec(m, precision=precision, slab_factor=0.0)
        torch.cuda.synchronize()
        report.append((1e3 * (time.perf_counter() - t0) / repeats, err, xi, ng, P, rc))
    ctx.close()
    ok = sorted((r for r in report if r[1] <= target), key=lambda r: r[0])
    if not ok:
        raise ValueError(f"no candidate reaches operator error {target:g} (best {min(r[1] for r in report):.2e})")
    ms, err, xi, ng, P, rc = ok[0]
    eta = P * (np.asarray(box, dtype=np.float64) / ng * xi) ** 2 / math.pi
    return xi, ng, P, eta, rc, ms, report
