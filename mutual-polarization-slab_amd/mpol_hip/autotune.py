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

coarse_operator(x, box, eps, ewald_tol, grid) / tune_inner(ctx, ...)
    The correction operator of the mixed solves: the same particles on a
    coarser grid at a looser rule tolerance (e.g. 128^3, P = 6 at 1e-4 for a
    256^3, P = 10 operator at 1e-6), set with Context.set_inner_operator.
    Its FP32 correction solves cost less per step; the FP64 residuals on the
    full operator still decide convergence, so the solve meets its tolerance
    for the full operator, in more refinement sweeps (each removes about the
    two operators' relative distance, the floor).  tune_inner times the
    solve with none and with each candidate and keeps the fastest.
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


# Correction-operator candidates of the mixed solves, (Ewald tolerance of the
# parameter rule, grid nodes along the longest axis), and the floors their
# correction solves stop at.
INNER_CANDIDATES = ((1e-4, 128), (1e-3, 128), (1e-4, 96), (1e-3, 96), (1e-5, 128), (1e-5, 160))
INNER_FLOORS = (1e-2, 3e-3, 1e-3)


def coarse_operator(x, box, eps, ewald_tol=1e-4, grid=128, device=None, tables=None, tune=True,
                    quadrupole=False):
    """(Context, (xi, ngrid, P, eta, rc)): the particles x with permittivities
    eps (one set) on the Ewald parameters of `grid` nodes along the longest
    box axis at rule tolerance ewald_tol (ewald_candidates), for
    Context.set_inner_operator; tune: its FP32 launch shapes by tune_launch
    (dipole operator); quadrupole: with the seven quadrupole tables, for
    dipole + quadrupole solves.  tables(xi) -> (perp, para, rvals), default
    the reference's table grid."""
    from .context import Context, default_tables, real_space_table_quad
    cands = ewald_candidates(box, ewald_tol, sizes=(int(grid),))
    if not cands:
        raise ValueError(f"no Ewald parameters with {grid} nodes at tol {ewald_tol} for box {box}")
    xi, ng, P, eta, rc = cands[0]
    perp, para, rv = (tables or default_tables)(xi)
    c = Context(device)
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, eta, rc)
    if quadrupole:
        # field_quad_1..3, grad_quad_1..4 on the same rvals (entry 0 the self term)
        c.set_quad_tables(np.stack(real_space_table_quad(rv[1:], xi)[1:]))
    c.set_permittivity(eps)
    if tune and not quadrupole:
        tune_launch(c, FP32)
    return c, (xi, ng, P, eta, rc)


def tune_inner(ctx, x, box, eps, b, x0, tols, accept=None, candidates=INNER_CANDIDATES, floors=INNER_FLOORS,
               repeats=3, slab_factor=1.0, restart=10, maxit=100, quadrupole=False):
    """Choose the correction operator of ctx's mixed solves by timing them:
    none (ctx's own FP32 operator) and each coarse candidate x floor.  For
    each setting the loosest tolerance of `tols` (loosest first) whose solve
    of b from x0 converges (its FP64 residual on ctx's operator meets it) and
    passes accept(result) (default: always) is timed over `repeats` solves;
    the fastest setting is left on ctx (quadrupole: the dipole + quadrupole
    solve, quad_solve, and coarse operators with quadrupole tables).  Returns
    (inner Context or None,
    {"ewald_tol", "grid", "P", "floor", "tol"} or None for none, report
    [(label, tol, ms, matvecs_fp32, matvecs_fp64, matvecs_inner)])."""
    import torch

    nrhs = b.shape[0] if b.ndim == 3 else 1

    def run(t):
        if quadrupole:
            return ctx.quad_solve(b, x0, t, restart=restart, maxit=maxit, mixed=True)
        return ctx.solve(b, x0, t, restart=restart, maxit=maxit, mixed=True, slab_factor=slab_factor)

    def timed():
        for t in tols:
            r = run(t)
            if r.converged == nrhs and (accept is None or accept(r)):
                break
        else:
            return None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            r = run(t)
        torch.cuda.synchronize()
        return t, 1e3 * (time.perf_counter() - t0) / repeats, r

    ctx.set_inner_operator(None)
    report = []
    best = (math.inf, None, None)
    got = timed()
    if got is not None:
        t, ms, r = got
        report.append(("none", t, ms, r.matvecs_fp32, r.matvecs_fp64, r.matvecs_inner))
        best = (ms, None, None)
    for etol, grid in candidates:
        try:
            inner, (xi, ng, P, eta, rc) = coarse_operator(x, box, eps, etol, grid, device=ctx.device,
                                                          quadrupole=quadrupole)
        except ValueError:
            continue
        for fl in floors:
            ctx.set_inner_operator(inner, fl)
            got = timed()
            if got is None:
                continue
            t, ms, r = got
            report.append((f"tol {etol:g} grid {grid} P {P} floor {fl:g}", t, ms, r.matvecs_fp32, r.matvecs_fp64,
                           r.matvecs_inner))
            if ms < best[0]:
                if best[1] is not None and best[1] is not inner:
                    best[1].close()
                best = (ms, inner, {"ewald_tol": etol, "grid": int(grid), "P": int(P), "floor": fl, "tol": t})
        ctx.set_inner_operator(None)
        if best[1] is not inner:
            inner.close()
    if best[1] is not None:
        ctx.set_inner_operator(best[1], best[2]["floor"])
    return best[1], best[2], report
