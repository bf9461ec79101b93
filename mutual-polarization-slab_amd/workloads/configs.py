"""The other full-size workloads of the solver, measured like the main one.

tensor:     N = 1e4 equal spheres, phi = 0.3 (FCC + hard-sphere Metropolis),
            beta = 0.8, the three unit fields of the capacitance tensor in one
            batched solve, 128^3, P = 10, dipole error <= 1e-6.
quadrupole: N = 2e4, phi = 0.35, body-centred tetragonal lattice c/a = 1.2
            (25 x 25 x 16 cells, seeded Gaussian jitter 0.05 a), half of the
            spheres at permittivity ratio 10 and half at 0.5 (seeded), field
            along z, dipole + quadrupole unknowns (8 per sphere), 160^3,
            P = 12, xi from the longest box side at tol 1e-6.

Each reports FP64 and mixed (FP32 operator inside FP64 refinement) solves/s
at the loosest tolerance of {1e-6, 5e-7, ..., 1e-8} whose solution is
within 1e-6 of a tight (1e-11) FP64 solve, the operator applications per
solve and the error; and, when autotune.tune_inner finds one faster, the
two-level mixed solve (correction solves on a coarser operator).

    python -m workloads.configs [--only tensor|quadrupole]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

TOL_LADDER = (1e-6, 5e-7, 2e-7, 1e-7, 5e-8, 2e-8, 1e-8)
ERR_BAR = 1e-6


def bct_lattice(cells=(25, 25, 16), c_over_a=1.2, phi=0.35, jitter=0.05, seed=3):
    """(x (N,3), box (3,)): a body-centred tetragonal lattice of unit spheres
    (two per cell) at volume fraction phi with Gaussian jitter*a per
    coordinate, wrapped into the box."""
    nc = np.asarray(cells)
    a = (2.0 * 4.0 / 3.0 * np.pi / (phi * c_over_a)) ** (1.0 / 3.0)
    cell = np.array([a, a, c_over_a * a])
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in nc], indexing="ij"), axis=-1).reshape(-1, 3)
    corners = idx * cell
    x = np.concatenate([corners, corners + 0.5 * cell])
    rng = np.random.default_rng(seed)
    x = x + rng.normal(0.0, jitter * a, x.shape)
    box = nc * cell
    return np.mod(x, box), box


def _time(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, r


def _ladder(solve, ref, reps):
    """(tol, seconds per solve, result, error) of the loosest ladder tolerance
    meeting the error bar."""
    for t in TOL_LADDER:
        r = solve(t)
        e = float(np.max(np.abs(r.x.cpu().numpy() - ref)) / np.max(np.abs(ref)))
        if e <= ERR_BAR or t == TOL_LADDER[-1]:
            s, r = _time(lambda: solve(t), reps)
            return t, s, r, e


def tensor(reps=3, sweeps=200):
    import torch

    import workloads
    from mpol_hip import FP32, FP64, Context, default_tables
    from mpol_hip.autotune import tune_launch
    cfg = workloads.configuration("n1e4_phi0.3", sweeps=sweeps)
    n = cfg["x"].shape[0]
    perp, para, rv = default_tables(cfg["xi"])
    ctx = Context()
    ctx.set_tables(perp, para, rv)
    ctx.set_geometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["eta"], cfg["rc"])
    ctx.set_permittivity(cfg["eps"])
    tune_launch(ctx, FP32, nrhs=3)
    tune_launch(ctx, FP64, nrhs=3)
    phi = 4.0 * np.pi * n / (3.0 * np.prod(cfg["box"]))
    mg = 4.0 * np.pi * cfg["beta"] / (1.0 - cfg["beta"] * phi)
    b = torch.as_tensor(np.stack([np.tile(e, n).reshape(n, 3) for e in np.eye(3)]), device=ctx.device)
    x0 = mg * b

    def solve(mixed, t):
        return ctx.solve(b, x0, t, restart=10, maxit=100, mixed=mixed, slab_factor=0.0)

    ref = solve(False, 1e-11).x.cpu().numpy()
    out = {"workload": "N=1e4 unit spheres, phi=0.3, beta=0.8, 3 fields batched, 128^3, P=10",
           "n": n, "rhs": 3, "xi": round(cfg["xi"], 6), "rc": round(cfg["rc"], 6)}
    for name, mixed in (("mixed", True), ("fp64", False)):
        t, s, r, e = _ladder(lambda tt: solve(mixed, tt), ref, reps)
        out[name] = {"tol": t, "rhs_per_s": round(3.0 / s, 2), "ms_per_batch": round(1e3 * s, 3), "rel_err": e,
                     "matvecs": {"fp32": r.matvecs_fp32, "fp64": r.matvecs_fp64}}
    out["speedup_mixed_vs_fp64"] = round(out["mixed"]["rhs_per_s"] / out["fp64"]["rhs_per_s"], 3)
    # two levels: correction solves on a coarser operator
    from mpol_hip.autotune import tune_inner
    inner, icfg, _ = tune_inner(
        ctx, cfg["x"], cfg["box"], cfg["eps"], b, x0, TOL_LADDER, repeats=reps, slab_factor=0.0,
        candidates=((1e-4, 64), (1e-3, 64), (1e-3, 32)),
        accept=lambda r: float(np.max(np.abs(r.x.cpu().numpy() - ref)) / np.max(np.abs(ref))) <= ERR_BAR)
    if inner is not None:
        t, s, r, e = _ladder(lambda tt: solve(True, tt), ref, reps)
        out["mixed_two_level"] = {"tol": t, "rhs_per_s": round(3.0 / s, 2), "ms_per_batch": round(1e3 * s, 3),
                                  "rel_err": e, "correction_operator": icfg,
                                  "matvecs": {"coarse_fp32": r.matvecs_inner, "fp32": r.matvecs_fp32 - r.matvecs_inner,
                                              "fp64": r.matvecs_fp64}}
        out["speedup_two_level_vs_fp64"] = round(out["mixed_two_level"]["rhs_per_s"] / out["fp64"]["rhs_per_s"], 3)
        ctx.set_inner_operator(None)
        inner.close()
    C = ref.mean(axis=1)  # capacitance tensor rows (field i -> mean dipole)
    out["capacitance_tensor_diag"] = [float(C[i, i]) for i in range(3)]
    ctx.close()
    return out


def quadrupole(reps=3):
    import torch

    import quadrupole as quad
    x, box = bct_lattice()
    n = x.shape[0]
    ng, P, tol = 160, 12, 1e-6
    s = np.sqrt(-np.log(tol))
    xi = (ng - 1) * np.pi / (2.0 * float(np.max(box)) * s)
    rc = s / xi
    ngrid = np.full(3, ng)
    eta = P * (box / ngrid * xi) ** 2 / np.pi
    r = np.concatenate(([0.0], np.arange(0.001, 10.0 + 1e-12, 0.001)))
    tabs = quad._driver_tables(r, xi)
    ctx = quad._context(x, box, ngrid, P, xi, eta, rc, r, tabs)
    rng = np.random.default_rng(17)
    eps = np.where(rng.random(n) < 0.5, 10.0, 0.5)
    ctx.set_permittivity(eps)
    ctx._eps_key = None
    phi = 4.0 * np.pi * n / (3.0 * np.prod(box))
    beta = (eps - 1.0) / (eps + 2.0)
    mg = 4.0 * np.pi * beta / (1.0 - beta * phi)
    b = np.zeros((1, n, 8))
    b[0, :, 2] = 1.0
    x0 = np.zeros_like(b)
    x0[0, :, 2] = mg
    bt = torch.as_tensor(b, device=ctx.device)
    x0t = torch.as_tensor(x0, device=ctx.device)

    def solve(mixed, t):
        return ctx.quad_solve(bt, x0t, t, restart=10, maxit=100, mixed=mixed)

    ref = solve(False, 1e-11).x.cpu().numpy()
    out = {"workload": "N=2e4 BCT c/a=1.2 jitter 0.05a, phi=0.35, eps 10 / 0.5, dipole+quadrupole, 160^3, P=12",
           "n": n, "box": [round(v, 4) for v in box], "xi": round(xi, 6), "rc": round(rc, 6)}
    for name, mixed in (("mixed", True), ("fp64", False)):
        t, s_, r_, e = _ladder(lambda tt: solve(mixed, tt), ref, reps)
        out[name] = {"tol": t, "solves_per_s": round(1.0 / s_, 2), "ms_per_solve": round(1e3 * s_, 3), "rel_err": e,
                     "matvecs": {"fp32": r_.matvecs_fp32, "fp64": r_.matvecs_fp64}}
    out["speedup_mixed_vs_fp64"] = round(out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"], 3)
    # two levels: correction solves on a coarser dipole + quadrupole operator
    from mpol_hip.autotune import tune_inner
    inner, icfg, _ = tune_inner(
        ctx, x, box, eps, bt, x0t, TOL_LADDER, quadrupole=True, repeats=reps,
        candidates=((1e-4, 96), (1e-4, 64), (1e-3, 64), (1e-5, 96)),
        accept=lambda r: float(np.max(np.abs(r.x.cpu().numpy() - ref)) / np.max(np.abs(ref))) <= ERR_BAR)
    if inner is not None:
        t, s_, r_, e = _ladder(lambda tt: solve(True, tt), ref, reps)
        out["mixed_two_level"] = {"tol": t, "solves_per_s": round(1.0 / s_, 2), "ms_per_solve": round(1e3 * s_, 3),
                                  "rel_err": e, "correction_operator": icfg,
                                  "matvecs": {"coarse_fp32": r_.matvecs_inner,
                                              "fp32": r_.matvecs_fp32 - r_.matvecs_inner,
                                              "fp64": r_.matvecs_fp64}}
        out["speedup_two_level_vs_fp64"] = round(out["mixed_two_level"]["solves_per_s"]
                                                 / out["fp64"]["solves_per_s"], 3)
        ctx.set_inner_operator(None)
        inner.close()
    out["mean_S_z"] = float(ref[0, :, 2].mean())
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=("tensor", "quadrupole"), default=None)
    args = ap.parse_args(argv)
    res = {}
    if args.only in (None, "tensor"):
        res["tensor"] = tensor()
        print("[configs] tensor done", file=sys.stderr, flush=True)
    if args.only in (None, "quadrupole"):
        res["quadrupole"] = quadrupole()
        print("[configs] quadrupole done", file=sys.stderr, flush=True)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
