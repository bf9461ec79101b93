"""Throughput of the many-configuration driver on the heterogeneous workload:
COUNT independent polydisperse configurations (N log-uniform in [500, 8000],
log-normal radii of mean 1 and sigma 0.2, phi uniform in [0.1, 0.45]), each
over a 16-point permittivity-ratio spectrum log-spaced in [0.1, 100] x 3
field directions (COUNT x 48 right-hand sides), FP32 operator inside FP64
GMRES and refinement, grids chosen per configuration among 32^3..128^3.

Prints one JSON line: right-hand sides per second over the whole set (packing
and the per-configuration grid choice timed apart; geometry set-up of every
configuration included), the chosen grids, matvec counts, the mixed-vs-FP64
dipole error on --check configurations, with --fp64 the same set solved with
the FP64 operator throughout, and with --plain-sample K, K configurations
solved one right-hand side at a time in FP64 on one stream (the way the
reference solves them), extrapolated to the set by particle count.

    python -m workloads.heterogeneous [--count 256] [--workers 8] [--check 4]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def plain_fp64(configs, eps, grids, tol, gmres_tol):
    """Each configuration's spectrum x fields solved one right-hand side at a
    time with the FP64 operator on one stream (geometry set once per
    configuration): (seconds, right-hand sides)."""
    import torch

    from mpol_hip import Context, batch
    ctx = Context()
    fields = np.eye(3)
    nrhs = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for cfg, G in zip(configs, grids):
        x, a, box = cfg["x"], cfg["a"], cfg["box"]
        n = x.shape[0]
        xi, ng, P, eta, rc = batch.grid_candidates(box, tol, (G,))[0]
        ctx.set_tables(np.zeros(2), np.zeros(2), np.array([0.0, 2.0 * rc]))
        ctx.set_geometry(x, box, ng, P, xi, eta, rc)
        ctx.set_sizes(a)
        phi = 4.0 * np.pi * float(np.sum(a ** 3)) / (3.0 * float(np.prod(box)))
        for e in eps:
            if abs(e - 1.0) <= 1e-12:
                continue
            ctx.set_permittivity(np.full(n, e))
            g = batch._guesses(a, np.array([e]), fields, phi)
            for f in range(3):
                b = np.broadcast_to(fields[f], (1, n, 3)).copy()
                ctx.sizes_solve(b, g[f:f + 1], gmres_tol, restart=min(3 * n, 10), maxit=min(3 * n, 100))
                nrhs += 1
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    ctx.close()
    return secs, nrhs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--check", type=int, default=4, help="configurations re-solved in FP64 for the error")
    ap.add_argument("--tol", type=float, default=1e-3, help="Ewald rule tolerance")
    ap.add_argument("--gmres-tol", type=float, default=1e-8,
                    help="GMRES tolerance (1e-8: mixed dipoles within 1e-6 of tight FP64 solves)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--fp64", action="store_true", help="also time the whole set with the FP64 operator throughout")
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="configurations also solved one right-hand side at a time in FP64 on one stream")
    args = ap.parse_args(argv)

    import workloads
    from mpol_hip import batch

    t0 = time.perf_counter()
    configs, eps = workloads.heterogeneous_set(count=args.count, seed=args.seed)
    pack_s = time.perf_counter() - t0
    print(f"[heterogeneous] {len(configs)} configurations packed in {pack_s:.1f} s", file=sys.stderr, flush=True)
    grids, tune_s = None, 0.0
    if not args.no_tune:
        grids, _, tune_s = batch.tune_grids(configs, eps, tol=args.tol)
        print(f"[heterogeneous] grids chosen in {tune_s:.1f} s", file=sys.stderr, flush=True)
    res = batch.solve_configurations(configs, eps, tol=args.tol, gmres_tol=args.gmres_tol, precision="mixed",
                                     workers=args.workers, tune=False, grids=grids)
    print(f"[heterogeneous] solved in {res.seconds:.2f} s", file=sys.stderr, flush=True)
    err = None
    if args.check > 0:
        # the largest configurations by N, at the grids the mixed run chose
        idx = sorted(range(len(configs)), key=lambda i: -configs[i]["x"].shape[0])[:args.check]
        sub = [configs[i] for i in idx]
        r64 = batch.solve_configurations(sub, eps, tol=args.tol, gmres_tol=1e-10, precision="fp64",
                                         workers=args.workers, grids=[int(res.ngrid[i]) for i in idx])
        err = max(float(np.max(np.abs(res.p[i] - r64.p[j])) / np.max(np.abs(r64.p[j]))) for j, i in enumerate(idx))
    plain = None
    if args.fp64:
        # the same set, grids and tolerance with the FP64 operator in every iteration
        rp = batch.solve_configurations(configs, eps, tol=args.tol, gmres_tol=args.gmres_tol, precision="fp64",
                                        workers=args.workers, grids=[int(g) for g in res.ngrid])
        print(f"[heterogeneous] FP64 throughout: {rp.seconds:.2f} s", file=sys.stderr, flush=True)
        live = int(np.sum(np.abs(eps - 1.0) > 1e-12))
        plain = {"seconds": round(rp.seconds, 3), "rhs_per_s": round(len(configs) * live * 3 / rp.seconds, 2),
                 "matvecs_per_config_mean": float(rp.matvecs[:, 0].mean()),
                 "speedup_mixed_vs_fp64": round(rp.seconds / res.seconds, 3)}
    plain1 = None
    if args.plain_sample > 0:
        idx = list(range(0, len(configs), max(1, len(configs) // args.plain_sample)))[:args.plain_sample]
        secs, nrhs = plain_fp64([configs[i] for i in idx], eps, [int(res.ngrid[i]) for i in idx], args.tol,
                                args.gmres_tol)
        share = sum(configs[i]["x"].shape[0] for i in idx) / sum(c["x"].shape[0] for c in configs)
        plain1 = {"configs": len(idx), "rhs": nrhs, "seconds": round(secs, 3), "rhs_per_s": round(nrhs / secs, 2),
                  "whole_set_seconds_extrapolated_by_particles": round(secs / share, 2),
                  "speedup_mixed_batched_vs_this": round(secs / share / res.seconds, 2)}
    live = int(np.sum(np.abs(eps - 1.0) > 1e-12))
    grids, counts = np.unique(res.ngrid, return_counts=True)
    out = {
        "workload": "heterogeneous polydisperse configurations",
        "configs": len(configs),
        "particles": int(sum(c["x"].shape[0] for c in configs)),
        "rhs": res.rhs,
        "rhs_solved": len(configs) * live * 3,
        "seconds": round(res.seconds, 3),
        "rhs_per_s": round(len(configs) * live * 3 / res.seconds, 2),  # solved right-hand sides only
        "grid_tuning_seconds": round(tune_s, 3),
        "workers": args.workers,
        "ewald_tol": args.tol,
        "gmres_tol": args.gmres_tol,
        "grids": {int(g): int(c) for g, c in zip(grids, counts)},
        "matvecs_per_config_mean": {"fp64": float(res.matvecs[:, 0].mean()), "fp32": float(res.matvecs[:, 1].mean())},
        "converged_fraction": float(res.converged.sum() / max(1, len(configs) * live * 3)),
        "mixed_rel_dipole_err_vs_fp64": err,
        "pack_seconds": round(pack_s, 2),
        "fp64_throughout": plain,
        "fp64_one_rhs_at_a_time": plain1,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
