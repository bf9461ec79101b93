"""Solve driver for kernel traces: the benchmark workload (N = 1e5 unit
spheres, phi = 0.4, beta = 0.9, 256^3, P = 10, field along z, slab factor 0),
a few mixed solves at tol 5e-7, two-level mixed solves (correction solves on
the 128^3, P = 6 operator at rule tolerance 1e-4, floor 1e-2) at tol 1e-6
and plain FP64 solves at tol 2e-7 (the tolerances that meet the 1e-6 dipole
bar), with default launch shapes.  Each solve is bracketed by device
synchronisations, a printed line and a 10 ms host pause, so a trace of this
run separates the solves.

    python -m workloads.solve_loop [--mixed 3] [--two-level 3] [--fp64 2] [--nrhs 1] [--sweeps 200]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mixed", type=int, default=3)
    ap.add_argument("--two-level", type=int, default=0)
    ap.add_argument("--fp64", type=int, default=2)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=200)
    args = ap.parse_args(argv)

    import numpy as np
    import torch

    import workloads
    from mpol_hip import Context, default_tables

    cfg = workloads.configuration("n1e5_phi0.4", sweeps=args.sweeps)
    n = cfg["x"].shape[0]
    perp, para, rv = default_tables(cfg["xi"])
    ctx = Context()
    ctx.set_tables(perp, para, rv)
    ctx.set_geometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["eta"], cfg["rc"])
    ctx.set_permittivity(cfg["eps"])
    phi = 4.0 * np.pi * n / (3.0 * np.prod(cfg["box"]))
    mg = 4.0 * np.pi * cfg["beta"] / (1.0 - cfg["beta"] * phi)
    fields = np.eye(3)[[2, 0, 1]][:args.nrhs]
    b = torch.as_tensor(np.stack([np.tile(f, n).reshape(n, 3) for f in fields]), device=ctx.device)
    x0 = mg * b
    print(f"workload n={n} grid={list(cfg['ngrid'])} P={cfg['P']} nrhs={args.nrhs}", flush=True)
    from mpol_hip.autotune import coarse_operator
    inner = None
    if args.two_level:
        inner, _ = coarse_operator(cfg["x"], cfg["box"], cfg["eps"], 1e-4, 128)
    for name, count, mixed, tol in (("mixed", args.mixed, True, 5e-7), ("two-level", args.two_level, True, 1e-6),
                                    ("fp64", args.fp64, False, 2e-7)):
        ctx.set_inner_operator(inner if name == "two-level" else None, 1e-2)
        for k in range(count):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = ctx.solve(b, x0, tol, restart=10, maxit=100, mixed=mixed, slab_factor=0.0)
            torch.cuda.synchronize()
            print(f"solve {name} {k + 1}/{count}: {1e3 * (time.perf_counter() - t0):.3f} ms, matvecs fp32 "
                  f"{r.matvecs_fp32} (inner {r.matvecs_inner}) fp64 {r.matvecs_fp64}, flag {r.flag}", flush=True)
            time.sleep(0.01)  # a host gap that separates the solves in a trace
    ctx.set_inner_operator(None)
    ctx.close()
    if inner is not None:
        inner.close()


if __name__ == "__main__":
    main()
