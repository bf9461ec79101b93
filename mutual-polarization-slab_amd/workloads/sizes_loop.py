"""Polydisperse operator driver for kernel profiles: one configuration of the
heterogeneous workload's kind (N spheres, log-normal radii sigma 0.2, volume
fraction phi), its 16-point permittivity spectrum x 3 fields batched (45
right-hand sides: the eps = 1 point has no solve), the grid of size G chosen
by the many-configuration driver's rule at tolerance 1e-3, then repeated
batched operator applications and one mixed solve on one stream.

    python -m workloads.sizes_loop [--n 3000] [--phi 0.3] [--grid 32] [--precision fp32|fp64]
                                   [--reps 20] [--engine 0|2] [--seed 11] [--copy-mb 0]

(run from mutual-polarization-slab_amd/, or with it on sys.path).  engine 2
selects the per-particle kernels (atomic spread, LDS-staged contraction) and
the separate y/z passes instead of the stored-entry kernels and the
one-kernel y/z passes.  Prints one line per step.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=3000)
    ap.add_argument("--phi", type=float, default=0.3)
    ap.add_argument("--grid", type=int, default=32)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--engine", type=int, default=0)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first copy a float32 device buffer of this many MiB (counter calibration)")
    args = ap.parse_args(argv)

    import torch

    import workloads
    from mpol_hip import FP32, FP64, batch

    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(f"copied {args.copy_mb} MiB device to device", flush=True)
    x, a, box = workloads.polydisperse(args.n, args.phi, args.seed)
    eps = np.logspace(-1.0, 2.0, 16)
    live = eps[np.abs(eps - 1.0) > 1e-12]
    fields = np.eye(3)
    cand = batch.grid_candidates(box, 1e-3, (args.grid,))
    if not cand:
        raise SystemExit(f"grid {args.grid} does not fit box {box}")
    xi, ng, P, eta, rc = cand[0]
    print(f"n={args.n} phi={args.phi} grid={list(ng)} P={P} rhs={live.size * 3}", flush=True)
    w = batch._Worker(torch.device("cuda", 0))
    sets = np.repeat(np.broadcast_to(live[:, None], (live.size, args.n)), 3, axis=0)
    prec = FP32 if args.precision == "fp32" else FP64
    with torch.cuda.stream(w.stream):
        w._setup(x, a, box, cand[0], sets)
        if args.engine:
            w.ctx.set_tuning(engine=args.engine)
        phi = 4.0 * np.pi * float(np.sum(a ** 3)) / (3.0 * float(np.prod(box)))
        g = torch.as_tensor(batch._guesses(a, live, fields, phi), device=w.ctx.device)
        w.ctx.sizes_matvec(g, precision=prec)
        w.stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.reps):
            w.ctx.sizes_matvec(g, precision=prec)
        w.stream.synchronize()
        print(f"batched operator: {1e3 * (time.perf_counter() - t0) / args.reps:.4f} ms", flush=True)
        b = np.ascontiguousarray(np.broadcast_to(np.repeat(fields[None], live.size, 0).reshape(-1, 3)[:, None, :],
                                                 (live.size * 3, args.n, 3)))
        t0 = time.perf_counter()
        r = w.ctx.sizes_solve(b, g, 1e-6, restart=10, maxit=100, mixed=True)
        w.stream.synchronize()
        print(f"mixed solve: {1e3 * (time.perf_counter() - t0):.2f} ms, operator applications "
              f"{r.matvecs_fp32} FP32 + {r.matvecs_fp64} FP64, converged {r.converged}/{live.size * 3}", flush=True)
    w.close()


if __name__ == "__main__":
    main()
