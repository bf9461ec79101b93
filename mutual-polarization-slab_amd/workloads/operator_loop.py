"""Operator driver for kernel profiles: builds a named workload, sets up one
context and applies the dipole operator a few times.

    python -m workloads.operator_loop [--workload n1e5_phi0.4] [--precision fp32|fp64]
                                      [--reps 3] [--nrhs 1] [--engine 0|1|2] [--sweeps 20]
                                      [--chunk 0] [--rhs-group 0] [--copy-mb 256]
                                      [--coarse 128 --coarse-tol 1e-4]

--coarse G: the workload's particles on the correction operator of the
two-level mixed solves instead (G nodes along the longest axis at rule
tolerance --coarse-tol, launch shapes tuned, as autotune.coarse_operator
builds it).

(run from mutual-polarization-slab_amd/, or with it on sys.path).  Prints
one line per step so a profiler run shows progress.
"""
from __future__ import annotations

import argparse
import os
import sys

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="n1e5_phi0.4")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--engine", type=int, default=0)
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--chunk", type=int, default=0, help="particles staged per chunk (0 = automatic)")
    ap.add_argument("--rhs-group", type=int, default=0, help="right-hand sides per tile workgroup (0 = automatic)")
    ap.add_argument("--coarse", type=int, default=0, help="grid of the coarse correction operator (0 = off)")
    ap.add_argument("--coarse-tol", type=float, default=1e-4)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first copy a float32 device buffer of this many MiB (counter calibration)")
    args = ap.parse_args(argv)

    import torch

    import workloads
    from mpol_hip import FP32, FP64, Context, default_tables

    cfg = workloads.configuration(args.workload, sweeps=args.sweeps)
    print(f"workload {args.workload}: n={cfg['x'].shape[0]} grid={list(cfg['ngrid'])} P={cfg['P']}", flush=True)
    if args.coarse:
        from mpol_hip.autotune import coarse_operator
        ctx, (xi, ng, P, _, rc) = coarse_operator(cfg["x"], cfg["box"], cfg["eps"], args.coarse_tol, args.coarse)
        print(f"coarse operator: grid={list(ng)} P={P} xi={xi:.4f} rc={rc:.3f}", flush=True)
    else:
        perp, para, rv = default_tables(cfg["xi"])
        ctx = Context()
        ctx.set_tables(perp, para, rv)
        ctx.set_geometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["eta"], cfg["rc"])
        ctx.set_permittivity(cfg["eps"])
        ctx.set_tuning(engine=args.engine, chunk=args.chunk, rhs_group=args.rhs_group)
    prec = FP32 if args.precision == "fp32" else FP64
    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {args.copy_mb} MiB device to device", flush=True)
        del src, dst
    g = torch.Generator().manual_seed(7)
    m = torch.randn((args.nrhs, cfg["x"].shape[0], 3), generator=g, dtype=torch.float64).to("cuda")
    for r in range(args.reps):
        out = ctx.matvec(m, precision=prec, slab_factor=0.0)
        torch.cuda.synchronize()
        print(f"matvec {r + 1}/{args.reps} ({args.precision}, {args.nrhs} rhs): |out| = {out.norm().item():.6e}",
              flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
