"""Benchmark: polarization solves/sec and matvecs/sec at N = 1e5 on one MI355X.

Workload (synthetic, seeded): N = 1e5 equal unit spheres at volume fraction
0.4 in a cubic periodic box (FCC lattice with random vacancies + hard-sphere
Monte Carlo sweeps), beta = 0.9 (eps = 28), unit field along z, 256^3 grid,
P = 10, xi = 1.0612 and rc = 3.503 (the largest xi that gives 256 nodes at
tol 1e-6), slab_factor 0.

One step = one complete polarization solve from the Maxwell-Garnett guess
with the mixed-precision path (FP32 operator inside FP64 iterative
refinement, FP64 Arnoldi/Givens), at the loosest residual tolerance of the
ladder 1e-6, 5e-7, 2e-7, 1e-7, 5e-8 whose dipoles are within 1e-6 (max-norm,
relative) of a tight (tol 1e-10) FP64 solve -- chosen untimed, per path.
The FP32 correction solves run either on the same operator (one level) or
on a coarser one (two levels: a smaller grid at a looser Ewald tolerance,
chosen untimed by mpol_hip.autotune.tune_inner); the FP64 residuals, which
decide convergence, are always on the 256^3, P = 10 operator.  The faster of
the two at the error bar is the step; both are reported.
Inputs are resident in HBM before the timed region.  Also reported: the plain
FP64 solve at its own such tolerance, a batched solve of the three field
directions (capacitance tensor), matvec throughput in both precisions,
per-stage kernel times (HIP events), the roofline of the dominant kernel (its
HBM traffic from the committed rocprofv3 counter summary when one matches),
and the FP64 NumPy oracle timed on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "mutual-polarization-slab_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "n1e5_phi0.4"


def stage_bytes(n, ng, P, npairs, nrhs, fbytes, rank1=True):
    """Algorithmic HBM bytes of each operator stage for one launch (nrhs
    right-hand sides): the half spectrum is 3 components x nz*ny*(nx/2+1)
    complex values; particle data is the stencil factors (3P values), the
    stencil start (3 int32), the permutation (int32) and the FP64 dipoles.
    rank1: the rank-one scheme of the y and z passes (x and y components as
    one grid between the y passes: 5/3, 4/3 and 5/3 of a read + write of the
    spectrum instead of 2 each)."""
    nx, ny, nz = ng
    kx = nx // 2 + 1
    spec = 3 * nrhs * nz * ny * kx * 2 * fbytes
    part = n * (3 * P * fbytes + 16)
    return {
        "slab_sum": nrhs * n * 24,
        "real_space": npairs * (4 + 5 * fbytes) + n * (4 + 8) + nrhs * n * (24 + 24),
        "spread_xfft": spec + part + nrhs * n * 24,
        "fft_y_fwd": 5 * spec // 3 if rank1 else 2 * spec,
        "fft_z_scale": 4 * spec // 3 if rank1 else 2 * spec,
        "fft_y_inv": 5 * spec // 3 if rank1 else 2 * spec,
        "xifft_contract": spec + part,
        "contract_reduce": nrhs * n * (4 + 24),
    }


def canonical_matvec_bytes(n, ng, s):
    """6*C*G*s + N*(12 + 2*C_p*s) with C = C_p = 3."""
    G = int(np.prod(ng))
    return 6 * 3 * G * s + n * (12 + 2 * 3 * s)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, perp, para, rv, gpu_fp64_matvec, full_solve_tol=None, guess=None):
    """FP64 NumPy oracle on the host (numpy FFT / bincount: one core): the
    precompute, then the median of three matvecs of the same workload, and
    optionally (full_solve_tol) one complete SciPy GMRES(10) solve from the
    guess, as the reference solves."""
    from oracle import capacitance_oracle as co
    from oracle import ewald_oracle as eo
    t0 = time.perf_counter()
    geo = eo.OperatorGeometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["rc"], perp, para, rv)
    t1 = time.perf_counter()
    rng = np.random.default_rng(5)
    times, err = [], None
    for k in range(3):
        m = rng.standard_normal((cfg["x"].shape[0], 3))
        ta = time.perf_counter()
        out = np.real(eo.field(geo, m, cfg["eps"], 0.0))
        times.append(time.perf_counter() - ta)
        if k == 0:
            gpu = gpu_fp64_matvec(m)
            err = float(np.max(np.abs(gpu - out)) / np.max(np.abs(out)))
    res = {"precompute_s": t1 - t0, "matvec_s": float(np.median(times)), "matvec_s_all": times,
           "gpu_fp64_matvec_relerr": err}
    if full_solve_tol is not None:
        log(f"CPU baseline: matvecs {', '.join(f'{t:.1f}' for t in times)} s; full FP64 oracle solve")
        ts = time.perf_counter()
        _, info = co.solve_dipoles(geo, cfg["eps"], np.array([0.0, 0.0, 1.0]), guess, full_solve_tol,
                                   slab_factor=0.0, use_scipy_gmres=True)
        res["full_solve_s"] = time.perf_counter() - ts
        res["full_solve_info"] = {k: v for k, v in info.items() if isinstance(v, (int, float))}
    return res


# GMRES tolerance of the heterogeneous set's solves, chosen so that the
# mixed dipoles are within 1e-6 of tight FP64 solves (at 1e-6 they differ by
# 1.3e-5: the set reaches eps = 100); the line reports the measured error
HET_GMRES_TOL = 1e-8


def heterogeneous_sample(count=256, plain=16, check=4, workers=8):
    """Many independent polydisperse configurations (N log-uniform in
    [500, 8000], phi in [0.1, 0.45], 16-point permittivity spectrum x 3
    fields each; 256 of them by default, 12,288 right-hand sides), every
    solve at GMRES tolerance HET_GMRES_TOL: the batched mixed-precision
    driver over the whole set against (a) the same driver with the FP64
    operator throughout (batched, eight workers) and (b) `plain` of the
    configurations (every count/plain-th) solved one right-hand side at a
    time in FP64 on one stream, compared by rate and, scaled by particle
    count, as the whole set's time; the mixed dipoles of the `check` largest
    configurations against tight FP64 solves.  Grid choice is timed apart."""
    import workloads
    from mpol_hip import batch
    from workloads.heterogeneous import plain_fp64
    gt = HET_GMRES_TOL
    configs, eps = workloads.heterogeneous_set(count=count, seed=7)
    grids, _, tune_s = batch.tune_grids(configs, eps, tol=1e-3)
    res = batch.solve_configurations(configs, eps, tol=1e-3, gmres_tol=gt, precision="mixed", workers=workers,
                                     tune=False, grids=grids)
    r64 = batch.solve_configurations(configs, eps, tol=1e-3, gmres_tol=gt, precision="fp64", workers=workers,
                                     tune=False, grids=grids)
    big = sorted(range(count), key=lambda i: -configs[i]["x"].shape[0])[:check]
    rt = batch.solve_configurations([configs[i] for i in big], eps, tol=1e-3, gmres_tol=1e-11, precision="fp64",
                                    workers=workers, tune=False, grids=[int(grids[i]) for i in big])

    def err(r, i, j):
        return float(np.max(np.abs(r.p[i] - rt.p[j])) / np.max(np.abs(rt.p[j])))

    idx = list(range(0, count, count // plain))[:plain]
    secs, nr = plain_fp64([configs[i] for i in idx], eps, [int(grids[i]) for i in idx], 1e-3, gt)
    live = int(np.sum(np.abs(eps - 1.0) > 1e-12))
    solved = count * live * 3  # eps = 1 points (zero dipoles, no solve) not counted
    rate = solved / res.seconds
    plain_rate = nr / secs
    # the sample's share of the set's particles scales its time to the whole set
    share = sum(configs[i]["x"].shape[0] for i in idx) / sum(c["x"].shape[0] for c in configs)
    return {"configs": count, "particles": int(sum(c["x"].shape[0] for c in configs)), "rhs": res.rhs,
            "rhs_solved": solved, "ewald_tol": 1e-3, "gmres_tol": gt, "workers": workers,
            "batched_mixed_rhs_per_s": round(rate, 2), "seconds": round(res.seconds, 3),
            "grid_tuning_s": round(tune_s, 3),
            "mixed_rel_dipole_err_vs_tight_fp64": max(err(res, i, j) for j, i in enumerate(big)),
            "fp64_batched_rhs_per_s": round(solved / r64.seconds, 2), "fp64_batched_seconds": round(r64.seconds, 3),
            "fp64_batched_rel_dipole_err_vs_tight_fp64": max(err(r64, i, j) for j, i in enumerate(big)),
            "speedup_batched_mixed_vs_fp64_batched": round(r64.seconds / res.seconds, 2),
            "fp64_one_rhs_at_a_time_rhs_per_s": round(plain_rate, 2), "fp64_one_rhs_sample_configs": len(idx),
            "speedup_batched_mixed_vs_fp64_one_rhs": round(rate / plain_rate, 2),
            "fp64_one_rhs_whole_set_seconds_extrapolated": round(secs / share, 2),
            "speedup_whole_set_extrapolated_by_particles": round(secs / share / res.seconds, 2),
            "grids": {int(g): int(c) for g, c in zip(*np.unique(grids, return_counts=True))},
            "matvecs_per_config_mean": {"mixed_fp32": float(res.matvecs[:, 1].mean()),
                                        "mixed_fp64": float(res.matvecs[:, 0].mean()),
                                        "fp64": float(r64.matvecs[:, 0].mean())},
            "converged_fraction": float(res.converged.sum() / max(1, solved))}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TOL_LADDER = (1e-6, 5e-7, 2e-7, 1e-7, 5e-8)
ERR_BAR = 1e-6
TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_n1e5_g256.json")
COARSE_TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_n1e5_g128_coarse.json")


def measured_traffic(kernel_name, path=TRAFFIC_FILE):
    """HBM bytes per launch of a kernel from the committed counter summary
    (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel), or None."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    ent = t.get("kernels", {}).get(kernel_name)
    return None if ent is None else ent.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full-solve", action="store_true",
                    help="also time one complete FP64 oracle solve on the host (about 2.5 minutes)")
    ap.add_argument("--no-heterogeneous", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-other-workloads", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); there is no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import workloads
    from mpol_hip import FP32, FP64, Context, default_tables

    dev = torch.device("cuda", 0)
    log("building workload")
    cfg = workloads.configuration(WORKLOAD, sweeps=args.sweeps)
    log("geometry precompute on the GPU")
    n = cfg["x"].shape[0]
    perp, para, rv = default_tables(cfg["xi"])
    ctx = Context(dev)
    ctx.set_tables(perp, para, rv)
    t0 = time.perf_counter()
    ctx.set_geometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["eta"], cfg["rc"])
    ctx.set_permittivity(cfg["eps"])
    torch.cuda.synchronize()
    geo_s = time.perf_counter() - t0
    ginfo = ctx.geometry_info()
    # the tile images (per precision) are built by the first operator
    # application after set_geometry: its time minus a repeat's
    image_ms = {}
    mprobe = torch.zeros((1, n, 3), device=dev, dtype=torch.float64)
    for name, prec in (("fp32", FP32), ("fp64", FP64)):
        tt = []
        for _ in range(2):
            torch.cuda.synchronize()
            ta = time.perf_counter()
            ctx.matvec(mprobe, precision=prec, slab_factor=0.0)
            torch.cuda.synchronize()
            tt.append(time.perf_counter() - ta)
        image_ms[name] = round(1e3 * (tt[0] - tt[1]), 3)
    log("autotuning launch shapes (untimed)")
    from mpol_hip.autotune import tune_launch
    knobs = {"fp32": tune_launch(ctx, FP32), "fp64": tune_launch(ctx, FP64)}
    log(f"launch knobs {knobs}")
    slab = 0.0
    phi = 4.0 * np.pi * n / (3.0 * np.prod(cfg["box"]))
    beta = cfg["beta"]
    mg = 4.0 * np.pi * beta / (1.0 - beta * phi)  # Maxwell-Garnett guess factor
    Hhat = np.array([0.0, 0.0, 1.0])
    b = torch.as_tensor(np.tile(Hhat, n).reshape(1, n, 3), device=dev)
    x0 = mg * b

    def solve(mixed, t, bb=b, xx=x0):
        return ctx.solve(bb, xx, t, restart=10, maxit=100, mixed=mixed, slab_factor=slab)

    def rel_err(x, ref):
        return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))

    # reference solution: FP64 at tight tolerance
    log("FP64 reference solve (tol 1e-10)")
    ref = solve(False, 1e-10)
    p_ref = ref.x.cpu().numpy()

    def calibrate(mixed):
        """Loosest ladder tolerance whose dipoles meet the error bar."""
        for t in TOL_LADDER:
            r = solve(mixed, t)
            e = rel_err(r.x.cpu().numpy(), p_ref)
            if e <= ERR_BAR:
                return t, e
        return TOL_LADDER[-1], e

    def timed_mixed(t, label):
        log(f"timed mixed solves ({label}) at tol {t:g}: {args.warmup} warmup + {args.steps}")
        for _ in range(args.warmup):
            solve(True, t)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = None
        for _ in range(args.steps):
            res = solve(True, t)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / args.steps, res

    # one-level mixed scheme: correction solves on this operator in FP32
    tol1, _ = calibrate(True)
    ms1, last1 = timed_mixed(tol1, "FP32 correction solves on the same operator")
    err1 = rel_err(last1.x.cpu().numpy(), p_ref)
    one_level = {"solves_per_s": round(1e3 / ms1, 3), "ms_per_solve": round(ms1, 3), "tol": tol1,
                 "rel_dipole_err_vs_fp64": err1,
                 "matvecs_per_solve": {"fp32": last1.matvecs_fp32, "fp64": last1.matvecs_fp64}}
    # two-level: correction solves on a coarser operator (autotuned, untimed),
    # FP64 residuals on this one
    log("correction operator: timing the coarse candidates at the error bar (untimed)")
    from mpol_hip.autotune import tune_inner
    inner, inner_cfg, inner_report = tune_inner(
        ctx, cfg["x"], cfg["box"], cfg["eps"], b, x0, TOL_LADDER,
        accept=lambda r: rel_err(r.x.cpu().numpy(), p_ref) <= ERR_BAR, slab_factor=slab)
    two_level = None
    tol, ms_step, last, mixed_err = tol1, ms1, last1, err1
    if inner is not None:
        tol2 = inner_cfg["tol"]
        ms2, last2 = timed_mixed(tol2, f"FP32 correction solves on the coarse operator {inner_cfg}")
        err2 = rel_err(last2.x.cpu().numpy(), p_ref)
        two_level = {"solves_per_s": round(1e3 / ms2, 3), "ms_per_solve": round(ms2, 3), "tol": tol2,
                     "rel_dipole_err_vs_fp64": err2, "correction_operator": inner_cfg,
                     "matvecs_per_solve": {"coarse_fp32": last2.matvecs_inner,
                                           "fp32": last2.matvecs_fp32 - last2.matvecs_inner,
                                           "fp64": last2.matvecs_fp64},
                     "refinements": last2.refinements}
        if ms2 < ms1 and err2 <= ERR_BAR:
            tol, ms_step, last, mixed_err = tol2, ms2, last2, err2
        if last is last1:
            ctx.set_inner_operator(None)
    inner_used = last is not last1

    # plain FP64 path at its own tolerance
    tol64, _ = calibrate(False)
    log(f"plain FP64 solves at tol {tol64:g}")
    n64 = max(1, min(3, args.steps))
    solve(False, tol64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n64):
        r64 = solve(False, tol64)
    torch.cuda.synchronize()
    ms64 = 1e3 * (time.perf_counter() - t0) / n64
    fp64_err = rel_err(r64.x.cpu().numpy(), p_ref)

    # the three field directions of the capacitance tensor in one batched solve
    log("batched three-field mixed solves")
    b3 = torch.as_tensor(np.stack([np.tile(e, n).reshape(n, 3) for e in np.eye(3)]), device=dev)
    x03 = mg * b3
    ref3 = solve(False, 1e-10, b3, x03).x.cpu().numpy()
    # its own loosest ladder tolerance at the error bar (the worst field decides)
    tol3 = TOL_LADDER[-1]
    for t in TOL_LADDER:
        if rel_err(solve(True, t, b3, x03).x.cpu().numpy(), ref3) <= ERR_BAR:
            tol3 = t
            break
    solve(True, tol3, b3, x03)
    torch.cuda.synchronize()
    nb = max(1, min(3, args.steps))
    t0 = time.perf_counter()
    for _ in range(nb):
        r3 = solve(True, tol3, b3, x03)
    torch.cuda.synchronize()
    ms3 = 1e3 * (time.perf_counter() - t0) / nb
    batched = {"rhs": 3, "solves_per_s": round(3e3 / ms3, 3), "ms_per_batch": round(ms3, 3), "tol": tol3,
               "rel_dipole_err_vs_fp64": rel_err(r3.x.cpu().numpy(), ref3),
               "matvecs_per_batch": {"coarse_fp32": r3.matvecs_inner, "fp32": r3.matvecs_fp32 - r3.matvecs_inner,
                                     "fp64": r3.matvecs_fp64}}

    # matvec throughput and per-stage times
    m = torch.randn((1, n, 3), device=dev, dtype=torch.float64)
    out = {}
    for name, prec in (("fp32", FP32), ("fp64", FP64)):
        for _ in range(3):
            ctx.matvec(m, precision=prec, slab_factor=slab)
        torch.cuda.synchronize()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.matvec(m, precision=prec, slab_factor=slab)
        torch.cuda.synchronize()
        out[name] = reps / (time.perf_counter() - t0)
    stages32, _ = ctx.matvec_timed(m, precision=FP32, slab_factor=slab, repeats=20)
    stages64, _ = ctx.matvec_timed(m, precision=FP64, slab_factor=slab, repeats=10)
    by32 = stage_bytes(n, cfg["ngrid"], cfg["P"], ginfo["npairs"], 1, 4)
    dom = max(stages32, key=stages32.get)
    achieved = by32[dom] / (stages32[dom] * 1e-3) / 1e9
    kernel_names = {"spread_xfft": "k_spread_mfma_xfft", "xifft_contract": "k_xifft_contract_mfma",
                    "fft_y_fwd": "k_fft_y_xy", "fft_y_inv": "k_fft_y_xy", "fft_z_scale": "k_fft_z_scale_reg256_rank1",
                    "real_space": "k_real_space", "slab_sum": "k_sum_z", "contract_reduce": "k_contract_reduce"}
    roofline_full = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(kernel_names.get(dom, dom)),
                     "algorithmic_bytes": by32[dom]}
    roofline, solve_stage_ms = roofline_full, None
    if inner_used:
        # the timed step's kernel time by (operator, stage): stage time x
        # applications per solve; the roofline is the largest one's
        stages_c, _ = inner.matvec_timed(m, precision=FP32, slab_factor=slab, repeats=20)
        gi = inner.geometry_info()
        n_in = last.matvecs_inner
        share = {}
        for tag, st_ms, cnt in (("coarse", stages_c, n_in), ("fp32", stages32, last.matvecs_fp32 - n_in),
                                ("fp64", stages64, last.matvecs_fp64)):
            for k, v in st_ms.items():
                share[f"{tag}:{k}"] = v * cnt
        solve_stage_ms = {k: round(v, 4) for k, v in sorted(share.items(), key=lambda kv: -kv[1])}
        tag, sname = max(share, key=share.get).split(":")
        if tag == "coarse":
            G, rows = int(gi["ngrid"][0]), int(gi["tile_rows"])
            byc = stage_bytes(n, gi["ngrid"], gi["P"], gi["npairs"], 1, 4, rank1=False)
            ach = byc[sname] / (stages_c[sname] * 1e-3) / 1e9
            cnames = {"spread_xfft": f"k_spread_mfma_xfft<float, {G}, 3, {rows}>",
                      "xifft_contract": f"k_xifft_contract_mfma<float, {G}, 3, {rows}>",
                      "fft_y_fwd": "k_fft_y", "fft_y_inv": "k_fft_y", "fft_z_scale": "k_fft_z_scale",
                      "real_space": "k_real_space", "contract_reduce": "k_contract_reduce"}
            roofline = {"bound": "hbm", "kernel": f"{sname} of the {G}^3, P={gi['P']} correction operator "
                                                  f"({cnames.get(sname, sname)}, FP32)",
                        "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(cnames.get(sname, sname), COARSE_TRAFFIC_FILE),
                        "algorithmic_bytes": byc[sname], "stage_ms": round(stages_c[sname], 4)}
        elif tag == "fp32":
            roofline = dict(roofline_full, kernel=f"{dom} of the full operator (FP32)")
    stage_gbs = {k: round(by32[k] / (stages32[k] * 1e-3) / 1e9, 1) for k in stages32 if stages32[k] > 0}
    bmv32 = canonical_matvec_bytes(n, cfg["ngrid"], 4)
    bmv64 = canonical_matvec_bytes(n, cfg["ngrid"], 8)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline: FP64 NumPy oracle (precompute + 3 matvecs, ~35 s"
            + ("; then one full solve, ~2.5 min)" if args.cpu_full_solve else ")"))

        def gpu_mv(mm):
            return ctx.matvec(mm[None], precision=FP64, slab_factor=0.0).cpu().numpy()[0]
        cb = cpu_baseline(cfg, perp, para, rv, gpu_mv, full_solve_tol=tol64 if args.cpu_full_solve else None,
                          guess=x0.cpu().numpy()[0])
        r64_mv = r64.matvecs_fp64
        if "full_solve_s" in cb:
            value = 1.0 / cb["full_solve_s"]
            sample = (f"one complete FP64 NumPy-oracle solve of this workload (SciPy GMRES(10) at tol {tol64:g} "
                      f"from the same guess, {cb['full_solve_s']:.1f} s, precompute {cb['precompute_s']:.1f} s "
                      "excluded)")
        else:
            value = 1.0 / (cb["matvec_s"] * r64_mv)
            sample = (f"median of 3 FP64 NumPy-oracle matvecs of this workload ({cb['matvec_s']:.2f} s each, "
                      f"precompute {cb['precompute_s']:.1f} s excluded) x {r64_mv} matvecs of the GPU FP64 solve "
                      f"at tol {tol64:g}: extrapolated (--cpu-full-solve times a complete solve)")
        cpu = {"value": round(value, 6), "unit": "solves/s", "cores": 1, "kind": "port", "sample": sample,
               "cpu_model": cpu_model(), "host_cpus": len(os.sched_getaffinity(0)),
               "threads_env": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                                                "MKL_NUM_THREADS")},
               "note": "numpy pocketfft, bincount and add.at run on one core",
               "precompute_s": round(cb["precompute_s"], 2),
               "matvec_s_median": round(cb["matvec_s"], 3),
               "matvec_s_all": [round(t, 3) for t in cb["matvec_s_all"]],
               "matvecs_per_s": round(1.0 / cb["matvec_s"], 5),
               "full_solve_s": round(cb["full_solve_s"], 2) if "full_solve_s" in cb else None,
               "gpu_fp64_matvec_relerr_vs_oracle": cb["gpu_fp64_matvec_relerr"]}

    adj = None
    if not args.no_adjoint:
        log("adjoint forces and dU/dbeta vs central finite differences (32 particles, h = 1e-4)")
        from workloads import adjoint_check
        adj, _ = adjoint_check.run(particles=32, h=1e-4, solve_tol=1e-12, sweeps=args.sweeps, verbose=False)
    others = None
    if not args.no_other_workloads:
        log("other full-size workloads: three-field tensor (N=1e4, 128^3), dipole+quadrupole (N=2e4, 160^3)")
        from workloads import configs as other_configs
        others = {"tensor": other_configs.tensor(sweeps=args.sweeps), "quadrupole": other_configs.quadrupole()}
    het = None
    if not args.no_heterogeneous:
        log("heterogeneous polydisperse configurations: batched mixed vs one right-hand side at a time in FP64")
        het = heterogeneous_sample()

    solves_per_s = 1e3 / ms_step
    line = {
        "metric": "polarization solves/sec & matvecs/sec at N=1e5, rel err <=1e-6 vs FP64 oracle",
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 operator / f64 Krylov+refinement",
        "data": "synthetic",
        "config": {"workload": "N=1e5 unit spheres, phi=0.4, beta=0.9, E0=z, 256^3, P=10, xi=1.0612, "
                               "rc=3.503, slab_factor=0, 1 rhs, dipole error <= 1e-6; positions: FCC lattice "
                               f"with random vacancies + {cfg['sweeps']} hard-sphere MC sweeps (the workload "
                               "names RSA + 1e4 relaxation sweeps; RSA saturates below phi = 0.4)"
                               + ("; mixed scheme: FP64 residuals on this operator, FP32 correction solves on "
                                  f"a coarser one ({inner_cfg['grid']}^3, P={inner_cfg['P']}, rule tol "
                                  f"{inner_cfg['ewald_tol']:g}, floor {inner_cfg['floor']:g})" if inner_used else
                                  "; mixed scheme: FP64 residuals, FP32 correction solves on this operator"),
                   "n": n, "ngrid": [int(v) for v in cfg["ngrid"]], "P": cfg["P"], "xi": round(cfg["xi"], 6),
                   "rc": round(cfg["rc"], 6), "tol": tol, "seed": cfg["seed"], "mc_sweeps": cfg["sweeps"],
                   "npairs": ginfo["npairs"]},
        "matvecs_per_solve": {"coarse_fp32": last.matvecs_inner, "fp32": last.matvecs_fp32 - last.matvecs_inner,
                              "fp64": last.matvecs_fp64},
        "mixed_rel_dipole_err_vs_fp64": mixed_err,
        "mixed_one_level": one_level,
        "mixed_two_level": two_level,
        "correction_operator_report": [{"setting": lab, "tol": t, "ms": round(ms, 3), "coarse_fp32": mi,
                                        "fp32": m32 - mi, "fp64": m64}
                                       for lab, t, ms, m32, m64, mi in inner_report],
        "fp64_solves_per_s": round(1e3 / ms64, 3),
        "fp64_tol": tol64,
        "fp64_rel_dipole_err_vs_tight": fp64_err,
        "fp64_matvecs_per_solve": r64.matvecs_fp64,
        "speedup_mixed_vs_fp64": round(ms64 / ms_step, 3),
        "speedup_one_level_mixed_vs_fp64": round(ms64 / ms1, 3),
        "batched_three_fields": batched,
        "speedup_batched_mixed_vs_fp64": round((3e3 / ms3) / (1e3 / ms64), 3),
        "matvecs_per_s": {k: round(v, 1) for k, v in out.items()},
        "matvec_hbm_frac": {"fp32": round(bmv32 * out["fp32"] / 8e12, 4), "fp64": round(bmv64 * out["fp64"] / 8e12, 4)},
        "stage_ms_fp32": {k: round(v, 4) for k, v in stages32.items()},
        "stage_ms_fp64": {k: round(v, 4) for k, v in stages64.items()},
        "stage_gbs_fp32": stage_gbs,
        "geometry_s": round(geo_s, 3),
        "tile_image_build_ms": image_ms,
        "launch_knobs": knobs,
        "batched_heterogeneous": het,
        "adjoint_vs_finite_differences": adj,
        "other_workloads": others,
        "roofline": roofline,
        "roofline_full_operator_fp32": roofline_full,
        "solve_kernel_ms_by_stage": solve_stage_ms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line, default=lambda o: o.item() if hasattr(o, "item") else str(o)))


if __name__ == "__main__":
    main()
