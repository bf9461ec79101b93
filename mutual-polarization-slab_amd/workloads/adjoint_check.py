"""Adjoint forces and dU/dbeta of the N=1e5 workload against central finite
differences.

U = -1/2 E0 . sum_i p_i for E0 = z.  The adjoint gradients come from
adjoint.polarization_energy (forward solve, adjoint solve, mpol_matvec_vjp).
The finite differences evaluate, at x +- h e (and beta +- h), the stationary
form U_var(p) = -b.p + 1/2 p.(A p), which equals U at the solution and whose
error is quadratic in the solve's error (a linear form would need the dipoles
to 1e-16 to difference at h = 1e-4); the sums are taken in extended
precision.  A coordinate whose step moves the particle's nearest grid node
(the spread's rounding, where U is not smooth) is skipped and counted.

The real-space couplings are interpolated linearly between table nodes (as
in the reference, "CPU - Ewald test/ewald.py":1009-1014), so U(x) has slope
jumps wherever a pair distance crosses a node; with the reference's 1e-3 grid
and ~50 neighbours per particle a step of 1e-4 crosses some node almost
always, and the jumps (~1e-3 of the pair force derivative) exceed 1e-5.  The
check therefore builds the tables on a grid of step --table-step (default
1e-6) up to rc + 0.1; the adjoint is exact for any table grid.

    python -m workloads.adjoint_check [--particles 32] [--h 1e-4] [--solve-tol 1e-12]

Prints one JSON line.
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


def _node(x, h):
    """Nearest grid node, rounding halves away from zero (the spread's rule)."""
    q = x / h
    return np.sign(q) * np.floor(np.abs(q) + 0.5)


def run(workload="n1e5_phi0.4", particles=32, h=1e-4, solve_tol=1e-12, seed=5, sweeps=200, table_step=1e-6,
        verbose=True):
    import torch

    import adjoint
    import workloads
    from mpol_hip import Context, real_space_table

    def log(msg):
        if verbose:
            print(f"[adjoint_check] {msg}", file=sys.stderr, flush=True)

    cfg = workloads.configuration(workload, sweeps=sweeps)
    x, box, eps = cfg["x"], cfg["box"], cfg["eps"]
    n = x.shape[0]
    beta = cfg["beta"]
    xi, ng, P, rc = cfg["xi"], cfg["ngrid"], cfg["P"], cfg["rc"]
    hg = box / ng
    dev = torch.device("cuda", 0)
    rt = np.arange(1, int(np.ceil((rc + 0.1) / table_step)) + 1, dtype=np.float64) * table_step
    tabs = real_space_table(rt, xi)
    tables = (tabs[2], tabs[3], np.concatenate(([0.0], rt)))
    # adjoint gradients
    t0 = time.perf_counter()
    xt = torch.as_tensor(x, device=dev).requires_grad_(True)
    et = torch.as_tensor(eps, device=dev).requires_grad_(True)
    U = adjoint.polarization_energy(xt, et, box, xi, errortol=cfg["tol"], ngrid=ng, P=P, rc=rc, solve_tol=solve_tol,
                                    tables=tables)
    gx, ge = torch.autograd.grad(U, (xt, et))
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    F_adj = -gx.cpu().numpy()
    deps_dbeta = 3.0 / (1.0 - beta) ** 2  # eps = (1 + 2 beta) / (1 - beta)
    dU_dbeta_adj = float(ge.sum().item()) * deps_dbeta
    log(f"adjoint forces in {t_adj:.2f} s, U = {U.item():.10g}")

    # finite differences of the stationary energy
    perp, para, rv = tables
    ctx = Context(dev)
    ctx.set_tables(perp, para, rv)
    eta = P * (hg * xi) ** 2 / np.pi
    b = torch.as_tensor(np.tile([0.0, 0.0, 1.0], n).reshape(1, n, 3), device=dev)
    ctx.set_geometry(x, box, ng, P, xi, eta, rc)
    ctx.set_permittivity(eps)
    p0 = ctx.solve(b, b * 0.0, solve_tol, restart=10, maxit=100, slab_factor=0.0).x

    def u_var(xp, ep):
        ctx.set_geometry(xp, box, ng, P, xi, eta, rc)
        ctx.set_permittivity(ep)
        r = ctx.solve(b, p0, solve_tol, restart=10, maxit=100, slab_factor=0.0)
        if r.flag != 0:
            raise RuntimeError(f"finite-difference solve did not converge ({r.max_relres:.2e})")
        p = r.x
        ap = ctx.matvec(p, slab_factor=0.0)
        pp = p.cpu().numpy().astype(np.longdouble)
        return float(-np.sum(pp[0, :, 2]) + 0.5 * np.sum(pp * ap.cpu().numpy().astype(np.longdouble)))

    rng = np.random.default_rng(seed)
    idx = rng.choice(n, size=particles, replace=False)
    t0 = time.perf_counter()
    rows, skipped = [], 0
    for i in idx:
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            if _node(xp[i, d], hg[d]) != _node(xm[i, d], hg[d]):
                skipped += 1
                continue
            f_fd = -(u_var(xp, eps) - u_var(xm, eps)) / (2.0 * h)
            rows.append((int(i), d, float(F_adj[i, d]), f_fd))
    epp = np.full(n, (1.0 + 2.0 * (beta + h)) / (1.0 - (beta + h)))
    epm = np.full(n, (1.0 + 2.0 * (beta - h)) / (1.0 - (beta - h)))
    dU_dbeta_fd = (u_var(x, epp) - u_var(x, epm)) / (2.0 * h)
    t_fd = time.perf_counter() - t0
    ctx.close()
    fa = np.array([r[2] for r in rows])
    ff = np.array([r[3] for r in rows])
    scale = float(np.max(np.abs(ff)))
    out = {
        "workload": workload, "n": n, "particles": particles, "components": len(rows), "skipped_node_switch": skipped,
        "h": h, "solve_tol": solve_tol, "table_step": table_step,
        "force_max_abs_diff": float(np.max(np.abs(fa - ff))), "force_max_abs": scale,
        "force_rel_err": float(np.max(np.abs(fa - ff)) / scale),
        "dU_dbeta_adjoint": dU_dbeta_adj, "dU_dbeta_fd": dU_dbeta_fd,
        "dU_dbeta_rel_err": abs(dU_dbeta_adj - dU_dbeta_fd) / abs(dU_dbeta_fd),
        "adjoint_seconds": round(t_adj, 3), "fd_seconds": round(t_fd, 3),
    }
    return out, rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="n1e5_phi0.4")
    ap.add_argument("--particles", type=int, default=32)
    ap.add_argument("--h", type=float, default=1e-4)
    ap.add_argument("--solve-tol", type=float, default=1e-12)
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--table-step", type=float, default=1e-6)
    args = ap.parse_args(argv)
    out, _ = run(args.workload, args.particles, args.h, args.solve_tol, sweeps=args.sweeps,
                 table_step=args.table_step)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
