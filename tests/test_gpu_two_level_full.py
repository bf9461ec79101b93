"""The two-level mixed solve at the benchmark workload (N = 1e5 unit spheres,
phi = 0.4, beta = 0.9, 256^3, P = 10, field along z, slab factor 0): the
correction operator at 128^3, P = 6 (rule tolerance 1e-4, floor 1e-2), FP64
residuals on the full operator.  At tol 1e-6 the dipoles are within 1e-6 of
a tight (1e-10) FP64 solve, the true FP64 residual meets the tolerance, and
the three field directions batched meet it too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_two_level_mixed_solve_at_full_size(gpu):
    import torch

    import workloads
    from mpol_hip import FP64, Context, default_tables
    from mpol_hip.autotune import coarse_operator
    print("case n1e5_phi0.4 two-level, 128^3 correction operator", flush=True)
    cfg = workloads.configuration("n1e5_phi0.4", sweeps=200)
    n = cfg["x"].shape[0]
    perp, para, rv = default_tables(cfg["xi"])
    ctx = Context(gpu)
    ctx.set_tables(perp, para, rv)
    ctx.set_geometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["eta"], cfg["rc"])
    ctx.set_permittivity(cfg["eps"])
    inner, (xi2, ng2, P2, _, rc2) = coarse_operator(cfg["x"], cfg["box"], cfg["eps"], 1e-4, 128, device=gpu)
    assert int(ng2[0]) == 128 and P2 == 6
    phi = 4.0 * np.pi * n / (3.0 * np.prod(cfg["box"]))
    mg = 4.0 * np.pi * cfg["beta"] / (1.0 - cfg["beta"] * phi)
    for fields, tol in ((np.eye(3)[2:], 1e-6), (np.eye(3), 2e-7)):
        b = torch.as_tensor(np.stack([np.tile(f, n).reshape(n, 3) for f in fields]), device=gpu)
        ref = ctx.solve(b, mg * b, 1e-10, restart=10, maxit=100, slab_factor=0.0).x
        ctx.set_inner_operator(inner, 1e-2)
        r = ctx.solve(b, mg * b, tol, restart=10, maxit=100, mixed=True, slab_factor=0.0)
        ctx.set_inner_operator(None)
        res = b - ctx.matvec(r.x, precision=FP64, slab_factor=0.0)
        rel = (torch.linalg.norm(res.reshape(b.shape[0], -1), dim=1)
               / torch.linalg.norm(b.reshape(b.shape[0], -1), dim=1)).cpu().numpy()
        err = float(torch.max(torch.abs(r.x - ref)) / torch.max(torch.abs(ref)))
        print("fields", b.shape[0], "tol", tol, "matvecs coarse", r.matvecs_inner, "fp32",
              r.matvecs_fp32 - r.matvecs_inner, "fp64", r.matvecs_fp64, "true relres", rel, "err", err, flush=True)
        assert r.converged == b.shape[0] and np.all(rel <= tol * 1.001)
        assert r.matvecs_inner > 0
        assert err <= 1e-6
    inner.close()
    ctx.close()
