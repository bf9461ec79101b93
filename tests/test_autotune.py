"""Autotuner: Ewald candidates follow the reference's parameter rule (CPU),
tuned launch shapes leave the operator unchanged and the tuned Ewald grid
gives the same operator to the Ewald tolerance (GPU)."""
import math

import numpy as np
import pytest

from helpers import relerr, tables


def _packing(n, phi, seed):
    import workloads
    return workloads.hard_spheres(n, phi, seed, sweeps=5)


def test_ewald_candidates_follow_the_reference_rule():
    from mpol_hip.autotune import CT_SIZES, ewald_candidates
    box = np.array([60.0, 60.0, 45.0])
    tol = 1e-6
    cands = ewald_candidates(box, tol)
    assert len(cands) >= 3
    s = math.sqrt(-math.log(tol))
    for xi, ng, P, eta, rc in cands:
        assert P == math.ceil(-2 * math.log(tol) / math.pi)
        assert abs(rc - s / xi) < 1e-12 and rc <= np.min(box) / 2
        kcut = 2 * xi * s
        need = np.ceil(1 + box * kcut / np.pi)
        assert np.all(ng >= need) and all(int(v) in CT_SIZES for v in ng)
        assert ng[0] == max(ng)  # the longest axis sets xi
        np.testing.assert_allclose(eta, P * (box / ng * xi) ** 2 / np.pi, rtol=1e-14)
    xis = [c[0] for c in cands]
    assert xis == sorted(xis)


@pytest.mark.gpu
def test_tuned_launch_shapes_keep_the_operator(gpu):
    from mpol_hip import FP32, FP64, Context
    from mpol_hip.autotune import tune_launch
    x, box = _packing(3000, 0.3, 5)
    xi, P, ng = 1.0, 9, np.array([64, 64, 64])
    rc = 3.3
    perp, para, rv = tables(xi)
    c = Context(gpu)
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    c.set_permittivity(np.full(3000, 9.0))
    m = np.random.default_rng(2).standard_normal((1, 3000, 3))
    ref64 = c.matvec(m).cpu().numpy()
    ref32 = c.matvec(m, precision=FP32).cpu().numpy()
    k32 = tune_launch(c, FP32)
    k64 = tune_launch(c, FP64)
    print("tuned", k32, k64, flush=True)
    assert relerr(c.matvec(m).cpu().numpy(), ref64) < 1e-13
    assert relerr(c.matvec(m, precision=FP32).cpu().numpy(), ref32) < 2e-6
    assert tune_launch(c, FP32) == k32  # cached


@pytest.mark.gpu
def test_tuned_ewald_grid_meets_the_error_target(gpu):
    """The chosen splitting is the fastest candidate whose measured operator
    error is within the target; its operator agrees with a tight FP64
    operator to that target (checked again independently)."""
    from mpol_hip import Context
    from mpol_hip.autotune import _tight_reference, tune_ewald
    x, box = _packing(1500, 0.25, 9)
    eps = np.full(1500, 6.0)
    target = 1e-5
    xi, ng, P, eta, rc, ms, report = tune_ewald(x, box, eps, target, tols=(1e-6, 1e-8), sizes=(64, 96, 128),
                                                tables=tables, device=gpu)
    print("chosen", xi, ng, P, rc, ms, flush=True)
    ok = [r for r in report if r[1] <= target]
    assert ms == min(r[0] for r in ok)
    m = np.random.default_rng(3).standard_normal((1, 1500, 3))
    outs = []
    for (xi_c, ng_c, P_c, eta_c, rc_c) in ((xi, ng, P, eta, rc), _tight_reference(box)):
        perp, para, rv = tables(xi_c)
        c = Context(gpu)
        c.set_tables(perp, para, rv)
        c.set_geometry(x, box, ng_c, P_c, xi_c, eta_c, rc_c)
        c.set_permittivity(eps)
        outs.append(c.matvec(m, slab_factor=0.0).cpu().numpy())
    assert relerr(outs[0], outs[1]) < 3 * target


def test_tune_inner_keeps_the_fastest_accepted_setting(monkeypatch):
    """tune_inner's selection on the host, with stand-in contexts: per setting
    the loosest tolerance whose solve converges and passes accept is timed;
    the fastest setting stays set on the context, every other candidate
    context is closed, and no candidate at all leaves none set."""
    import types

    from mpol_hip import autotune

    cost = {None: 9.0, (1e-4, 128, 1e-2): 6.0, (1e-4, 128, 3e-3): 7.0, (1e-3, 96, 1e-2): 5.0,
            (1e-3, 96, 3e-3): 8.0}
    rejected = {(1e-3, 96, 1e-2)}  # fastest, but its dipoles miss the bar at every tolerance

    class Inner:
        def __init__(self, etol, grid):
            self.key, self.closed = (etol, grid), False

        def close(self):
            self.closed = True

    made = []

    def fake_coarse(x, box, eps, etol, grid, device=None, quadrupole=False):
        made.append(Inner(etol, grid))
        return made[-1], (0.5, np.array([grid] * 3), 6, np.zeros(3), 4.0)

    class Ctx:
        device = None

        def __init__(self):
            self.inner, self.floor, self.clock = None, None, 0.0

        def set_inner_operator(self, inner, floor=1e-3):
            self.inner, self.floor = inner, floor

        def solve(self, b, x0, t, **kw):
            key = None if self.inner is None else (*self.inner.key, self.floor)
            self.clock += cost[key] * 1e-3
            return types.SimpleNamespace(converged=1, matvecs_fp32=1, matvecs_fp64=1, matvecs_inner=0,
                                         bad=key in rejected)

    ctx = Ctx()
    monkeypatch.setattr(autotune, "coarse_operator", fake_coarse)
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: ctx.clock)
    fake_torch = types.SimpleNamespace(cuda=types.SimpleNamespace(synchronize=lambda: None))
    monkeypatch.setitem(__import__("sys").modules, "torch", fake_torch)
    b = np.zeros((1, 4, 3))
    inner, cfg, report = autotune.tune_inner(ctx, None, None, None, b, b, (1e-6, 5e-7),
                                             accept=lambda r: not r.bad,
                                             candidates=((1e-4, 128), (1e-3, 96)), floors=(1e-2, 3e-3), repeats=2)
    assert cfg == {"ewald_tol": 1e-4, "grid": 128, "P": 6, "floor": 1e-2, "tol": 1e-6}
    assert inner is made[0] and not made[0].closed and made[1].closed
    assert ctx.inner is made[0] and ctx.floor == 1e-2
    labels = [r[0] for r in report]
    assert labels[0] == "none" and len(report) == 4  # the rejected setting is not timed
    assert [round(r[2], 6) for r in report] == [9.0, 6.0, 7.0, 8.0]
    # nothing faster than none: none stays
    cost.update({(1e-4, 128, 1e-2): 10.0, (1e-4, 128, 3e-3): 11.0, (1e-3, 96, 3e-3): 12.0})
    made.clear()
    ctx = Ctx()
    inner, cfg, _ = autotune.tune_inner(ctx, None, None, None, b, b, (1e-6,), accept=lambda r: not r.bad,
                                        candidates=((1e-4, 128), (1e-3, 96)), floors=(1e-2, 3e-3))
    assert inner is None and cfg is None and ctx.inner is None and all(m.closed for m in made)
