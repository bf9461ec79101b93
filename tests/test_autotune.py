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
