"""Host-side pieces of the many-configuration driver: the polydisperse
packings it is fed with and the grid candidates of the Ewald rule."""
import math

import numpy as np
import pytest

import workloads
from mpol_hip import batch


@pytest.mark.parametrize("n,phi,seed", [(300, 0.45, 1), (800, 0.1, 2), (1500, 0.3, 3)])
def test_polydisperse_packing_has_no_overlaps(n, phi, seed):
    x, a, box = workloads.polydisperse(n, phi, seed)
    assert x.shape == (n, 3) and a.shape == (n,)
    assert np.all((x >= 0) & (x < box[0]))
    vol = 4.0 / 3.0 * math.pi * np.sum(a ** 3)
    assert abs(vol / np.prod(box) - phi) < 1e-12
    assert workloads.min_gap_ratio(x, a, box) >= 1.0
    # brute force on the first 200 spheres against all others
    d = x[:200, None, :] - x[None, :, :]
    d -= box[0] * np.rint(d / box[0])
    r = np.sqrt(np.sum(d * d, axis=2)) / (a[:200, None] + a[None, :])
    r[np.arange(200), np.arange(200)] = np.inf
    assert r.min() >= 1.0


def test_polydisperse_radii_are_lognormal_mean_one():
    _, a, _ = workloads.polydisperse(4000, 0.2, 5)
    assert abs(a.mean() - 1.0) < 0.02
    assert abs(np.std(np.log(a)) - 0.2) < 0.01


def test_grid_candidates_follow_the_rule():
    box = np.full(3, 40.0)
    tol = 1e-3
    cands = batch.grid_candidates(box, tol)
    assert [int(c[1][0]) for c in cands] == [32, 64, 96, 128]
    for xi, ng, P, eta, rc in cands:
        rc2, ng2, P2, eta2 = batch.rule_parameters(box, xi, tol)
        assert np.all(ng == ng2) and P == P2 and rc == rc2
        # a slightly larger xi must give a larger grid (xi is the largest for G)
        assert batch.rule_parameters(box, xi * (1 + 1e-9), tol)[1][0] == ng[0] + 1
        assert rc <= 20.0
    # a small box rules out the coarse grids (rc > L/2)
    small = batch.grid_candidates(np.full(3, 8.0), tol)
    assert all(c[4] <= 4.0 for c in small)


def test_heterogeneous_set_ranges():
    configs, eps = workloads.heterogeneous_set(count=6, seed=3, n_range=(200, 600))
    assert len(configs) == 6 and eps.size == 16
    assert eps[0] == pytest.approx(0.1) and eps[-1] == pytest.approx(100.0)
    for c in configs:
        n = c["x"].shape[0]
        assert 200 <= n <= 600
        assert 0.1 <= c["phi"] <= 0.45
        assert workloads.min_gap_ratio(c["x"], c["a"], c["box"]) >= 1.0
