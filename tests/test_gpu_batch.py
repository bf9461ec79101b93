"""Many-configuration driver on the GPU: every configuration's spectrum x
fields batch equals the single-configuration polydisperse spectrum solve
(sizes.CapacitanceTensorSpectrum, closed-form couplings) at the splitting the
driver chose, to 1e-10 in FP64; mixed precision to 1e-5 of FP64 (GMRES
tolerance 1e-7); a spectrum point at eps = 1 gives zero dipoles."""
import numpy as np
import pytest

import sizes
import workloads
from mpol_hip import batch

pytestmark = pytest.mark.gpu


def _configs():
    out = []
    for k, (n, phi) in enumerate([(260, 0.3), (400, 0.2), (180, 0.4)]):
        x, a, box = workloads.polydisperse(n, phi, 40 + k)
        out.append({"x": x, "a": a, "box": box})
    return out


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_batch_matches_single_configuration_solves(gpu):
    configs = _configs()
    eps = np.array([0.5, 1.0, 4.0, 20.0])
    res = batch.solve_configurations(configs, eps, precision="fp64", workers=2, device=gpu)
    assert res.rhs == 3 * 4 * 3
    assert res.C.shape == (3, 4, 3, 3)
    for i, c in enumerate(configs):
        n = c["x"].shape[0]
        assert np.all(res.p[i][1] == 0.0)  # eps = 1
        live = [0, 2, 3]
        C, p = sizes.CapacitanceTensorSpectrum(c["x"], c["box"], c["a"], np.tile(eps[live], (n, 1)), res.xi[i],
                                               precision="fp64", closed_form=True)
        ref = np.real(p[:, :, :, 0]).reshape(n, 3, 3, len(live)).transpose(3, 1, 0, 2)  # (k, f, N, 3)
        assert _rel(res.p[i][live], ref) < 1e-10, i
        assert _rel(res.C[i][live], ref.mean(axis=2)) < 1e-10


def test_batch_mixed_precision_close_to_fp64(gpu):
    configs = _configs()[:2]
    eps = np.array([0.2, 8.0, 60.0])
    r64 = batch.solve_configurations(configs, eps, gmres_tol=1e-9, precision="fp64", workers=2, tune=False,
                                     device=gpu)
    r32 = batch.solve_configurations(configs, eps, gmres_tol=1e-7, precision="mixed", workers=2, tune=False,
                                     device=gpu)
    for i in range(2):
        assert _rel(r32.p[i], r64.p[i]) < 1e-5
        assert r32.matvecs[i][1] > 0  # the bulk ran on the FP32 operator


def test_tune_grids_reports_and_early_stop(gpu):
    """tune_grids returns, per configuration, the fastest timed candidate;
    without early_stop every candidate is timed, with it the search ends at
    the first candidate slower than the best so far."""
    configs = _configs()[:2]
    eps = np.array([0.5, 4.0])
    full, rep_full, _ = batch.tune_grids(configs, eps, early_stop=False, device=gpu)
    short, rep_short, _ = batch.tune_grids(configs, eps, device=gpu)
    for i, c in enumerate(configs):
        sizes_ok = [int(cand[1][0]) for cand in batch.grid_candidates(c["box"], 1e-3)]
        assert [g for g, _ in rep_full[i]] == sizes_ok
        assert full[i] == min(rep_full[i], key=lambda t: t[1])[0]
        times = [ms for _, ms in rep_short[i]]
        assert [g for g, _ in rep_short[i]] == sizes_ok[:len(times)]
        # every candidate timed before the last one improved on the best so far
        assert all(times[k] < min(times[:k], default=np.inf) for k in range(len(times) - 1))
        assert short[i] == min(rep_short[i], key=lambda t: t[1])[0]


def test_batch_is_deterministic(gpu):
    """The polydisperse kernels and the GMRES reductions have fixed
    summation orders (no atomics): two runs of the driver on concurrent
    worker streams give bitwise the same dipoles."""
    configs = _configs()
    eps = np.array([0.3, 5.0, 40.0])
    grids = [32] * len(configs)
    r1 = batch.solve_configurations(configs, eps, gmres_tol=1e-7, precision="mixed", workers=3, tune=False,
                                    grids=grids, device=gpu)
    r2 = batch.solve_configurations(configs, eps, gmres_tol=1e-7, precision="mixed", workers=3, tune=False,
                                    grids=grids, device=gpu)
    for a, b in zip(r1.p, r2.p):
        assert np.array_equal(a, b)
    assert np.array_equal(r1.matvecs, r2.matvecs)
