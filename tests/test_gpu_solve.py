"""Parity of the device-resident GMRES and the capacitance workflow with the
reference's golden solves and the FP64 oracle.

FP64 mode: dipoles within 1e-10 (relative to max |p|) of the reference solved
at tol 1e-12.  Mixed mode (FP32 operator, FP64 refinement): within 1e-6 of the
FP64 solution.
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from helpers import relerr, rsa, tables

pytestmark = pytest.mark.gpu

SOLVE_FILES = sorted(glob.glob(os.path.join(GOLDEN, "solve_*.npz")))


def solve_ctx(d, op):
    from mpol_hip import Context
    xi = float(op["xi"])
    perp, para, rv = tables(xi)
    c = Context()
    c.set_tables(perp, para, rv)
    h = d["box"] / op["ngrid"]
    c.set_geometry(d["x"], d["box"], op["ngrid"], int(op["P"]), xi, int(op["P"]) * (h * xi) ** 2 / np.pi,
                   float(op["rc"]))
    c.set_permittivity(d["eps"])
    return c


@pytest.mark.parametrize("path", SOLVE_FILES, ids=os.path.basename)
def test_solve_fp64_matches_reference(gpu, path):
    d = np.load(path)
    op = golden("operator_n64_g21.npz")
    print("case", os.path.basename(path), flush=True)
    c = solve_ctx(d, op)
    slab = 1.0 if "slab1" in path else 0.0
    n = d["x"].shape[0]
    b = np.tile(d["H"], n).reshape(1, n, 3)
    res = c.solve(b, d["guess"][None], 1e-12, restart=10, maxit=100, slab_factor=slab)
    p = res.x.cpu().numpy()[0]
    assert res.flag == 0
    assert res.max_relres <= 1e-12
    assert relerr(p, np.real(d["p_scipy"])) < 1e-10
    assert relerr(p, np.real(d["p_python"])) < 1e-10


@pytest.mark.parametrize("path", SOLVE_FILES, ids=os.path.basename)
def test_solve_mixed_within_1e6(gpu, path):
    d = np.load(path)
    op = golden("operator_n64_g21.npz")
    print("case", os.path.basename(path), flush=True)
    c = solve_ctx(d, op)
    slab = 1.0 if "slab1" in path else 0.0
    n = d["x"].shape[0]
    b = np.tile(d["H"], n).reshape(1, n, 3)
    res = c.solve(b, d["guess"][None], 1e-8, restart=10, maxit=100, mixed=True, slab_factor=slab)
    p = res.x.cpu().numpy()[0]
    assert res.flag == 0
    assert res.matvecs_fp32 > 0
    assert relerr(p, np.real(d["p_scipy"])) < 1e-6


def test_solve_long_krylov_cycle(gpu):
    """One GMRES cycle of 150 Arnoldi steps (tolerance out of reach, so every
    step runs): the Gram-Schmidt passes then orthogonalise against more than
    the 64 vectors whose wave sums one LDS group of the dot-product kernels
    holds, and the subtraction kernels finish up to 150 coefficients each from
    the partial sums.  The dipoles still meet the reference's solve."""
    path = SOLVE_FILES[0]
    d = np.load(path)
    op = golden("operator_n64_g21.npz")
    print("case", os.path.basename(path), flush=True)
    c = solve_ctx(d, op)
    slab = 1.0 if "slab1" in path else 0.0
    n = d["x"].shape[0]
    b = np.tile(d["H"], n).reshape(1, n, 3)
    res = c.solve(b, d["guess"][None], 1e-300, restart=150, maxit=1, slab_factor=slab)
    p = res.x.cpu().numpy()[0]
    assert res.matvecs_fp64 >= 150
    assert res.max_relres <= 1e-12
    assert relerr(p, np.real(d["p_scipy"])) < 1e-10


def test_dropin_magnetic_dipole_and_run_capacitance(gpu):
    import capacitance
    for slab in (0, 1):
        d = golden(f"run_n64_slab{slab}.npz")
        C, p = capacitance.run_capacitance(d["x"], d["box"], d["eps"], float(d["xi"]), d["H0"],
                                           errortol=float(d["errortol"]), slab_factor=float(slab))
        # both stop GMRES at tol 1e-5 from the same start; solutions agree to
        # the tolerance times the conditioning
        assert relerr(p, d["p"]) < 1e-4
        assert abs(C[0] - d["C"][0]) / abs(d["C"][0]) < 1e-5
        info = capacitance.magnetic_dipole.last_info
        assert info["flag"] == 0 and info["relres"] <= 1e-5


def test_tensor_spectrum_batched_equals_single_solves(gpu):
    import capacitance
    x, box = rsa(200, 0.25, 5)
    rng = np.random.default_rng(1)
    eps = np.stack((np.full(200, 4.0), 1.0 + 30 * rng.random(200)), axis=1)
    C, p, info = capacitance.capacitance_tensor_spectrum(x, box, eps, 0.9, errortol=1e-8, return_info=True)
    assert info["flag"] == 0
    # symmetric tensor for real permittivities
    for k in range(2):
        assert np.max(np.abs(C[:, :, k] - C[:, :, k].T)) < 1e-6 * np.max(np.abs(C[:, :, k]))
    # column k, direction 2 equals a single run_capacitance-style solve
    import ewald
    rc, _, ng, h, P, eta = capacitance.ewald_parameters(box, 0.9, 1e-8)
    perp, para, rv = tables(0.9)
    phi = 4 * np.pi * 200 / (3 * np.prod(box))
    guess = capacitance.maxwell_garnett_guess(eps[:, 1], phi, np.array([0, 0, 1.0]))
    S = capacitance.magnetic_dipole(x, eps[:, 1], np.array([0, 0, 1.0]), box, None, None, ng, h, P, 0.9, eta, rc,
                                    guess, None, None, perp, para, rv, 1e-10, slab_factor=0.0)
    assert relerr(p[:, :, 2, 1], S.real) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_run_capacitance_trajectory_frames_concurrent(gpu, precision):
    """A trajectory (N, 3, F): the frames solved concurrently on worker
    streams give the dipoles and capacitances of the one-frame loop
    ("CPU - Ewald test/capacitance.py":495-513) -- to rounding in FP64; in
    mixed precision the FP32 operator's rounding (the 17^3 grid is off the
    matrix-core path, where the spread's atomics order varies) moves the
    iterates within the solve tolerance -- and its first frame (the golden
    run's positions) matches the reference's run_capacitance."""
    import capacitance
    d = golden("run_n64_slab1.npz")
    x0, box = d["x"], d["box"]
    rng = np.random.default_rng(12)
    frames = [x0] + [np.mod(x0 + 0.05 * k * rng.standard_normal(x0.shape), box) for k in range(1, 6)]
    x = np.stack(frames, axis=2)
    args = (x, box, d["eps"], float(d["xi"]), d["H0"])
    kw = dict(errortol=float(d["errortol"]), slab_factor=1.0, precision=precision)
    C1, p1 = capacitance.run_capacitance(*args, frame_workers=1, **kw)
    C4, p4 = capacitance.run_capacitance(*args, frame_workers=4, **kw)
    assert p4.shape == (64, 3, 6) and C4.shape == (6,)
    bar = 1e-12 if precision == "fp64" else 1e-5
    assert relerr(p4, p1) < bar, relerr(p4, p1)
    assert np.max(np.abs(C4 - C1) / np.abs(C1)) < bar
    assert relerr(p4[:, :, 0], d["p"][:, :, 0]) < 1e-4
    assert abs(C4[0] - d["C"][0]) / abs(d["C"][0]) < 1e-5
    assert len(set(np.round(np.real(C4), 12))) == 6  # the frames differ


def test_mixed_solve_with_a_coarse_correction_operator(gpu):
    """Two-level mixed scheme: correction solves on a coarser operator (32^3,
    P = 6 at rule tolerance 1e-4) while the FP64 residuals stay on the full
    one (64^3, P = 8).  The result meets tol for the full operator (its true
    FP64 residual), agrees with the full operator's tight FP64 solve to the
    tolerance's reach, in more refinement sweeps than one level, for one and
    for three right-hand sides; misuse is refused."""
    import torch

    import workloads
    from mpol_hip import FP64, Context
    from mpol_hip.autotune import coarse_operator
    n = 3000
    x, box = workloads.hard_spheres(n, 0.3, 11, sweeps=20)
    xi, P, ng = 1.0, 8, np.array([64, 64, 64])
    rc = float(np.sqrt(-np.log(1e-6)) / xi)
    eps = np.full(n, 28.0)
    perp, para, rv = tables(xi)
    c = Context(gpu)
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    c.set_permittivity(eps)
    inner, (xi2, ng2, P2, _, rc2) = coarse_operator(x, box, eps, 1e-4, 32, device=gpu)
    print("coarse", xi2, ng2, P2, rc2, flush=True)
    assert int(ng2[0]) == 32 and P2 == 6
    phi = 4 * np.pi * n / (3 * np.prod(box))
    mg = 4 * np.pi * (9.0 / 10.0) / (1 - 0.9 * phi)
    tol = 1e-8
    for fields in (np.eye(3)[2:], np.eye(3)):
        b = torch.as_tensor(np.stack([np.tile(e, n).reshape(n, 3) for e in fields]), device=gpu)
        ref = c.solve(b, mg * b, 1e-12, restart=10, maxit=200).x
        c.set_inner_operator(None)
        one = c.solve(b, mg * b, tol, restart=10, maxit=100, mixed=True)
        c.set_inner_operator(inner, 1e-3)
        two = c.solve(b, mg * b, tol, restart=10, maxit=100, mixed=True)
        c.set_inner_operator(None)
        res = b - c.matvec(two.x, precision=FP64)
        rel = (torch.linalg.norm(res.reshape(b.shape[0], -1), dim=1)
               / torch.linalg.norm(b.reshape(b.shape[0], -1), dim=1)).cpu().numpy()
        err = relerr(two.x.cpu().numpy(), ref.cpu().numpy())
        print("rhs", b.shape[0], "one level", one.matvecs_fp32, one.matvecs_fp64, one.refinements,
              "two levels", two.matvecs_fp32, two.matvecs_fp64, two.refinements, "true relres", rel,
              "err vs tight FP64", err, flush=True)
        assert two.converged == b.shape[0]
        assert np.all(rel <= tol * 1.001)
        assert err < 1e-6
        assert two.refinements > one.refinements
        assert one.matvecs_inner == 0 and 0 < two.matvecs_inner <= two.matvecs_fp32
    with pytest.raises(ValueError):
        c.set_inner_operator(c)
    other = Context(gpu)
    other.set_tables(perp, para, rv)
    other.set_geometry(x[:100], box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    with pytest.raises(ValueError):
        c.set_inner_operator(other)
    inner.set_permittivity(np.stack([eps, eps]))
    c.set_inner_operator(inner)
    b = torch.as_tensor(np.tile([0.0, 0.0, 1.0], n).reshape(1, n, 3), device=gpu)
    with pytest.raises(ValueError):
        c.solve(b, mg * b, tol, mixed=True)  # permittivity sets differ
    c.set_inner_operator(None)
    # the tuner: the fastest of none and the candidate's floors, left on c
    from mpol_hip.autotune import tune_inner
    inner.set_permittivity(eps)
    picked, cfg, report = tune_inner(c, x, box, eps, b, mg * b, (1e-8,), candidates=((1e-4, 32),),
                                     floors=(1e-3, 3e-4), repeats=2, slab_factor=1.0)
    print("tuner", cfg, report, flush=True)
    assert [r[0] for r in report][0] == "none" and len(report) == 3
    assert min(r[2] for r in report) == (report[0][2] if cfg is None else
                                         [r[2] for r in report if r[0] != "none" and f"floor {cfg['floor']:g}" in r[0]][0])
    r = c.solve(b, mg * b, 1e-8, mixed=True)
    assert r.converged == 1
    c.set_inner_operator(None)


def test_two_level_solve_with_complex_permittivities(gpu):
    """Two-level mixed solve with complex permittivities (each system a
    (re, im) pair of right-hand sides): the inner operator carries the same
    complex permittivities; the result meets tol on the full operator's true
    FP64 residual and agrees with its tight FP64 solve.  An inner operator
    left at real permittivities is refused."""
    import torch

    import workloads
    from mpol_hip import FP64, Context
    from mpol_hip.autotune import coarse_operator
    n = 2000
    x, box = workloads.hard_spheres(n, 0.3, 13, sweeps=20)
    xi, P, ng = 1.0, 8, np.array([64, 64, 64])
    rc = float(np.sqrt(-np.log(1e-6)) / xi)
    rng = np.random.default_rng(5)
    eps = (4.0 + 20.0 * rng.random(n)) + 1j * (0.5 + 2.0 * rng.random(n))
    perp, para, rv = tables(xi)
    c = Context(gpu)
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    c.set_permittivity(eps)
    inner, _ = coarse_operator(x, box, eps.real, 1e-4, 32, device=gpu)
    b = np.zeros((1, n, 3), dtype=np.complex128)
    b[0, :, 2] = 1.0
    x0 = np.zeros_like(b)
    ref = c.solve(b, x0, 1e-12, restart=10, maxit=300).x
    c.set_inner_operator(inner, 1e-3)
    with pytest.raises(ValueError):
        c.solve(b, x0, 1e-8, mixed=True)  # the inner operator's permittivities are real
    inner.set_permittivity(eps)
    two = c.solve(b, x0, 1e-8, restart=10, maxit=100, mixed=True)
    c.set_inner_operator(None)
    res = torch.as_tensor(b, device=gpu) - c.matvec(two.x, precision=FP64)
    rel = float(torch.linalg.norm(res) / np.linalg.norm(b))
    err = relerr(two.x.cpu().numpy(), ref.cpu().numpy())
    print("complex two levels", two.matvecs_inner, two.matvecs_fp32, two.matvecs_fp64, two.refinements,
          "true relres", rel, "err", err, flush=True)
    assert two.converged == 1 and rel <= 1e-8 * 1.001 and err < 1e-6 and two.matvecs_inner > 0
    inner.close()
