"""The reference's step-by-step fast-path functions through the drop-in
ewald.py on the GPU, against golden vectors the reference wrote
(tests/golden/dropin_n24.npz: 24 spheres, non-cubic box, 8 x 10 x 7 grid,
P = 5, random complex dipoles and grids).

Integer outputs (node indices, kept pairs) must be identical; floating-point
outputs agree to rounding (1e-13 relative to the array's largest entry: the
GPU forms the reference's bincount / add.at sums with atomics, in another
order).  Then magnetic_field fed with the tuples these functions return, and
the reference's errors.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def case():
    return golden("dropin_n24.npz")


@pytest.fixture(scope="module")
def tables(case):
    import ewald
    r = np.arange(0.001, 10.0 + 1e-12, 0.001)
    t = ewald.real_space_table(r, float(case["xi"]))
    return t[2], t[3], np.concatenate(([0.0], r))


def test_prepare_grid_kernel(gpu, case):
    import ewald
    lin, ws, wc = ewald.prepare_grid_kernel(case["x"], case["Ngrid"], case["h"], float(case["xi"]), case["eta"],
                                            int(case["P"]), case["offset"], case["offsetxyz"])
    assert lin.dtype == np.int64 and lin.shape == case["linnode"].shape
    assert np.array_equal(lin, case["linnode"])
    assert _rel(ws, case["w_spread"]) < TOL
    assert _rel(wc, case["w_contract"]) < TOL


def test_spread_and_contract(gpu, case):
    import ewald
    args = (case["Ngrid"], case["h"], float(case["xi"]), case["eta"], int(case["P"]))
    H = ewald.spread_fast(case["m"], case["Ngrid"], case["linnode"], case["w_spread"])
    assert H.shape == case["H_fast"].shape and H.dtype == np.complex128
    assert _rel(H, case["H_fast"]) < TOL
    Hx, Hy, Hz = ewald.spread(case["x"], case["m"], *args, case["offset"], case["offsetxyz"])
    assert _rel(np.stack((Hx, Hy, Hz), axis=3), case["H_loop"]) < TOL
    Hk = ewald.contract_fast(case["Htilde"], case["Ngrid"], case["linnode"], case["w_contract"])
    assert Hk.shape == (24, 3) and _rel(Hk, case["Hk_fast"]) < TOL
    Hk2 = ewald.contract(case["x"], *args, case["Htilde"], case["offset"], case["offsetxyz"])
    assert _rel(Hk2, case["Hk_loop"]) < TOL


def test_scale_and_kspace(gpu, case):
    import ewald
    out = ewald.scale(case["fH"], case["k"], case["Ngrid"], float(case["xi"]), case["eta"])
    assert out.shape == case["fHt"].shape and _rel(out, case["fHt"]) < TOL
    khat, coeff = ewald.precompute_kspace(case["Ngrid"], case["box"], float(case["xi"]), case["eta"])
    assert khat.dtype == np.complex128 and coeff.dtype == np.float64
    assert _rel(khat, case["khat"]) < TOL
    assert _rel(coeff, case["coeff"]) < TOL
    out2 = ewald.scale_fast(case["fH"], case["khat"], case["coeff"])
    assert _rel(out2, case["fHt_fast"]) < TOL


def test_real_space_pairs_and_sums(gpu, case, tables):
    import ewald
    perp, para, rv = tables
    p1f, p2f, rhat, A, B = ewald.prepare_real_space_pairs(case["x"], case["box"], case["p1"], case["p2"],
                                                          float(case["rc"]), perp, para, rv)
    assert np.array_equal(p1f, case["p1f"]) and np.array_equal(p2f, case["p2f"])
    assert _rel(rhat, case["rhat"]) < TOL
    assert _rel(A, case["A"]) < 1e-12 and _rel(B, case["B"]) < 1e-12
    Hr = ewald.real_space_fast(case["m"], case["lam"], case["p1f"], case["p2f"], case["rhat"], case["A"],
                               case["B"], float(perp[0]))
    assert _rel(Hr, case["Hr_fast"]) < 1e-12
    Hr2 = ewald.real_space(case["x"], case["m"], case["lam"], case["box"], case["p1"], case["p2"], float(case["rc"]),
                           perp, para, rv)
    assert _rel(Hr2, case["Hr_loop"]) < 1e-12
    # no candidate pairs: the reference's empty results
    e = ewald.prepare_real_space_pairs(case["x"], case["box"], np.array([], int), np.array([], int),
                                       float(case["rc"]), perp, para, rv)
    assert e[0].size == 0 and e[2].shape == (0, 3)


def test_reference_errors(gpu, case, tables):
    import ewald
    perp, para, rv = tables
    with pytest.raises(ValueError, match="offset/offsetxyz"):
        ewald.prepare_grid_kernel(case["x"], case["Ngrid"], case["h"], float(case["xi"]), case["eta"],
                                  int(case["P"]) + 1, case["offset"], case["offsetxyz"])
    short = rv <= 3.0  # kept separations beyond the table
    with pytest.raises(ValueError, match="outside rvals"):
        ewald.prepare_real_space_pairs(case["x"], case["box"], case["p1"], case["p2"], float(case["rc"]),
                                       perp[short], para[short], rv[short])
    with pytest.raises(ValueError, match="exceeds the maximum"):
        ewald.real_space(case["x"], case["m"], case["lam"], case["box"], case["p1"], case["p2"], float(case["rc"]),
                         perp[short], para[short], rv[short])


def test_magnetic_field_takes_the_precomputed_tuples(gpu):
    """A reference caller builds real_precomp / grid_precomp / k_precomp with
    these functions and hands them to magnetic_field."""
    import ewald
    g = golden("operator_n64_g32.npz")
    x, box, xi, P, rc = g["x"], g["box"], float(g["xi"]), int(g["P"]), float(g["rc"])
    Ngrid = g["ngrid"]
    h = box / Ngrid
    eta = P * (h * xi) ** 2 / np.pi
    r = np.arange(0.001, 10.0 + 1e-12, 0.001)
    t = ewald.real_space_table(r, xi)
    perp, para, rv = t[2], t[3], np.concatenate(([0.0], r))
    cell, ncell = ewald.cell_list(x, box, rc)
    p1, p2 = ewald.neighbor_list(cell, ncell)
    offset, offsetxyz = ewald.pre_calculations(P, h)
    real = ewald.prepare_real_space_pairs(x, box, p1, p2, rc, perp, para, rv)
    grid = ewald.prepare_grid_kernel(x, Ngrid, h, xi, eta, P, offset, offsetxyz)
    ks = ewald.precompute_kspace(Ngrid, box, xi, eta)
    assert real[0].size == int(g["npairs"])
    for slab in (0, 1):
        out = ewald.magnetic_field(x, g["m"], g["eps"], box, p1, p2, Ngrid, h, P, xi, eta, rc, offset, offsetxyz,
                                   perp, para, rv, real_precomp=real, grid_precomp=grid, k_precomp=ks,
                                   slab_factor=float(slab))
        ref = g[f"Hk_slab{slab}"] + g[f"Hr_slab{slab}"] - g[f"Hslab_slab{slab}"]
        assert _rel(out, ref) < 5e-11
