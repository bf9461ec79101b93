"""The FP64 oracle against golden vectors written by the reference itself
(tests/golden/make_golden.py imports "CPU - Ewald test/ewald.py" and
capacitance.py).  CPU only."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from helpers import relerr
from oracle import capacitance_oracle as co
from oracle import ewald_oracle as eo

TABLE_FILES = sorted(glob.glob(os.path.join(GOLDEN, "tables_*.npz")))
OPERATOR_FILES = sorted(glob.glob(os.path.join(GOLDEN, "operator_*.npz")))
SOLVE_FILES = sorted(glob.glob(os.path.join(GOLDEN, "solve_*.npz")))


@pytest.mark.parametrize("path", TABLE_FILES, ids=os.path.basename)
def test_tables_match_reference(path):
    d = np.load(path)
    r = d["r"]
    tabs = eo.real_space_table(r, float(d["xi"]))
    for k in range(6):
        assert tabs[k][0] == pytest.approx(d["self_terms"][k], rel=1e-14, abs=1e-300)
        ref = d["tables"][k]
        scale = np.max(np.abs(ref))
        # r >= 0.1: both evaluations are accurate to rounding
        big = r >= 0.1
        assert np.max(np.abs(tabs[k][1:][big] - ref[big])) / scale < 1e-10, eo.TABLE_NAMES[k]
        # r < 0.1 (deep overlap, never met by spheres that do not overlap):
        # the 1/r^3..1/r^4 terms cancel catastrophically in both evaluations;
        # only the couplings used by the solver are held to a bound there
        if k in (2, 3):
            assert np.max(np.abs(tabs[k][1:] - ref)) / scale < 1e-5


@pytest.mark.parametrize("path", OPERATOR_FILES, ids=os.path.basename)
def test_operator_matches_reference(path):
    d = np.load(path)
    xi = float(d["xi"])
    perp, para, rv = co.default_tables(xi)
    geo = eo.OperatorGeometry(d["x"], d["box"], d["ngrid"], int(d["P"]), xi, float(d["rc"]), perp, para, rv)
    assert geo.real[0].size == int(d["npairs"])
    for s in (0, 1):
        Hk, Hr, Hs = eo.field_components(geo, d["m"], d["eps"], s)
        assert relerr(Hk, d[f"Hk_slab{s}"]) < 1e-13
        assert relerr(Hr, d[f"Hr_slab{s}"]) < 1e-13
        np.testing.assert_allclose(Hs, d[f"Hslab_slab{s}"], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("path", SOLVE_FILES, ids=os.path.basename)
def test_solve_matches_reference(path):
    d = np.load(path)
    op = golden("operator_n64_g21.npz")
    assert np.array_equal(op["x"], d["x"])
    xi = float(op["xi"])
    perp, para, rv = co.default_tables(xi)
    geo = eo.OperatorGeometry(d["x"], d["box"], op["ngrid"], int(op["P"]), xi, float(op["rc"]), perp, para, rv)
    slab = 1.0 if "slab1" in path else 0.0
    p, info = co.solve_dipoles(geo, d["eps"], d["H"], d["guess"], 1e-12, slab, use_scipy_gmres=True)
    assert relerr(p, d["p_scipy"]) < 1e-10
    p2, _ = co.solve_dipoles(geo, d["eps"], d["H"], d["guess"], 1e-12, slab, use_scipy_gmres=False)
    assert relerr(p2, d["p_python"]) < 1e-10
    # both reference paths agree with each other at the tolerance
    assert relerr(d["p_scipy"], d["p_python"]) < 1e-10


@pytest.mark.parametrize("slab", [0, 1])
def test_run_capacitance_matches_reference(slab):
    d = golden(f"run_n64_slab{slab}.npz")
    C, p = co.run_capacitance(d["x"], d["box"], d["eps"], float(d["xi"]), d["H0"], errortol=float(d["errortol"]),
                              slab_factor=float(slab))
    # both are GMRES iterates at tol 1e-5 from the same start, same algorithm
    assert relerr(p, d["p"]) < 1e-9
    assert abs(C[0] - d["C"][0]) / abs(d["C"][0]) < 1e-9
