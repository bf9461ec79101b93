"""Pins of the polydisperse oracle (oracle/sizes_oracle.py).

The reference's MATLAB (Capacitance_Sizes/) cannot run here; the restatement
is pinned by:
  * unit radii: its real-space tables equal the pinned dipole tables
    (ewald_oracle, checked against the reference's own golden vectors), and
    the whole operator equals the dipole operator to the spectral tolerance
    (different grids, P and splitting of the Gaussian);
  * mixed radii: the total operator does not depend on xi (real and wave
    space each change by ~1e-2), the operator is symmetric, the tables are
    continuous at contact (r = a_i + a_j) and containment (r = |a_i - a_j|),
    and their r -> 0 limit is the tabulated self value;
  * the small-d branch of the spread kernel is its d -> 0 limit (for unit
    radii, where the reference's exp(-2 a xi^2/eta) equals exp(-2 a^2 xi^2/eta));
  * uniquetol groups radii like MATLAB's (tolerance 1e-6 of the largest).
"""
import numpy as np
import pytest

from oracle import ewald_oracle as eo
from oracle import sizes_oracle as so

from helpers import rsa


def _mixed_config(n, phi, seed, L=None):
    x, box = rsa(n, phi, seed, L=L)
    rng = np.random.default_rng(seed + 7)
    a = rng.choice([0.8, 1.0, 1.25], size=n)
    return x, box, a


def _tables(a, xi, dr=1e-3, rmax=10.0):
    au, _ = so.uniquetol(a)
    r = np.concatenate(([0.0], np.arange(dr, rmax + dr / 2, dr)))
    perp, para = so.real_space_table(r, au, xi)
    return r, perp, para


def test_uniquetol_groups_like_matlab():
    a = np.array([1.0, 1.0 + 5e-7, 2.0, 0.5, 2.0 - 1e-6, 0.5 + 3e-6])
    u, ind = so.uniquetol(a)
    assert len(u) == 4  # 0.5+3e-6 differs from 0.5 by more than 1e-6 * 2.0 = 2e-6
    assert ind[0] == ind[1] and ind[2] == ind[4] and ind[3] != ind[5]
    assert np.all(np.isin(u, a))


def test_unit_radius_tables_are_the_dipole_tables():
    xi = 0.9
    r = np.concatenate(([0.0], np.linspace(0.05, 9.0, 300)))
    perp, para = so.real_space_table(r, [1.0], xi)
    t = eo.real_space_table(r[1:], xi)
    np.testing.assert_allclose(perp[:, 0], t[2], rtol=0, atol=1e-11 * np.max(np.abs(t[2])))
    np.testing.assert_allclose(para[:, 0], t[3], rtol=0, atol=1e-11 * np.max(np.abs(t[3])))


@pytest.mark.parametrize("ai,aj", [(1.0, 1.3), (1.3, 0.7), (0.8, 0.8)])
def test_tables_continuous_and_self_limit(ai, aj):
    xi = 0.9
    au = sorted({ai, aj})
    col = 1 if ai != aj else 0
    for p in (ai + aj, abs(ai - aj)):
        if p <= 0:
            continue
        pe, pa = so.real_space_table(np.array([0.0, p - 1e-9, p + 1e-9]), au, xi)
        assert abs(pe[2, col] - pe[1, col]) < 1e-9 and abs(pa[2, col] - pa[1, col]) < 1e-9
    pe, pa = so.real_space_table(np.array([0.0, 0.04, 0.08, 0.12]), au, xi)
    for t in (pe, pa):
        lim = 3 * t[1, col] - 3 * t[2, col] + t[3, col]
        assert abs(lim - t[0, col]) < 2e-4 * abs(t[0, col])


def test_small_d_kernel_branch_is_the_limit():
    """Full kernel / d at d = 3e-3 .. 1e-3 tends to the small-d branch
    (relative O(d^2)); below ~1e-4 the full form loses digits to cancellation,
    which is why the reference switches branches."""
    xi, eta = 0.9, 0.6
    u = np.array([0.6, -0.4, 0.2])
    u /= np.linalg.norm(u)
    branch = so.kernel(np.array([u * 1e-7]), 1.0, xi, eta)[0] / 1e-7
    prev = None
    for d in (1e-1, 1e-2, 1e-3):
        dev = np.max(np.abs(so.kernel(np.array([u * d]), 1.0, xi, eta)[0] / d / branch - 1.0))
        if prev is not None:
            assert dev < prev / 50  # quadratic convergence
        prev = dev
    assert prev < 1e-6


def test_unit_radii_operator_equals_dipole_operator():
    x, box = rsa(10, 0.04, 4, L=11.0)
    a = np.ones(10)
    xi, tol = 1.0, 1e-10
    r, perp, para = _tables(a, xi, dr=1e-4, rmax=6.0)
    geo = so.SizesGeometry(x, a, box, xi, tol, r, perp, para)
    rng = np.random.default_rng(2)
    p = rng.standard_normal((10, 3))
    eps = rng.uniform(2, 20, 10)
    Es = so.field(geo, p, eps)
    rc = np.sqrt(-np.log(tol)) / xi
    ng = np.ceil(1 + box * 2 * xi * np.sqrt(-np.log(tol)) / np.pi).astype(int)
    P = int(np.ceil(-2 * np.log(tol) / np.pi))
    dgeo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp[:, 0], para[:, 0], r)
    Ed = eo.field(dgeo, p, eps, 0.0)
    assert np.max(np.abs(Es - Ed)) < 1e-8 * np.max(np.abs(Ed))


def test_mixed_radii_xi_independence():
    x, box, a = _mixed_config(10, 0.04, 5, L=13.0)
    rng = np.random.default_rng(3)
    p = rng.standard_normal((10, 3))
    eps = rng.uniform(2, 20, 10)
    out = {}
    for xi in (1.0, 1.3):
        r, perp, para = _tables(a, xi, dr=1e-5, rmax=6.0)
        geo = so.SizesGeometry(x, a, box, xi, 1e-12, r, perp, para)
        out[xi] = so.field(geo, p, eps, parts=True)
    tot = {k: v[0] + v[1] for k, v in out.items()}
    rel = np.max(np.abs(tot[1.0] - tot[1.3])) / np.max(np.abs(tot[1.0]))
    moved = np.max(np.abs(out[1.0][0] - out[1.3][0])) / np.max(np.abs(tot[1.0]))
    assert rel < 1e-6, rel
    assert moved > 1e-3, moved


def test_mixed_radii_operator_symmetric():
    x, box, a = _mixed_config(8, 0.05, 6, L=12.0)
    r, perp, para = _tables(a, 1.0, dr=1e-4, rmax=6.0)
    geo = so.SizesGeometry(x, a, box, 1.0, 1e-8, r, perp, para)
    A = so.dense_operator(geo, np.full(8, 5.0))
    assert np.max(np.abs(A.imag)) < 1e-13 * np.max(np.abs(A))
    A = A.real
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.max(np.abs(A))


def test_solve_matches_dense_solve():
    x, box, a = _mixed_config(8, 0.05, 8, L=12.0)
    r, perp, para = _tables(a, 1.0)
    geo = so.SizesGeometry(x, a, box, 1.0, 1e-6, r, perp, para)
    eps = np.full(8, 7.0)
    p = so.solve(geo, eps, [0.0, 0.0, 1.0], np.zeros((8, 3)), 1e-12)
    A = so.dense_operator(geo, eps)
    sol = np.linalg.solve(A, np.tile([0.0, 0.0, 1.0], 8)).reshape(8, 3)
    assert np.max(np.abs(p - sol)) < 1e-10 * np.max(np.abs(sol))


# A lone sphere of radius a in a cubic periodic box is a simple-cubic lattice
# of one sphere per cell.  At the dipole level the field of the periodic
# images at a lattice site (k = 0 excluded) is the Lorentz field p/(3V), so
# p = 4 pi a^3 beta (E0 + p/(3V)), i.e. p = 4 pi a^3 beta/(1 - beta phi) E0
# with beta = (eps-1)/(eps+2) and phi = 4 pi a^3/(3V): the Clausius-Mossotti
# polarisability of a sphere of radius a.  This fixes the diagonal of the
# operator -- the dielectric term 3p/(4 pi a^3 (eps-1))
# (Capacitance_Sizes/RealSpace.m:36) together with the radius-dependent self
# terms -- at every radius: a wrong power of a, or a wrong coefficient, is
# off by O(1).  No pairs: the images lie beyond rc.
LONE_CASES = [(0.6, 4.0), (1.0, 4.0), (1.7, 4.0), (1.3, 0.3), (2.2, 40.0)]


def lone_sphere_expected(a, eps, L):
    beta = (eps - 1.0) / (eps + 2.0)
    phi = 4.0 * np.pi * a ** 3 / (3.0 * L ** 3)
    return 4.0 * np.pi * a ** 3 * beta / (1.0 - beta * phi)


@pytest.mark.parametrize("a,eps", LONE_CASES)
def test_lone_sphere_clausius_mossotti(a, eps):
    L, xi, tol = 20.0, 0.6, 1e-10
    x = np.array([[3.1, 7.4, 11.9]])
    geo = so.SizesGeometry(x, np.array([a]), np.array([L, L, L]), xi, tol, None, None, None, exact=True)
    assert geo.rc < L  # the periodic images are outside the real-space sum
    A = so.dense_operator(geo, np.array([eps]))
    p = np.linalg.solve(A, np.array([0.0, 0.0, 1.0])).real
    ref = lone_sphere_expected(a, eps, L)
    assert abs(p[2] - ref) <= 2e-6 * abs(ref), (p, ref)
    assert np.max(np.abs(p[:2])) <= 1e-9 * abs(ref)
