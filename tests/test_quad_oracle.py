"""Pins of the dipole + quadrupole oracle (oracle/quad_oracle.py).

The reference's MATLAB (Capacitance_Quadrupole/) cannot run here, so the
restatement is pinned by what the mathematics fixes:
  * the six couplings it shares with the dipole method are the pinned
    ewald_oracle tables (Quad/RealSpaceTable.m repeats those expressions);
  * the eight new tables: continuity at contact r = 2 (the overlap
    regularisation), the r -> 0 limit equal to the self terms, and the bare
    multipole limit at small xi of the field/gradient-dipole couplings
    (gradient of the pinned dipole field coupling);
  * the whole operator, all four blocks: independence of the Ewald
    splitting parameter xi (real and wave space change by ~1e-2 between the
    two xi, their sum by < 2e-6), the symmetry of diag(I, W/2) A with W the
    metric of the 5-component traceless quadrupole, and agreement of the
    dipole block with the dipole oracle;
  * the Hermitian identity that the GPU's real grids use;
  * the solve against a dense solve of the assembled matrix.
"""
import numpy as np
import pytest

from oracle import ewald_oracle as eo
from oracle import quad_oracle as qo

from helpers import rsa

PI = np.pi
W5 = np.array([[2, 0, 0, 1, 0], [0, 2, 0, 0, 0], [0, 0, 2, 0, 0], [1, 0, 0, 2, 0], [0, 0, 0, 0, 2.0]])


def _params(box, xi, tol, odd=False, even=False):
    rc = np.sqrt(-np.log(tol)) / xi
    kc = 2 * xi * np.sqrt(-np.log(tol))
    ng = np.ceil(1 + np.asarray(box) * kc / PI).astype(np.int64)
    if odd:
        ng = ng + (1 - ng % 2)
    if even:
        ng = ng + ng % 2
    P = int(np.ceil(-2 * np.log(tol) / PI))
    return rc, ng, P


def _geometry(x, box, xi, tol, dr=1e-4, odd=False, even=False):
    rc, ng, P = _params(box, xi, tol, odd, even)
    r = np.arange(dr, 7.0, dr)
    tabs, rv = qo.operator_tables(r, xi)
    return qo.QuadGeometry(x, box, ng, P, xi, rc, tabs, rv)


def test_shared_tables_are_the_dipole_tables():
    r = np.array([0.0, 0.5, 1.7, 2.3, 4.0])
    t = dict(zip(qo.TABLE_NAMES, qo.real_space_table(r, 0.8)))
    d = dict(zip(eo.TABLE_NAMES, eo.real_space_table(r[1:], 0.8)))
    for name in eo.TABLE_NAMES:
        assert t[name][0] == d[name][0]
        assert np.array_equal(t[name][2:], d[name][1:])
        assert np.isnan(t[name][1])  # the coupling at r = 0 itself is not finite


@pytest.mark.parametrize("xi", [0.7, 1.2])
def test_quad_tables_continuous_at_contact(xi):
    r = np.array([2.0 - 1e-9, 2.0 + 1e-9])
    t = dict(zip(qo.TABLE_NAMES, qo.real_space_table(r, xi)))
    for name in qo.OPERATOR_TABLES:
        assert abs(t[name][1] - t[name][0]) < 1e-8 * max(1.0, abs(t[name][0])), name


@pytest.mark.parametrize("xi", [0.7, 1.2])
def test_quad_tables_small_r_limit_is_self_term(xi):
    """Quadratic extrapolation from r = 0.04, 0.08, 0.12 to r = 0 (error
    O(r^3) ~ 1e-5) equals the prepended self term."""
    r = np.array([0.0, 0.04, 0.08, 0.12])
    t = dict(zip(qo.TABLE_NAMES, qo.real_space_table(r, xi)))
    for name in qo.OPERATOR_TABLES:
        v = t[name]
        lim = 3 * v[2] - 3 * v[3] + v[4]
        assert abs(lim - v[0]) < 3e-4, (name, lim, v[0])
    assert t["grad_quad_2"][0] == pytest.approx(qo.grad_quad_2_self(xi), rel=0, abs=0)
    assert t["grad_quad_2"][0] > 0.3


def test_field_quad_bare_limit_is_gradient_of_dipole_field():
    """At small xi and r > 2 the couplings approach their unscreened forms:
    A = 1/(4 pi r^3), B = -2/(4 pi r^3) for the field of a dipole, and its
    gradient C = -5 (B - A)/r = 15/(4 pi r^4), D = dA/dr = -3/(4 pi r^4),
    E = 0."""
    r = np.array([2.5, 3.0, 3.5])
    t = dict(zip(qo.TABLE_NAMES, qo.real_space_table(r, 0.02)))
    A = 1 / (4 * PI * r ** 3)
    B = -2 / (4 * PI * r ** 3)
    np.testing.assert_allclose(t["field_dip_1"], A, rtol=1e-3)
    np.testing.assert_allclose(t["field_dip_2"], B, rtol=1e-3)
    np.testing.assert_allclose(t["field_quad_1"], -5 * (B - A) / r, rtol=1e-3)
    np.testing.assert_allclose(t["field_quad_2"], -3 * A / r, rtol=1e-3)
    assert np.max(np.abs(t["field_quad_3"])) < 1e-5 * np.max(np.abs(t["field_quad_2"]))


def test_operator_xi_independence_all_blocks():
    """Hk + Hr of every block (S->E, Q->E, S->G, Q->G) at xi = 1.0 and 1.3."""
    x, box = rsa(8, 0.02, 3, L=12.0)
    rng = np.random.default_rng(1)
    S = rng.standard_normal((8, 3))
    Q = rng.standard_normal((8, 5))
    lam = np.full(8, 5.0)
    res = {}
    for xi in (1.0, 1.3):
        geo = _geometry(x, box, xi, 1e-13, dr=1e-5)
        res[xi] = {}
        for blk, (S2, Q2) in {"S": (S, 0 * Q), "Q": (0 * S, Q)}.items():
            (Ek, Gk), (Er, Gr) = qo.field(geo, S2, Q2, lam, parts=True)
            res[xi][blk] = (Ek + Er, Gk + Gr, Ek, Gk)
    for blk in ("S", "Q"):
        for i, name in enumerate(("E", "G")):
            a, b = res[1.0][blk][i], res[1.3][blk][i]
            rel = np.max(np.abs(a - b)) / np.max(np.abs(a))
            moved = np.max(np.abs(res[1.0][blk][i + 2] - res[1.3][blk][i + 2])) / np.max(np.abs(a))
            assert rel < 2e-6, (blk, name, rel)
            assert moved > 1e-3, (blk, name, moved)  # the split really changed


def test_operator_symmetric_in_quadrupole_metric():
    """diag(I_3N, W/2 per particle) A is symmetric (odd grid: no Nyquist bins)."""
    x, box = rsa(5, 0.02, 3, L=10.0)
    geo = _geometry(x, box, 1.0, 1e-8, odd=True)
    A = qo.dense_operator(geo, np.full(5, 4.0))
    assert np.max(np.abs(A.imag)) == 0.0
    M = np.block([[np.eye(15), np.zeros((15, 25))], [np.zeros((25, 15)), 0.5 * np.kron(np.eye(5), W5)]])
    B = M @ A.real
    assert np.max(np.abs(B - B.T)) < 1e-13 * np.max(np.abs(B))
    M1 = np.block([[np.eye(15), np.zeros((15, 25))], [np.zeros((25, 15)), np.kron(np.eye(5), W5)]])
    B1 = M1 @ A.real
    assert np.max(np.abs(B1 - B1.T)) > 1e-4 * np.max(np.abs(B1))  # the factor 1/2 matters


def test_dipole_block_equals_dipole_operator():
    """With Q = 0 the field part is the dipole operator of ewald_oracle
    (odd grid, slab term off)."""
    x, box = rsa(12, 0.05, 7, L=12.0)
    xi, tol = 0.9, 1e-6
    rc, ng, P = _params(box, xi, tol, odd=True)
    r = eo.default_rtable()
    tabs, rv = qo.operator_tables(r, xi)
    geo = qo.QuadGeometry(x, box, ng, P, xi, rc, tabs, rv)
    dgeo = eo.OperatorGeometry(x, box, ng, P, xi, rc, tabs[0], tabs[1], rv)
    rng = np.random.default_rng(3)
    S = rng.standard_normal((12, 3))
    lam = rng.uniform(2.0, 30.0, 12)
    E, _ = qo.field(geo, S, np.zeros((12, 5)), lam)
    Ed = eo.field(dgeo, S, lam, 0.0)
    assert np.max(np.abs(E - Ed)) < 1e-13 * np.max(np.abs(Ed))


def _hermitian_field(geo, S, Q, lam):
    """The operator with each Nyquist bin's k-space matrix averaged over the
    sign flip of its Nyquist components (what real grids compute)."""
    m8 = np.concatenate([S, Q], axis=1).astype(np.complex128)
    H = qo.spread(m8, geo.Ngrid, geo.linnode, geo.w_spread)
    fH = np.fft.fftn(H, axes=(0, 1, 2))
    kbar = geo.k.copy()
    for d in range(3):
        n = int(geo.Ngrid[d])
        if n % 2 == 0:
            idx = [slice(None)] * 3
            idx[d] = n // 2
            kbar[tuple(idx) + (d,)] *= -1.0
    G = 0.5 * (qo.scale(fH, geo.k, geo.xi, geo.eta) + qo.scale(fH, kbar, geo.xi, geo.eta))
    K = qo.contract(np.fft.ifftn(G, axes=(0, 1, 2)), geo.Ngrid, geo.linnode, geo.w_contract)
    (_, _), (Er, Gr) = qo.field(geo, S, Q, lam, parts=True)
    return K[:, :3] + Er, K[:, 3:] + Gr


def test_real_part_equals_hermitian_operator():
    """For real moments on an even grid the reference's result has a small
    imaginary part from the Nyquist bins; its real part is the Hermitian
    operator (the sign-averaged Nyquist k), whose output is real."""
    x, box = rsa(10, 0.05, 5, L=11.0)
    geo = _geometry(x, box, 0.8, 1e-5, even=True)
    assert np.all(geo.Ngrid % 2 == 0)
    rng = np.random.default_rng(9)
    S = rng.standard_normal((10, 3))
    Q = rng.standard_normal((10, 5))
    lam = np.full(10, 6.0)
    E, G = qo.field(geo, S, Q, lam)
    Eh, Gh = _hermitian_field(geo, S, Q, lam)
    assert np.max(np.abs(Eh.imag)) < 1e-15 and np.max(np.abs(Gh.imag)) < 1e-15
    assert np.max(np.abs(E.real - Eh.real)) < 1e-14 * np.max(np.abs(E))
    assert np.max(np.abs(G.real - Gh.real)) < 1e-14 * np.max(np.abs(G))
    assert np.max(np.abs(E.imag)) > 0.0  # the reference's Nyquist bins do leave an imaginary part


def test_solve_matches_dense_solve():
    x, box = rsa(6, 0.03, 2, L=10.0)
    geo = _geometry(x, box, 1.0, 1e-6)
    lam = np.full(6, 8.0)
    S, Q = qo.solve(geo, lam, [0.0, 0.0, 1.0], np.zeros((6, 3)), np.zeros((6, 5)), 1e-12)
    A = qo.dense_operator(geo, lam)
    b = np.concatenate([np.tile([0.0, 0.0, 1.0], 6), np.zeros(30)])
    sol = np.linalg.solve(A, b)
    assert np.max(np.abs(S.reshape(-1) - sol[:18])) < 1e-10 * np.max(np.abs(sol))
    assert np.max(np.abs(Q.reshape(-1) - sol[18:])) < 1e-10 * np.max(np.abs(sol))
    assert np.max(np.abs(Q)) > 1e-3 * np.max(np.abs(S))  # quadrupoles are induced


# A lone sphere in a cubic periodic box (a simple-cubic lattice) under an
# imposed traceless field gradient g (right-hand side [0; g]): the inversion
# symmetry of the lattice decouples S from Q (S = 0), and the quadrupole is
# the l = 2 response of a dielectric sphere, Q = g / D2 with
# D2 = 3/(4 pi) (eps + 3/2)/(eps - 1), whose eps-dependence is the l = 2
# Clausius-Mossotti factor (eps - 1)/(eps + 3/2).  Only the images' rank-4
# lattice sum, O((a/L)^5) (the cubic splitting of the diagonal and the
# off-diagonal components, 2e-5 at L = 16), perturbs it.  This fixes the
# dielectric term 15/(8 pi (eps-1)) (Capacitance_Quadrupole/RealSpace.m:29)
# against the quadrupole self terms: with 15/(4 pi (eps-1)) D2 would be
# 30-50 % off.
def _lone_quad_geometry(L=16.0, xi=0.7, tol=1e-10):
    rc, ng, P = _params(np.full(3, L), xi, tol)
    r = np.arange(1e-3, 11.0, 1e-3)
    tabs, rv = qo.operator_tables(r, xi)
    return qo.QuadGeometry(np.array([[3.1, 7.4, 11.9]]), np.full(3, L), ng, P, xi, rc, tabs, rv)


def lone_quad_expected(eps):
    return 4.0 * PI / 3.0 * (eps - 1.0) / (eps + 1.5)


def test_lone_sphere_quadrupole_clausius_mossotti():
    geo = _lone_quad_geometry()
    assert geo.rc < 16.0
    g = np.array([1.0, 0.3, -0.2, -0.5, 0.4])
    resp = {}
    for eps in (4.0, 0.3, 40.0, 1.5):
        A = qo.dense_operator(geo, np.array([eps]))
        sol = np.linalg.solve(A, np.concatenate([np.zeros(3), g]))
        S, Q = sol[:3], sol[3:].real
        assert np.max(np.abs(S)) < 1e-10
        ref = lone_quad_expected(eps) * g
        assert np.max(np.abs(Q - ref)) <= 1e-4 * np.max(np.abs(ref)), (eps, Q / g, lone_quad_expected(eps))
        resp[eps] = Q
    # the ratio between permittivities needs no normalisation
    ratio = resp[4.0] / resp[40.0]
    assert np.max(np.abs(ratio / (lone_quad_expected(4.0) / lone_quad_expected(40.0)) - 1.0)) < 3e-5
