"""Field-map oracle (Compute_Field/FieldMap.m restated) pinned by the
physics: outside a lone sphere in a large periodic box the field is the
point-dipole field (3 (p.r) r - p) / (4 pi d^3), also between d = 1 and 2,
where only the overlap-free tables give it; and the library's overlap-free
table routine against the oracle's."""
import numpy as np
import pytest

from oracle import ewald_oracle as eo


def _dipole_field(p, u, d):
    return (3.0 * (p @ u) * u - p) / (4.0 * np.pi * d ** 3)


@pytest.mark.parametrize("d", [1.2, 1.6, 2.5, 4.0])
def test_lone_sphere_field_is_point_dipole_field(d):
    box = np.full(3, 40.0)
    xp = np.array([[20.0, 20.0, 20.0]])
    p = np.array([[0.3, -0.2, 1.0]])
    u = np.array([0.48, 0.6, 0.64])
    u /= np.linalg.norm(u)
    E = eo.field_map(xp + d * u, xp, p, box, 0.6)[0]
    ref = _dipole_field(p[0], u, d)
    # real dipoles: only the reference's even-grid Nyquist-plane term is imaginary
    assert np.max(np.abs(E.imag)) < 1e-5 * np.max(np.abs(E.real))
    assert np.max(np.abs(E.real - ref)) / np.max(np.abs(ref)) < 5e-3  # Ewald tolerance 1e-3, periodic images
    if d < 2.0:
        Ereg = eo.field_map(xp + d * u, xp, p, box, 0.6, regularized_tables=True)[0]
        assert np.max(np.abs(Ereg.real - ref)) / np.max(np.abs(ref)) > 5e-2


def test_field_is_linear_in_the_dipoles():
    rng = np.random.default_rng(3)
    box = np.full(3, 16.0)
    xp = rng.random((6, 3)) * box
    xe = rng.random((5, 3)) * box
    p1, p2 = rng.standard_normal((6, 3)), rng.standard_normal((6, 3))
    a = eo.field_map(xe, xp, p1, box, 0.8)
    b = eo.field_map(xe, xp, p2, box, 0.8)
    c = eo.field_map(xe, xp, 2.0 * p1 - 1j * p2, box, 0.8)
    assert np.max(np.abs(c - (2.0 * a - 1j * b))) < 1e-12 * np.max(np.abs(c))


def test_library_field_tables_match_oracle():
    from mpol_hip import real_space_table, real_space_table_field
    r = np.array([0.05, 0.5, 1.2, 1.9, 2.0, 2.5, 5.0, 9.9])
    for xi in (0.5, 1.0):
        a = real_space_table_field(r, xi)
        b = eo.real_space_table(r, xi, regularize=False)
        reg = real_space_table(r, xi)
        for k in range(6):
            assert np.max(np.abs(a[k] - b[k])) <= 1e-12 * max(1.0, np.max(np.abs(b[k])))
            # identical to the sphere-sphere tables from contact on
            assert np.all(a[k][1:][r >= 2.0] == reg[k][1:][r >= 2.0])


# ---------------------------------------------------------------------------
# dipole + quadrupole and polydisperse field maps
# (Compute_Field_Quadrupole/FieldMap.m, Compute_Field_Sizes/FieldMap.m)
# ---------------------------------------------------------------------------
def _quad_full(q5):
    xx, xy, xz, yy, yz = q5
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, -xx - yy]])


def _quadrupole_field(q5, r, h=1e-5):
    """Field of a point quadrupole, the second-order term of the multipole
    expansion: E = -1/2 Q_ij d_i E_dip(e_j) (central differences of the
    point-dipole field)."""
    Q = _quad_full(q5)
    d = np.linalg.norm(r)
    E = np.zeros(3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        for j in range(3):
            ej = np.eye(3)[j]
            dp = _dipole_field(ej, (r + e) / np.linalg.norm(r + e), np.linalg.norm(r + e))
            dm = _dipole_field(ej, (r - e) / np.linalg.norm(r - e), np.linalg.norm(r - e))
            E += -0.5 * Q[i, j] * (dp - dm) / (2 * h)
    del d
    return E


LONE_BOX = np.full(3, 40.0)
LONE_X = np.array([[20.0, 20.0, 20.0]])
LONE_U = np.array([0.48, 0.6, 0.64]) / np.linalg.norm([0.48, 0.6, 0.64])


@pytest.mark.parametrize("d", [2.2, 3.0])
def test_quadrupole_map_lone_sphere_multipole_fields(d):
    """Outside a lone sphere (inside rc): its dipole gives the point-dipole
    field and its quadrupole the point-quadrupole field -1/2 Q:grad E_dip, to
    the map's Ewald tolerance (1e-3) -- a wrong factor or sign of the
    quadrupole term is off by 50 % or more."""
    from oracle import quad_oracle as qo
    xe = LONE_X + d * LONE_U
    p = np.array([[0.3, -0.2, 1.0]])
    E = qo.field_map(xe, LONE_X, p, np.zeros((1, 5)), LONE_BOX, 0.6)[0]
    ref = _dipole_field(p[0], LONE_U, d)
    assert np.max(np.abs(E.real - ref)) / np.max(np.abs(ref)) < 5e-3
    q = np.array([[0.4, -0.3, 0.2, -0.1, 0.5]])
    E = qo.field_map(xe, LONE_X, np.zeros((1, 3)), q, LONE_BOX, 0.6)[0]
    ref = _quadrupole_field(q[0], d * LONE_U)
    assert np.max(np.abs(E.real - ref)) / np.max(np.abs(ref)) < 1.5e-2


@pytest.mark.parametrize("a,d", [(0.7, 1.64), (0.7, 3.5), (1.3, 2.96), (1.3, 3.5)])
def test_sizes_map_lone_sphere_dipole_field(a, d):
    """Outside a lone sphere of radius a, with the probe treated as a sphere
    of the smallest radius (here a) not overlapping it, the field is the
    point-dipole field."""
    from oracle import sizes_oracle as so
    xe = LONE_X + d * LONE_U
    p = np.array([[0.3, -0.2, 1.0]])
    E = so.field_map(xe, LONE_X, np.array([a]), p, LONE_BOX, 0.6)[0]
    ref = _dipole_field(p[0], LONE_U, d)
    assert np.max(np.abs(E.real - ref)) / np.max(np.abs(ref)) < 5e-3


def test_sizes_map_two_radii():
    """Two spheres of radii 0.6 and 1.5 far apart, only the large one
    polarised: at a probe 2.3 from it (probes take the smallest radius, 0.6:
    no overlap) the field is its point-dipole field."""
    from oracle import sizes_oracle as so
    xp = np.array([[20.0, 20.0, 20.0], [8.0, 30.0, 10.0]])
    p = np.array([[0.3, -0.2, 1.0], [0.0, 0.0, 0.0]])
    xe = xp[:1] + 2.3 * LONE_U
    E = so.field_map(xe, xp, np.array([1.5, 0.6]), p, LONE_BOX, 0.6)[0]
    ref = _dipole_field(p[0], LONE_U, 2.3)
    assert np.max(np.abs(E.real - ref)) / np.max(np.abs(ref)) < 5e-3
    with pytest.raises(ValueError, match="closer"):
        so.field_map(xp[:1] + 0.3 * LONE_U, xp, np.array([1.5, 0.6]), p, LONE_BOX, 0.6)


def test_library_quad_and_sizes_field_tables_match_oracle():
    from mpol_hip import (real_space_table_quad, real_space_table_quad_field, real_space_table_sizes,
                          real_space_table_sizes_field)
    from oracle import quad_oracle as qo
    from oracle import sizes_oracle as so
    r = np.array([0.05, 0.5, 1.2, 1.9, 2.0, 2.5, 5.0, 9.9])
    for xi in (0.5, 1.0):
        a = real_space_table_quad_field(r, xi)  # pot_quad, field_quad_1..3, grad_quad_1..4 (self first)
        b = dict(zip(qo.TABLE_NAMES, qo.real_space_table(r, xi, regularize=False)))
        reg = real_space_table_quad(r, xi)
        names = ("pot_quad", "field_quad_1", "field_quad_2", "field_quad_3", "grad_quad_1", "grad_quad_2",
                 "grad_quad_3", "grad_quad_4")
        deep = r < 0.1  # deep overlap: the 1/r^n terms cancel in both evaluations
        for k, nm in enumerate(names):
            assert np.max(np.abs(a[k][1:][~deep] - b[nm][~deep])) <= 1e-12 * max(1.0, np.max(np.abs(b[nm])))
            assert np.all(np.abs(a[k][1:][deep] - b[nm][deep]) <= 1e-5 * np.abs(b[nm][deep]))
            assert np.all(a[k][1:][r >= 2.0] == reg[k][1:][r >= 2.0])
        radii = np.array([0.6, 1.0, 1.5])
        pf, qf = real_space_table_sizes_field(r, radii, xi)
        po, qo_ = so.real_space_table_field(r, radii, xi)
        assert np.max(np.abs(pf - po)) <= 1e-12 * max(1.0, np.max(np.abs(po)))
        assert np.max(np.abs(qf - qo_)) <= 1e-12 * max(1.0, np.max(np.abs(qo_)))
        rr = np.concatenate(([0.0], r))
        pr, qr = real_space_table_sizes(rr, radii, xi)
        col = 0
        for i in range(3):
            for j in range(i, 3):
                far = r >= radii[i] + radii[j]
                assert np.all(pf[far, col] == pr[1:][far, col]) and np.all(qf[far, col] == qr[1:][far, col])
                col += 1
