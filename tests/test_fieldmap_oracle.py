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
