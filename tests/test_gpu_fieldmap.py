"""FieldMap on the GPU against the oracle's restatement of
Compute_Field/FieldMap.m (1e-11 relative), complex dipoles, and a probe at a
particle centre raising ValueError.  The GPU grids are real, so for real
dipoles it gives the real part of the reference's result (which carries an
even-grid Nyquist-plane imaginary part of ~1e-8); complex dipoles are
compared part by part."""
import numpy as np
import pytest

from helpers import rsa
from oracle import ewald_oracle as eo

pytestmark = pytest.mark.gpu


def test_fieldmap_matches_oracle(gpu):
    import fieldmap
    x, box = rsa(40, 0.15, 8)
    rng = np.random.default_rng(5)
    probes = []
    while len(probes) < 25:
        c = rng.random(3) * box
        d = c - x
        d -= box * np.rint(d / box)
        if np.min(np.sum(d * d, axis=1)) > 1.0:  # outside every sphere
            probes.append(c)
    xe = np.array(probes)
    p = rng.standard_normal((40, 3)) + 1j * rng.standard_normal((40, 3))
    print("case fieldmap n_p=40 n_e=25", flush=True)
    E = fieldmap.FieldMap(xe, x, p, box, 0.8)
    ref = (np.real(eo.field_map(xe, x, p.real, box, 0.8)) + 1j * np.real(eo.field_map(xe, x, p.imag, box, 0.8)))
    assert np.max(np.abs(E - ref)) / np.max(np.abs(ref)) < 1e-11


def test_fieldmap_probe_at_particle_centre_raises(gpu):
    import fieldmap
    x, box = rsa(20, 0.1, 2)
    with pytest.raises(ValueError):
        fieldmap.FieldMap(x[:1], x, np.ones((20, 3)), box, 0.8)


def _probe_config(seed, n_p=6, n_e=5, L=16.0, min_gap=0.7):
    rng = np.random.default_rng(seed)
    box = np.full(3, L)
    xp = rng.random((n_p, 3)) * box
    xe = []
    while len(xe) < n_e:
        c = rng.random(3) * box
        d = xp - c
        d -= box * np.rint(d / box)
        if np.min(np.sqrt(np.sum(d * d, axis=1))) > min_gap:
            xe.append(c)
    return xp, np.array(xe), box


def test_quadrupole_field_map_matches_oracle(gpu):
    """Compute_Field_Quadrupole/FieldMap.m on the GPU (the dipole +
    quadrupole operator with zero-moment probes) against the oracle's
    restatement: real moments, so the GPU's Hermitian-part grids match the
    oracle's complex result's real part to rounding."""
    import fieldmap
    from oracle import quad_oracle as qo
    xp, xe, box = _probe_config(21)
    rng = np.random.default_rng(22)
    S = rng.standard_normal((xp.shape[0], 3))
    Q = rng.standard_normal((xp.shape[0], 5))
    E = fieldmap.FieldMapQuadrupole(xe, xp, S, Q, box, 0.8)
    ref = qo.field_map(xe, xp, S, Q, box, 0.8)
    assert E.shape == (xe.shape[0], 3) and E.dtype == np.complex128
    assert np.max(np.abs(E.real - ref.real)) < 1e-11 * np.max(np.abs(ref.real))
    # lone sphere: the point-quadrupole field (the oracle's pin) through the GPU map
    from test_fieldmap_oracle import LONE_BOX, LONE_U, LONE_X, _quadrupole_field
    q = np.array([[0.4, -0.3, 0.2, -0.1, 0.5]])
    Eq = fieldmap.FieldMapQuadrupole(LONE_X + 2.2 * LONE_U, LONE_X, np.zeros((1, 3)), q, LONE_BOX, 0.6)[0]
    refq = _quadrupole_field(q[0], 2.2 * LONE_U)
    assert np.max(np.abs(Eq.real - refq)) / np.max(np.abs(refq)) < 1.5e-2


def test_sizes_field_map_matches_oracle(gpu):
    """Compute_Field_Sizes/FieldMap.m on the GPU (the polydisperse operator
    with zero-dipole probes of the smallest radius, overlap-free pair tables)
    against the oracle's restatement."""
    import fieldmap
    from oracle import sizes_oracle as so
    xp, xe, box = _probe_config(31)
    rng = np.random.default_rng(32)
    a = rng.choice([0.7, 1.0, 1.3], size=xp.shape[0])
    p = rng.standard_normal((xp.shape[0], 3))
    E = fieldmap.FieldMapSizes(xe, xp, a, p, box, 0.8)
    ref = so.field_map(xe, xp, a, p, box, 0.8)
    assert np.max(np.abs(E.real - ref.real)) < 1e-11 * np.max(np.abs(ref.real))
    with pytest.raises(ValueError, match="closer than 0.5"):
        fieldmap.FieldMapSizes(xp[:1] + 0.2, xp, a, p, box, 0.8)
