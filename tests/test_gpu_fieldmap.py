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
