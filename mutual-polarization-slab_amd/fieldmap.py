"""Field at probe points from particle dipoles on the MI355X: the interface
of the reference's Compute_Field/FieldMap.m as a Python function of the same
name (0-based arrays, C order).

    FieldMap(x_E, x_p, p, box, xi, p1=None, p2=None) -> E (N_E, 3) complex128

As there (FieldMap.m:17-55): Ewald tolerance 1e-3 and the parameter rule for
rc, the grid, P and eta; the real-space couplings from tables without the
(r < 2) overlap terms (the field of a sphere's dipole at a point,
Compute_Field/RealSpaceTable.m); the particles' dipoles spread to the grid,
scaled and contracted at the probes (MagneticField.m:33-60) plus the real-
space sum over probe-particle pairs with 0 < d < rc (RealSpace.m:23-57); and
E = -(Hk + Hr).  No slab term, no self term.

The GPU evaluates it with the dipole operator on the particles and the probes
together, the probes carrying zero dipoles: a probe's row of the operator is
then exactly the wave-space field contracted at the probe plus the real-space
sum over the particles within rc (pairs of two probes and the particles' own
rows contribute nothing to it).  p1, p2 (the reference's optional neighbour
list) are accepted and not needed.  A probe at a particle's centre (d = 0,
which the reference skips) raises ValueError.
"""
from __future__ import annotations

import numpy as np

from mpol_hip import context as _ctxmod

__all__ = ["FieldMap", "field_map"]

_TABLES: dict = {}


def _field_tables(xi):
    key = float(xi)
    if key not in _TABLES:
        r = np.arange(0.001, 10.0 + 1e-12, 0.001, dtype=np.float64)
        t = _ctxmod.real_space_table_field(r, xi)
        _TABLES[key] = (t[2], t[3], np.concatenate(([0.0], r)))
    return _TABLES[key]


def FieldMap(x_E, x_p, p, box, xi, p1=None, p2=None):
    """E (N_E, 3) complex128: the field at x_E (N_E, 3) of the dipoles p
    (N_p, 3) at x_p (N_p, 3) in the periodic box (FieldMap.m:1-57)."""
    del p1, p2
    x_E = np.asarray(x_E, dtype=np.float64).reshape(-1, 3)
    x_p = np.asarray(x_p, dtype=np.float64).reshape(-1, 3)
    p = np.asarray(p).reshape(-1, 3)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    if p.shape[0] != x_p.shape[0]:
        raise ValueError("p must have one dipole per particle")
    errortol = 1e-3
    rc = np.sqrt(-np.log(errortol)) / xi
    kcut = 2 * xi * np.sqrt(-np.log(errortol))
    ngrid = np.ceil(1 + box * kcut / np.pi).astype(np.int64)
    h = box / ngrid
    P = int(np.ceil(-2 * np.log(errortol) / np.pi))
    eta = P * (h * xi) ** 2 / np.pi
    if np.any(rc > box / 2):
        raise ValueError("Real space cutoff (rc = %.3f) larger than half the box size (%.3f, %.3f, %.3f). "
                         "Accuracy to specified error tolerance not guaranteed." % (rc, box[0], box[1], box[2]))
    n_p, n_e = x_p.shape[0], x_E.shape[0]
    if n_e == 0:
        return np.zeros((0, 3), dtype=np.complex128)
    perp, para, rv = _field_tables(xi)
    ctx = _ctxmod.Context()
    try:
        ctx.set_tables(perp, para, rv)
        try:
            ctx.set_geometry(np.vstack([x_p, x_E]), box, ngrid, P, xi, eta, rc)
        except ValueError as e:
            raise ValueError(f"{e} (a probe at a particle centre?)") from None
        ctx.set_permittivity(np.full(n_p + n_e, 2.0))  # any eps != 1: the probes' dipoles are zero
        m = np.zeros((2, n_p + n_e, 3))
        m[0, :n_p] = np.real(p)
        m[1, :n_p] = np.imag(p) if np.iscomplexobj(p) else 0.0
        out = ctx.matvec(m, slab_factor=0.0).cpu().numpy()
    finally:
        ctx.close()
    return -(out[0, n_p:] + 1j * out[1, n_p:])


field_map = FieldMap
