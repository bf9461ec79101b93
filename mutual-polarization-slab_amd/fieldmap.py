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

__all__ = ["FieldMap", "field_map", "FieldMapQuadrupole", "FieldMapSizes"]

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


# ---------------------------------------------------------------------------
# dipoles and quadrupoles (Compute_Field_Quadrupole/FieldMap.m)
# ---------------------------------------------------------------------------
_QTABLES: dict = {}


def _quad_field_tables(xi):
    """The nine operator tables without overlap terms on the field map's
    r = [0, 1e-4:1e-4:10] (Compute_Field_Quadrupole/FieldMap.m:23-25; its
    RealSpaceTable prepends no self terms, so entry k is the coupling at
    r_vals[k]; entry 0, never read for a probe, is the self term here)."""
    key = float(xi)
    if key not in _QTABLES:
        r = np.arange(1, 100001, dtype=np.float64) * 1e-4
        d = _ctxmod.real_space_table_field(r, xi)
        q = _ctxmod.real_space_table_quad_field(r, xi)
        _QTABLES[key] = (np.stack([d[2], d[3]] + [q[k] for k in range(1, 8)]), np.concatenate(([0.0], r)))
    return _QTABLES[key]


def FieldMapQuadrupole(x_E, x_p, p, Q, box, xi):
    """E (N_E, 3) complex128: the field at x_E (N_E, 3) of the dipoles p
    (N_p, 3) and quadrupoles Q (N_p, 5; xx, xy, xz, yy, yz) at x_p in the
    periodic box (Compute_Field_Quadrupole/FieldMap.m:1-54): tolerance 1e-3
    and its parameter rule, tables without overlap terms, the wave-space
    field contracted at the probes plus the real-space sum over probe-
    particle pairs with d < rc (RealSpace.m:27-100), E = -(Ek + Er).  On the
    GPU: the dipole + quadrupole operator on particles and probes together,
    the probes carrying zero moments."""
    x_E = np.asarray(x_E, dtype=np.float64).reshape(-1, 3)
    x_p = np.asarray(x_p, dtype=np.float64).reshape(-1, 3)
    p = np.asarray(p).reshape(-1, 3)
    Q = np.asarray(Q).reshape(-1, 5)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    if p.shape[0] != x_p.shape[0] or Q.shape[0] != x_p.shape[0]:
        raise ValueError("p and Q must have one entry per particle")
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
    tabs, rv = _quad_field_tables(xi)
    ctx = _ctxmod.Context()
    try:
        ctx.set_tables(tabs[0], tabs[1], rv)
        try:
            ctx.set_geometry(np.vstack([x_p, x_E]), box, ngrid, P, xi, eta, rc)
        except ValueError as e:
            raise ValueError(f"{e} (a probe at a particle centre?)") from None
        ctx.set_quad_tables(tabs[2:])
        ctx.set_permittivity(np.full(n_p + n_e, 2.0))  # any eps != 1: the probes' moments are zero
        m = np.zeros((2, n_p + n_e, 8))
        m[0, :n_p, :3], m[0, :n_p, 3:] = np.real(p), np.real(Q)
        m[1, :n_p, :3], m[1, :n_p, 3:] = np.imag(p), np.imag(Q)
        out = ctx.quad_matvec(m).cpu().numpy()
    finally:
        ctx.close()
    return -(out[0, n_p:, :3] + 1j * out[1, n_p:, :3])


# ---------------------------------------------------------------------------
# spheres of different radii (Compute_Field_Sizes/FieldMap.m)
# ---------------------------------------------------------------------------
def FieldMapSizes(x_E, x_p, a, p, box, xi):
    """E (N_E, 3) complex128: the field at x_E (N_E, 3) of the dipoles p
    (N_p, 3) of spheres of radii a (N_p,) at x_p
    (Compute_Field_Sizes/FieldMap.m:1-43): tolerance 1e-3 and the method's
    isotropic parameters (ComputeField.m:22-29), pair tables without overlap
    terms on r = [0.5 : 1e-3 : r_c + 1e-2] for the unique radii, every probe a
    sphere of the smallest radius (its contraction kernel and its table
    column), E = -(E_q + E_r).  On the GPU: the polydisperse operator on
    particles and probes together, the probes carrying zero dipoles.  A probe
    closer than 0.5 to a particle has no table interval (the reference fails
    there) and raises ValueError."""
    from sizes import uniquetol
    x_E = np.asarray(x_E, dtype=np.float64).reshape(-1, 3)
    x_p = np.asarray(x_p, dtype=np.float64).reshape(-1, 3)
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    p = np.asarray(p).reshape(-1, 3)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    if p.shape[0] != x_p.shape[0] or a.size != x_p.shape[0]:
        raise ValueError("a and p must have one entry per particle")
    tol = 1e-3
    rc = np.sqrt(-np.log(tol)) / xi
    if np.any(rc > box / 2):
        raise ValueError("Real space cutoff (r_c = %.3f) larger than half the box size (%.3f, %.3f, %.3f)."
                         % (rc, box[0], box[1], box[2]))
    n_p, n_e = x_p.shape[0], x_E.shape[0]
    if n_e == 0:
        return np.zeros((0, 3), dtype=np.complex128)
    if n_p:
        from scipy.spatial import cKDTree
        dmin, _ = cKDTree(np.mod(x_p, box), boxsize=box).query(np.mod(x_E, box), k=1)
        if np.min(dmin) < 0.5:
            raise ValueError("a probe lies closer than 0.5 to a particle: the field-map table starts at 0.5")
    qc = 2 * xi * np.sqrt(-np.log(tol))
    ng = np.ceil(1 + box * qc / np.pi).astype(np.int64)
    h = box / ng
    P = int(np.ceil(2 / np.mean(h) - 2 * np.log(tol) / np.pi))
    eta = P * (np.mean(h) * xi) ** 2 / np.pi
    a_uniq, a_ind = uniquetol(a)
    a_min_ind = int(a_ind[int(np.argmin(a))])
    r_table = np.arange(0.5, rc + 10 * tol + 1e-12, tol)
    perp, para = _ctxmod.real_space_table_sizes_field(r_table, a_uniq, xi)
    # the library's tables start with an r -> 0 row: a zero row at r = 0,
    # read only for pairs closer than 0.5 (excluded above)
    rv = np.concatenate(([0.0], r_table))
    perp = np.vstack([np.zeros((1, perp.shape[1])), perp])
    para = np.vstack([np.zeros((1, para.shape[1])), para])
    ctx = _ctxmod.Context()
    try:
        ctx.set_tables(perp[:, 0], para[:, 0], rv)
        ctx.set_geometry(np.vstack([x_p, x_E]), box, ng, P, xi, np.full(3, eta), rc)
        radius = np.concatenate([a, np.full(n_e, float(np.min(a)))])
        types = np.concatenate([a_ind, np.full(n_e, a_min_ind)]).astype(np.int32)
        ctx.set_sizes(radius, types, perp, para)
        ctx.set_permittivity(np.full(n_p + n_e, 2.0))  # any eps != 1: the probes' dipoles are zero
        m = np.zeros((2, n_p + n_e, 3))
        m[0, :n_p], m[1, :n_p] = np.real(p), np.imag(p)
        out = ctx.sizes_matvec(m).cpu().numpy()
    finally:
        ctx.close()
    return -(out[0, n_p:] + 1j * out[1, n_p:])
