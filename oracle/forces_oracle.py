"""FP64 CPU oracle for the truncated point-dipole force kernel.

TEST INFRASTRUCTURE ONLY (see ewald_oracle.py's header).

Restates compute_forces of ``CPU - HOOMD simulation test/dipole_forces_ext/
_dipole_forces_ext.cpp`` as an O(N^2) NumPy loop over i < j:
  U_ij = pref [ (mi.mj)/r^3 - 3 (mi.r)(mj.r)/r^5 ]            (:213)
  F_i  = 3 pref/r^5 [ (mi.mj) r + (mj.r) mi + (mi.r) mj
                      - 5 (mi.r)(mj.r) r / r^2 ],  F_j = -F_i  (:215-222)
with r = r_i - r_j, minimum image by rint in x, y (and z if periodic_z)
(:41-44, :199-203), pairs with r^2 > r_cut^2 or r^2 < 1e-24 skipped (:205-206),
and the energy split 1/2 per particle (:227, :236).  Validation errors follow
:131-155 (raised as RuntimeError).
"""
from __future__ import annotations

import numpy as np


def _validate(pos, dip, box, r_cut, periodic_z):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    dip = np.ascontiguousarray(dip, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)
    if pos.ndim != 2 or pos.shape[1] != 3 or dip.shape != pos.shape or box.shape != (3,):
        raise RuntimeError("Invalid shapes: pos(N,3), dip(N,3), box(3) required.")
    if not np.all(box > 0.0):
        raise RuntimeError("Box lengths must be positive.")
    if not r_cut > 0.0:
        raise RuntimeError("r_cut must be > 0.")
    if r_cut >= min(0.5 * box[0], 0.5 * box[1]):
        raise RuntimeError("r_cut too large for minimum-image periodic x,y. Require r_cut < 0.5*min(Lx,Ly).")
    if periodic_z and r_cut >= 0.5 * box[2]:
        raise RuntimeError("r_cut too large for minimum-image periodic z. Require r_cut < 0.5*Lz when "
                           "periodic_z=True.")
    return pos, dip, box


def compute_forces(pos, dip, box, r_cut, prefactor=1.0, periodic_z=True, use_cell_list=True):
    pos, dip, box = _validate(pos, dip, box, r_cut, periodic_z)
    n = pos.shape[0]
    f = np.zeros((n, 3))
    e = np.zeros(n)
    rc2 = r_cut * r_cut
    for i in range(n - 1):
        d = pos[i][None, :] - pos[i + 1:]
        d[:, 0] -= box[0] * np.rint(d[:, 0] / box[0])
        d[:, 1] -= box[1] * np.rint(d[:, 1] / box[1])
        if periodic_z:
            d[:, 2] -= box[2] * np.rint(d[:, 2] / box[2])
        r2 = np.sum(d * d, axis=1)
        ok = (r2 <= rc2) & (r2 >= 1e-24)
        if not np.any(ok):
            continue
        js = np.nonzero(ok)[0] + i + 1
        d = d[ok]
        r2 = r2[ok]
        inv2 = 1.0 / r2
        inv3 = inv2 * np.sqrt(inv2)
        inv5 = inv3 * inv2
        mi = dip[i][None, :]
        mj = dip[js]
        mimj = np.sum(mi * mj, axis=1)
        mir = np.sum(mi * d, axis=1)
        mjr = np.sum(mj * d, axis=1)
        U = prefactor * (mimj * inv3 - 3.0 * mir * mjr * inv5)
        coef = prefactor * 3.0 * inv5
        F = coef[:, None] * (mimj[:, None] * d + mjr[:, None] * mi + mir[:, None] * mj
                             - 5.0 * (mir * mjr * inv2)[:, None] * d)
        f[i] += F.sum(axis=0)
        np.add.at(f, js, -F)
        e[i] += 0.5 * U.sum()
        np.add.at(e, js, 0.5 * U)
    return f, e
