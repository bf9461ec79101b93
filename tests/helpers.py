"""Shared test helpers: seeded configurations and oracle geometry builders."""
import numpy as np

from oracle import ewald_oracle as eo


def rsa(n, phi, seed, L=None):
    """Seeded random sequential addition of unit spheres (cubic periodic box)."""
    if L is None:
        L = (n * 4.0 / 3.0 * np.pi / phi) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pts = np.zeros((0, 3))
    while pts.shape[0] < n:
        c = rng.random(3) * L
        if pts.shape[0]:
            d = pts - c
            d -= L * np.rint(d / L)
            if np.min(np.sum(d * d, axis=1)) < 4.0:
                continue
        pts = np.vstack((pts, c))
    return pts, np.array([L, L, L])


def tables(xi):
    r = eo.default_rtable()
    t = eo.real_space_table(r, xi)
    return t[2], t[3], np.concatenate(([0.0], r))


def oracle_geometry(x, box, ngrid, P, xi, rc):
    perp, para, rv = tables(xi)
    return eo.OperatorGeometry(x, box, ngrid, P, xi, rc, perp, para, rv)


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
