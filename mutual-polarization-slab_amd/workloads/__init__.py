"""Seeded synthetic workloads (positions, permittivities) for the benchmark
and the tests.  Not part of the solver.

configuration(name) returns the inputs of one named workload:
  "n64_phi0.2"    N=64 equal spheres, RSA at phi=0.2 (CPU-runnable case)
  "n1e4_phi0.3"   N=1e4, phi=0.3, FCC + hard-sphere MC (RSA at 0.3 is slow in Python)
  "n1e5_phi0.4"   N=1e5, phi=0.4, FCC + hard-sphere MC (RSA saturates near 0.38)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libhs_packing.so")
_SRC = os.path.join(_HERE, "hs_packing.c")
_lib = None


def build() -> str:
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", _SO, _SRC, "-lm"], check=True)
    return _SO


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        lib = ctypes.CDLL(_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.fcc_mc.restype = ctypes.c_double
        lib.fcc_mc.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_int, dp, dp]
        lib.min_distance.restype = ctypes.c_double
        lib.min_distance.argtypes = [ctypes.c_int, dp, ctypes.c_double]
        lib.poly_pack.restype = ctypes.c_double
        lib.poly_pack.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_uint64, ctypes.c_int, dp, dp]
        lib.poly_min_ratio.restype = ctypes.c_double
        lib.poly_min_ratio.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double]
        _lib = lib
    return _lib


def hard_spheres(n: int, phi: float, seed: int, sweeps: int = 200):
    """(x (n,3), box (3,)) of n non-overlapping unit spheres at volume fraction phi."""
    lib = _load()
    x = np.zeros((n, 3))
    box = np.zeros(3)
    dp = ctypes.POINTER(ctypes.c_double)
    acc = lib.fcc_mc(int(n), float(phi), int(seed), int(sweeps), x.ctypes.data_as(dp), box.ctypes.data_as(dp))
    if acc < 0:
        raise ValueError("cannot build a hard-sphere configuration at this volume fraction")
    return x, box


def min_distance(x, box) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return _load().min_distance(x.shape[0], x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(box[0]))


def polydisperse(n: int, phi: float, seed: int, sigma: float = 0.2, max_iter: int = 20000):
    """(x (n,3), a (n,), box (3,)): n non-overlapping spheres with log-normal
    radii (mean 1, log-standard deviation sigma) at volume fraction phi in a
    cubic periodic box (uniform random centres, then pairwise overlap
    removal)."""
    rng = np.random.default_rng(seed)
    a = np.exp(rng.normal(-0.5 * sigma * sigma, sigma, n))
    x = np.zeros((n, 3))
    box = np.zeros(3)
    dp = ctypes.POINTER(ctypes.c_double)
    it = _load().poly_pack(int(n), a.ctypes.data_as(dp), float(phi), int(seed), int(max_iter),
                           x.ctypes.data_as(dp), box.ctypes.data_as(dp))
    if it < 0:
        raise ValueError("cannot build a polydisperse hard-sphere configuration at this volume fraction")
    return x, a, box


def min_gap_ratio(x, a, box) -> float:
    """Smallest d_ij / (a_i + a_j) over all pairs (>= 1: no overlaps)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    return _load().poly_min_ratio(x.shape[0], x.ctypes.data_as(dp), a.ctypes.data_as(dp), float(box[0]))


def heterogeneous_set(count: int = 256, seed: int = 7, n_range=(500, 8000), phi_range=(0.1, 0.45),
                      sigma: float = 0.2, n_eps: int = 16, eps_range=(0.1, 100.0)):
    """Independent polydisperse configurations: N log-uniform in n_range,
    log-normal radii (mean 1, sigma), phi uniform in phi_range, and an
    n_eps-point permittivity-ratio spectrum log-spaced over eps_range.
    Returns (configs [{"x", "a", "box", "phi", "seed"}], eps (n_eps,))."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count)).astype(np.int64)
    phis = rng.uniform(phi_range[0], phi_range[1], count)

    def one(k):
        s = int(seed) * 100003 + k
        x, a, box = polydisperse(int(ns[k]), float(phis[k]), s, sigma)
        return {"x": x, "a": a, "box": box, "phi": float(phis[k]), "seed": s}

    _load()
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:  # the packer releases the GIL
        configs = list(ex.map(one, range(count)))
    eps = np.logspace(np.log10(eps_range[0]), np.log10(eps_range[1]), n_eps)
    return configs, eps


def rsa(n: int, phi: float, seed: int):
    """Seeded random sequential addition (cubic periodic box); small n only."""
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


def ewald_for_grid(box, ngrid, P, tol):
    """xi giving `ngrid` nodes by the reference rule Ngrid = ceil(1 + L kcut/pi)
    at tolerance tol (largest such xi), and rc = sqrt(-ln tol)/xi."""
    L = float(np.min(box))
    s = np.sqrt(-np.log(tol))
    xi = (ngrid - 1) * np.pi / (2.0 * L * s)
    rc = s / xi
    h = np.asarray(box, dtype=np.float64) / ngrid
    eta = P * (h * xi) ** 2 / np.pi
    return xi, rc, eta


WORKLOADS = {
    # name: (n, phi, beta, ngrid, P, tol, generator)
    "n64_phi0.2": (64, 0.2, 0.5, 32, 8, 1e-8, "rsa"),
    "n1e4_phi0.3": (10_000, 0.3, 0.8, 128, 10, 1e-6, "fcc_mc"),
    "n1e5_phi0.4": (100_000, 0.4, 0.9, 256, 10, 1e-6, "fcc_mc"),
}


def configuration(name: str, seed: int = 2024, sweeps: int = 200):
    n, phi, beta, ng, P, tol, gen = WORKLOADS[name]
    if gen == "rsa":
        x, box = rsa(n, phi, seed)
        xi, rc = 0.5, 8.0
        eta = P * (box / ng * xi) ** 2 / np.pi
    else:
        x, box = hard_spheres(n, phi, seed, sweeps)
        xi, rc, eta = ewald_for_grid(box, ng, P, tol)
    eps = (1.0 + 2.0 * beta) / (1.0 - beta)
    return {"name": name, "x": x, "box": box, "eps": np.full(n, eps), "beta": beta, "phi": phi, "xi": xi,
            "rc": rc, "ngrid": np.array([ng, ng, ng]), "P": P, "eta": eta, "tol": tol, "seed": seed,
            "sweeps": sweeps}
