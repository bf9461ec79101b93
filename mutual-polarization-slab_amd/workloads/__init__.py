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
