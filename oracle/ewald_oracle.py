"""FP64 CPU oracle for the spectral-Ewald dipole operator.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the test reference (or the timed CPU baseline),
never as the thing measured or shipped.

This is a clean NumPy restatement of the reference operator
(``CPU - Ewald test/ewald.py``) -- same arithmetic, same data flow as the
reference fast path (precomputed N x P^3 node indices and weights, bincount
spread, complex128 ``fftn``/``ifftn`` in fftshift order, gather contract,
``np.add.at`` real-space scatter).  Each function cites the reference lines it
follows.

How each part is pinned (beyond restatement):
  * real_space_table        -- golden vectors written by the reference itself
                               (tests/golden), far-field limit, continuity at
                               r = 2, r -> 0 limit equals the self term.
  * cell/neighbor lists     -- brute-force all-pairs on small N.
  * grid kernel / spread /
    contract                -- Gaussian normalisation, adjointness, golden
                               operator vectors.
  * k-space coefficient     -- J_{3/2} against scipy spherical Bessel j1.
  * magnetic_field          -- golden vectors from the reference, xi-invariance
                               of Hk + Hr, symmetry of the assembled matrix.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfc as _erfc
from scipy.special import spherical_jn as _spherical_jn  # noqa: F401  (used by tests)

PI = np.pi
PI32 = np.pi ** 1.5


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def round_half_away(v):
    """MATLAB round: nearest integer, ties away from zero ([E]ewald.py:71-78)."""
    v = np.asarray(v, dtype=np.float64)
    return np.where(v >= 0.0, np.floor(v + 0.5), -np.floor(-v + 0.5))


def j32(k):
    """Bessel J_{3/2}(k) with J(0)=0 ([E]ewald.py:51-65)."""
    k = np.asarray(k, dtype=np.float64)
    out = np.zeros_like(k)
    nz = k != 0
    kk = k[nz]
    out[nz] = np.sqrt(2.0 / (PI * kk)) * (np.sin(kk) / kk - np.cos(kk))
    return out


# ---------------------------------------------------------------------------
# cell list / neighbor list  ([E]ewald.py:84-199, Het/CellList.m, Het/NeighborList.m)
# ---------------------------------------------------------------------------
def cell_list(x, box, rmax):
    """Wrap x into the box and bin into max(3, floor(L/rmax)) cells per axis
    ([E]ewald.py:84-111).  Returns 0-based (N,3) cell coords and Ncell."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    xw = np.mod(np.asarray(x, dtype=np.float64), box)
    ncell = np.maximum(np.floor(box / rmax).astype(np.int64), 3)
    cell = np.floor((xw / box) * ncell).astype(np.int64)
    return cell, ncell


def neighbor_list(cell, ncell):
    """All directed pairs (i, j), i != j, whose cells are periodic neighbours
    (27-cell stencil) ([E]ewald.py:114-199).  Order: grouped by the cell of i;
    the order only affects floating-point summation order."""
    cell = np.asarray(cell, dtype=np.int64)
    ncell = np.asarray(ncell, dtype=np.int64).reshape(3)
    n = cell.shape[0]
    if n == 0:
        e = np.zeros(0, dtype=np.int64)
        return e, e
    cid = cell[:, 0] + ncell[0] * (cell[:, 1] + ncell[1] * cell[:, 2])
    order = np.lexsort((np.arange(n), cid))
    ncells = int(np.prod(ncell))
    counts = np.bincount(cid, minlength=ncells)
    start = np.cumsum(counts) - counts
    sorted_ids = order
    p1s, p2s = [], []
    for d in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T:
        nb = np.mod(cell + d[None, :], ncell[None, :])
        nbid = nb[:, 0] + ncell[0] * (nb[:, 1] + ncell[1] * nb[:, 2])
        cnt = counts[nbid]
        i_rep = np.repeat(np.arange(n), cnt)
        offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        j_rep = sorted_ids[np.repeat(start[nbid], cnt) + offs]
        p1s.append(i_rep)
        p2s.append(j_rep)
    p1 = np.concatenate(p1s)
    p2 = np.concatenate(p2s)
    keep = p1 != p2
    return p1[keep], p2[keep]


# ---------------------------------------------------------------------------
# P^3 stencil  ([E]ewald.py:205-223, Het/PreCalculations.m:18-23)
# ---------------------------------------------------------------------------
def pre_calculations(P, h):
    """Stencil offsets, x fastest, centred at floor((P-1)/2)."""
    P = int(P)
    lin = np.arange(P ** 3)
    c = (P - 1) // 2
    off = np.stack((lin % P - c, (lin // P) % P - c, lin // (P * P) - c), axis=1).astype(np.int64)
    return off, off.astype(np.float64) * np.asarray(h, dtype=np.float64).reshape(1, 3)


# ---------------------------------------------------------------------------
# grid kernel, spread, contract  ([E]ewald.py:233-368)
# ---------------------------------------------------------------------------
def prepare_grid_kernel(x, Ngrid, h, xi, eta, P, offset, offsetxyz):
    """Node indices (F-order, shifted by -1 node as in the reference) and the
    Gaussian spread / contract weights ([E]ewald.py:233-296)."""
    x = np.asarray(x, dtype=np.float64)
    ng = np.asarray(Ngrid, dtype=np.int64).reshape(3)
    h = np.asarray(h, dtype=np.float64).reshape(3)
    eta = np.asarray(eta, dtype=np.float64).reshape(3)
    offset = np.asarray(offset, dtype=np.int64)
    offsetxyz = np.asarray(offsetxyz, dtype=np.float64)
    if offset.shape[0] != P ** 3 or offsetxyz.shape[0] != P ** 3:
        raise ValueError("offset/offsetxyz shapes do not match P^3")
    node0 = round_half_away(x / h[None, :]).astype(np.int64)
    node = np.mod(node0[:, None, :] + offset[None, :, :] - 1, ng[None, None, :])
    lin = node[..., 0] + ng[0] * (node[..., 1] + ng[1] * node[..., 2])
    norm = (2.0 * xi * xi / PI) ** 1.5 / np.sqrt(np.prod(eta))
    rel = (node0 * h[None, :] - x)[:, None, :] + offsetxyz[None, :, :]
    w = norm * np.exp(-2.0 * xi * xi * np.sum(rel * rel / eta[None, None, :], axis=2))
    return lin, w, w * float(np.prod(h))


def spread_fast(m, Ngrid, linnode, w_spread):
    """Grid (Nx,Ny,Nz,3) complex = sum of w * m_c at linnode ([E]ewald.py:299-331)."""
    m = np.asarray(m, dtype=np.complex128)
    ng = [int(v) for v in np.asarray(Ngrid).reshape(3)]
    G = ng[0] * ng[1] * ng[2]
    flat = np.asarray(linnode, dtype=np.int64).reshape(-1)
    out = np.zeros(tuple(ng) + (3,), dtype=np.complex128)
    for c in range(3):
        wc = (w_spread * m[:, c][:, None]).reshape(-1)
        re = np.bincount(flat, weights=wc.real, minlength=G)
        im = np.bincount(flat, weights=wc.imag, minlength=G)
        out[..., c] = (re + 1j * im).reshape(ng, order="F")
    return out


def contract_fast(Ht, Ngrid, linnode, w_contract):
    """Hk_i,c = sum w_contract * Ht_c[linnode] ([E]ewald.py:334-368)."""
    Ht = np.asarray(Ht, dtype=np.complex128)
    out = np.zeros((linnode.shape[0], 3), dtype=np.complex128)
    for c in range(3):
        flat = Ht[..., c].ravel(order="F")
        out[:, c] = np.sum(w_contract * flat[linnode], axis=1)
    return out


# ---------------------------------------------------------------------------
# k space  ([E]ewald.py:523-599)
# ---------------------------------------------------------------------------
def kvectors_shifted(Ngrid, box):
    """Wave vectors in fftshift order ([E]ewald.py:548-552)."""
    ng = np.asarray(Ngrid, dtype=np.int64).reshape(3)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    ks = []
    for d in range(3):
        n = int(ng[d])
        lo = -((n - 1 + 1) // 2)  # -ceil((n-1)/2)
        hi = (n - 1) // 2        # floor((n-1)/2)
        ks.append(np.arange(lo, hi + 1) * (2.0 * PI / box[d]))
    return np.meshgrid(ks[0], ks[1], ks[2], indexing="ij")


def precompute_kspace(Ngrid, box, xi, eta):
    """khat and the scalar coefficient c(k) = 9 pi/(2|k|) J_{3/2}(|k|)^2
    exp(-sum (1-eta_i) k_i^2/(4 xi^2)) / |k|^2, c(0)=0 ([E]ewald.py:523-582)."""
    eta = np.asarray(eta, dtype=np.float64).reshape(3)
    KX, KY, KZ = kvectors_shifted(Ngrid, box)
    k2 = KX * KX + KY * KY + KZ * KZ
    kn = np.sqrt(k2)
    safe = np.where(kn == 0.0, 1.0, kn)
    khat = np.stack((KX / safe, KY / safe, KZ / safe), axis=3)
    zero = kn == 0.0
    khat[zero] = 0.0
    ek = (1.0 - eta[0]) * KX ** 2 + (1.0 - eta[1]) * KY ** 2 + (1.0 - eta[2]) * KZ ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        coef = 9.0 * PI / (2.0 * kn) * j32(kn) ** 2 * np.exp(-ek / (4.0 * xi * xi)) / k2
    coef[zero] = 0.0
    coef = np.nan_to_num(coef)
    return khat.astype(np.complex128), coef


def scale_fast(fH, khat, coef):
    """F~ = c(k) (khat . F) khat ([E]ewald.py:585-599)."""
    s = np.sum(khat * fH, axis=3) * coef
    return s[..., None] * khat


# ---------------------------------------------------------------------------
# real-space tables ([E]ewald.py:604-931, Het/RealSpaceTable.m)
# ---------------------------------------------------------------------------
# Every coupling has the form
#   sum_{s in (+2,-2,0)} [ E_s(r) exp(-(r+s)^2 xi^2) + F_s(r) erfc((r+s) xi) ]
#   + [r < 2] R(r)
# with E_s, F_s = scale * poly(r) / r^k.  Each entry below gives
# (scale(xi), k, ascending polynomial coefficients(xi)).
def _coupling_defs(x):
    x2, x4, x6 = x * x, x ** 4, x ** 6
    a = 3.0 + 4.0 * x2 - 32.0 * x4
    b = 3.0 - 12.0 * x2 + 32.0 * x4
    return {
        "pot_charge": {
            "Ep": (1.0 / (32 * PI32 * x), 1, [-2.0, -1.0]),
            "Em": (1.0 / (32 * PI32 * x), 1, [2.0, -1.0]),
            "E0": (1.0 / (16 * PI32 * x), 0, [1.0]),
            "Fp": (1.0 / (64 * PI * x2), 1, [8 * x2 + 1, 8 * x2, 2 * x2]),
            "Fm": (1.0 / (64 * PI * x2), 1, [8 * x2 + 1, -8 * x2, 2 * x2]),
            "F0": (1.0 / (32 * PI * x2), 1, [-1.0, 0.0, -2 * x2]),
            "R": lambda r: -1.0 / (4 * PI * r) + (4.0 - r) / (16 * PI),
        },
        "pot_dip": {
            "Ep": (1.0 / (256 * PI32 * x ** 3), 2, [2 * (1 - 8 * x2), -3 + 8 * x2, -4 * x2, -6 * x2]),
            "Em": (1.0 / (256 * PI32 * x ** 3), 2, [-2 * (1 - 8 * x2), -3 + 8 * x2, 4 * x2, -6 * x2]),
            "E0": (3.0 / (128 * PI32 * x ** 3), 1, [1.0, 0.0, 2 * x2]),
            "Fp": (1.0 / (512 * PI * x4), 2, [-3 + 64 * x4, 0.0, 12 * x2, 32 * x4, 12 * x4]),
            "Fm": (1.0 / (512 * PI * x4), 2, [-3 + 64 * x4, 0.0, 12 * x2, -32 * x4, 12 * x4]),
            "F0": (3.0 / (256 * PI * x4), 2, [1.0, 0.0, -4 * x2, 0.0, -4 * x4]),
            "R": lambda r: -1.0 / (4 * PI * r * r) + r / (8 * PI) * (1.0 - 3.0 * r / 8.0),
        },
        "field_dip_1": {
            "Ep": (1.0 / (1024 * PI32 * x ** 5), 3, [2 * a, b, -8 * x2 * (3 + 2 * x2), 8 * x2 * (2 - 7 * x2), -8 * x4, 4 * x4]),
            "Em": (1.0 / (1024 * PI32 * x ** 5), 3, [-2 * a, b, 8 * x2 * (3 + 2 * x2), 8 * x2 * (2 - 7 * x2), 8 * x4, 4 * x4]),
            "E0": (1.0 / (512 * PI32 * x ** 5), 2, [-3 + 36 * x2, 0.0, -8 * x2 * (2 - 9 * x2), 0.0, -4 * x4]),
            "Fp": (1.0 / (2048 * PI * x6), 3, [3 - 36 * x2 + 256 * x6, 0.0, -18 * x2 * (1 - 8 * x2), 256 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -8 * x6]),
            "Fm": (1.0 / (2048 * PI * x6), 3, [3 - 36 * x2 + 256 * x6, 0.0, -18 * x2 * (1 - 8 * x2), -256 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -8 * x6]),
            "F0": (1.0 / (1024 * PI * x6), 3, [-3 + 36 * x2, 0.0, 18 * x2 * (1 - 8 * x2), 0.0, 36 * x4 * (1 - 4 * x2), 0.0, 8 * x6]),
            "R": lambda r: -1.0 / (4 * PI * r ** 3) + (1.0 - 9.0 * r / 16.0 + r ** 3 / 32.0) / (4 * PI),
        },
        "field_dip_2": {
            "Ep": (1.0 / (512 * PI32 * x ** 5), 3, [-2 * a, -b, -4 * x2 * (3 - 4 * x2), 2 * x2 * (7 - 20 * x2), -16 * x4, 8 * x4]),
            "Em": (1.0 / (512 * PI32 * x ** 5), 3, [2 * a, -b, 4 * x2 * (3 - 4 * x2), 2 * x2 * (7 - 20 * x2), 16 * x4, 8 * x4]),
            "E0": (1.0 / (256 * PI32 * x ** 5), 2, [3 - 36 * x2, 0.0, -2 * x2 * (7 - 36 * x2), 0.0, -8 * x4]),
            "Fp": (1.0 / (1024 * PI * x6), 3, [-3 + 36 * x2 - 256 * x6, 0.0, 0.0, 128 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -16 * x6]),
            "Fm": (1.0 / (1024 * PI * x6), 3, [-3 + 36 * x2 - 256 * x6, 0.0, 0.0, -128 * x6, -36 * x4 * (1 - 4 * x2), 0.0, -16 * x6]),
            "F0": (1.0 / (512 * PI * x6), 3, [3 - 36 * x2, 0.0, 0.0, 0.0, 36 * x4 * (1 - 4 * x2), 0.0, 16 * x6]),
            "R": lambda r: 1.0 / (2 * PI * r ** 3) + (1.0 - 9.0 * r / 8.0 + r ** 3 / 8.0) / (4 * PI),
        },
        "field_dip_force_1": {
            "Ep": (3.0 / (1024 * PI32 * x ** 5), 4, [-2 * a, -b, 16 * x4, 4 * x2 * (1 - 2 * x2), -8 * x4, 4 * x4]),
            "Em": (3.0 / (1024 * PI32 * x ** 5), 4, [2 * a, -b, -16 * x4, 4 * x2 * (1 - 2 * x2), 8 * x4, 4 * x4]),
            "E0": (3.0 / (512 * PI32 * x ** 5), 3, [3 - 36 * x2, 0.0, -4 * x2 * (1 - 6 * x2), 0.0, -4 * x4]),
            "Fp": (3.0 / (2048 * PI * x6), 4, [-3 + 36 * x2 - 256 * x6, 0.0, 6 * x2 * (1 - 8 * x2), 0.0, -12 * x4 * (1 - 4 * x2), 0.0, -8 * x6]),
            "Fm": (3.0 / (2048 * PI * x6), 4, [-3 + 36 * x2 - 256 * x6, 0.0, 6 * x2 * (1 - 8 * x2), 0.0, -12 * x4 * (1 - 4 * x2), 0.0, -8 * x6]),
            "F0": (3.0 / (1024 * PI * x6), 4, [3 - 36 * x2, 0.0, -6 * x2 * (1 - 8 * x2), 0.0, 12 * x4 * (1 - 4 * x2), 0.0, 8 * x6]),
            "R": lambda r: 3.0 / (4 * PI * r ** 4) - 3.0 / (64 * PI) * (3.0 - r * r / 2.0),
        },
        "field_dip_force_2": {
            "Ep": (9.0 / (1024 * PI32 * x ** 5), 4, [2 * a, b, 8 * x2 * (1 - 2 * x2), 8 * x4, -8 * x4, 4 * x4]),
            "Em": (9.0 / (1024 * PI32 * x ** 5), 4, [-2 * a, b, -8 * x2 * (1 - 2 * x2), 8 * x4, 8 * x4, 4 * x4]),
            "E0": (9.0 / (512 * PI32 * x ** 5), 3, [-3 + 36 * x2, 0.0, 8 * x4, 0.0, -4 * x4]),
            "Fp": (9.0 / (2048 * PI * x6), 4, [3 - 36 * x2 + 256 * x6, 0.0, -2 * x2 * (1 - 8 * x2), 0.0, -4 * x4 * (1 - 4 * x2), 0.0, -8 * x6]),
            "Fm": (9.0 / (2048 * PI * x6), 4, [3 - 36 * x2 + 256 * x6, 0.0, -2 * x2 * (1 - 8 * x2), 0.0, -4 * x4 * (1 - 4 * x2), 0.0, -8 * x6]),
            "F0": (9.0 / (1024 * PI * x6), 4, [-3 + 36 * x2, 0.0, 2 * x2 * (1 - 8 * x2), 0.0, 4 * x4 * (1 - 4 * x2), 0.0, 8 * x6]),
            "R": lambda r: -9.0 / (4 * PI * r ** 4) - 9.0 / (64 * PI) * (1.0 - r * r / 2.0),
        },
    }


TABLE_NAMES = ("pot_charge", "pot_dip", "field_dip_1", "field_dip_2",
               "field_dip_force_1", "field_dip_force_2")


def _rational(defn, r):
    scale, k, coeffs = defn
    return scale * np.polyval(coeffs[::-1], r) / r ** k


def real_space_table(r, xi, regularize=True):
    """The six real-space couplings of unit spheres at separations r (r > 0),
    each with its r = 0 self term prepended ([E]ewald.py:604-931).
    regularize=False leaves out the (r < 2) overlap terms, as the field-map
    tables do (Compute_Field/RealSpaceTable.m, regpoly lines commented out)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    x = float(xi)
    defs = _coupling_defs(x)
    ep = np.exp(-(r + 2.0) ** 2 * x * x)
    em = np.exp(-(r - 2.0) ** 2 * x * x)
    e0 = np.exp(-r * r * x * x)
    fp = _erfc((r + 2.0) * x)
    fm = _erfc((r - 2.0) * x)
    f0 = _erfc(r * x)
    overlap = (r < 2.0) & bool(regularize)
    out = {}
    for name in TABLE_NAMES:
        d = defs[name]
        val = (_rational(d["Ep"], r) * ep + _rational(d["Em"], r) * em + _rational(d["E0"], r) * e0
               + _rational(d["Fp"], r) * fp + _rational(d["Fm"], r) * fm + _rational(d["F0"], r) * f0)
        val = val + np.where(overlap, d["R"](np.where(overlap, r, 1.0)), 0.0)
        out[name] = val
    e4 = np.exp(-4.0 * x * x)
    tail = _erfc(2.0 * x) / (4.0 * PI)
    self_pc = (1.0 - e4) / (8.0 * PI32 * x) + tail
    self_fd = (-1.0 + 6.0 * x * x + (1.0 - 2.0 * x * x) * e4) / (16.0 * PI32 * x ** 3) + tail
    selfs = {"pot_charge": self_pc, "pot_dip": 0.0, "field_dip_1": self_fd, "field_dip_2": self_fd,
             "field_dip_force_1": 0.0, "field_dip_force_2": 0.0}
    return tuple(np.concatenate(([selfs[n]], out[n])) for n in TABLE_NAMES)


def default_rtable():
    """r grid used by run_capacitance: 0, then 0.001:0.001:10 ([E]capacitance.py:474-476)."""
    r = np.arange(0.001, 10.0 + 1e-12, 0.001, dtype=np.float64)
    return r


# ---------------------------------------------------------------------------
# real space  ([E]ewald.py:939-1072)
# ---------------------------------------------------------------------------
def prepare_real_space_pairs(x, box, p1, p2, rc, perp, para, rvals):
    """Filter candidate pairs by min-image d < rc and interpolate the tables
    linearly (interval = first rvals >= d) ([E]ewald.py:939-1016)."""
    x = np.asarray(x, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    p1 = np.asarray(p1, dtype=np.int64).reshape(-1)
    p2 = np.asarray(p2, dtype=np.int64).reshape(-1)
    rvals = np.asarray(rvals, dtype=np.float64)
    if p1.size == 0:
        e = np.zeros(0)
        return p1, p2, e.reshape(0, 3), e, e
    r = x[p1] - x[p2]
    r = r - box[None, :] * np.trunc(2.0 * r / box[None, :])
    d = np.sqrt(np.sum(r * r, axis=1))
    keep = d < float(rc)
    p1, p2, r, d = p1[keep], p2[keep], r[keep], d[keep]
    if d.size == 0:
        e = np.zeros(0)
        return p1, p2, e.reshape(0, 3), e, e
    rhat = r / d[:, None]
    idx = np.searchsorted(rvals, d, side="left")
    if np.any(idx <= 0) or np.any(idx >= rvals.size):
        raise ValueError("RealSpace interpolation: pair separations fall outside rvals table.")
    lo, hi = rvals[idx - 1], rvals[idx]
    t = (d - lo) / (hi - lo)
    A = perp[idx - 1] * (1.0 - t) + perp[idx] * t
    B = para[idx - 1] * (1.0 - t) + para[idx] * t
    return p1, p2, rhat, A, B


def real_space_fast(m, lambdap, p1, p2, rhat, A, B, self_perp, complex_eps=False):
    """Hr = (-3/(4 pi (1-eps)) + perp[0]) m_i + sum_j [A m_j + (B-A)(m_j.r)r]
    ([E]ewald.py:1019-1072).  The Python reference casts eps to float64
    ([E]ewald.py:1185); complex_eps keeps it complex as the MATLAB version
    does (Het/RealSpace.m:22)."""
    m = np.asarray(m, dtype=np.complex128)
    lam = np.asarray(lambdap, dtype=np.complex128 if complex_eps else np.float64).reshape(-1)
    out = ((-3.0 / (4.0 * PI * (1.0 - lam)))[:, None]) * m + m * float(self_perp)
    if np.asarray(p1).size == 0:
        return out
    mj = m[p2]
    proj = np.sum(mj * rhat, axis=1)
    np.add.at(out, p1, A[:, None] * mj + ((B - A) * proj)[:, None] * rhat)
    return out


# ---------------------------------------------------------------------------
# the operator  ([E]ewald.py:1146-1327)
# ---------------------------------------------------------------------------
class OperatorGeometry:
    """Everything the reference precomputes once per solve
    ([E]capacitance.py:370-385)."""

    def __init__(self, x, box, Ngrid, P, xi, rc, perp, para, rvals):
        self.x = np.asarray(x, dtype=np.float64)
        self.box = np.asarray(box, dtype=np.float64).reshape(3)
        self.Ngrid = np.asarray(Ngrid, dtype=np.int64).reshape(3)
        self.h = self.box / self.Ngrid
        self.P = int(P)
        self.xi = float(xi)
        self.eta = self.P * (self.h * self.xi) ** 2 / PI
        self.rc = float(rc)
        self.perp, self.para, self.rvals = perp, para, rvals
        cell, ncell = cell_list(self.x, self.box, self.rc)
        self.p1, self.p2 = neighbor_list(cell, ncell)
        self.offset, self.offsetxyz = pre_calculations(self.P, self.h)
        self.real = prepare_real_space_pairs(self.x, self.box, self.p1, self.p2, self.rc, perp, para, rvals)
        self.grid = prepare_grid_kernel(self.x, self.Ngrid, self.h, self.xi, self.eta, self.P,
                                        self.offset, self.offsetxyz)
        self.kspace = precompute_kspace(self.Ngrid, self.box, self.xi, self.eta)


def field_components(geo: OperatorGeometry, m, lambdap, slab_factor=1.0, complex_eps=False):
    """(Hk, Hr, H_slab) as in magnetic_field_components ([E]ewald.py:1239-1327)."""
    m = np.asarray(m, dtype=np.complex128)
    lin, ws, wc = geo.grid
    H = spread_fast(m, geo.Ngrid, lin, ws)
    fH = np.fft.fftshift(np.fft.fftn(H, axes=(0, 1, 2)), axes=(0, 1, 2))
    khat, coef = geo.kspace
    Ht = np.fft.ifftn(np.fft.ifftshift(scale_fast(fH, khat, coef), axes=(0, 1, 2)), axes=(0, 1, 2))
    Hk = contract_fast(Ht, geo.Ngrid, lin, wc)
    p1f, p2f, rhat, A, B = geo.real
    Hr = real_space_fast(m, lambdap, p1f, p2f, rhat, A, B, float(geo.perp[0]), complex_eps)
    slab = np.zeros((m.shape[0], 3), dtype=np.complex128)
    slab[:, 2] = -float(slab_factor) * np.sum(m[:, 2]) / np.prod(geo.box)
    return Hk, Hr, slab


def field(geo: OperatorGeometry, m, lambdap, slab_factor=1.0, complex_eps=False):
    """A m = Hk + Hr - H_slab ([E]ewald.py:1236)."""
    Hk, Hr, slab = field_components(geo, m, lambdap, slab_factor, complex_eps)
    return Hk + Hr - slab


# ---------------------------------------------------------------------------
# field at probe points (Compute_Field/FieldMap.m, MagneticField.m, RealSpace.m)
# ---------------------------------------------------------------------------
def field_map(x_E, x_p, p, box, xi, regularized_tables=False):
    """E (N_E,3) complex at the probe points x_E from the dipoles p at x_p:
    errortol 1e-3 and the parameter rule (FieldMap.m:17-36), tables without
    overlap terms (FieldMap.m:23-25), the particles' grid contracted at the
    probes (MagneticField.m:33-60) plus the real-space sum over probe-particle
    pairs with 0 < d < rc (RealSpace.m:23-57, brute force over all pairs here
    in place of NeighborList_Types), and E = -(Hk + Hr) (FieldMap.m:55).
    regularized_tables=True swaps in the sphere-sphere tables (a check that
    the choice matters, not the reference's field map)."""
    x_E = np.asarray(x_E, dtype=np.float64).reshape(-1, 3)
    x_p = np.asarray(x_p, dtype=np.float64).reshape(-1, 3)
    p = np.asarray(p, dtype=np.complex128).reshape(-1, 3)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    errortol = 1e-3
    r = default_rtable()[1:]
    tabs = real_space_table(r, xi, regularize=bool(regularized_tables))
    perp, para, rvals = tabs[2], tabs[3], np.concatenate(([0.0], r))
    rc = np.sqrt(-np.log(errortol)) / xi
    kcut = 2 * xi * np.sqrt(-np.log(errortol))
    Ngrid = np.ceil(1 + box * kcut / PI).astype(np.int64)
    h = box / Ngrid
    P = int(np.ceil(-2 * np.log(errortol) / PI))
    eta = P * (h * xi) ** 2 / PI
    if np.any(rc > box / 2):
        raise ValueError("Real space cutoff larger than half the box size")
    offset, offsetxyz = pre_calculations(P, h)
    lin_p, w_p, _ = prepare_grid_kernel(x_p, Ngrid, h, xi, eta, P, offset, offsetxyz)
    H = spread_fast(p, Ngrid, lin_p, w_p)
    fH = np.fft.fftshift(np.fft.fftn(H, axes=(0, 1, 2)), axes=(0, 1, 2))
    khat, coef = precompute_kspace(Ngrid, box, xi, eta)
    Ht = np.fft.ifftn(np.fft.ifftshift(scale_fast(fH, khat, coef), axes=(0, 1, 2)), axes=(0, 1, 2))
    lin_e, _, wc_e = prepare_grid_kernel(x_E, Ngrid, h, xi, eta, P, offset, offsetxyz)
    Hk = contract_fast(Ht, Ngrid, lin_e, wc_e)
    Hr = np.zeros((x_E.shape[0], 3), dtype=np.complex128)
    for i in range(x_E.shape[0]):
        rr = x_E[i][None, :] - x_p
        rr = rr - box[None, :] * np.trunc(2.0 * rr / box[None, :])
        d = np.sqrt(np.sum(rr * rr, axis=1))
        keep = (d < rc) & (d > 0)
        if not np.any(keep):
            continue
        d, rr, mj = d[keep], rr[keep], p[keep]
        rhat = rr / d[:, None]
        idx = np.searchsorted(rvals, d, side="left")
        lo, hi = rvals[idx - 1], rvals[idx]
        t = (d - lo) / (hi - lo)
        A = perp[idx - 1] * (1.0 - t) + perp[idx] * t
        B = para[idx - 1] * (1.0 - t) + para[idx] * t
        proj = np.sum(mj * rhat, axis=1)
        Hr[i] = np.sum(A[:, None] * (mj - proj[:, None] * rhat) + (B * proj)[:, None] * rhat, axis=0)
    return -(Hk + Hr)
