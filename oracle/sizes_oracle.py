"""FP64 CPU oracle for the polydisperse dipole operator of Capacitance_Sizes/.

TEST INFRASTRUCTURE ONLY (see ewald_oracle.py's header): only tests/ may use
it, as the test reference.

A NumPy restatement of the reference's MATLAB, same data flow: the dipoles
are spread "as charges" onto ONE scalar grid through a radius-dependent
kernel (Siz/Spread.m:39-72), the grid is scaled by exp(-(1-eta) q^2/(4 xi^2))/q^2
(Siz/Scale.m:16-21), and contracted with the vector kernel
(Siz/Contract.m:39-73); the real-space sum interpolates one table column per
pair of distinct radii (Siz/RealSpace.m:19-76, Siz/RealSpaceTable.m:1-150).
Ewald parameters follow Siz/ComputeField.m:22-29 (isotropic eta from the mean
grid spacing, P = ceil(2/mean(h) - 2 ln tol / pi)).

Pinned by (tests/test_sizes_oracle.py): for unit radii the tables equal the
pinned dipole tables of ewald_oracle and the whole operator equals the
dipole operator to the spectral tolerance; for mixed radii the total operator
is independent of xi and symmetric; the tables are
continuous at contact and at containment; the small-d branch of the kernel
is its d -> 0 limit; and a lone sphere of radius a (one per cubic cell) has
the Clausius-Mossotti dipole 4 pi a^3 beta/(1 - beta phi) E0 at several radii
and permittivities, which fixes the dielectric term 3p/(4 pi a^3 (eps-1))
(RealSpace.m:36) and the self terms.  The MATLAB cannot run here, so there are no golden
vectors.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres
from scipy.special import erf as _erf
from scipy.special import erfc as _erfc

from . import ewald_oracle as eo

PI = np.pi
P32 = np.pi ** 1.5


# ---------------------------------------------------------------------------
# radii types (MATLAB uniquetol with its default tolerance)
# ---------------------------------------------------------------------------
def uniquetol(a, tol=1e-6):
    """(unique values, index of each element's value): values within
    tol*max|a| of the first (smallest) member of a group are one value."""
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    order = np.argsort(a, kind="stable")
    scale = tol * np.max(np.abs(a)) if a.size else 0.0
    uniq = []
    ind = np.zeros(a.size, dtype=np.int64)
    for k in order:
        if not uniq or abs(a[k] - uniq[-1]) > scale:
            uniq.append(a[k])
        ind[k] = len(uniq) - 1
    return np.array(uniq), ind


def pair_columns(M):
    """col[i, j]: the table column of radius types (i, j) (RealSpace.m:26-37)."""
    col = np.zeros((M, M), dtype=np.int64)
    lin = 0
    for i in range(M):
        for j in range(i, M):
            col[i, j] = col[j, i] = lin
            lin += 1
    return col


# ---------------------------------------------------------------------------
# real-space couplings (Siz/RealSpaceTable.m)
# ---------------------------------------------------------------------------
def _couplings(r, ai, aj, x, regularize=True):
    """(perp, para) at separations r > 0 for radii ai, aj (:34-135);
    regularize=False leaves out the overlap terms f_reg, as
    Compute_Field_Sizes/RealSpaceTable.m does."""
    r = np.asarray(r, dtype=np.float64)
    s, d = ai + aj, ai - aj
    x2, x4, x6 = x * x, x ** 4, x ** 6
    cp = 1.0 / (1024 * P32 * ai ** 3 * aj ** 3 * r ** 3 * x ** 5)
    c3p = (4 * ai ** 3 - 3 * ai ** 2 * aj - 3 * ai * aj ** 2 + 4 * aj ** 3)
    c3m = (4 * ai ** 3 + 3 * ai ** 2 * aj - 3 * ai * aj ** 2 - 4 * aj ** 3)
    qp = (ai ** 2 - 4 * ai * aj + aj ** 2)
    qm = (ai ** 2 + 4 * ai * aj + aj ** 2)
    f1 = cp * (4 * r ** 5 * x4 - 4 * s * r ** 4 * x4 + r ** 3 * (16 * x2 + 8 * (-4 * ai ** 2 + ai * aj - 4 * aj ** 2) * x4)
               + r ** 2 * (-12 * s * x2 - 8 * c3p * x4)
               + r * (3 - 12 * (ai ** 2 - ai * aj + aj ** 2) * x2 - 4 * s ** 2 * qp * x4)
               + 3 * s + 4 * c3p * x2 + 4 * s ** 3 * qp * x4)
    f2 = cp * (4 * r ** 5 * x4 + 4 * s * r ** 4 * x4 + r ** 3 * (16 * x2 + 8 * (-4 * ai ** 2 + ai * aj - 4 * aj ** 2) * x4)
               + r ** 2 * (12 * s * x2 + 8 * c3p * x4)
               + r * (3 - 12 * (ai ** 2 - ai * aj + aj ** 2) * x2 - 4 * s ** 2 * qp * x4)
               - 3 * s - 4 * c3p * x2 - 4 * s ** 3 * qp * x4)
    f3 = cp * (-4 * r ** 5 * x4 + 4 * (-ai + aj) * r ** 4 * x4
               + r ** 3 * (-16 * x2 + 8 * (4 * ai ** 2 + ai * aj + 4 * aj ** 2) * x4)
               + r ** 2 * (12 * (-ai + aj) * x2 - 8 * c3m * x4)
               + r * (-3 + 12 * (ai ** 2 + ai * aj + aj ** 2) * x2 + 4 * d ** 2 * qm * x4)
               + 3 * d + 4 * c3m * x2 + 4 * d ** 3 * qm * x4)
    f4 = cp * (-4 * r ** 5 * x4 + 4 * d * r ** 4 * x4
               + r ** 3 * (-16 * x2 + 8 * (4 * ai ** 2 + ai * aj + 4 * aj ** 2) * x4)
               + r ** 2 * (12 * d * x2 + 8 * c3m * x4)
               + r * (-3 + 12 * (ai ** 2 + ai * aj + aj ** 2) * x2 + 4 * d ** 2 * qm * x4)
               - 3 * d - 4 * c3m * x2 - 4 * d ** 3 * qm * x4)
    ce = 1.0 / (2048 * PI * ai ** 3 * aj ** 3 * r ** 3 * x6)
    base = (36 * r ** 4 * x4 * (-1 + 2 * (ai ** 2 + aj ** 2) * x2)
            + r ** 2 * (-18 * x2 + 72 * (ai ** 2 + aj ** 2) * x4 + 72 * (ai ** 2 - aj ** 2) ** 2 * x6))
    f5 = ce * (-8 * r ** 6 * x6 + base + 128 * (ai ** 3 + aj ** 3) * r ** 3 * x6
               + 3 - 18 * (ai ** 2 + aj ** 2) * x2 - 36 * (ai ** 2 - aj ** 2) ** 2 * x4 - 8 * s ** 4 * qp * x6)
    f6 = ce * (-8 * r ** 6 * x6 + base - 128 * (ai ** 3 + aj ** 3) * r ** 3 * x6
               + 3 - 18 * (ai ** 2 + aj ** 2) * x2 - 36 * (ai ** 2 - aj ** 2) ** 2 * x4 - 8 * s ** 4 * qp * x6)
    f7 = ce * (8 * r ** 6 * x6 - base + 128 * (ai ** 3 - aj ** 3) * r ** 3 * x6
               - 3 + 18 * (ai ** 2 + aj ** 2) * x2 + 36 * (ai ** 2 - aj ** 2) ** 2 * x4 + 8 * d ** 4 * qm * x6)
    f8 = ce * (8 * r ** 6 * x6 - base - 128 * (ai ** 3 - aj ** 3) * r ** 3 * x6
               - 3 + 18 * (ai ** 2 + aj ** 2) * x2 + 36 * (ai ** 2 - aj ** 2) ** 2 * x4 + 8 * d ** 4 * qm * x6)
    reg = np.zeros_like(r)
    m1 = (r < s) & (r >= ai - aj) & (r >= aj - ai)
    rr = r[m1]
    reg[m1] = (s - rr) ** 3 / (128 * PI * ai ** 3 * aj ** 3 * rr ** 3) * (
        -rr ** 3 - 3 * s * rr ** 2 + 3 * qp * rr + ai ** 3 - 3 * ai ** 2 * aj - 3 * ai * aj ** 2 + aj ** 3)
    m2 = (r < aj - ai) & (r > ai - aj)
    reg[m2] = (r[m2] ** 3 - aj ** 3) / (4 * PI * aj ** 3 * r[m2] ** 3)
    m3 = (r < ai - aj) & (r > aj - ai)
    reg[m3] = (r[m3] ** 3 - ai ** 3) / (4 * PI * ai ** 3 * r[m3] ** 3)
    perp = (f1 * np.exp(-(r + s) ** 2 * x2) + f2 * np.exp(-(r - s) ** 2 * x2) + f3 * np.exp(-(r - ai + aj) ** 2 * x2)
            + f4 * np.exp(-(r + ai - aj) ** 2 * x2) + f5 * _erfc((r + s) * x) + f6 * _erfc((r - s) * x)
            + f7 * _erfc((r - ai + aj) * x) + f8 * _erfc((r + ai - aj) * x) + reg * regularize)
    # rr component
    cq = 1.0 / (512 * P32 * ai ** 3 * aj ** 3 * r ** 3 * x ** 5)
    b3p = (ai ** 3 - 3 * ai ** 2 * aj - 3 * ai * aj ** 2 + aj ** 3)
    b3m = (ai ** 3 + 3 * ai ** 2 * aj - 3 * ai * aj ** 2 - aj ** 3)
    g1 = cq * (8 * r ** 5 * x4 - 8 * s * r ** 4 * x4 + r ** 3 * (14 * x2 - 4 * (7 * ai ** 2 - 4 * ai * aj + 7 * aj ** 2) * x4)
               + r ** 2 * (-6 * s * x2 - 4 * b3p * x4)
               + r * (-3 + 12 * (ai ** 2 - ai * aj + aj ** 2) * x2 + 4 * s ** 2 * qp * x4)
               - 3 * s - 4 * c3p * x2 - 4 * s ** 3 * qp * x4)
    g2 = cq * (8 * r ** 5 * x4 + 8 * s * r ** 4 * x4 + r ** 3 * (14 * x2 - 4 * (7 * ai ** 2 - 4 * ai * aj + 7 * aj ** 2) * x4)
               + r ** 2 * (6 * s * x2 + 4 * b3p * x4)
               + r * (-3 + 12 * (ai ** 2 - ai * aj + aj ** 2) * x2 + 4 * s ** 2 * qp * x4)
               + 3 * s + 4 * c3p * x2 + 4 * s ** 3 * qp * x4)
    g3 = cq * (-8 * r ** 5 * x4 + 8 * (-ai + aj) * r ** 4 * x4
               + r ** 3 * (-14 * x2 + 4 * (7 * ai ** 2 + 4 * ai * aj + 7 * aj ** 2) * x4)
               + r ** 2 * (6 * (-ai + aj) * x2 - 4 * b3m * x4)
               + r * (3 - 12 * (ai ** 2 + ai * aj + aj ** 2) * x2 - 4 * d ** 2 * qm * x4)
               + 3 * (-ai + aj) - 4 * c3m * x2 - 4 * d ** 3 * qm * x4)
    g4 = cq * (-8 * r ** 5 * x4 + 8 * d * r ** 4 * x4
               + r ** 3 * (-14 * x2 + 4 * (7 * ai ** 2 + 4 * ai * aj + 7 * aj ** 2) * x4)
               + r ** 2 * (6 * d * x2 + 4 * b3m * x4)
               + r * (3 - 12 * (ai ** 2 + ai * aj + aj ** 2) * x2 - 4 * d ** 2 * qm * x4)
               + 3 * d + 4 * c3m * x2 + 4 * d ** 3 * qm * x4)
    cf = 1.0 / (1024 * PI * ai ** 3 * aj ** 3 * r ** 3 * x6)
    base2 = 36 * r ** 4 * x4 * (-1 + 2 * (ai ** 2 + aj ** 2) * x2)
    tail = -3 + 18 * (ai ** 2 + aj ** 2) * x2 + 36 * (ai ** 2 - aj ** 2) ** 2 * x4
    g5 = cf * (-16 * r ** 6 * x6 + base2 + 64 * (ai ** 3 + aj ** 3) * r ** 3 * x6 + tail + 8 * s ** 4 * qp * x6)
    g6 = cf * (-16 * r ** 6 * x6 + base2 - 64 * (ai ** 3 + aj ** 3) * r ** 3 * x6 + tail + 8 * s ** 4 * qp * x6)
    g7 = cf * (16 * r ** 6 * x6 - base2 + 64 * (ai ** 3 - aj ** 3) * r ** 3 * x6 - tail - 8 * d ** 4 * qm * x6)
    g8 = cf * (16 * r ** 6 * x6 - base2 + 64 * (-ai ** 3 + aj ** 3) * r ** 3 * x6 - tail - 8 * d ** 4 * qm * x6)
    reg = np.zeros_like(r)
    rr = r[m1]
    reg[m1] = -1 / (64 * PI * ai ** 3 * aj ** 3 * rr ** 3) * (
        s ** 4 * qp - 8 * (ai ** 3 + aj ** 3) * rr ** 3 + 9 * (ai ** 2 + aj ** 2) * rr ** 4 - 2 * rr ** 6)
    reg[m2] = (r[m2] ** 3 + 2 * aj ** 3) / (4 * PI * aj ** 3 * r[m2] ** 3)
    reg[m3] = (r[m3] ** 3 + 2 * ai ** 3) / (4 * PI * ai ** 3 * r[m3] ** 3)
    para = (g1 * np.exp(-(r + s) ** 2 * x2) + g2 * np.exp(-(r - s) ** 2 * x2) + g3 * np.exp(-(r - ai + aj) ** 2 * x2)
            + g4 * np.exp(-(r + ai - aj) ** 2 * x2) + g5 * _erfc((r + s) * x) + g6 * _erfc((r - s) * x)
            + g7 * _erfc((r - ai + aj) * x) + g8 * _erfc((r + ai - aj) * x) + reg * regularize)
    return perp, para


def self_coupling(ai, aj, x):
    """The r -> 0 value of both components (:96-99)."""
    return (1 / (16 * P32 * ai ** 3 * aj ** 3 * x ** 3)
            * ((1 - 2 * ai ** 2 * x ** 2 + 2 * ai * aj * x ** 2 - 2 * aj ** 2 * x ** 2) * np.exp(-(ai + aj) ** 2 * x ** 2)
               + (-1 + 2 * ai ** 2 * x ** 2 + 2 * ai * aj * x ** 2 + 2 * aj ** 2 * x ** 2) * np.exp(-(ai - aj) ** 2 * x ** 2))
            + 1 / (8 * PI * ai ** 3 * aj ** 3) * ((ai ** 3 - aj ** 3) * _erf((ai - aj) * x)
                                                  - (ai ** 3 + aj ** 3) * _erf((ai + aj) * x))
            + min(ai ** 3, aj ** 3) / (4 * PI * ai ** 3 * aj ** 3))


def real_space_table(r, a_uniq, xi):
    """(Ep_perp, Ep_para), each (len(r), M(M+1)/2); row 0 is the r -> 0 value
    (r[0] must be 0), columns in the upper-triangular order of the radius types."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    a = np.asarray(a_uniq, dtype=np.float64).reshape(-1)
    M = a.size
    perp = np.zeros((r.size, M * (M + 1) // 2))
    para = np.zeros_like(perp)
    lin = 0
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(M):
            for j in range(i, M):
                p1, p2 = _couplings(r, a[i], a[j], xi)
                perp[:, lin], para[:, lin] = p1, p2
                perp[0, lin] = para[0, lin] = self_coupling(a[i], a[j], xi)
                lin += 1
    return perp, para


# ---------------------------------------------------------------------------
# wave space (Siz/Spread.m, Scale.m, Contract.m)
# ---------------------------------------------------------------------------
def ewald_parameters(box, xi, tol):
    """r_c, q_c, N_grid, h, P, eta (Siz/ComputeField.m:23-29)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    rc = np.sqrt(-np.log(tol)) / xi
    qc = 2 * xi * np.sqrt(-np.log(tol))
    ng = np.ceil(1 + box * qc / PI).astype(np.int64)
    h = box / ng
    P = int(np.ceil(2 / np.mean(h) - 2 * np.log(tol) / PI))
    eta = P * (np.mean(h) * xi) ** 2 / PI
    return rc, qc, ng, h, P, eta


def kernel(rvec, a, xi, eta):
    """Scalar kernel K with spread weight K (r.p)/|r| (vector r from the
    particle to the node), and its small-d form (Spread.m:56-61), which the
    reference writes with exp(-2 a xi^2/eta)."""
    d = np.sqrt(np.sum(rvec * rvec, axis=-1))
    with np.errstate(divide="ignore", invalid="ignore"):
        K = 3.0 / (8 * np.sqrt(2) * P32 * a ** 3 * np.sqrt(eta) * xi * d ** 2) * (
            (eta + 4 * a * xi ** 2 * d) * np.exp(-2 * (d + a) ** 2 * xi ** 2 / eta)
            - (eta - 4 * a * xi ** 2 * d) * np.exp(-2 * (d - a) ** 2 * xi ** 2 / eta))
        w = K[..., None] * rvec / d[..., None]
    small = d < 1e-6
    w[small] = (8 * np.sqrt(2) * xi ** 5 * np.exp(-2 * a * xi ** 2 / eta) / (P32 * eta ** 2.5)) * rvec[small]
    return w


class SizesGeometry:
    """Everything of ComputeField.m that depends only on positions and radii;
    tables (Ep_perp, Ep_para) on r_table with radius types from uniquetol,
    or (exact=True) the couplings of each pair's own radii at its separation."""

    def __init__(self, x, a, box, xi, tol, r_table, Ep_perp, Ep_para, Ngrid=None, P=None, exact=False):
        self.x = np.asarray(x, dtype=np.float64)
        self.a = np.asarray(a, dtype=np.float64).reshape(-1)
        self.box = np.asarray(box, dtype=np.float64).reshape(3)
        self.xi = float(xi)
        rc, _, ng, h, Pd, eta = ewald_parameters(self.box, xi, tol)
        if Ngrid is not None:
            ng = np.asarray(Ngrid, dtype=np.int64)
            h = self.box / ng
        if P is not None:
            Pd = int(P)
        if Ngrid is not None or P is not None:
            eta = Pd * (np.mean(h) * xi) ** 2 / PI
        self.rc, self.Ngrid, self.h, self.P, self.eta = rc, ng, h, Pd, eta
        n = self.x.shape[0]
        lin = np.arange(self.P ** 3)
        c = (self.P - 1) // 2
        off = np.stack((lin % self.P - c, (lin // self.P) % self.P - c, lin // (self.P * self.P) - c), axis=1)
        node0 = eo.round_half_away(self.x / h[None, :]).astype(np.int64)
        node = np.mod(node0[:, None, :] + off[None, :, :] - 1, ng[None, None, :])
        self.linnode = node[..., 0] + ng[0] * (node[..., 1] + ng[1] * node[..., 2])
        rvec = (node0 * h)[:, None, :] + (off * h)[None, :, :] - self.x[:, None, :]
        self.w = np.stack([kernel(rvec[i], self.a[i], self.xi, self.eta) for i in range(n)])  # (N, P^3, 3)
        ks = [np.fft.fftfreq(int(m), d=1.0 / int(m)) * 2 * PI / L for m, L in zip(ng, self.box)]
        KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
        q2 = KX ** 2 + KY ** 2 + KZ ** 2
        with np.errstate(divide="ignore"):
            self.scale = np.where(q2 == 0, 0.0, np.exp(-(1 - self.eta) * q2 / (4 * self.xi ** 2)) / q2)
        # real space
        self.r_table = np.asarray(r_table, dtype=np.float64)
        self.Ep_perp = np.asarray(Ep_perp, dtype=np.float64)
        self.Ep_para = np.asarray(Ep_para, dtype=np.float64)
        _, self.a_ind = uniquetol(self.a)
        M = int(self.a_ind.max()) + 1
        self.col = pair_columns(M)
        cell, ncell = eo.cell_list(self.x, self.box, self.rc)
        p1, p2 = eo.neighbor_list(cell, ncell)
        r = self.x[p1] - self.x[p2]
        r = r - self.box * np.trunc(2 * r / self.box)
        d = np.sqrt(np.sum(r * r, axis=1))
        keep = d < self.rc
        self.p1, self.p2, d, r = p1[keep], p2[keep], d[keep], r[keep]
        self.rhat = r / d[:, None]
        self.exact = exact
        if exact:
            # couplings of each pair's own radii evaluated at d (no table)
            self.A = np.zeros(d.size)
            self.B = np.zeros(d.size)
            for k in range(d.size):
                pa, pb = _couplings(np.array([d[k]]), self.a[self.p1[k]], self.a[self.p2[k]], self.xi)
                self.A[k], self.B[k] = pa[0], pb[0]
            self.self_val = np.array([self_coupling(ai, ai, self.xi) for ai in self.a])
            return
        idx = np.searchsorted(self.r_table, d, side="left")
        if np.any(idx <= 0) or np.any(idx >= self.r_table.size):
            raise ValueError("pair separation outside the real-space table range")
        cols = self.col[self.a_ind[self.p1], self.a_ind[self.p2]]
        lo, hi = self.r_table[idx - 1], self.r_table[idx]
        wl, wh = (hi - d) / (hi - lo), (d - lo) / (hi - lo)
        self.A = self.Ep_perp[idx - 1, cols] * wl + self.Ep_perp[idx, cols] * wh
        self.B = self.Ep_para[idx - 1, cols] * wl + self.Ep_para[idx, cols] * wh
        self.self_col = self.col[self.a_ind, self.a_ind]
        self.self_val = self.Ep_perp[0, self.self_col]


def field(geo: SizesGeometry, p, eps_p, parts=False):
    """E_0 (N,3) complex: ComputeField.m for dipoles p."""
    p = np.asarray(p, dtype=np.complex128)
    eps = np.asarray(eps_p).reshape(-1)
    n = p.shape[0]
    G = int(np.prod(geo.Ngrid))
    vals = np.einsum("nkc,nc->nk", geo.w, p)
    grid = (np.bincount(geo.linnode.reshape(-1), weights=np.real(vals).reshape(-1), minlength=G)
            + 1j * np.bincount(geo.linnode.reshape(-1), weights=np.imag(vals).reshape(-1), minlength=G))
    grid = grid.reshape(tuple(geo.Ngrid[::-1])).transpose(2, 1, 0)
    Et = np.fft.ifftn(np.fft.fftn(grid) * geo.scale)
    flat = Et.transpose(2, 1, 0).reshape(-1)
    Eq = np.prod(geo.h) * np.einsum("nkc,nk->nc", geo.w, flat[geo.linnode])
    Er = 3 * p / (4 * PI * geo.a[:, None] ** 3 * (eps[:, None] - 1)) + p * geo.self_val[:, None]
    pj = p[geo.p2]
    proj = np.sum(pj * geo.rhat, axis=1)
    contrib = geo.A[:, None] * (pj - proj[:, None] * geo.rhat) + (geo.B * proj)[:, None] * geo.rhat
    Er = Er + np.stack([np.bincount(geo.p1, weights=np.real(contrib[:, c]), minlength=n)
                        + 1j * np.bincount(geo.p1, weights=np.imag(contrib[:, c]), minlength=n)
                        for c in range(3)], axis=1)
    if parts:
        return Eq, Er
    return Eq + Er


def solve(geo: SizesGeometry, eps_p, E0, p_guess, tol):
    """ComputeDipole.m:33-38: GMRES(min(3N,10)), at most min(3N,100) cycles."""
    n = geo.x.shape[0]
    A = LinearOperator((3 * n, 3 * n), matvec=lambda v: field(geo, v.reshape(n, 3), eps_p).reshape(-1),
                       dtype=np.complex128)
    b = np.tile(np.asarray(E0, dtype=np.complex128), n)
    sol, _ = gmres(A, b, x0=np.asarray(p_guess, dtype=np.complex128).reshape(-1), rtol=tol, atol=0.0,
                   restart=min(3 * n, 10), maxiter=min(3 * n, 100))
    return sol.reshape(n, 3)


def dense_operator(geo: SizesGeometry, eps_p):
    n = geo.x.shape[0]
    A = np.zeros((3 * n, 3 * n), dtype=np.complex128)
    for k in range(3 * n):
        v = np.zeros(3 * n)
        v[k] = 1.0
        A[:, k] = field(geo, v.reshape(n, 3), eps_p).reshape(-1)
    return A


# ---------------------------------------------------------------------------
# field at probe points (Compute_Field_Sizes/FieldMap.m, ComputeField.m, RealSpace.m)
# ---------------------------------------------------------------------------
def real_space_table_field(r, a_uniq, xi):
    """(Ep_perp, Ep_para), each (len(r), M(M+1)/2), at every r > 0 given,
    without the overlap terms and without an r -> 0 row
    (Compute_Field_Sizes/RealSpaceTable.m)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    a = np.asarray(a_uniq, dtype=np.float64).reshape(-1)
    M = a.size
    perp = np.zeros((r.size, M * (M + 1) // 2))
    para = np.zeros_like(perp)
    lin = 0
    for i in range(M):
        for j in range(i, M):
            perp[:, lin], para[:, lin] = _couplings(r, a[i], a[j], xi, regularize=False)
            lin += 1
    return perp, para


def field_map(x_E, x_p, a, p, box, xi):
    """E (N_E,3) complex at the probe points x_E from the dipoles p of the
    spheres of radii a at x_p: tolerance 1e-3 (FieldMap.m:17-21), pair tables
    on r = [0.5 : 1e-3 : r_c + 1e-2] without overlap terms for the unique
    radii (:26-31), every probe treated as a sphere of the smallest radius in
    the wave-space contraction and in its table column (ComputeField.m:46-51,
    RealSpace.m:22-25, 56-58), the real-space sum over probe-particle pairs
    with d < r_c (brute force over all pairs in place of NeighborList_Types;
    a pair closer than the table start has no interval, as in the
    reference), and E = -ComputeField (FieldMap.m:40)."""
    x_E = np.asarray(x_E, dtype=np.float64).reshape(-1, 3)
    x_p = np.asarray(x_p, dtype=np.float64).reshape(-1, 3)
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    p = np.asarray(p, dtype=np.complex128).reshape(-1, 3)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    tol = 1e-3
    rc, _, ng, h, P, eta = ewald_parameters(box, xi, tol)
    if np.any(rc > box / 2):
        raise ValueError("Real space cutoff larger than half the box size")
    a_uniq, a_ind = uniquetol(a)
    r_table = np.arange(0.5, rc + 10 * tol + 1e-12, tol)
    Ep_perp, Ep_para = real_space_table_field(r_table, a_uniq, xi)
    col = pair_columns(a_uniq.size)
    a_min_ind = a_ind[int(np.argmin(a))]
    amin = float(np.min(a))
    # wave space: spread at the particles, contract at the probes with radius min(a)
    lin = np.arange(P ** 3)
    c = (P - 1) // 2
    off = np.stack((lin % P - c, (lin // P) % P - c, lin // (P * P) - c), axis=1)

    def stencil(x, rad):
        node0 = eo.round_half_away(x / h[None, :]).astype(np.int64)
        node = np.mod(node0[:, None, :] + off[None, :, :] - 1, ng[None, None, :])
        linnode = node[..., 0] + ng[0] * (node[..., 1] + ng[1] * node[..., 2])
        rvec = (node0 * h)[:, None, :] + (off * h)[None, :, :] - x[:, None, :]
        return linnode, np.stack([kernel(rvec[k], rad[k], xi, eta) for k in range(x.shape[0])])

    G = int(np.prod(ng))
    lin_p, w_p = stencil(x_p, a)
    vals = np.einsum("nkc,nc->nk", w_p, p)
    grid = (np.bincount(lin_p.reshape(-1), weights=np.real(vals).reshape(-1), minlength=G)
            + 1j * np.bincount(lin_p.reshape(-1), weights=np.imag(vals).reshape(-1), minlength=G))
    grid = grid.reshape(tuple(ng[::-1])).transpose(2, 1, 0)
    ks = [np.fft.fftfreq(int(m), d=1.0 / int(m)) * 2 * PI / L for m, L in zip(ng, box)]
    KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
    q2 = KX ** 2 + KY ** 2 + KZ ** 2
    with np.errstate(divide="ignore"):
        sc = np.where(q2 == 0, 0.0, np.exp(-(1 - eta) * q2 / (4 * xi ** 2)) / q2)
    Et = np.fft.ifftn(np.fft.fftn(grid) * sc)
    flat = Et.transpose(2, 1, 0).reshape(-1)
    lin_e, w_e = stencil(x_E, np.full(x_E.shape[0], amin))
    Eq = np.prod(h) * np.einsum("nkc,nk->nc", w_e, flat[lin_e])
    # real space
    Er = np.zeros((x_E.shape[0], 3), dtype=np.complex128)
    for i in range(x_E.shape[0]):
        rr = x_E[i][None, :] - x_p
        rr = rr - box[None, :] * np.trunc(2.0 * rr / box[None, :])
        d = np.sqrt(np.sum(rr * rr, axis=1))
        keep = d < rc
        if not np.any(keep):
            continue
        d, rr, pj = d[keep], rr[keep], p[keep]
        idx = np.searchsorted(r_table, d, side="left")
        if np.any(idx <= 0):
            raise ValueError("a probe lies closer to a particle than the field-map table starts (0.5)")
        cols = col[a_min_ind, a_ind[keep]]
        lo, hi = r_table[idx - 1], r_table[idx]
        wl, wh = (hi - d) / (hi - lo), (d - lo) / (hi - lo)
        A = Ep_perp[idx - 1, cols] * wl + Ep_perp[idx, cols] * wh
        B = Ep_para[idx - 1, cols] * wl + Ep_para[idx, cols] * wh
        rhat = rr / d[:, None]
        proj = np.sum(pj * rhat, axis=1)
        Er[i] = np.sum(A[:, None] * (pj - proj[:, None] * rhat) + (B * proj)[:, None] * rhat, axis=0)
    return -(Eq + Er)
