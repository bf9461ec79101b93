"""Properties the mathematics fixes, checked on the FP64 oracle (CPU only).

Each is chosen so that a plausible mistake fails it: a wrong coefficient in
the tables breaks the xi-invariance or the far-field identity; a shifted
stencil or wrong weight breaks normalisation/adjointness; a transposed or
sign-flipped coupling breaks symmetry or the lattice closed form.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn

from helpers import rsa, tables
from oracle import capacitance_oracle as co
from oracle import ewald_oracle as eo


def test_bessel_j32_matches_spherical_bessel():
    k = np.linspace(0.01, 30, 500)
    ref = spherical_jn(1, k) * np.sqrt(2 * k / np.pi)
    np.testing.assert_allclose(eo.j32(k), ref, rtol=1e-9, atol=1e-14)
    assert eo.j32(np.array([0.0]))[0] == 0.0


@pytest.mark.parametrize("xi", [0.5, 1.0])
def test_tables_far_field_and_limits(xi):
    """Beyond the overlap and the Ewald range the real-space couplings vanish;
    at r = 2 the overlap correction is continuous; as r -> 0 the I-rr and rr
    couplings tend to the self term."""
    r = np.array([2.0 - 1e-9, 2.0 + 1e-9])
    t = eo.real_space_table(r, xi)
    for k in (2, 3):
        assert abs(t[k][1] - t[k][2]) < 1e-7 * abs(t[k][1]) + 1e-12
    far = np.array([12.0 / xi])
    t = eo.real_space_table(far, xi)
    for k in range(6):
        assert abs(t[k][1]) < 1e-20
    # linear extrapolation from r = 0.01, 0.02 to r = 0 recovers the self term
    small = np.array([0.01, 0.02])
    t = eo.real_space_table(small, xi)
    self_fd = t[2][0]
    for k in (2, 3):
        assert abs(2 * t[k][1] - t[k][2] - self_fd) < 2e-4 * abs(self_fd)


def test_gaussian_weights_normalised_and_adjoint():
    x, box = rsa(30, 0.1, 2)
    ng = np.array([30, 30, 30])
    P, xi = 12, 0.6
    h = box / ng
    eta = P * (h * xi) ** 2 / np.pi
    off, offxyz = eo.pre_calculations(P, h)
    lin, ws, wc = eo.prepare_grid_kernel(x, ng, h, xi, eta, P, off, offxyz)
    # sum of contract weights of a particle = integral of a unit Gaussian
    np.testing.assert_allclose(wc.sum(axis=1), 1.0, rtol=1e-6)
    # <Spread m, G> * prod(h) = <m, Contract G>
    rng = np.random.default_rng(0)
    m = rng.standard_normal((30, 3))
    G = rng.standard_normal(tuple(ng) + (3,))
    lhs = np.sum(np.real(eo.spread_fast(m, ng, lin, ws)) * G) * np.prod(h)
    rhs = np.sum(m * np.real(eo.contract_fast(G, ng, lin, wc)))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_neighbor_list_equals_brute_force():
    x, box = rsa(150, 0.3, 4)
    rc = 3.0
    cell, nc = eo.cell_list(x, box, rc)
    p1, p2 = eo.neighbor_list(cell, nc)
    perp, para, rv = tables(1.0)
    p1f, p2f, _, _, _ = eo.prepare_real_space_pairs(x, box, p1, p2, rc, perp, para, rv)
    got = set(zip(p1f.tolist(), p2f.tolist()))
    d = x[:, None, :] - x[None, :, :]
    d -= box * np.trunc(2 * d / box)
    dist = np.sqrt(np.sum(d * d, axis=2))
    want = {(i, j) for i in range(150) for j in range(150) if i != j and dist[i, j] < rc}
    assert got == want


def test_field_xi_invariant():
    """Hk + Hr does not depend on the Ewald split, to the linear table
    interpolation error (rc < L/2 so that minimum image holds)."""
    x, box = rsa(40, 0.15, 9)
    rng = np.random.default_rng(1)
    m = rng.standard_normal((40, 3))
    eps = np.full(40, 5.0)
    outs = []
    for xi in (1.1, 1.3):
        tol = 1e-12
        rc = np.sqrt(-np.log(tol)) / xi
        assert rc < box[0] / 2
        ng = np.ceil(1 + box * 2 * xi * np.sqrt(-np.log(tol)) / np.pi).astype(int)
        ng = np.array([int(v) + (int(v) % 2) for v in ng])
        P = int(np.ceil(-2 * np.log(tol) / np.pi))
        perp, para, rv = tables(xi)
        geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp, para, rv)
        outs.append(np.real(eo.field(geo, m, eps, 0.0)))
    assert np.max(np.abs(outs[0] - outs[1])) / np.max(np.abs(outs[0])) < 2e-6


def test_operator_symmetric_dense():
    x, box = rsa(12, 0.1, 3)
    xi = 0.7
    perp, para, rv = tables(xi)
    geo = eo.OperatorGeometry(x, box, np.array([20, 20, 20]), 8, xi, 5.0, perp, para, rv)
    A = co.dense_operator(geo, np.full(12, 3.0), slab_factor=1.0)
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.max(np.abs(A))


def test_simple_cubic_lattice_is_maxwell_garnett():
    """For a perfect simple-cubic lattice (Lorentz local field) the exact
    dipole-only solution is 4 pi beta/(1 - beta phi) along the field
    (to the Ewald tolerance), with no slab term."""
    a = 3.0
    k = 3
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * a
    box = np.array([k * a] * 3)
    eps = np.full(x.shape[0], 6.0)
    xi, tol = 0.9, 1e-10
    rc, _, ng, _, P, _ = co.ewald_parameters(box, xi, tol)
    rc = min(rc, 4.4)
    perp, para, rv = tables(xi)
    geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp, para, rv)
    n = x.shape[0]
    phi = 4 * np.pi * n / (3 * np.prod(box))
    beta = 5.0 / 8.0
    p, info = co.solve_dipoles(geo, eps, [0, 0, 1.0], np.zeros((n, 3)), 1e-12, slab_factor=0.0)
    mg = 4 * np.pi * beta / (1 - beta * phi)
    np.testing.assert_allclose(np.real(p[:, 2]), mg, rtol=2e-6)
    assert np.max(np.abs(p[:, :2])) < 1e-6 * mg


def test_real_part_equals_hermitian_scale():
    """The reference's result for real dipoles has an imaginary part only
    from its Nyquist planes (k = -N/2 on both sides of the spectrum); its
    real part is the operator averaged over the sign of the Nyquist
    components, i.e. khat khat^T without the cross terms that pair exactly one
    Nyquist axis with a regular one (the GPU's formulation)."""
    d = np.load(__import__("conftest").GOLDEN + "/operator_n512_g42.npz")
    xi = float(d["xi"])
    perp, para, rv = tables(xi)
    geo = eo.OperatorGeometry(d["x"], d["box"], d["ngrid"], int(d["P"]), xi, float(d["rc"]), perp, para, rv)
    m = np.real(d["m"])
    Hk, _, _ = eo.field_components(geo, m, d["eps"], 0.0)
    lin, ws, wc = geo.grid
    H = eo.spread_fast(m, geo.Ngrid, lin, ws)
    F = np.fft.fftn(H, axes=(0, 1, 2))
    khat, coef = geo.kspace
    khat = np.real(np.fft.ifftshift(khat, axes=(0, 1, 2)))
    coef = np.fft.ifftshift(coef)
    ng = geo.Ngrid
    nyq = [np.zeros(int(v), bool) for v in ng]
    for a in range(3):
        if ng[a] % 2 == 0:
            nyq[a][ng[a] // 2] = True
    NY = np.meshgrid(*nyq, indexing="ij")
    M = khat[..., :, None] * khat[..., None, :]
    for a in range(3):
        for b in range(3):
            if a != b:
                M[..., a, b] = np.where(NY[a] ^ NY[b], 0.0, M[..., a, b])
    G = coef[..., None] * np.einsum("...ab,...b->...a", M, F)
    Ht = np.fft.ifftn(G, axes=(0, 1, 2))
    assert np.max(np.abs(Ht.imag)) < 1e-14 * np.max(np.abs(Ht.real))
    Hk2 = eo.contract_fast(np.real(Ht), geo.Ngrid, lin, wc)
    assert np.max(np.abs(np.real(Hk2) - np.real(Hk))) < 1e-13 * np.max(np.abs(Hk))
    assert np.max(np.abs(np.imag(Hk))) > 1e-12 * np.max(np.abs(Hk))  # the part that is dropped


@pytest.mark.parametrize("xi", [0.8, 1.3])
def test_force_tables_are_the_coupling_derivatives(xi):
    """field_dip_force_1/2 ("CPU - Ewald test/ewald.py":798-916), tabulated
    by the reference but used nowhere there, are the r-derivatives of the
    pair couplings: f1 = dA/dr = (B - A)/r (A and B derive from one scalar
    potential) and f2 = d(B - A)/dr, inside and outside the overlap region --
    so the pair force they give is the gradient of the real-space pair energy
    A (p_i.p_j) + (B - A)(p_i.u)(p_j.u)."""
    r = np.array([0.5, 1.0, 1.9, 2.1, 2.5, 3.0, 4.0])
    h = 1e-6
    t = eo.real_space_table(r, xi)
    tp = eo.real_space_table(r + h, xi)
    tm = eo.real_space_table(r - h, xi)
    A, B, f1, f2 = (t[k][1:] for k in (2, 3, 4, 5))
    dA = (tp[2][1:] - tm[2][1:]) / (2 * h)
    dB = (tp[3][1:] - tm[3][1:]) / (2 * h)
    scale = np.max(np.abs(f1))
    assert np.max(np.abs(f1 - dA)) < 1e-7 * scale
    assert np.max(np.abs(f1 - (B - A) / r)) < 1e-12 * scale
    assert np.max(np.abs(f2 - (dB - dA))) < 1e-7 * np.max(np.abs(f2))
