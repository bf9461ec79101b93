"""Adjoint (vector-Jacobian product) of the dipole operator and the
differentiable solve, on the GPU.

* mpol_matvec_vjp against central differences of lam^T A(x) p through the
  forward operator itself (fine tables so that no difference step crosses a
  table node often; positions away from the grid-node rounding switches);
* energy gradients of polarization_energy (position and permittivity)
  against central differences of the energy -- the bar is 1e-5 relative;
* the forces of well separated spheres in a large box against
  dipole_forces_ext.compute_forces with prefactor 1/(4 pi) on the solved
  dipoles (frozen-dipole forces; they differ by the periodic images only).
"""
import numpy as np
import pytest

from oracle import ewald_oracle as eo

from helpers import rsa

pytestmark = pytest.mark.gpu


def _fine_tables(xi, rmax, dr):
    r = np.arange(dr, rmax + dr / 2, dr)
    t = eo.real_space_table(r, xi)
    return t[2], t[3], np.concatenate(([0.0], r))


def _safe_positions(x, h, margin=1e-3):
    """Shift particles whose x/h fraction is within margin of a rounding switch."""
    f = x / h - np.floor(x / h)
    bad = np.abs(f - 0.5) < margin
    return x + bad * 4 * margin * h


def _context(gpu, x, box, xi, tol, tables, eps):
    from mpol_hip import Context
    import adjoint
    rc, ng, P, eta = adjoint.ewald_setup(box, xi, tol)
    c = Context(gpu)
    c.set_tables(*tables)
    c.set_geometry(x, box, ng, P, xi, eta, rc)
    c.set_permittivity(eps)
    return c, (rc, ng, P, eta)


def test_matvec_vjp_matches_finite_differences(gpu):
    import ctypes
    import torch
    from mpol_hip import _lib
    x, box = rsa(14, 0.06, 21, L=12.0)
    xi, tol = 1.0, 1e-8
    tables = _fine_tables(xi, 5.0, 1e-5)
    rc, ng, P, eta = __import__("adjoint").ewald_setup(box, xi, tol)
    x = _safe_positions(x, box / ng)
    eps = np.full(14, 9.0)
    c, _ = _context(gpu, x, box, xi, tol, tables, eps)
    rng = np.random.default_rng(4)
    p = torch.as_tensor(rng.standard_normal((2, 14, 3)), device=gpu)
    lam = torch.as_tensor(rng.standard_normal((2, 14, 3)), device=gpu)
    gx = torch.empty((14, 3), dtype=torch.float64, device=gpu)
    ge = torch.empty((14,), dtype=torch.float64, device=gpu)
    _lib.check(c._lib.mpol_matvec_vjp(c._ctx, ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(lam.data_ptr()), 2,
                                      ctypes.c_void_p(gx.data_ptr()), ctypes.c_void_p(ge.data_ptr())))
    gx = gx.cpu().numpy()
    ge = ge.cpu().numpy()

    def form(xx, ee):
        c.set_geometry(xx, box, ng, P, xi, eta, rc)
        c.set_permittivity(ee)
        return float((lam * c.matvec(p, slab_factor=0.0)).sum().item())

    h = 1e-6
    for i in (0, 5, 11):
        for d in range(3):
            xp = x.copy()
            xp[i, d] += h
            xm = x.copy()
            xm[i, d] -= h
            fd = (form(xp, eps) - form(xm, eps)) / (2 * h)
            assert abs(fd - gx[i, d]) <= 2e-6 * np.max(np.abs(gx)), (i, d, fd, gx[i, d])
    for i in (2, 9):
        ep = eps.copy()
        ep[i] += 1e-4
        em = eps.copy()
        em[i] -= 1e-4
        fd = (form(x, ep) - form(x, em)) / 2e-4
        assert abs(fd - ge[i]) <= 1e-7 * max(1.0, abs(ge[i]))


@pytest.mark.parametrize("slab", [0.0, 1.0])
def test_energy_gradient_matches_finite_differences(gpu, slab):
    """Forces and permittivity derivatives of U = -1/2 E0 . sum p within
    1e-5 of central differences of the energy."""
    import torch
    import adjoint
    x, box = rsa(12, 0.08, 5, L=11.0)
    xi, tol = 1.0, 1e-10
    tables = _fine_tables(xi, 5.0, 1e-5)
    rc, ng, P, eta = adjoint.ewald_setup(box, xi, tol)
    x = _safe_positions(x, box / ng)
    rng = np.random.default_rng(8)
    eps = rng.uniform(4.0, 30.0, 12)
    xt = torch.as_tensor(x, device=gpu).requires_grad_(True)
    et = torch.as_tensor(eps, device=gpu).requires_grad_(True)
    U = adjoint.polarization_energy(xt, et, box, xi, tol, slab_factor=slab, tables=tables)
    gx, ge = torch.autograd.grad(U, (xt, et))
    F = -gx.cpu().numpy()
    ge = ge.cpu().numpy()

    def energy(xx, ee):
        return float(adjoint.polarization_energy(xx, ee, box, xi, tol, slab_factor=slab, tables=tables).item())

    h = 1e-6
    scale = np.max(np.abs(F))
    for i in (1, 7):
        for d in range(3):
            xp = x.copy()
            xp[i, d] += h
            xm = x.copy()
            xm[i, d] -= h
            fd = -(energy(xp, eps) - energy(xm, eps)) / (2 * h)
            assert abs(fd - F[i, d]) <= 1e-5 * scale, (i, d, fd, F[i, d])
    for i in (0, 4):
        ep = eps.copy()
        ep[i] += 1e-4
        em = eps.copy()
        em[i] -= 1e-4
        fd = (energy(x, ep) - energy(x, em)) / 2e-4
        assert abs(fd - ge[i]) <= 1e-5 * np.max(np.abs(ge))
    # the forces sum to zero (translation invariance), up to the grid discretisation
    # (a Gaussian-spread grid is translation invariant only to its accuracy)
    assert np.max(np.abs(F.sum(axis=0))) < 1e-5 * scale


def test_forces_of_separated_spheres_match_frozen_dipole_kernel(gpu):
    import adjoint
    import dipole_forces_ext
    box = np.array([40.0, 40.0, 40.0])
    x = np.array([[20.0, 20.0, 17.0], [20.3, 19.6, 20.2], [22.9, 20.4, 21.1]])
    xi, tol = 0.5, 1e-8
    eps = np.full(3, 12.0)
    F, U = adjoint.polarization_forces(x, eps, box, xi, tol)
    p = adjoint.solve_dipoles(x, eps, box, xi, tol)[0].detach().cpu().numpy()
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    assert np.min(d[d > 0]) > 2.0
    Fk, _ = dipole_forces_ext.compute_forces(x - 20.0, p, box, 19.0, 1.0 / (4 * np.pi))
    assert np.max(np.abs(F - Fk)) < 2e-3 * np.max(np.abs(Fk)), (F, Fk)
    assert U < 0


def test_forward_ewald_forces_match_adjoint_forces(gpu):
    """mpol_ewald_forces (real space from the tabulated pair-force couplings
    field_dip_force_1/2, wave space from the Gaussian derivative) against the
    adjoint forces, which differentiate the interpolated couplings instead.
    The slope of a linear interpolant is the derivative at the middle of its
    interval, so the two differ by A''(d - mid), first order in the table
    spacing: about 7e-4 of the largest force on the default 1e-3 grid, and a
    hundred times less on a 1e-5 grid (both converge to the derivative of the
    exact couplings)."""
    import adjoint
    x, box = rsa(40, 0.15, 11)
    xi, tol = 1.0, 1e-10
    rng = np.random.default_rng(3)
    eps = rng.uniform(3.0, 25.0, 40)
    Fa, Ua = adjoint.polarization_forces(x, eps, box, xi, tol)
    Ff, Uf, p = adjoint.ewald_forces(x, eps, box, xi, tol)
    scale = np.max(np.abs(Fa))
    assert abs(Uf - Ua) <= 1e-12 * abs(Ua)
    coarse = np.max(np.abs(Ff - Fa)) / scale
    assert coarse < 2e-3, coarse
    # fine tables for both: couplings and force couplings on a 1e-5 grid
    r = np.arange(1e-5, 7.0, 1e-5)
    t = eo.real_space_table(r, xi)
    fine = (t[2], t[3], np.concatenate(([0.0], r)))
    Fa2, _ = adjoint.polarization_forces(x, eps, box, xi, tol, tables=fine)
    Ff2, _, _ = adjoint.ewald_forces(x, eps, box, xi, tol, tables=fine, force_tables=(t[4], t[5]))
    fine_err = np.max(np.abs(Ff2 - Fa2)) / scale
    assert fine_err < 2e-5 and fine_err < 0.05 * coarse, (fine_err, coarse)
    assert np.max(np.abs(Ff2 - Fa)) < 2e-3 * scale


def test_forward_ewald_forces_of_separated_spheres_match_frozen_dipole_kernel(gpu):
    import adjoint
    import dipole_forces_ext
    box = np.array([40.0, 40.0, 40.0])
    x = np.array([[20.0, 20.0, 17.0], [20.3, 19.6, 20.2], [22.9, 20.4, 21.1]])
    F, U, p = adjoint.ewald_forces(x, np.full(3, 12.0), box, 0.5, 1e-8)
    Fk, _ = dipole_forces_ext.compute_forces(x - 20.0, p, box, 19.0, 1.0 / (4 * np.pi))
    assert np.max(np.abs(F - Fk)) < 2e-3 * np.max(np.abs(Fk)), (F, Fk)
    assert U < 0
