"""Differentiable polarization solve: energies, forces and permittivity
gradients through a hand-written adjoint.

PolarizationSolve is a torch.autograd.Function around the device GMRES of
libmpol.so: forward solves A(x, eps) p = b; backward solves the adjoint
system A lam = dL/dp (A is symmetric: real-space couplings, the Hermitian
wave-space operator, the slab term and the diagonal all are) and returns

    dL/dx   = -d/dx   (lam^T A p)   (mpol_matvec_vjp: pair-sum and Gaussian
    dL/deps = -d/deps (lam^T A p)    weight derivatives on the GPU)
    dL/db   = lam

so any scalar built from the dipoles differentiates with respect to the
particle positions and permittivities.  polarization_energy is the
electrostatic energy U = -1/2 E0 . sum_i p_i of the induced dipoles in a
unit field, polarization_forces its negative position gradient: the forces
the reference's HOOMD driver approximates with truncated point dipoles
(dipole_forces_ext/hoomd_force.py:192-226; for well separated spheres in a
large box they agree with compute_forces(..., prefactor=1/(4 pi)) on the
solved dipoles).

Positions enter the operator also through the nearest grid node of each
particle (a rounding), so U(x) is smooth only between those switches; the
gradient is that of the smooth piece containing x.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from mpol_hip import _lib
from mpol_hip.context import Context, default_force_tables, default_tables, require_device

__all__ = ["PolarizationSolve", "solve_dipoles", "polarization_energy", "polarization_forces", "ewald_forces",
           "ewald_setup"]


def ewald_setup(box, xi, errortol):
    """rc, Ngrid, P, eta of the dipole method ("CPU - Ewald test/capacitance.py":357-362)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    rc = np.sqrt(-np.log(errortol)) / xi
    kcut = 2 * xi * np.sqrt(-np.log(errortol))
    ng = np.ceil(1 + box * kcut / np.pi).astype(np.int64)
    h = box / ng
    P = int(np.ceil(-2 * np.log(errortol) / np.pi))
    eta = P * (h * xi) ** 2 / np.pi
    return rc, ng, P, eta


class _Setup:
    def __init__(self, box, xi, errortol, ngrid=None, P=None, tables=None, slab_factor=0.0, restart=10, maxit=100,
                 solve_tol=None, rc=None):
        self.box = np.asarray(box, dtype=np.float64).reshape(3)
        self.xi = float(xi)
        # errortol sets the Ewald parameters (as in the reference) and, unless
        # solve_tol is given, the GMRES tolerance
        self.tol = float(errortol if solve_tol is None else solve_tol)
        rc_rule, ng, Pd, eta = ewald_setup(self.box, self.xi, float(errortol))
        rc = rc_rule if rc is None else float(rc)
        if ngrid is not None:
            ng = np.asarray(ngrid, dtype=np.int64).reshape(3)
        if P is not None:
            Pd = int(P)
        if ngrid is not None or P is not None:
            eta = Pd * (self.box / ng * self.xi) ** 2 / np.pi
        self.rc, self.ng, self.P, self.eta = rc, ng, Pd, eta
        if np.any(self.rc > self.box / 2):
            raise ValueError("Real-space cutoff exceeds half the box size; tolerance unmet.")
        self.tables = tables if tables is not None else default_tables(self.xi)
        self.slab = float(slab_factor)
        self.restart, self.maxit = int(restart), int(maxit)


class PolarizationSolve(torch.autograd.Function):
    """p = A(x, eps)^{-1} b on the GPU, differentiable in x (N,3), eps (N,)
    and b (nrhs, N, 3); all FP64 device tensors."""

    @staticmethod
    def forward(ctx, x, eps, b, setup: _Setup):
        c = Context(x.device)
        perp, para, rv = setup.tables
        c.set_tables(perp, para, rv)
        c.set_geometry(x.detach(), setup.box, setup.ng, setup.P, setup.xi, setup.eta, setup.rc)
        c.set_permittivity(eps.detach())
        res = c.solve(b.detach(), torch.zeros_like(b), setup.tol, restart=setup.restart, maxit=setup.maxit,
                      slab_factor=setup.slab)
        if res.flag != 0:
            raise RuntimeError(f"polarization solve did not converge (relres {res.max_relres:.2e})")
        p = res.x
        ctx.save_for_backward(p)
        ctx.c = c
        ctx.setup = setup
        return p

    @staticmethod
    def backward(ctx, grad_p):
        (p,) = ctx.saved_tensors
        c, setup = ctx.c, ctx.setup
        g = grad_p.contiguous().to(torch.float64)
        res = c.solve(g, torch.zeros_like(g), setup.tol, restart=setup.restart, maxit=setup.maxit,
                      slab_factor=setup.slab)
        lam = res.x.contiguous()
        gx = torch.empty((c.n, 3), dtype=torch.float64, device=p.device)
        geps = torch.empty((c.n,), dtype=torch.float64, device=p.device)
        c._sync_stream()
        _lib.check(c._lib.mpol_matvec_vjp(c._ctx, ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(lam.data_ptr()),
                                          p.shape[0], ctypes.c_void_p(gx.data_ptr()),
                                          ctypes.c_void_p(geps.data_ptr())))
        return -gx, -geps, lam, None


def _as_dev(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=device)


def solve_dipoles(x, eps, box, xi, errortol=1e-10, H0=(0.0, 0.0, 1.0), slab_factor=0.0, ngrid=None, P=None,
                  tables=None, solve_tol=None, rc=None):
    """Induced dipoles (nrhs, N, 3) for unit fields H0 ((3,) or (nrhs, 3)),
    differentiable with respect to x and eps when they require grad.
    errortol derives rc, the grid, P and eta by the reference's rule (ngrid,
    P and rc, when given, replace the derived ones); solve_tol (default
    errortol) is the GMRES relative tolerance of the forward and adjoint
    solves."""
    dev = x.device if isinstance(x, torch.Tensor) and x.is_cuda else require_device()
    xt = _as_dev(x, dev)
    n = xt.shape[0]
    et = _as_dev(eps, dev).reshape(-1)
    if et.numel() == 1:
        et = et.expand(n)
    H = torch.as_tensor(np.asarray(H0, dtype=np.float64).reshape(-1, 3), device=dev)
    b = H[:, None, :].expand(H.shape[0], n, 3).contiguous()
    setup = _Setup(box, xi, errortol, ngrid, P, tables, slab_factor, solve_tol=solve_tol, rc=rc)
    return PolarizationSolve.apply(xt, et.contiguous(), b, setup)


def polarization_energy(x, eps, box, xi, errortol=1e-10, H0=(0.0, 0.0, 1.0), slab_factor=0.0, ngrid=None, P=None,
                        tables=None, solve_tol=None, rc=None):
    """U = -1/2 H0 . sum_i p_i for the dipoles induced by the unit field H0
    (a differentiable scalar tensor)."""
    p = solve_dipoles(x, eps, box, xi, errortol, H0, slab_factor, ngrid, P, tables, solve_tol, rc)
    H = torch.as_tensor(np.asarray(H0, dtype=np.float64).reshape(1, 1, 3), device=p.device)
    return -0.5 * (p * H).sum()


def polarization_forces(x, eps, box, xi, errortol=1e-10, H0=(0.0, 0.0, 1.0), slab_factor=0.0, ngrid=None, P=None,
                        tables=None, solve_tol=None, rc=None):
    """(F = -dU/dx (N,3), U) as numpy arrays."""
    dev = x.device if isinstance(x, torch.Tensor) and x.is_cuda else require_device()
    xt = _as_dev(x, dev).detach().clone().requires_grad_(True)
    U = polarization_energy(xt, eps, box, xi, errortol, H0, slab_factor, ngrid, P, tables, solve_tol, rc)
    (g,) = torch.autograd.grad(U, xt)
    return (-g).cpu().numpy(), float(U.item())


def ewald_forces(x, eps, box, xi, errortol=1e-10, H0=(0.0, 0.0, 1.0), slab_factor=0.0, ngrid=None, P=None,
                 tables=None, force_tables=None, solve_tol=None, rc=None):
    """(F = -dU/dx (N,3), U, p (N,3)) as numpy arrays, computed forward: the
    dipoles solved once, then mpol_ewald_forces -- the real-space part from
    the tabulated pair-force couplings field_dip_force_1/2 ("CPU - Ewald
    test/ewald.py":798-916, which the reference tabulates and never uses),
    the wave-space part from the Gaussian weights' derivative.  Equal to
    polarization_forces (the adjoint) up to the interpolation of the tables
    (the adjoint differentiates the interpolated couplings, this interpolates
    their derivatives).  force_tables: (f1, f2) on the rvals of tables
    (default: on default_tables' grid)."""
    dev = x.device if isinstance(x, torch.Tensor) and x.is_cuda else require_device()
    xt = _as_dev(x, dev)
    n = xt.shape[0]
    et = _as_dev(eps, dev).reshape(-1)
    if et.numel() == 1:
        et = et.expand(n).contiguous()
    setup = _Setup(box, xi, errortol, ngrid, P, tables, slab_factor, solve_tol=solve_tol, rc=rc)
    if force_tables is None:
        if tables is not None:
            raise ValueError("force_tables are needed on the rvals of custom tables")
        force_tables = default_force_tables(setup.xi)
    H = torch.as_tensor(np.asarray(H0, dtype=np.float64).reshape(-1, 3), device=dev)
    b = H[:, None, :].expand(H.shape[0], n, 3).contiguous()
    c = Context(dev)
    try:
        perp, para, rv = setup.tables
        c.set_tables(perp, para, rv)
        c.set_geometry(xt, setup.box, setup.ng, setup.P, setup.xi, setup.eta, setup.rc)
        c.set_permittivity(et)
        res = c.solve(b, torch.zeros_like(b), setup.tol, restart=setup.restart, maxit=setup.maxit,
                      slab_factor=setup.slab)
        if res.flag != 0:
            raise RuntimeError(f"polarization solve did not converge (relres {res.max_relres:.2e})")
        F = c.ewald_forces(res.x, force_tables[0], force_tables[1])
        U = -0.5 * float(torch.sum(b * res.x).item())
        return F.cpu().numpy(), U, res.x.cpu().numpy()[0]
    finally:
        c.close()
