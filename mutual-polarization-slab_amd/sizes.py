"""Polydisperse mutual polarization (spheres of different radii) on the MI355X.

The interface of the reference's MATLAB package Capacitance_Sizes/, as Python
functions of the same names and argument order (0-based indices where MATLAB
is 1-based; C-order arrays with MATLAB's dimension order):

  RealSpaceTable(r, a, xi)                         Siz/RealSpaceTable.m:1-150
  ComputeField(x, a, p, eps_p, box, n1, n2, xi,
               r_table, Ep_perp, Ep_para, tol)     Siz/ComputeField.m:1-58
  ComputeDipole(x, a, eps_p, E0, box, p_guess, ...) Siz/ComputeDipole.m:1-53
  Dipole(x, a, eps_p, box, E_0, p_guess, xi, ...)   Siz/Dipole.m:1-39
  DipoleSpectrum(x, box, a, eps_p, E_0, xi)         Siz/DipoleSpectrum.m:1-68
  CapacitanceTensor(x, a, eps_p, box, p_guess, ...) Siz/CapacitanceTensor.m:1-55
  CapacitanceTensorSpectrum(x, box, a, eps_p, xi)   Siz/CapacitanceTensorSpectrum.m:1-62

The operator and the GMRES solve run in the HIP kernels of libmpol.so
(mpol_sizes_matvec / mpol_sizes_solve); there is no CPU path.  n1, n2 are
accepted for signature compatibility (the GPU builds the same 27-cell pair
list from x).  Radius types follow MATLAB's uniquetol (tolerance 1e-6 of the
largest radius) as in RealSpace.m:21.

Beyond the reference: closed_form=True evaluates the pair couplings of each
pair's own radii on the GPU instead of interpolating per-radius-pair tables
(the tables need M(M+1)/2 columns for M distinct radii, which a continuous
radius distribution makes impractical), and complex permittivities are
solved as complex systems.
"""
from __future__ import annotations

import numpy as np

import ewald as _ewald
from mpol_hip import FP32, FP64
from mpol_hip import context as _ctxmod

__all__ = ["RealSpaceTable", "ComputeField", "ComputeDipole", "Dipole", "DipoleSpectrum", "CapacitanceTensor",
           "CapacitanceTensorSpectrum", "ewald_parameters", "uniquetol", "real_space_table", "compute_field",
           "compute_dipole", "dipole", "dipole_spectrum", "capacitance_tensor", "capacitance_tensor_spectrum"]


def uniquetol(a, tol=1e-6):
    """(unique radii, type index of every particle) with MATLAB uniquetol's
    rule: values within tol*max|a| of a group's first (smallest) member
    belong to it."""
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


def RealSpaceTable(r, a, xi):
    """(Ep_perp, Ep_para), each (len(r), M(M+1)/2) for the M radii a; row 0
    is the r -> 0 value (the reference expects r[0] == 0)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    if r.size and r[0] != 0:
        import warnings
        warnings.warn("First entry of real space table is the r->0 limit, but r(1) = %.3g" % r[0])
    return _ctxmod.real_space_table_sizes(r, np.asarray(a, dtype=np.float64).reshape(-1), xi)


def ewald_parameters(box, xi, tol):
    """r_c, q_c, N_grid, h, P, eta (Siz/ComputeField.m:23-29): isotropic eta
    from the mean grid spacing and P = ceil(2/mean(h) - 2 ln tol/pi)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    rc = np.sqrt(-np.log(tol)) / xi
    qc = 2 * xi * np.sqrt(-np.log(tol))
    ng = np.ceil(1 + box * qc / np.pi).astype(np.int64)
    h = box / ng
    P = int(np.ceil(2 / np.mean(h) - 2 * np.log(tol) / np.pi))
    eta = P * (np.mean(h) * xi) ** 2 / np.pi
    return rc, qc, ng, h, P, eta


_SCACHE: dict = {}
_SCACHE_MAX = 4


def _context(x, a, box, xi, tol, r_table=None, Ep_perp=None, Ep_para=None, closed_form=False):
    x = np.asarray(x, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    if np.any(a <= 0):
        raise ValueError("radii must be positive")
    rc, _, ng, h, P, eta = ewald_parameters(box, xi, tol)
    if closed_form:
        rv = np.array([0.0, 2.0 * rc])
        tab_key = b"closed"
    else:
        rv = np.ascontiguousarray(np.asarray(r_table, dtype=np.float64).reshape(-1))
        Ep_perp = np.real(np.asarray(Ep_perp)).reshape(rv.size, -1)
        Ep_para = np.real(np.asarray(Ep_para)).reshape(rv.size, -1)
        tab_key = _ewald._digest(rv, np.ascontiguousarray(Ep_perp), np.ascontiguousarray(Ep_para))
    key = (_ewald._digest(x), _ewald._digest(a), _ewald._digest(box), float(xi), float(tol), tab_key)
    ctx = _SCACHE.get(key)
    if ctx is None:
        ctx = _ctxmod.Context()
        if closed_form:
            ctx.set_tables(np.zeros(2), np.zeros(2), rv)
        else:
            ctx.set_tables(Ep_perp[:, 0], Ep_para[:, 0], rv)
        ctx.set_geometry(x, box, ng, P, xi, np.full(3, eta), rc)
        if closed_form:
            ctx.set_sizes(a)
        else:
            au, ind = uniquetol(a)
            M = au.size
            if Ep_perp.shape[1] != M * (M + 1) // 2:
                raise ValueError(f"tables have {Ep_perp.shape[1]} columns; {M} distinct radii need {M * (M + 1) // 2}")
            ctx.set_sizes(a, ind, Ep_perp, Ep_para)
        ctx._eps_key = None
        _SCACHE[key] = ctx
        while len(_SCACHE) > _SCACHE_MAX:
            _SCACHE.pop(next(iter(_SCACHE))).close()
    return ctx, rc


def _set_eps(ctx, eps):
    eps = np.asarray(eps)
    eps = eps.astype(np.complex128) if np.iscomplexobj(eps) and np.any(np.imag(eps) != 0) else np.real(eps).astype(
        np.float64)
    k = (_ewald._digest(np.ascontiguousarray(eps).view(np.float64)), eps.dtype.str, eps.shape)
    if getattr(ctx, "_eps_key", None) != k:
        ctx.set_permittivity(eps)
        ctx._eps_key = k


def _check_rc(rc, box):
    box = np.asarray(box, dtype=np.float64).reshape(3)
    if np.any(rc > box / 2):
        raise ValueError("Real space cutoff (r_c = %.3f) larger than half the box size (%.3f, %.3f, %.3f)."
                         % (rc, box[0], box[1], box[2]))


def ComputeField(x, a, p, eps_p, box, n1, n2, xi, r_table, Ep_perp, Ep_para, tol, *, precision="fp64",
                 closed_form=False):
    """E_0 (N,3) complex128: the field at the particles consistent with the
    dipoles p (wave + real space + dielectric term)."""
    del n1, n2
    ctx, _ = _context(x, a, box, xi, tol, r_table, Ep_perp, Ep_para, closed_form)
    _set_eps(ctx, np.asarray(eps_p).reshape(1, -1))
    prec = FP32 if precision == "fp32" else FP64
    p = np.asarray(p)
    if ctx.complex:
        return ctx.sizes_matvec(p[None].astype(np.complex128), precision=prec)[0].cpu().numpy()
    if np.iscomplexobj(p) and np.any(np.imag(p) != 0):
        out = ctx.sizes_matvec(np.stack([np.real(p), np.imag(p)]), precision=prec).cpu().numpy()
        return out[0] + 1j * out[1]
    return ctx.sizes_matvec(np.real(p)[None], precision=prec)[0].cpu().numpy().astype(np.complex128)


def _solve(ctx, n, E0s, guesses, tol, precision):
    """Batched solve for fields E0s (k,3) from guesses (k,N,3)."""
    cplx = ctx.complex
    dt = np.complex128 if cplx else np.float64
    b = np.zeros((E0s.shape[0], n, 3), dtype=dt)
    b[:] = E0s[:, None, :]
    x0 = np.asarray(guesses).astype(dt) if cplx else np.real(np.asarray(guesses)).astype(dt)
    res = ctx.sizes_solve(b, x0, float(tol), restart=min(3 * n, 10), maxit=min(3 * n, 100),
                          mixed=(precision == "mixed"))
    return res.x.cpu().numpy().astype(np.complex128), res


def ComputeDipole(x, a, eps_p, E0, box, p_guess, n1, n2, xi, r_table, Ep_perp, Ep_para, tol, *,
                  precision="fp64", closed_form=False):
    """p (N,3): GMRES(min(3N,10)) with at most min(3N,100) cycles to the
    relative tolerance tol (ComputeDipole.m:25-38)."""
    del n1, n2
    n = np.asarray(x).shape[0]
    ctx, _ = _context(x, a, box, xi, tol, r_table, Ep_perp, Ep_para, closed_form)
    _set_eps(ctx, np.asarray(eps_p).reshape(1, -1))
    sol, _ = _solve(ctx, n, np.asarray(E0, dtype=np.float64).reshape(1, 3),
                    np.asarray(p_guess).reshape(1, n, 3), tol, precision)
    return sol[0]


def Dipole(x, a, eps_p, box, E_0, p_guess, xi, tol, r_table, Ep_perp, Ep_para, *, precision="fp64",
           closed_form=False):
    """(p_ave (3,), p (N,3)) (Dipole.m:21-37)."""
    rc = np.sqrt(-np.log(tol)) / xi
    _check_rc(rc, box)
    p = ComputeDipole(x, a, eps_p, E_0, box, p_guess, None, None, xi, r_table, Ep_perp, Ep_para, tol,
                      precision=precision, closed_form=closed_form)
    return p.mean(axis=0), p


def CapacitanceTensor(x, a, eps_p, box, p_guess, xi, tol, r_table, Ep_perp, Ep_para, *, precision="fp64",
                      closed_form=False):
    """(C (9,), p (N,9)): for unit fields along x, y, z, C[3i:3i+3] is the
    mean dipole and p[:, 3i:3i+3] the dipoles (CapacitanceTensor.m:24-52).
    The three fields are one batched device solve, each from p_guess; the
    reference warm-starts fields 2 and 3 from the permuted previous solution,
    which changes only the iterate path."""
    rc = np.sqrt(-np.log(tol)) / xi
    _check_rc(rc, box)
    n = np.asarray(x).shape[0]
    ctx, _ = _context(x, a, box, xi, tol, r_table, Ep_perp, Ep_para, closed_form)
    _set_eps(ctx, np.asarray(eps_p).reshape(1, -1))
    g = np.broadcast_to(np.asarray(p_guess).reshape(-1, n, 3), (3, n, 3))
    sol, _ = _solve(ctx, n, np.eye(3), g, tol, precision)
    p = np.concatenate([sol[i] for i in range(3)], axis=1)
    return p.mean(axis=0), p


def _frames(x):
    x = np.asarray(x, dtype=np.float64)
    return x[:, :, None] if x.ndim == 2 else x


def _spectrum(x, box, a, eps_p, fields, xi, guesses, precision, closed_form):
    """Per frame, every (spectrum point, field) in one batched solve:
    S (N_frames, N_k, n_fields, N, 3)."""
    x = _frames(x)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    n, _, nfr = x.shape
    eps = np.asarray(eps_p).reshape(n, -1)
    nk = eps.shape[1]
    nf = fields.shape[0]
    tol = 1e-3
    phi = 4 * np.pi * np.sum(a ** 3) / (3 * np.prod(box))
    rc = np.sqrt(-np.log(tol)) / xi
    _check_rc(rc, box)
    if closed_form:
        r_table = Ep_perp = Ep_para = None
    else:
        au, _ = uniquetol(a)
        r_table = np.concatenate(([0.0], 1.0 + np.arange(0, 9001) * tol))
        Ep_perp, Ep_para = RealSpaceTable(r_table, au, xi)
    cplx = np.iscomplexobj(eps) and np.any(np.imag(eps) != 0)
    out = np.zeros((nfr, nk, nf, n, 3), dtype=np.complex128)
    for j in range(nfr):
        ctx, _ = _context(x[:, :, j], a, box, xi, tol, r_table, Ep_perp, Ep_para, closed_form)
        sets = np.repeat(eps.T, nf, axis=0)
        _set_eps(ctx, sets if cplx else np.real(sets))
        g = np.zeros((nk * nf, n, 3), dtype=np.complex128)
        for k in range(nk):
            e = eps[:, k]
            with np.errstate(invalid="ignore", divide="ignore"):
                alpha = (e - 1) / (e + 2)
            alpha = np.where(np.isinf(e), 1.0, alpha)
            mg = 4 * np.pi * a ** 3 * alpha / (1 - alpha * phi)
            for f in range(nf):
                if guesses is not None:
                    g[k * nf + f] = guesses[:, :, k, j]
                else:
                    g[k * nf + f] = mg[:, None] * fields[f][None, :]
        b_fields = np.tile(fields, (nk, 1))
        sol, _ = _solve(ctx, n, b_fields, g, tol, precision)
        out[j] = sol.reshape(nk, nf, n, 3)
    return out


def DipoleSpectrum(x, box, a, eps_p, E_0, xi, p_guess=None, *, precision="fp64", closed_form=False):
    """(p_ave (N_k,3), p (N,3,N_k,N_frames)) with tol 1e-3 and
    r_table = [0, 1:1e-3:10] (DipoleSpectrum.m:17-66)."""
    S = _spectrum(x, box, a, eps_p, np.asarray(E_0, dtype=np.float64).reshape(1, 3), xi, p_guess, precision,
                  closed_form)
    p = np.transpose(S[:, :, 0], (2, 3, 1, 0))  # (N, 3, N_k, N_frames)
    p_ave = S[:, :, 0].mean(axis=2).mean(axis=0)  # (N_k, 3)
    return p_ave, p


def CapacitanceTensorSpectrum(x, box, a, eps_p, xi, *, precision="fp64", closed_form=False):
    """(C (N_k,9), p (N,9,N_k,N_frames)) with tol 1e-3 and
    r_table = [0, 1:1e-3:10]; the 9 components are [xx, yx, zx, xy, yy, zy,
    xz, yz, zz] (field direction major), averaged over frames
    (CapacitanceTensorSpectrum.m:17-60)."""
    S = _spectrum(x, box, a, eps_p, np.eye(3), xi, None, precision, closed_form)
    nfr, nk, _, n, _ = S.shape
    p = np.transpose(S, (3, 2, 4, 1, 0)).reshape(n, 9, nk, nfr)
    C = p.mean(axis=0).mean(axis=2).T  # (N_k, 9)
    return C, p


# snake_case aliases
real_space_table = RealSpaceTable
compute_field = ComputeField
compute_dipole = ComputeDipole
dipole = Dipole
dipole_spectrum = DipoleSpectrum
capacitance_tensor = CapacitanceTensor
capacitance_tensor_spectrum = CapacitanceTensorSpectrum
