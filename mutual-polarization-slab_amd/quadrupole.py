"""Dipole + quadrupole mutual polarization on the MI355X.

The interface of the reference's MATLAB package Capacitance_Quadrupole/, as
Python functions of the same names and argument order (0-based indices where
MATLAB is 1-based; C-order arrays with MATLAB's dimension order):

  RealSpaceTable(r, xi)                         Quad/RealSpaceTable.m:1-380
  MatrixVectorMultiply(x, S, Q, lambda_p, ...)  Quad/MatrixVectorMultiply.m:1-78
  MatrixVectorSolve(x, lambda_p, E0, ...)       Quad/MatrixVectorSolve.m:1-62
  Capacitance(x, lambda_p, box, ...)            Quad/Capacitance.m:1-52
  CapacitanceTensor(x, lambda_p, box, ...)      Quad/CapacitanceTensor.m:1-66
  CapacitanceSpectrum(x, box, eps_p, xi)        Quad/CapacitanceSpectrum.m:1-58
  CapacitanceTensorSpectrum(x, box, eps_p, xi)  Quad/CapacitanceTensorSpectrum.m:1-64

The operator and the GMRES solve run in the HIP kernels of libmpol.so
(mpol_quad_matvec / mpol_quad_solve); there is no CPU path.  The neighbour
list p1, p2 and the stencil offsets offset, offsetxyz are accepted for
signature compatibility; the GPU precompute builds its own from x (the same
27-cell rule and the same stencil).  Complex permittivities (dielectric
functions of a spectrum) are supported: the system is then solved as a
complex system.

The unknowns per particle are the dipole S (3) and the traceless quadrupole
Q (5: xx, xy, xz, yy, yz); the operator returns the field E (3) and the field
gradient G (5) at every particle.
"""
from __future__ import annotations

import numpy as np

import ewald as _ewald
from mpol_hip import FP32, FP64
from mpol_hip import context as _ctxmod

TABLE_NAMES = ("pot_charge", "pot_dip", "pot_quad", "field_dip_1", "field_dip_2", "field_dip_force_1",
               "field_dip_force_2", "field_quad_1", "field_quad_2", "field_quad_3", "grad_quad_1",
               "grad_quad_2", "grad_quad_3", "grad_quad_4")
_DIPOLE_NAMES = ("pot_charge", "pot_dip", "field_dip_1", "field_dip_2", "field_dip_force_1", "field_dip_force_2")
_QUAD_NAMES = ("pot_quad", "field_quad_1", "field_quad_2", "field_quad_3", "grad_quad_1", "grad_quad_2",
               "grad_quad_3", "grad_quad_4")

__all__ = ["RealSpaceTable", "MatrixVectorMultiply", "MatrixVectorSolve", "Capacitance", "CapacitanceTensor",
           "CapacitanceSpectrum", "CapacitanceTensorSpectrum", "ewald_parameters", "TABLE_NAMES",
           "real_space_table", "matrix_vector_multiply", "matrix_vector_solve", "capacitance",
           "capacitance_tensor", "capacitance_spectrum", "capacitance_tensor_spectrum"]


def _precision(name):
    if name in ("fp64", "double", None):
        return FP64, False
    if name == "mixed":
        return FP64, True
    if name == "fp32":
        return FP32, False
    raise ValueError(f"unknown precision {name!r} (fp64, mixed or fp32)")


# ---------------------------------------------------------------------------
# tables
# ---------------------------------------------------------------------------
def RealSpaceTable(r, xi):
    """The 14 real-space couplings at separations r, in the output order of
    Quad/RealSpaceTable.m.  As there, the couplings are evaluated at every r
    given (r = 0 gives a non-finite value) and, when r[0] == 0, the self
    terms are prepended, so every table then has len(r) + 1 entries."""
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    pos = r > 0
    rp = r[pos]
    dip = dict(zip(_DIPOLE_NAMES, _ctxmod.real_space_table(rp, xi)))
    quad = dict(zip(_QUAD_NAMES, _ctxmod.real_space_table_quad(rp, xi)))
    out = []
    for name in TABLE_NAMES:
        t = dip[name] if name in dip else quad[name]
        v = np.full(r.size, np.nan)
        v[pos] = t[1:]
        if r.size and r[0] == 0.0:
            v = np.concatenate(([t[0]], v))
        out.append(v)
    return tuple(out)


def ewald_parameters(box, xi, errortol):
    """rc, kcut, Ngrid, h, P, eta (Quad/Capacitance.m:25-31)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    rc = np.sqrt(-np.log(errortol)) / xi
    kcut = 2 * xi * np.sqrt(-np.log(errortol))
    Ngrid = np.ceil(1 + box * kcut / np.pi).astype(np.int64)
    h = box / Ngrid
    P = int(np.ceil(-2 * np.log(errortol) / np.pi))
    eta = P * (h * xi) ** 2 / np.pi
    return rc, kcut, Ngrid, h, P, eta


# ---------------------------------------------------------------------------
# device context for one geometry + table set
# ---------------------------------------------------------------------------
_QCACHE: "dict" = {}
_QCACHE_MAX = 4


def _context(x, box, Ngrid, P, xi, eta, rc, r_vals, tabs9):
    """Context with the dipole pair (field_dip_1, field_dip_2) and the seven
    quadrupole tables on r_vals.  Tables longer than r_vals (the drivers'
    tables carry one extra leading entry) are cut to len(r_vals): the
    interpolation reads entries k-1 and k for the first r_vals[k] >= d, so
    the cut keeps every entry the reference reads."""
    rv = np.ascontiguousarray(np.asarray(r_vals, dtype=np.float64).reshape(-1))
    L = rv.size
    tabs = [np.real(np.asarray(t, dtype=np.complex128)).reshape(-1) for t in tabs9]
    if any(t.size < L for t in tabs):
        raise ValueError("real-space tables are shorter than r_vals")
    tabs = np.stack([t[:L] for t in tabs])
    x = np.asarray(x, dtype=np.float64)
    key = (_ewald._digest(x), _ewald._digest(np.asarray(box, dtype=np.float64)),
           _ewald._digest(np.asarray(Ngrid, dtype=np.int64)), int(P), float(xi),
           _ewald._digest(np.asarray(eta, dtype=np.float64)), float(rc), _ewald._digest(rv, tabs))
    ctx = _QCACHE.get(key)
    if ctx is None:
        ctx = _ctxmod.Context()
        ctx.set_tables(tabs[0], tabs[1], rv)
        ctx.set_geometry(x, box, Ngrid, P, xi, eta, rc)
        ctx.set_quad_tables(tabs[2:])
        ctx._eps_key = None
        _QCACHE[key] = ctx
        while len(_QCACHE) > _QCACHE_MAX:
            _QCACHE.pop(next(iter(_QCACHE))).close()
    return ctx


def _set_eps(ctx, lam):
    lam = np.asarray(lam)
    lam = lam.astype(np.complex128) if np.iscomplexobj(lam) and np.any(np.imag(lam) != 0) else np.real(lam).astype(
        np.float64)
    k = _ewald._digest(np.ascontiguousarray(lam).view(np.float64))
    if getattr(ctx, "_eps_key", None) != (k, lam.dtype.str, lam.shape):
        ctx.set_permittivity(lam)
        ctx._eps_key = (k, lam.dtype.str, lam.shape)


def _moments(S, Q):
    S = np.asarray(S)
    Q = np.asarray(Q)
    if S.ndim != 2 or S.shape[1] != 3 or Q.shape != (S.shape[0], 5):
        raise ValueError("S must be (N, 3) and Q (N, 5)")
    return np.concatenate([S, Q], axis=1)


# ---------------------------------------------------------------------------
# operator and solve
# ---------------------------------------------------------------------------
def MatrixVectorMultiply(x, S, Q, lambda_p, box, p1, p2, Ngrid, h, P, xi, eta, rc, offset, offsetxyz, r_vals,
                         field_dip_1, field_dip_2, field_quad_1, field_quad_2, field_quad_3, grad_quad_1,
                         grad_quad_2, grad_quad_3, grad_quad_4, *, precision="fp64"):
    """(E0 (N,3), G0 (N,5)) complex128: field and field gradient at the
    particles from their dipoles S and quadrupoles Q (wave + real space)."""
    del p1, p2, h, offset, offsetxyz  # rebuilt on the GPU from x, Ngrid, P and rc
    prec, _ = _precision(precision)
    ctx = _context(x, box, Ngrid, P, xi, eta, rc, r_vals,
                   (field_dip_1, field_dip_2, field_quad_1, field_quad_2, field_quad_3, grad_quad_1, grad_quad_2,
                    grad_quad_3, grad_quad_4))
    _set_eps(ctx, np.asarray(lambda_p).reshape(1, -1))
    m = _moments(S, Q)
    if ctx.complex or np.iscomplexobj(m) and np.any(np.imag(m) != 0):
        if not ctx.complex:
            # complex moments with real permittivities: real and imaginary parts separately
            mm = np.stack([np.real(m), np.imag(m)])
            out = ctx.quad_matvec(mm, precision=prec).cpu().numpy()
            res = out[0] + 1j * out[1]
        else:
            res = ctx.quad_matvec(m[None].astype(np.complex128), precision=prec)[0].cpu().numpy()
    else:
        res = ctx.quad_matvec(np.real(m)[None], precision=prec)[0].cpu().numpy().astype(np.complex128)
    return res[:, :3], res[:, 3:]


def MatrixVectorSolve(x, lambda_p, E0, S_guess, Q_guess, box, p1, p2, Ngrid, h, P, xi, eta, rc, errortol,
                      offset, offsetxyz, r_vals, field_dip_1, field_dip_2, field_quad_1, field_quad_2, field_quad_3,
                      grad_quad_1, grad_quad_2, grad_quad_3, grad_quad_4, *, precision="fp64",
                      return_info=False):
    """(S (N,3), Q (N,5)) solving [M_ES M_EQ; M_GS M_GQ][S; Q] = [E0; 0] by
    restarted GMRES (restart 10, at most 100 cycles, relative tolerance
    errortol).  E0 may be (3,) or (k, 3) for k fields at once (then S is
    (k, N, 3) and Q (k, N, 5))."""
    del p1, p2, h, offset, offsetxyz
    _, mixed = _precision(precision)
    ctx = _context(x, box, Ngrid, P, xi, eta, rc, r_vals,
                   (field_dip_1, field_dip_2, field_quad_1, field_quad_2, field_quad_3, grad_quad_1, grad_quad_2,
                    grad_quad_3, grad_quad_4))
    n = np.asarray(x).shape[0]
    E0 = np.asarray(E0)
    single = E0.ndim == 1
    E0 = E0.reshape(-1, 3)
    k = E0.shape[0]
    lam = np.asarray(lambda_p).reshape(-1)
    _set_eps(ctx, lam.reshape(1, -1))
    b = np.zeros((k, n, 8), dtype=np.complex128 if ctx.complex else np.float64)
    b[:, :, :3] = E0[:, None, :]
    x0 = np.zeros_like(b)
    Sg = np.asarray(S_guess).reshape((-1, n, 3))
    Qg = np.asarray(Q_guess).reshape((-1, n, 5))
    x0[:, :, :3] = Sg if ctx.complex else np.real(Sg)
    x0[:, :, 3:] = Qg if ctx.complex else np.real(Qg)
    res = ctx.quad_solve(b, x0, float(errortol), restart=min(8 * n, 10), maxit=min(8 * n, 100), mixed=mixed)
    sol = res.x.cpu().numpy().astype(np.complex128)
    S, Q = sol[:, :, :3], sol[:, :, 3:]
    if single:
        S, Q = S[0], Q[0]
    if return_info:
        return S, Q, res
    return S, Q


def _check_rc(rc, box):
    box = np.asarray(box, dtype=np.float64).reshape(3)
    if np.any(rc > box / 2):
        raise ValueError("Real space cutoff (rc = %.3f) larger than half the box size (%.3f, %.3f, %.3f). "
                         "Accuracy to specified error tolerance not guaranteed." % (rc, box[0], box[1], box[2]))


def Capacitance(x, lambda_p, box, S_guess, Q_guess, xi, errortol, offset, offsetxyz, r_vals, field_dip_1,
                field_dip_2, field_quad_1, field_quad_2, field_quad_3, grad_quad_1, grad_quad_2, grad_quad_3,
                grad_quad_4, *, precision="fp64"):
    """(C, S, Q) for a unit field along z: C = mean of the z dipoles
    (Quad/Capacitance.m:21-50)."""
    rc, _, Ngrid, h, P, eta = ewald_parameters(box, xi, errortol)
    _check_rc(rc, box)
    S, Q = MatrixVectorSolve(x, lambda_p, np.array([0.0, 0.0, 1.0]), S_guess, Q_guess, box, None, None, Ngrid, h, P,
                             xi, eta, rc, errortol, offset, offsetxyz, r_vals, field_dip_1, field_dip_2,
                             field_quad_1, field_quad_2, field_quad_3, grad_quad_1, grad_quad_2, grad_quad_3,
                             grad_quad_4, precision=precision)
    return np.mean(S[:, 2]), S, Q


def CapacitanceTensor(x, lambda_p, box, S_guess, Q_guess, xi, errortol, offset, offsetxyz, r_vals, field_dip_1,
                      field_dip_2, field_quad_1, field_quad_2, field_quad_3, grad_quad_1, grad_quad_2, grad_quad_3,
                      grad_quad_4, *, precision="fp64"):
    """(C (3,3), S (N,3,3), Q (N,5,3)) for unit fields along x, y, z:
    column i of C is the mean dipole for field i (Quad/CapacitanceTensor.m:59).
    The three fields are solved together in one batched device solve, each
    from S_guess, Q_guess; the reference warm-starts fields 2 and 3 from the
    permuted previous solution (:62-63), which changes only the iterate path."""
    rc, _, Ngrid, h, P, eta = ewald_parameters(box, xi, errortol)
    _check_rc(rc, box)
    n = np.asarray(x).shape[0]
    Sg = np.broadcast_to(np.asarray(S_guess).reshape(-1, n, 3), (3, n, 3))
    Qg = np.broadcast_to(np.asarray(Q_guess).reshape(-1, n, 5), (3, n, 5))
    S3, Q3 = MatrixVectorSolve(x, lambda_p, np.eye(3), Sg, Qg, box, None, None, Ngrid, h, P, xi, eta, rc, errortol,
                               offset, offsetxyz, r_vals, field_dip_1, field_dip_2, field_quad_1, field_quad_2,
                               field_quad_3, grad_quad_1, grad_quad_2, grad_quad_3, grad_quad_4,
                               precision=precision)
    S = np.transpose(S3, (1, 2, 0))
    Q = np.transpose(Q3, (1, 2, 0))
    C = S.mean(axis=0)
    return C, S, Q


def _driver_tables(r_table, xi):
    t = dict(zip(TABLE_NAMES, RealSpaceTable(r_table, xi)))
    return tuple(t[n] for n in ("field_dip_1", "field_dip_2", "field_quad_1", "field_quad_2", "field_quad_3",
                                "grad_quad_1", "grad_quad_2", "grad_quad_3", "grad_quad_4"))


def _spectrum_solve(x, box, eps, xi, errortol, r_table, fields, precision):
    """Every (spectrum point, field) pair of one frame in one batched solve;
    returns S (Nk, nf, N, 3) and Q (Nk, nf, N, 5)."""
    n = x.shape[0]
    nk = eps.shape[1]
    nf = fields.shape[0]
    phi = 4 * np.pi * n / (3 * np.prod(box))  # the reference calls it eta (volume fraction)
    rc, _, Ngrid, h, P, eta = ewald_parameters(box, xi, errortol)
    _check_rc(rc, box)
    tabs = _driver_tables(r_table, xi)
    ctx = _context(x, box, Ngrid, P, xi, eta, rc, r_table, tabs)
    cplx = np.iscomplexobj(eps) and np.any(np.imag(eps) != 0)
    eps_sets = np.repeat(eps.T, nf, axis=0)  # (nk*nf, n), k-major
    _set_eps(ctx, eps_sets if cplx else np.real(eps_sets))
    dt = np.complex128 if cplx else np.float64
    b = np.zeros((nk * nf, n, 8), dtype=dt)
    x0 = np.zeros_like(b)
    for k in range(nk):
        beta = (eps[:, k] - 1) / (eps[:, k] + 2)
        mg = 4 * np.pi * beta / (1 - beta * phi)
        for f in range(nf):
            b[k * nf + f, :, :3] = fields[f]
            # Maxwell-Garnett guess along the applied field
            x0[k * nf + f, :, :3] = (mg if cplx else np.real(mg))[:, None] * fields[f][None, :]
    _, mixed = _precision(precision)
    res = ctx.quad_solve(b, x0, float(errortol), restart=min(8 * n, 10), maxit=min(8 * n, 100), mixed=mixed)
    sol = res.x.cpu().numpy().astype(np.complex128).reshape(nk, nf, n, 8)
    return sol[..., :3], sol[..., 3:]


def _frames(x):
    x = np.asarray(x, dtype=np.float64)
    return x[:, :, None] if x.ndim == 2 else x


def CapacitanceSpectrum(x, box, eps_p, xi, *, precision="fp64"):
    """C (N_k,) complex: the z capacitance averaged over frames at each
    spectrum point, errortol 1e-3 (Quad/CapacitanceSpectrum.m:16-58).  The
    tables are built on r_table = 0:1e-4:10 with the driver's alignment."""
    x = _frames(x)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    eps = np.asarray(eps_p)
    eps = eps.reshape(x.shape[0], -1)
    r_table = np.arange(0, 100001) * 1e-4
    C = np.zeros((x.shape[2], eps.shape[1]), dtype=np.complex128)
    for j in range(x.shape[2]):
        S, _ = _spectrum_solve(x[:, :, j], box, eps, xi, 1e-3, r_table, np.array([[0.0, 0.0, 1.0]]), precision)
        C[j] = S[:, 0, :, 2].mean(axis=1)
    return C.mean(axis=0)


def CapacitanceTensorSpectrum(x, box, eps_p, xi, *, precision="fp64"):
    """(C (3,3,N_k), p (N,3,3,N_k,N_frames), Q (N,5,3,N_k,N_frames)) with
    errortol 1e-3 and r_table = [0, 1:1e-4:10]
    (Quad/CapacitanceTensorSpectrum.m:16-62): column i of C(:,:,k) is the
    mean dipole for a unit field along axis i, averaged over frames."""
    x = _frames(x)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    n, _, nfr = x.shape
    eps = np.asarray(eps_p).reshape(n, -1)
    nk = eps.shape[1]
    r_table = np.concatenate(([0.0], 1.0 + np.arange(0, 90001) * 1e-4))
    C = np.zeros((3, 3, nfr, nk), dtype=np.complex128)
    p = np.zeros((n, 3, 3, nk, nfr), dtype=np.complex128)
    Q = np.zeros((n, 5, 3, nk, nfr), dtype=np.complex128)
    for j in range(nfr):
        S, Qs = _spectrum_solve(x[:, :, j], box, eps, xi, 1e-3, r_table, np.eye(3), precision)
        # S (nk, field, n, 3) -> p (n, 3, field, nk)
        p[:, :, :, :, j] = np.transpose(S, (2, 3, 1, 0))
        Q[:, :, :, :, j] = np.transpose(Qs, (2, 3, 1, 0))
        C[:, :, j, :] = np.transpose(S.mean(axis=2), (2, 1, 0))
    return C.mean(axis=2), p, Q


# snake_case aliases
real_space_table = RealSpaceTable
matrix_vector_multiply = MatrixVectorMultiply
matrix_vector_solve = MatrixVectorSolve
capacitance = Capacitance
capacitance_tensor = CapacitanceTensor
capacitance_spectrum = CapacitanceSpectrum
capacitance_tensor_spectrum = CapacitanceTensorSpectrum
