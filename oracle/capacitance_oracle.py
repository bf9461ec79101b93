"""FP64 CPU oracle for the GMRES solve and the capacitance workflow.

TEST INFRASTRUCTURE ONLY (see ewald_oracle.py's header): used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.

Restates ``CPU - Ewald test/capacitance.py`` (restarted GMRES with modified
Gram-Schmidt and complex Givens rotations, the SciPy GMRES default path, the
Ewald parameter derivation and the Maxwell-Garnett initial guess) and the
MATLAB batching semantics of ``Capacitance_Hetero/Capacitance_Tensor.m`` and
``CapacitanceTensorSpectrum.m`` that have no Python counterpart.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres as scipy_gmres

from . import ewald_oracle as eo


def ewald_parameters(box, xi, errortol):
    """rc, kcut, Ngrid, h, P, eta from one tolerance ([E]capacitance.py:357-362)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    s = np.sqrt(-np.log(errortol))
    rc = float(s / xi)
    kcut = 2.0 * xi * s
    ngrid = np.ceil(1.0 + box * kcut / np.pi).astype(np.int64)
    h = box / ngrid
    P = int(np.ceil(-2.0 * np.log(errortol) / np.pi))
    eta = P * (h * xi) ** 2 / np.pi
    return rc, kcut, ngrid, h, P, eta


def gmres_restarted(matvec, b, x0, tol, restart, maxit):
    """Restarted GMRES, MGS Arnoldi, complex Givens, best-so-far least squares
    ([E]capacitance.py:60-199).  Returns (x, flag, relres, nmatvec)."""
    b = np.asarray(b, dtype=np.complex128).reshape(-1)
    x = np.asarray(x0, dtype=np.complex128).reshape(-1).copy()
    nb = np.linalg.norm(b)
    nmv = 0
    if nb == 0.0:
        return x, 0, 0.0, nmv

    def mv(v):
        nonlocal nmv
        nmv += 1
        return np.asarray(matvec(v), dtype=np.complex128).reshape(-1)

    for _ in range(maxit):
        r = b - mv(x)
        beta = np.linalg.norm(r)
        rel = float(beta / nb)
        if rel <= tol:
            return x, 0, rel, nmv
        n = b.size
        V = np.zeros((n, restart + 1), dtype=np.complex128)
        H = np.zeros((restart + 1, restart), dtype=np.complex128)
        V[:, 0] = r / beta
        g = np.zeros(restart + 1, dtype=np.complex128)
        g[0] = beta
        cs = np.zeros(restart, dtype=np.complex128)
        sn = np.zeros(restart, dtype=np.complex128)
        done = False
        ybest = None
        relbest = rel
        for j in range(restart):
            w = mv(V[:, j])
            for i in range(j + 1):
                H[i, j] = np.vdot(V[:, i], w)
                w = w - H[i, j] * V[:, i]
            H[j + 1, j] = np.linalg.norm(w)
            V[:, j + 1] = w / H[j + 1, j] if H[j + 1, j] != 0.0 else 0.0
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            a, c = H[j, j], H[j + 1, j]
            den = np.sqrt(abs(a) ** 2 + abs(c) ** 2)
            cs[j], sn[j] = (1.0, 0.0) if den == 0.0 else (a / den, c / den)
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0.0
            gj = g[j]
            g[j] = cs[j] * gj + sn[j] * g[j + 1]
            g[j + 1] = -np.conj(sn[j]) * gj + cs[j] * g[j + 1]
            rin = float(abs(g[j + 1]) / nb)
            if rin < relbest:
                ybest = np.linalg.lstsq(H[:j + 1, :j + 1], g[:j + 1], rcond=None)[0]
                relbest = rin
            if rin <= tol:
                y = np.linalg.lstsq(H[:j + 1, :j + 1], g[:j + 1], rcond=None)[0]
                x = x + V[:, :j + 1] @ y
                done = True
                break
        if not done:
            if ybest is None:
                ybest = np.linalg.lstsq(H[:restart, :restart], g[:restart], rcond=None)[0]
            x = x + V[:, :ybest.size] @ ybest
    r = b - mv(x)
    return x, 1, float(np.linalg.norm(r) / nb), nmv


def solve_dipoles(geo, lambdap, Hhat, x0, errortol, slab_factor=1.0, use_scipy_gmres=True):
    """magnetic_dipole ([E]capacitance.py:205-321): restart=min(3N,10),
    maxit=min(3N,100), one extra matvec for the true residual.
    Returns (p (N,3), info dict)."""
    n = geo.x.shape[0]
    b = np.tile(np.asarray(Hhat, dtype=np.complex128).reshape(3), n)
    x0 = np.asarray(x0, dtype=np.complex128).reshape(n, 3).ravel()
    count = [0]

    def fun(v):
        count[0] += 1
        return eo.field(geo, np.asarray(v).reshape(n, 3), lambdap, slab_factor).reshape(-1)

    restart = min(3 * n, 10)
    maxit = min(3 * n, 100)
    if use_scipy_gmres:
        op = LinearOperator((3 * n, 3 * n), matvec=fun, dtype=np.complex128)
        sol, info = scipy_gmres(op, b, x0=x0, restart=restart, maxiter=maxit, rtol=errortol, atol=0.0)
        r = b - fun(sol)
        rel = float(np.linalg.norm(r) / np.linalg.norm(b))
        flag = 0 if (info == 0 and rel <= errortol) else 1
    else:
        sol, flag, _, _ = gmres_restarted(fun, b, x0, errortol, restart, maxit)
        r = b - fun(sol)
        rel = float(np.linalg.norm(r) / np.linalg.norm(b))
    return np.asarray(sol).reshape(n, 3), {"flag": flag, "relres": rel, "matvecs": count[0]}


def maxwell_garnett_guess(eps_p, phi, Hhat):
    """4 pi beta/(1 - beta phi) Hhat, beta=(eps-1)/(eps+2), beta(inf)=1
    ([E]capacitance.py:495-501)."""
    eps_p = np.asarray(eps_p, dtype=np.float64).reshape(-1)
    with np.errstate(invalid="ignore"):
        beta = (eps_p - 1.0) / (eps_p + 2.0)
    beta[np.isinf(eps_p)] = 1.0
    amp = 4.0 * np.pi * beta / (1.0 - beta * phi)
    return amp[:, None] * np.asarray(Hhat, dtype=np.float64).reshape(1, 3)


_TABLE_CACHE: dict = {}


def default_tables(xi):
    """field_dip_1/2 on r = [0, 0.001:0.001:10] ([E]capacitance.py:469-477)."""
    key = float(xi)
    if key not in _TABLE_CACHE:
        r = eo.default_rtable()
        tabs = eo.real_space_table(r, xi)
        _TABLE_CACHE[key] = (tabs[2], tabs[3], np.concatenate(([0.0], r)))
    return _TABLE_CACHE[key]


def capacitance(x, eps_p, box, Sguess, xi, errortol, perp, para, rvals, Hhat,
                slab_factor=1.0, use_scipy_gmres=True):
    """capacitance ([E]capacitance.py:327-418): derive parameters, refuse
    rc > L/2, solve, C = mean(S . Hhat)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    rc, _, ngrid, _, P, _ = ewald_parameters(box, xi, errortol)
    if np.any(rc > box / 2.0):
        raise ValueError("Real-space cutoff exceeds half the box size; tolerance unmet.")
    geo = eo.OperatorGeometry(x, box, ngrid, P, xi, rc, perp, para, rvals)
    S, info = solve_dipoles(geo, eps_p, Hhat, Sguess, errortol, slab_factor, use_scipy_gmres)
    return np.mean(S @ np.asarray(Hhat, dtype=np.float64).reshape(3)), S, info


def run_capacitance(x, box, eps_p, xi, H0, errortol=1e-5, slab_factor=1.0, use_scipy_gmres=True):
    """run_capacitance ([E]capacitance.py:424-515)."""
    x = np.asarray(x, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    eps_p = np.asarray(eps_p, dtype=np.float64).reshape(-1)
    if x.ndim == 2:
        x = x[:, :, None]
    n, nf = x.shape[0], x.shape[2]
    phi = 4.0 * np.pi * n / (3.0 * np.prod(box))
    perp, para, rvals = default_tables(xi)
    H0 = np.asarray(H0, dtype=np.float64).reshape(3)
    Hhat = H0 / np.linalg.norm(H0)
    C = np.zeros(nf, dtype=np.complex128)
    p = np.zeros((n, 3, nf), dtype=np.complex128)
    for f in range(nf):
        guess = maxwell_garnett_guess(eps_p, phi, Hhat)
        C[f], p[:, :, f], _ = capacitance(x[:, :, f], eps_p, box, guess, xi, errortol, perp, para,
                                          rvals, Hhat, slab_factor, use_scipy_gmres)
    return C, p


def capacitance_tensor(geo, eps_p, x0, errortol, slab_factor=0.0, use_scipy_gmres=True):
    """Het/Capacitance_Tensor.m:22-59: solve for fields along x, y, z in turn,
    warm-starting each from the previous solution permuted [3,1,2];
    C(i,:) = mean(S_i)."""
    n = geo.x.shape[0]
    C = np.zeros((3, 3), dtype=np.complex128)
    S = np.zeros((n, 3, 3), dtype=np.complex128)
    guess = np.asarray(x0, dtype=np.complex128).reshape(n, 3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = 1.0
        Si, _ = solve_dipoles(geo, eps_p, e, guess, errortol, slab_factor, use_scipy_gmres)
        S[:, :, i] = Si
        C[i, :] = Si.mean(axis=0)
        guess = Si[:, [2, 0, 1]]
    return C, S


def dense_operator(geo, lambdap, slab_factor=1.0):
    """Assemble the 3N x 3N matrix column by column (tiny N only)."""
    n = geo.x.shape[0]
    A = np.zeros((3 * n, 3 * n))
    for k in range(3 * n):
        e = np.zeros(3 * n)
        e[k] = 1.0
        A[:, k] = np.real(eo.field(geo, e.reshape(n, 3), lambdap, slab_factor)).reshape(-1)
    return A
