"""capacitance.py -- drop-in for the reference's GMRES solve and capacitance
workflow, on MI355X.

Same entry points and argument meaning as "CPU - Ewald test/capacitance.py".
magnetic_dipole solves A p = tile(H, N) with the library's device-resident
restarted GMRES (restart = min(3N, 10), maxit = min(3N, 100) restart cycles,
relative tolerance errortol, FP64 Arnoldi/Givens); both values of
use_scipy_gmres select it.  The convergence flag the reference drops is kept:
pass return_info=True, or read magnetic_dipole.last_info.

Extensions (keyword-only, defaults reproduce the reference):
  precision="mixed"  FP32 operator inside FP64 iterative refinement;
  capacitance_tensor / capacitance_tensor_spectrum -- the three field
  directions (and a permittivity spectrum) batched into one solve
  (Capacitance_Hetero/Capacitance_Tensor.m, CapacitanceTensorSpectrum.m).
"""
from __future__ import annotations

import time
from typing import Callable, Tuple

import numpy as np
import torch

import ewald as _ewald
from ewald import (  # noqa: F401  (re-exported like the reference's imports)
    cell_list,
    magnetic_field,
    neighbor_list,
    pre_calculations,
    real_space_table,
)
from mpol_hip import context as _ctxmod


# ---------------------------------------------------------------------------
# parameters ("CPU - Ewald test/capacitance.py":357-362, 480-484)
# ---------------------------------------------------------------------------
def ewald_parameters(box, xi, errortol):
    """rc, kcut, Ngrid, h, P, eta derived from one tolerance."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    rc = float(np.sqrt(-np.log(errortol)) / xi)
    kcut = 2.0 * xi * np.sqrt(-np.log(errortol))
    ngrid = np.ceil(1.0 + box * kcut / np.pi).astype(int)
    h = box / ngrid
    P = int(np.ceil(-2.0 * np.log(errortol) / np.pi))
    eta = P * (h * xi) ** 2 / np.pi
    return rc, kcut, ngrid, h, P, eta


# ---------------------------------------------------------------------------
# restarted GMRES with a user matvec ("CPU - Ewald test/capacitance.py":60-199)
# ---------------------------------------------------------------------------
def gmres_restarted(matvec: Callable[[np.ndarray], np.ndarray], b, x0, tol: float, restart: int, maxit: int,
                    verbose: int = 0, *, device_matvec: bool = False) -> Tuple[np.ndarray, int, float]:
    """Restarted GMRES (MGS Arnoldi, complex Givens, best-so-far least squares)
    for a Python matvec; the Krylov vectors live on the GPU.  Returns
    (x, flag, relres) like the reference.

    The reference's matvec takes and returns NumPy vectors, so by default
    every application copies the vector to the host and back.  With
    device_matvec=True the matvec takes and returns complex128 torch tensors
    on the GPU (for example a wrapper of Context.matvec), and no vector
    leaves the device."""
    dev = _ctxmod.require_device()
    b_t = torch.as_tensor(np.asarray(b, dtype=np.complex128).reshape(-1), device=dev)
    x = torch.as_tensor(np.asarray(x0, dtype=np.complex128).reshape(-1), device=dev).clone()
    nb = float(torch.linalg.vector_norm(b_t))
    if nb == 0.0:
        return x.cpu().numpy(), 0, 0.0

    def mv(v):
        if device_matvec:
            out = matvec(v)
            return torch.as_tensor(out, device=dev).to(torch.complex128).reshape(-1)
        out = matvec(v.cpu().numpy())
        return torch.as_tensor(np.asarray(out, dtype=np.complex128).reshape(-1), device=dev)

    t0 = time.perf_counter()
    calls = 0
    for outer in range(maxit):
        r = b_t - mv(x)
        calls += 1
        beta = float(torch.linalg.vector_norm(r))
        rel = beta / nb
        if verbose >= 2:
            print(f"    GMRES outer {outer + 1}/{maxit}: relres={rel:.3e}", flush=True)
        if rel <= tol:
            if verbose >= 1:
                print(f"    GMRES converged: relres={rel:.3e} matvecs={calls} time={time.perf_counter() - t0:.2f}s")
            return x.cpu().numpy(), 0, rel
        V = torch.zeros((restart + 1, b_t.numel()), dtype=torch.complex128, device=dev)
        H = np.zeros((restart + 1, restart), dtype=np.complex128)
        V[0] = r / beta
        g = np.zeros(restart + 1, dtype=np.complex128)
        g[0] = beta
        cs = np.zeros(restart, dtype=np.complex128)
        sn = np.zeros(restart, dtype=np.complex128)
        done = False
        ybest = None
        relbest = rel
        for j in range(restart):
            w = mv(V[j])
            calls += 1
            for i in range(j + 1):
                hij = torch.vdot(V[i], w)
                H[i, j] = complex(hij)
                w = w - hij * V[i]
            H[j + 1, j] = float(torch.linalg.vector_norm(w))
            V[j + 1] = w / H[j + 1, j] if H[j + 1, j] != 0 else 0.0
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            a, c = H[j, j], H[j + 1, j]
            den = np.sqrt(abs(a) ** 2 + abs(c) ** 2)
            cs[j], sn[j] = (1.0, 0.0) if den == 0 else (a / den, c / den)
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0.0
            gj = g[j]
            g[j] = cs[j] * gj + sn[j] * g[j + 1]
            g[j + 1] = -np.conj(sn[j]) * gj + cs[j] * g[j + 1]
            rin = abs(g[j + 1]) / nb
            if verbose >= 3:
                print(f"      inner {j + 1}/{restart}: relres~{rin:.3e}", flush=True)
            if rin < relbest:
                ybest = np.linalg.lstsq(H[:j + 1, :j + 1], g[:j + 1], rcond=None)[0]
                relbest = rin
            if rin <= tol:
                y = np.linalg.lstsq(H[:j + 1, :j + 1], g[:j + 1], rcond=None)[0]
                x = x + torch.as_tensor(y, device=dev) @ V[:j + 1]
                done = True
                break
        if not done:
            if ybest is None:
                ybest = np.linalg.lstsq(H[:restart, :restart], g[:restart], rcond=None)[0]
            x = x + torch.as_tensor(ybest, device=dev) @ V[:ybest.size]
    r = b_t - mv(x)
    rel = float(torch.linalg.vector_norm(r)) / nb
    if verbose >= 1:
        print(f"    GMRES NOT converged: relres={rel:.3e} matvecs={calls} time={time.perf_counter() - t0:.2f}s")
    return x.cpu().numpy(), 1, rel


# ---------------------------------------------------------------------------
# device solve ("CPU - Ewald test/capacitance.py":205-321)
# ---------------------------------------------------------------------------
def _solve_on(ctx, b, x0, errortol, restart, maxit, slab_factor, precision):
    """Solve A x = b for a stack b, x0 of shape (nrhs, N, 3) (complex allowed)."""
    b = np.asarray(b)
    x0 = np.asarray(x0)
    parts_b = [np.real(b)]
    parts_x = [np.real(x0)]
    cplx = (np.iscomplexobj(b) and np.any(b.imag != 0)) or (np.iscomplexobj(x0) and np.any(x0.imag != 0))
    if cplx:
        parts_b.append(np.imag(b))
        parts_x.append(np.imag(x0) if np.iscomplexobj(x0) else np.zeros_like(np.real(x0)))
    bb = np.concatenate(parts_b, axis=0).astype(np.float64)
    xx = np.concatenate(parts_x, axis=0).astype(np.float64)
    mixed = str(precision).lower() in ("mixed", "fp32")
    res = ctx.solve(bb, xx, errortol, restart=restart, maxit=maxit, mixed=mixed, slab_factor=slab_factor)
    sol = res.x.detach().cpu().numpy()
    k = b.shape[0]
    out = sol[:k].astype(np.complex128)
    if cplx:
        out = out + 1j * sol[k:]
    info = {"flag": res.flag, "relres": res.max_relres, "matvecs_fp64": res.matvecs_fp64,
            "matvecs_fp32": res.matvecs_fp32, "refinements": res.refinements}
    return out, info


def magnetic_dipole(x, lambdap, H, box, p1, p2, Ngrid, h, P, xi, eta, rc, mold, offset, offsetxyz, Hperp, Hpara,
                    rvals, errortol, real_precomp=None, grid_precomp=None, k_precomp=None, slab_factor: float = 1.0,
                    use_scipy_gmres: bool = True, verbose: int = 0, *, precision: str = "fp64",
                    return_info: bool = False):
    """Self-consistent dipoles (N,3) complex128 for the applied field H."""
    n = np.asarray(x).shape[0]
    ctx = _ewald.geometry_context(x, box, Ngrid, P, xi, eta, rc, Hperp, Hpara, rvals)
    _ewald._with_permittivity(ctx, lambdap)
    Hv = np.asarray(H, dtype=np.complex128).reshape(3)
    b = np.tile(Hv, n).reshape(1, n, 3)
    x0 = np.asarray(mold, dtype=np.complex128).reshape(1, n, 3)
    restart = min(3 * n, 10)
    maxit = min(3 * n, 100)
    p, info = _solve_on(ctx, b, x0, errortol, restart, maxit, slab_factor, precision)
    magnetic_dipole.last_info = info
    if verbose >= 1:
        print(f"  MagneticDipole(GPU): flag={info['flag']} relres={info['relres']:.3e} "
              f"matvecs={info['matvecs_fp64'] + info['matvecs_fp32']}", flush=True)
    return (p[0], info) if return_info else p[0]


magnetic_dipole.last_info = None


# ---------------------------------------------------------------------------
# capacitance / run_capacitance ("CPU - Ewald test/capacitance.py":327-515)
# ---------------------------------------------------------------------------
def capacitance(x, lambdap, box, Sguess, xi, errortol, dipoletable1, dipoletable2, rtable, offset, offsetxyz, Hhat,
                slab_factor: float = 1.0, use_scipy_gmres: bool = True, verbose: int = 0, *,
                precision: str = "fp64"):
    """(C, S): mean dipole along Hhat and the (N,3) dipoles."""
    x = np.asarray(x, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    Hhat = np.asarray(Hhat, dtype=np.float64).reshape(3)
    rc, _, ngrid, h, P, eta = ewald_parameters(box, xi, errortol)
    if np.any(rc > box / 2.0):
        raise ValueError("Real-space cutoff exceeds half the box size; tolerance unmet.")
    t0 = time.perf_counter()
    S = magnetic_dipole(x, lambdap, Hhat, box, None, None, ngrid, h, P, xi, eta, rc, Sguess, offset, offsetxyz,
                        dipoletable1, dipoletable2, rtable, errortol, slab_factor=slab_factor,
                        use_scipy_gmres=use_scipy_gmres, verbose=verbose, precision=precision)
    C = np.mean(S @ Hhat)
    if verbose >= 1:
        print(f"Capacitance: done C={np.real(C):.6g} total_time={time.perf_counter() - t0:.2f}s", flush=True)
    return C, S


def maxwell_garnett_guess(eps_p, phi, Hhat):
    """4 pi beta/(1 - beta phi) Hhat with beta = (eps-1)/(eps+2), beta(inf) = 1."""
    eps_p = np.asarray(eps_p, dtype=np.float64).reshape(-1)
    with np.errstate(invalid="ignore"):
        beta = (eps_p - 1.0) / (eps_p + 2.0)
    beta[np.isinf(eps_p)] = 1.0
    return (4.0 * np.pi * beta / (1.0 - beta * phi))[:, None] * np.asarray(Hhat, dtype=np.float64).reshape(1, 3)


def run_capacitance(x, box, eps_p, xi, H0, errortol: float = 1e-5, slab_factor: float = 1.0,
                    use_scipy_gmres: bool = True, verbose: int = 0, *, precision: str = "fp64", p_guess=None,
                    frame_workers: int = 4):
    """(C (F,) complex128, p (N,3,F) complex128) for one frame or a trajectory.
    p_guess ((N,3) or (N,3,F); not in the reference) replaces the
    Maxwell-Garnett initial guess, e.g. the previous step's dipoles.  The
    frames of a trajectory are solved concurrently by frame_workers contexts
    on their own streams (mpol_hip.frames), each frame exactly as the
    one-frame loop solves it; frame_workers=1 runs that loop."""
    x = np.asarray(x, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    eps_p = np.asarray(eps_p, dtype=np.float64).reshape(-1)
    if x.ndim == 2:
        x = x[:, :, None]
    n, nf = x.shape[0], x.shape[2]
    phi = 4.0 * np.pi * n / (3.0 * np.prod(box))
    perp, para, rtable = _ctxmod.default_tables(xi)
    _, _, _, h, P, _ = ewald_parameters(box, xi, errortol)
    offset, offsetxyz = pre_calculations(P, h)
    H0 = np.asarray(H0, dtype=np.float64).reshape(3)
    Hhat = H0 / np.linalg.norm(H0)
    C = np.zeros(nf, dtype=np.complex128)
    p = np.zeros((n, 3, nf), dtype=np.complex128)
    if verbose >= 1:
        print(f"RunCapacitance: N={n} frames={nf} xi={xi} errortol={errortol:g} slab_factor={slab_factor:g}",
              flush=True)
    if p_guess is not None:
        p_guess = np.asarray(p_guess)
        if p_guess.ndim == 2:
            p_guess = p_guess[:, :, None]
        if p_guess.shape[0] != n or p_guess.shape[1] != 3 or p_guess.shape[2] not in (1, nf):
            raise ValueError(f"p_guess must have shape (N,3) or (N,3,F); got {p_guess.shape}")
    def guess_of(f):
        if p_guess is None:
            return maxwell_garnett_guess(eps_p, phi, Hhat)
        return p_guess[:, :, min(f, p_guess.shape[2] - 1)]

    cplx_guess = p_guess is not None and np.iscomplexobj(p_guess) and bool(np.any(np.imag(p_guess) != 0))
    if nf > 1 and frame_workers > 1 and not cplx_guess:
        from mpol_hip import frames as _frames
        rc, _, ngrid, h, P, eta = ewald_parameters(box, xi, errortol)
        if np.any(rc > box / 2.0):
            raise ValueError("Real-space cutoff exceeds half the box size; tolerance unmet.")
        guesses = np.stack([np.real(guess_of(f)) for f in range(nf)], axis=2)
        p, infos = _frames.solve_frames(x, box, eps_p, (rc, ngrid, P, xi, eta), errortol, Hhat, guesses,
                                        slab_factor=slab_factor, precision=precision, workers=frame_workers)
        for f in range(nf):
            C[f] = np.mean(p[:, :, f] @ Hhat)
            if verbose >= 1:
                inf = infos[f]
                print(f"  frame {f + 1}/{nf}: flag={inf['flag']} relres={inf['relres']:.3e} "
                      f"matvecs={inf['matvecs_fp64'] + inf['matvecs_fp32']} C={np.real(C[f]):.6g}", flush=True)
        magnetic_dipole.last_info = infos[-1]
        return C, p
    for f in range(nf):
        C[f], p[:, :, f] = capacitance(x[:, :, f], eps_p, box, guess_of(f), xi, errortol, perp, para, rtable,
                                       offset, offsetxyz, Hhat, slab_factor=slab_factor,
                                       use_scipy_gmres=use_scipy_gmres, verbose=verbose, precision=precision)
    return C, p


# ---------------------------------------------------------------------------
# batched tensor / spectrum (Capacitance_Hetero/Capacitance_Tensor.m,
# CapacitanceTensorSpectrum.m)
# ---------------------------------------------------------------------------
def capacitance_tensor_spectrum(x, box, eps_p, xi, errortol: float = 1e-3, slab_factor: float = 0.0, *,
                                precision: str = "fp64", return_info: bool = False):
    """Capacitance tensors C (3,3,Nk) and dipoles p (N,3,3,Nk) for a
    permittivity spectrum eps_p (N,Nk), all 3*Nk right-hand sides in one
    batched device solve.  C[i,:,k] = mean over particles of the dipoles for a
    unit field along axis i (row convention of Capacitance_Tensor.m:54).
    Each rhs starts from the Maxwell-Garnett guess of its own field; the
    reference warm-starts directions 2 and 3 from the permuted previous
    solution (Capacitance_Tensor.m:57), which changes only the iterate path,
    not the converged solution.  eps_p may be complex (a dielectric function,
    as the MATLAB drivers allow): each rhs is then one complex system."""
    x = np.asarray(x, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    eps = np.asarray(eps_p)
    cplx = np.iscomplexobj(eps) and bool(np.any(np.imag(eps) != 0))
    eps = eps.astype(np.complex128) if cplx else np.real(eps).astype(np.float64)
    if eps.ndim == 1:
        eps = eps[:, None]
    n, nk = eps.shape
    phi = 4.0 * np.pi * n / (3.0 * np.prod(box))
    rc, _, ngrid, h, P, eta = ewald_parameters(box, xi, errortol)
    if np.any(rc > box / 2.0):
        raise ValueError("Real-space cutoff exceeds half the box size; tolerance unmet.")
    perp, para, rtable = _ctxmod.default_tables(xi)
    ctx = _ewald.geometry_context(x, box, ngrid, P, xi, eta, rc, perp, para, rtable)
    # rhs order: k-major, direction-minor
    eps_sets = np.repeat(eps.T, 3, axis=0)  # (3*nk, n)
    ctx.set_permittivity(eps_sets)
    ctx._eps_key = None
    dt = np.complex128 if cplx else np.float64
    b = np.zeros((3 * nk, n, 3), dtype=dt)
    x0 = np.zeros((3 * nk, n, 3), dtype=dt)
    for k in range(nk):
        beta = (eps[:, k] - 1.0) / (eps[:, k] + 2.0)
        mg = 4.0 * np.pi * beta / (1.0 - beta * phi)
        for d in range(3):
            b[3 * k + d, :, d] = 1.0
            x0[3 * k + d, :, d] = mg
    restart = min(3 * n, 10)
    maxit = min(3 * n, 100)
    if cplx:
        # complex permittivities couple the real and imaginary parts: one complex system per rhs
        mixed = str(precision).lower() in ("mixed", "fp32")
        res = ctx.solve(b, x0, errortol, restart=restart, maxit=maxit, mixed=mixed, slab_factor=slab_factor)
        sol = res.x.detach().cpu().numpy()
        info = {"flag": res.flag, "relres": res.max_relres, "matvecs_fp64": res.matvecs_fp64,
                "matvecs_fp32": res.matvecs_fp32, "refinements": res.refinements}
    else:
        sol, info = _solve_on(ctx, b, x0, errortol, restart, maxit, slab_factor, precision)
    p = sol.reshape(nk, 3, n, 3).transpose(2, 3, 1, 0)  # (n, comp, dir, k)
    C = p.mean(axis=0).transpose(1, 0, 2)                # (dir, comp, k)
    return (C, p, info) if return_info else (C, p)
