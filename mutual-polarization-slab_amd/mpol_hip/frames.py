"""Whole trajectories: the frames of run_capacitance solved concurrently on
one GPU.

The reference solves a trajectory's frames one after the other
("CPU - Ewald test/capacitance.py":495-513: for each frame, capacitance()
with the Maxwell-Garnett guess).  Every frame is its own geometry (cell list,
pair list, tile images), and a solver context holds one geometry, so here the
frames go to a pool of worker threads, each owning a context on its own HIP
stream: while one frame's GMRES waits on its stream, the others' kernels
fill the GPU.  Every frame runs exactly the solve capacitance() runs
(restart min(3N, 10), maxit min(3N, 100), relative tolerance errortol, FP64
Arnoldi; FP32 operator inside FP64 refinement with precision="mixed"), so
its dipoles are the ones the one-frame-at-a-time loop gives, to rounding
(the vector-loop spread used off the matrix-core path accumulates with LDS
atomics, whose order varies from run to run).

solve_frames(x (N,3,F), box, eps (N,), (rc, ngrid, P, xi, eta), errortol,
Hhat, guesses (N,3,F), slab_factor, precision, workers)
-> (p (N,3,F) complex128, infos)
"""
from __future__ import annotations

import queue
import threading

import numpy as np


class _FrameWorker:
    def __init__(self, device):
        import torch

        from .context import Context
        self.torch = torch
        self.stream = torch.cuda.Stream(device)  # default priority
        with torch.cuda.stream(self.stream):
            self.ctx = Context(device)
        self.have_tables = False

    def solve(self, x, box, eps, params, tables, guess, b, tol, slab, mixed):
        rc, ngrid, P, xi, eta = params
        with self.torch.cuda.stream(self.stream):
            ctx = self.ctx
            if not self.have_tables:
                ctx.set_tables(*tables)
                self.have_tables = True
            ctx.set_geometry(x, box, ngrid, P, xi, eta, rc)
            ctx.set_permittivity(eps)
            n = x.shape[0]
            res = ctx.solve(b, guess.reshape(1, n, 3), tol, restart=min(3 * n, 10), maxit=min(3 * n, 100),
                            mixed=mixed, slab_factor=slab)
            p = res.x.cpu().numpy()[0]
        info = {"flag": res.flag, "relres": res.max_relres, "matvecs_fp64": res.matvecs_fp64,
                "matvecs_fp32": res.matvecs_fp32, "refinements": res.refinements}
        return p, info

    def close(self):
        self.ctx.close()


def solve_frames(x, box, eps, params, errortol, Hhat, guesses, slab_factor=1.0, precision="fp64", workers=4,
                 device=None):
    """Dipoles (N,3,F) complex128 of every frame of x (N,3,F) for the applied
    field Hhat, and one info dict per frame.  params = (rc, ngrid (3,), P, xi,
    eta (3,)) as capacitance() derives them; guesses (N,3,F): each frame's
    initial dipoles; errortol: GMRES relative tolerance."""
    import torch

    from .context import default_tables, require_device
    dev = require_device() if device is None else torch.device(device)
    x = np.asarray(x, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    n, nf = x.shape[0], x.shape[2]
    xi = float(params[3])
    tables = default_tables(xi)
    eps = np.real(np.asarray(eps)).astype(np.float64).reshape(-1)
    Hhat = np.asarray(Hhat, dtype=np.float64).reshape(3)
    b = np.ascontiguousarray(np.broadcast_to(Hhat, (1, n, 3)))
    mixed = str(precision).lower() in ("mixed", "fp32")
    guesses = np.real(np.asarray(guesses)).astype(np.float64)
    p = np.zeros((n, 3, nf), dtype=np.complex128)
    infos = [None] * nf
    work = queue.Queue()
    for f in range(nf):
        work.put(f)
    errors = []
    pool = [_FrameWorker(dev) for _ in range(max(1, min(int(workers), nf)))]

    def loop(w):
        while not errors:
            try:
                f = work.get_nowait()
            except queue.Empty:
                return
            try:
                pf, infos[f] = w.solve(np.ascontiguousarray(x[:, :, f]), box, eps, params, tables,
                                       np.ascontiguousarray(guesses[:, :, f]), b, errortol, slab_factor, mixed)
                p[:, :, f] = pf
            except Exception as e:  # surfaced on the calling thread
                errors.append(e)
                return

    torch.cuda.synchronize(dev)
    threads = [threading.Thread(target=loop, args=(w,)) for w in pool]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize(dev)
    for w in pool:
        w.close()
    if errors:
        raise errors[0]
    return p, infos
