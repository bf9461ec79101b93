"""Many independent polydisperse configurations, each over a permittivity
spectrum and several field directions, solved concurrently on one GPU.

The reference solves a configuration's spectrum point by point and field by
field (Siz/CapacitanceTensorSpectrum.m:42-56, a parfor over spectrum points
with three serial solves each).  Here a configuration's whole spectrum x
fields is ONE batched device solve (one right-hand side per (spectrum point,
field), each with its own permittivity set: the operator's grids, pair list
and Krylov state are shared by the batch), and configurations are spread over
worker threads, each owning a HIP stream and a solver context, so that solves
of different configurations overlap on the GPU (a context's GMRES waits on
its own stream only).

Per configuration the Ewald splitting is chosen among the compile-time FFT
grid sizes (default 32, 64, 96, 128 nodes per side): for each size G, the
largest xi whose grid by the polydisperse parameter rule
(Siz/ComputeField.m:23-29: rc = sqrt(-ln tol)/xi, qc = 2 xi sqrt(-ln tol),
Ngrid = ceil(1 + L qc/pi), isotropic eta, P = ceil(2/mean(h) - 2 ln tol/pi))
is G, subject to rc <= L/2.  With tune=True each candidate's batched FP32
operator application is timed on the configuration itself and the fastest is
kept; the rule's tolerance, and so the accuracy class, is the same for all
candidates.  Candidates are timed in increasing size and, with early_stop,
the search ends at the first one slower than the best so far (a larger grid
trades real-space pairs for grid work, so the time is usually unimodal in
the size).

solve_configurations(configs, eps, fields) -> BatchResult
    configs: [{"x": (N,3), "a": (N,), "box": (3,)}], eps: (Nk,) permittivity
    ratios (every particle of a configuration at the same ratio), fields:
    (nf,3) unit fields (default the three axes).  Pair couplings are the
    closed-form polydisperse ones (each pair's own radii).
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import FP32

GRID_SIZES = (32, 64, 96, 128)


@dataclass
class BatchResult:
    p: list                    # per configuration: (Nk, nf, N, 3) dipoles
    C: np.ndarray              # (count, Nk, nf, 3) mean dipole per (spectrum point, field)
    xi: np.ndarray             # (count,) chosen Ewald splitting
    ngrid: np.ndarray          # (count,) chosen grid nodes per side
    P: np.ndarray              # (count,) Gaussian support
    converged: np.ndarray      # (count,) right-hand sides that met the tolerance (eps = 1 points not counted)
    matvecs: np.ndarray        # (count, 2) FP64 / FP32 operator applications per batched solve
    seconds: float             # wall time of the solves (set-up and tuning included)
    rhs: int                   # right-hand sides solved in total
    tune_seconds: float = 0.0  # summed over workers
    report: list = field(default_factory=list)


def rule_parameters(box, xi, tol):
    """rc, Ngrid (3,), P, eta (3,) by the polydisperse rule
    (Siz/ComputeField.m:23-29)."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    s = math.sqrt(-math.log(tol))
    rc = s / xi
    ng = np.ceil(1.0 + box * (2.0 * xi * s) / math.pi).astype(np.int64)
    h = box / ng
    P = int(math.ceil(2.0 / float(np.mean(h)) - 2.0 * math.log(tol) / math.pi))
    eta = P * (float(np.mean(h)) * xi) ** 2 / math.pi
    return rc, ng, P, np.full(3, eta)


def grid_candidates(box, tol, sizes=GRID_SIZES):
    """[(xi, ngrid (3,), P, eta (3,), rc)]: per grid size G the largest xi
    whose rule grid on the longest side is G (rc <= L/2, P < min(ngrid))."""
    box = np.asarray(box, dtype=np.float64).reshape(3)
    s = math.sqrt(-math.log(tol))
    out = []
    for G in sizes:
        xi = (G - 1) * math.pi / (2.0 * float(np.max(box)) * s) * (1.0 - 1e-12)
        rc, ng, P, eta = rule_parameters(box, xi, tol)
        if rc > float(np.min(box)) / 2.0 or P >= int(np.min(ng)):
            continue
        out.append((xi, ng, P, eta, rc))
    return out


def _guesses(a, eps, fields, phi):
    """Maxwell-Garnett guesses 4 pi a^3 alpha / (1 - alpha phi) E0 per
    (spectrum point, field) (Siz/CapacitanceTensorSpectrum.m:47-49)."""
    nk, nf, n = eps.size, fields.shape[0], a.size
    g = np.empty((nk * nf, n, 3))
    for k in range(nk):
        e = eps[k]
        alpha = 1.0 if np.isinf(e) else (e - 1.0) / (e + 2.0)
        mg = 4.0 * np.pi * a ** 3 * alpha / (1.0 - alpha * phi)
        for f in range(nf):
            g[k * nf + f] = mg[:, None] * fields[f][None, :]
    return g


class _Worker:
    def __init__(self, device):
        import torch

        from .context import Context
        self.torch = torch
        self.stream = torch.cuda.Stream(device)  # default priority
        with torch.cuda.stream(self.stream):
            self.ctx = Context(device)
        self.tune_s = 0.0

    def _setup(self, x, a, box, cand, sets):
        xi, ng, P, eta, rc = cand
        ctx = self.ctx
        ctx.set_tables(np.zeros(2), np.zeros(2), np.array([0.0, 2.0 * rc]))
        ctx.set_geometry(x, box, ng, P, xi, eta, rc)
        ctx.set_sizes(a)
        ctx.set_permittivity(sets)

    def run(self, cfg, eps, fields, tol, gmres_tol, precision, tune, sizes, repeats, grid=None, tune_only=False,
            early_stop=True):
        torch = self.torch
        x = np.asarray(cfg["x"], dtype=np.float64)
        a = np.asarray(cfg["a"], dtype=np.float64).reshape(-1)
        box = np.asarray(cfg["box"], dtype=np.float64).reshape(3)
        n, nf = x.shape[0], fields.shape[0]
        # no contrast (eps = 1): the dipoles are exactly zero, which the
        # dielectric term 1/(eps - 1) of the operator cannot represent
        live = np.abs(eps - 1.0) > 1e-12
        p_all = np.zeros((eps.size, nf, n, 3))
        eps = eps[live]
        nk = eps.size
        cands = grid_candidates(box, tol, sizes if grid is None else (int(grid),))
        if not cands:
            raise ValueError(f"no grid in {sizes} gives rc <= L/2 for box {box}")
        if nk == 0:
            return p_all, cands[0], None, []
        sets = np.repeat(np.broadcast_to(eps[:, None], (nk, n)), nf, axis=0)  # (nk*nf, N)
        phi = 4.0 * np.pi * float(np.sum(a ** 3)) / (3.0 * float(np.prod(box)))
        g = _guesses(a, eps, fields, phi)
        report = []
        with torch.cuda.stream(self.stream):
            best, current = cands[0], None
            if tune and len(cands) > 1:
                t0 = time.perf_counter()
                gt = torch.as_tensor(g, device=self.ctx.device)
                best_ms = math.inf
                for cand in cands:
                    self._setup(x, a, box, cand, sets)
                    current = cand
                    self.ctx.sizes_matvec(gt, precision=FP32)  # warm-up
                    self.stream.synchronize()
                    t1 = time.perf_counter()
                    for _ in range(repeats):
                        self.ctx.sizes_matvec(gt, precision=FP32)
                    self.stream.synchronize()
                    ms = 1e3 * (time.perf_counter() - t1) / repeats
                    report.append((int(cand[1][0]), ms))
                    if ms < best_ms:
                        best, best_ms = cand, ms
                    elif early_stop:
                        break
                self.tune_s += time.perf_counter() - t0
            if tune_only:
                return None, best, None, report
            if best is not current:
                self._setup(x, a, box, best, sets)
            b = np.repeat(np.asarray(fields, dtype=np.float64)[None], nk, axis=0).reshape(nk * nf, 3)
            bb = np.broadcast_to(b[:, None, :], (nk * nf, n, 3))
            res = self.ctx.sizes_solve(np.ascontiguousarray(bb), g, gmres_tol, restart=min(3 * n, 10),
                                       maxit=min(3 * n, 100), mixed=(precision == "mixed"))
            p_all[live] = res.x.cpu().numpy().reshape(nk, nf, n, 3)
        return p_all, best, res, report

    def close(self):
        self.ctx.close()


def solve_configurations(configs, eps, fields=None, *, tol=1e-3, gmres_tol=None, precision="mixed", workers=8,
                         tune=True, sizes=GRID_SIZES, repeats=2, grids=None, device=None, early_stop=True,
                         _tune_only=False):
    """Solve every configuration over the spectrum eps (Nk,) x fields (nf,3)
    (default the three axes) with the polydisperse operator; see the module
    docstring.  tol is the Ewald rule's tolerance, gmres_tol (default tol) the
    relative residual of the solves; precision "mixed" (FP32 operator inside
    FP64 refinement) or "fp64".  grids (count,), if given, fixes each
    configuration's grid size instead of choosing it; early_stop ends a
    configuration's grid search at the first candidate slower than the best."""
    import torch

    from .context import require_device
    dev = require_device() if device is None else torch.device(device)
    eps = np.asarray(eps, dtype=np.float64).reshape(-1)
    fields = np.eye(3) if fields is None else np.asarray(fields, dtype=np.float64).reshape(-1, 3)
    gmres_tol = tol if gmres_tol is None else gmres_tol
    count = len(configs)
    # largest first, so that the last configurations to start are the cheap ones
    order = sorted(range(count), key=lambda i: -np.asarray(configs[i]["x"]).shape[0])
    work = queue.Queue()
    for i in order:
        work.put(i)
    results = [None] * count
    errors = []
    pool = [_Worker(dev) for _ in range(max(1, min(workers, count)))]

    def loop(w):
        while not errors:
            try:
                i = work.get_nowait()
            except queue.Empty:
                return
            try:
                results[i] = w.run(configs[i], eps, fields, tol, gmres_tol, precision, tune, sizes, repeats,
                                   None if grids is None else grids[i], _tune_only, early_stop)
            except Exception as e:  # surfaced on the calling thread
                errors.append(e)
                return

    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    threads = [threading.Thread(target=loop, args=(w,)) for w in pool]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize(dev)
    seconds = time.perf_counter() - t0
    tune_s = sum(w.tune_s for w in pool)
    for w in pool:
        w.close()
    if errors:
        raise errors[0]
    if _tune_only:
        return np.array([int(r[1][1][0]) for r in results]), [r[3] for r in results], seconds
    nk, nf = eps.size, fields.shape[0]
    C = np.stack([r[0].mean(axis=2) for r in results])  # (count, nk, nf, 3)
    return BatchResult(
        p=[r[0] for r in results], C=C,
        xi=np.array([r[1][0] for r in results]), ngrid=np.array([int(r[1][1][0]) for r in results]),
        P=np.array([r[1][2] for r in results]),
        converged=np.array([r[2].converged if r[2] is not None else 0 for r in results]),
        matvecs=np.array([[r[2].matvecs_fp64, r[2].matvecs_fp32] if r[2] is not None else [0, 0] for r in results]),
        seconds=seconds, rhs=count * nk * nf, tune_seconds=tune_s, report=[r[3] for r in results])


def tune_grids(configs, eps, fields=None, *, tol=1e-3, workers=1, sizes=GRID_SIZES, repeats=2, device=None,
               early_stop=True):
    """Grid size per configuration (fastest batched FP32 operator among the
    candidates timed), without solving: (grids (count,), timings [(G, ms)]
    per configuration, seconds).  One worker by default, so that the timings
    are not shared with other configurations' work.  Pass grids to
    solve_configurations."""
    return solve_configurations(configs, eps, fields, tol=tol, workers=workers, tune=True, sizes=sizes,
                                repeats=repeats, device=device, early_stop=early_stop, _tune_only=True)
