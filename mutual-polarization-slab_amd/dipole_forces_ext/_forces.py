"""compute_forces: truncated point-dipole forces and energies on the GPU.

Replaces the compiled compute_forces of
"CPU - HOOMD simulation test/dipole_forces_ext/_dipole_forces_ext.cpp":454-497
(brute force :119-254, cell list :256-438) with the HIP kernel behind
mpol_compute_forces.  Same arguments, defaults, return values and errors:
RuntimeError for bad shapes, box lengths or cutoffs (the reference's
std::runtime_error), ValueError for arrays of the wrong rank.

Host numpy arrays in -> host numpy arrays out (as the reference).  CUDA
tensors in -> CUDA tensors out, with no host round trip.
"""
from __future__ import annotations

import threading

import numpy as np
import torch

from mpol_hip import _lib
from mpol_hip.context import Context, require_device

_CTX: dict = {}
_LOCK = threading.Lock()


def _context(device: torch.device) -> Context:
    key = (device.type, device.index)
    with _LOCK:
        ctx = _CTX.get(key)
        if ctx is None:
            ctx = Context(device)
            _CTX[key] = ctx
        return ctx


def _check_rank(name, a, ndim):
    if a.ndim != ndim:
        raise ValueError(f"array has incorrect number of dimensions: {a.ndim}; expected {ndim} ({name})")


def compute_forces(pos, dip, box, r_cut, prefactor=1.0, periodic_z=True, use_cell_list=True):
    """forces (N, 3) float64 and per-particle energies (N,) float64.

    pos: (N, 3) positions (HOOMD-centred box or any periodic image);
    dip: (N, 3) dipoles; box: (3,) lengths; r_cut < 0.5 min(Lx, Ly)
    (and < 0.5 Lz when periodic_z); energy of each pair is split 1/2, 1/2.
    """
    on_device = isinstance(pos, torch.Tensor) and pos.is_cuda
    if on_device:
        device = pos.device
        pt = pos.to(torch.float64)
        dt = torch.as_tensor(dip, dtype=torch.float64, device=device)
        _check_rank("pos", pt, 2)
        _check_rank("dip", dt, 2)
    else:
        device = require_device()
        pn = np.asarray(pos, dtype=np.float64)
        dn = np.asarray(dip, dtype=np.float64)
        _check_rank("pos", pn, 2)
        _check_rank("dip", dn, 2)
        pt = torch.as_tensor(np.ascontiguousarray(pn), device=device)
        dt = torch.as_tensor(np.ascontiguousarray(dn), device=device)
    if isinstance(box, torch.Tensor):
        box = box.detach().cpu().numpy()
    bn = np.asarray(box, dtype=np.float64)
    _check_rank("box", bn, 1)
    ctx = _context(device)
    try:
        f, pe = ctx.compute_forces(pt, dt, bn, float(r_cut), float(prefactor), bool(periodic_z),
                                   bool(use_cell_list))
    except ValueError as e:  # argument errors of the C ABI: the reference raises RuntimeError
        raise RuntimeError(str(e)) from None
    except _lib.MpolError as e:
        raise RuntimeError(str(e)) from None
    if on_device:
        return f, pe
    return f.cpu().numpy(), pe.cpu().numpy()
