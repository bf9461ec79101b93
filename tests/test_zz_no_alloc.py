"""The allocation-free truncated-force entry point (mpol_compute_forces_noalloc)
against mpol_compute_forces, through the C ABI.

Both run the same kernels on the same cell rule, so the outputs must be equal
bit for bit.  Two calls in a row share the device scratch; the second has
fewer particles and fewer cells, so stale counts, starts or members from the
first call would show.  Sizes beyond the scratch are rejected before any
launch.
"""
import ctypes

import numpy as np
import pytest

from mpol_hip import _lib


def _config(n, density, seed):
    L = (n / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * L - 0.5 * L
    dip = rng.standard_normal((n, 3))
    return pos, dip, np.array([L, L, 1.1 * L])


def _call(fn, ctx, pos_t, dip_t, box, r_cut, pref, periodic_z, cells):
    import torch
    n = pos_t.shape[0]
    f = torch.full((n, 3), np.nan, dtype=torch.float64, device=pos_t.device)
    pe = torch.full((n,), np.nan, dtype=torch.float64, device=pos_t.device)
    box = np.ascontiguousarray(box, dtype=np.float64)
    ctx._sync_stream()
    _lib.check(fn(ctx._ctx, ctypes.c_void_p(pos_t.data_ptr()), ctypes.c_void_p(dip_t.data_ptr()), n, _lib.dptr(box),
                  float(r_cut), float(pref), periodic_z, cells, ctypes.c_void_p(f.data_ptr()),
                  ctypes.c_void_p(pe.data_ptr())))
    torch.cuda.synchronize()
    return f.cpu().numpy(), pe.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_forces_match_allocating_entry_point(gpu):
    import torch

    from mpol_hip import Context
    lib = _lib.load()
    ctx = Context(gpu)
    cases = [  # (n, density, seed, r_cut factor of L, prefactor, periodic_z, use_cell_list)
        (6000, 0.3, 11, 0.2, 1.0, 1, 1),      # many cells
        (700, 0.05, 12, 0.45, 0.25, 0, 1),    # fewer particles and cells than before, open z
        (1500, 0.3, 13, 0.12, 1.0, 1, 1),     # back to more cells (stale starts/counts would show)
        (300, 0.3, 14, 0.3, 2.0, 1, 0),       # all-pairs loop (no scratch)
    ]
    for n, dens, seed, fr, pref, pz, cells in cases:
        pos, dip, box = _config(n, dens, seed)
        r_cut = fr * min(box[0], box[1])
        pt = torch.as_tensor(pos, device=gpu)
        dt = torch.as_tensor(dip, device=gpu)
        f0, e0 = _call(lib.mpol_compute_forces, ctx, pt, dt, box, r_cut, pref, pz, cells)
        f1, e1 = _call(lib.mpol_compute_forces_noalloc, ctx, pt, dt, box, r_cut, pref, pz, cells)
        assert np.all(np.isfinite(f0)) and np.all(np.isfinite(e0))
        assert np.array_equal(f0, f1), (n, np.max(np.abs(f0 - f1)))
        assert np.array_equal(e0, e1), (n, np.max(np.abs(e0 - e1)))
        # a second call in a row on the same scratch gives the same result
        f2, e2 = _call(lib.mpol_compute_forces_noalloc, ctx, pt, dt, box, r_cut, pref, pz, cells)
        assert np.array_equal(f1, f2) and np.array_equal(e1, e2)
    ctx.close()


@pytest.mark.gpu
def test_noalloc_forces_reject_sizes_beyond_scratch(gpu):
    from mpol_hip import Context
    lib = _lib.load()
    ctx = Context(gpu)
    box = np.ascontiguousarray([100.0, 100.0, 100.0])
    null = ctypes.c_void_p(0)
    st = lib.mpol_compute_forces_noalloc(ctx._ctx, null, null, (1 << 18) + 1, _lib.dptr(box), 3.0, 1.0, 1, 1,
                                         null, null)
    assert st == _lib.MPOL_EINVAL
    assert b"exceeds its maximum" in lib.mpol_last_error()
    ctx.close()
