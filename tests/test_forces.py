"""Truncated fixed-dipole forces (dipole_forces_ext.compute_forces).

CPU part: the oracle against closed forms for a dipole pair, the reference's
own kernel tests T1-T5 ("CPU - HOOMD simulation test/force_energy_diagnostics.py":830-867,
restated without HOOMD), Newton's third law, energy/force consistency by
central differences, and the reference's C++ kernel compiled from its source
(oracle/_ref, when it was built).

GPU part: the HIP kernel through the package entry point against the oracle
and the compiled reference kernel (FP64: agreement to 1e-12 relative, the
order of the pair sums being the only difference), T1-T5 on the product,
the reference's argument errors, device-tensor inputs, and the HOOMD wrappers
driven by a minimal stand-in for hoomd.md.force.Custom.
"""
import contextlib
import importlib
import os
import sys
import types

import numpy as np
import pytest

from oracle import forces_oracle as fo

from conftest import REPO

BOX80 = np.array([80.0, 80.0, 80.0])
R_CUT = 25.0  # the diagnostics' r_cut in a box of 80


def _pair(r, mode):
    if mode == "head":
        return np.array([[0.0, 0.0, -0.5 * r], [0.0, 0.0, 0.5 * r]]), 2
    return np.array([[-0.5 * r, 0.0, 0.0], [0.5 * r, 0.0, 0.0]]), 0


def _pair_eval(cf, r, mode, mu=1.0, pref=1.0, **kw):
    pos, comp = _pair(r, mode)
    dip = np.tile([0.0, 0.0, mu], (2, 1))
    f, e = cf(pos, dip, BOX80, R_CUT, pref, **kw)
    return float(np.asarray(f)[1, comp]), float(np.sum(e))


def _random_config(n, density, seed, centred=True):
    L = (n / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * L
    if centred:
        pos -= 0.5 * L
    dip = rng.standard_normal((n, 3))
    return pos, dip, np.array([L, L, L])


def _ref_module():
    """The reference's compiled kernel from oracle/_ref, or None if not built."""
    path = os.path.join(REPO, "oracle", "_ref")
    if not os.path.isdir(path):
        return None
    sys.path.insert(0, path)
    try:
        return importlib.import_module("_dipole_forces_ext")
    except ImportError:
        return None
    finally:
        sys.path.remove(path)


def _check_t1_t5(cf):
    Fh, Uh = _pair_eval(cf, 2.0, "head")
    Fs, Us = _pair_eval(cf, 2.0, "side")
    assert Fh < 0.0  # T1: head-to-tail dipoles attract
    assert Fs > 0.0  # T2: side-by-side dipoles repel
    F3, _ = _pair_eval(cf, 2.0, "head", pref=3.0)
    assert abs(F3 / Fh - 3.0) < 1e-12  # T3
    F2, _ = _pair_eval(cf, 2.0, "head", mu=2.0)
    assert abs(F2 / Fh - 4.0) < 1e-12  # T4
    dr = 1e-4
    _, Up = _pair_eval(cf, 2.0 + dr, "head")
    _, Um = _pair_eval(cf, 2.0 - dr, "head")
    ref = -(Up - Um) / (2.0 * dr)
    assert abs((Fh - ref) / ref) < 2e-3  # T5 (the reference's bound)
    # the same central difference is far tighter than that bound for this kernel
    assert abs((Fh - ref) / ref) < 1e-7
    return Fh, Uh, Fs, Us


# ----------------------------------------------------------------- CPU / oracle


def test_oracle_pair_closed_forms():
    """Coaxial dipoles: F_z(top) = -6 pref mu^2 / r^4, U = -2 pref mu^2 / r^3;
    side by side: F_x(right) = 3 pref mu^2 / r^4, U = pref mu^2 / r^3."""
    for r in (2.0, 3.7, 11.0):
        for mu, pref in ((1.0, 1.0), (1.7, 0.3)):
            Fh, Uh = _pair_eval(fo.compute_forces, r, "head", mu, pref)
            Fs, Us = _pair_eval(fo.compute_forces, r, "side", mu, pref)
            assert Fh == pytest.approx(-6.0 * pref * mu * mu / r ** 4, rel=1e-14)
            assert Uh == pytest.approx(-2.0 * pref * mu * mu / r ** 3, rel=1e-14)
            assert Fs == pytest.approx(3.0 * pref * mu * mu / r ** 4, rel=1e-14)
            assert Us == pytest.approx(pref * mu * mu / r ** 3, rel=1e-14)


def test_oracle_t1_t5():
    _check_t1_t5(fo.compute_forces)


def test_oracle_pair_across_boundary_and_cutoff():
    """Minimum image in x, y always and in z only when periodic_z; pairs past
    r_cut contribute nothing."""
    box = np.array([10.0, 10.0, 10.0])
    dip = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    across_z = np.array([[0.0, 0.0, -4.5], [0.0, 0.0, 4.5]])  # image distance 1
    f, e = fo.compute_forces(across_z, dip, box, 3.0, periodic_z=True)
    assert np.sum(e) == pytest.approx(-2.0, rel=1e-14)
    f, e = fo.compute_forces(across_z, dip, box, 3.0, periodic_z=False)
    assert np.all(f == 0.0) and np.all(e == 0.0)
    across_x = np.array([[-4.5, 0.0, 0.0], [4.5, 0.0, 0.0]])
    f, e = fo.compute_forces(across_x, dip, box, 3.0, periodic_z=False)
    assert np.sum(e) == pytest.approx(1.0, rel=1e-14)
    far = np.array([[0.0, 0.0, 0.0], [3.01, 0.0, 0.0]])
    f, e = fo.compute_forces(far, dip, box, 3.0)
    assert np.all(f == 0.0) and np.all(e == 0.0)


def test_oracle_newton_and_energy_gradient():
    """Sum of forces vanishes; F_i = -dU/dx_i for pairs away from the cutoff."""
    pos, dip, box = _random_config(60, 0.05, 3)
    rc = 4.0
    f, e = fo.compute_forces(pos, dip, box, rc)
    assert np.max(np.abs(f.sum(axis=0))) < 1e-12 * np.max(np.abs(f))
    d = pos[:, None, :] - pos[None, :, :]
    d -= box * np.rint(d / box)
    r = np.sqrt(np.sum(d * d, axis=2))
    near_cut = np.abs(r - rc) < 1e-3
    for i in range(0, 60, 7):
        if np.any(near_cut[i]):
            continue
        for c in range(3):
            h = 1e-5
            p1 = pos.copy()
            p1[i, c] += h
            p2 = pos.copy()
            p2[i, c] -= h
            U1 = np.sum(fo.compute_forces(p1, dip, box, rc)[1])
            U2 = np.sum(fo.compute_forces(p2, dip, box, rc)[1])
            g = -(U1 - U2) / (2 * h)
            assert abs(g - f[i, c]) <= 1e-6 * max(1.0, abs(f[i, c]))


def test_oracle_errors():
    pos = np.zeros((2, 3))
    with pytest.raises(RuntimeError, match="Invalid shapes"):
        fo.compute_forces(np.zeros((2, 4)), pos, BOX80, 1.0)
    with pytest.raises(RuntimeError, match="Box lengths"):
        fo.compute_forces(pos, pos, [1.0, -1.0, 1.0], 0.1)
    with pytest.raises(RuntimeError, match="r_cut must be"):
        fo.compute_forces(pos, pos, BOX80, 0.0)
    with pytest.raises(RuntimeError, match="periodic x,y"):
        fo.compute_forces(pos, pos, BOX80, 40.0)
    with pytest.raises(RuntimeError, match="periodic z"):
        fo.compute_forces(pos, pos, [80.0, 80.0, 20.0], 12.0, periodic_z=True)
    fo.compute_forces(pos + [[0, 0, 0], [0, 0, 1]], pos, [80.0, 80.0, 20.0], 12.0, periodic_z=False)


def test_oracle_vs_reference_kernel():
    """The restatement against the reference's own C++ kernel (compiled from
    its source into oracle/_ref), brute force and cell list, both z modes."""
    ref = _ref_module()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    assert ref.__version__ == "0.5.1"
    pos, dip, box = _random_config(400, 0.08, 11)
    for pz in (True, False):
        for cells in (True, False):
            f0, e0 = fo.compute_forces(pos, dip, box, 3.3, 0.7, periodic_z=pz)
            f1, e1 = ref.compute_forces(pos, dip, box, 3.3, 0.7, periodic_z=pz, use_cell_list=cells)
            assert np.max(np.abs(f1 - f0)) <= 1e-13 * np.max(np.abs(f0))
            assert np.max(np.abs(e1 - e0)) <= 1e-13 * np.max(np.abs(e0))


# ----------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def dfe(gpu):
    import dipole_forces_ext
    return dipole_forces_ext


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.gpu
def test_gpu_forces_version_and_pair(dfe):
    assert dfe.__version__ == "0.5.1"
    Fh, Uh, Fs, Us = _check_t1_t5(dfe.compute_forces)
    assert Fh == pytest.approx(-6.0 / 16.0, rel=1e-14)
    assert Uh == pytest.approx(-2.0 / 8.0, rel=1e-14)
    assert Fs == pytest.approx(3.0 / 16.0, rel=1e-14)
    Fh2, _ = _pair_eval(dfe.compute_forces, 2.0, "head", use_cell_list=False)
    assert Fh2 == Fh


@pytest.mark.gpu
@pytest.mark.parametrize("periodic_z", [True, False])
@pytest.mark.parametrize("cells", [True, False])
def test_gpu_forces_vs_oracle(dfe, periodic_z, cells):
    pos, dip, box = _random_config(700, 0.1, 5)
    f0, e0 = fo.compute_forces(pos, dip, box, 3.1, 1.3, periodic_z=periodic_z)
    f1, e1 = dfe.compute_forces(pos, dip, box, 3.1, 1.3, periodic_z=periodic_z, use_cell_list=cells)
    assert isinstance(f1, np.ndarray) and f1.shape == (700, 3) and e1.shape == (700,)
    assert f1.dtype == np.float64 and e1.dtype == np.float64
    assert _rel(f1, f0) < 1e-12
    assert _rel(e1, e0) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["tiny_box", "slab", "uncentred"])
def test_gpu_forces_cell_edge_cases(dfe, shape):
    """Few cells per axis (1-3, where neighbour cells repeat), a tall
    non-cubic box with open z, and coordinates outside the centred box."""
    rng = np.random.default_rng(17)
    if shape == "tiny_box":
        box = np.array([6.1, 6.3, 9.2])
        pos = (rng.random((50, 3)) - 0.5) * box
        rc = 2.9
    elif shape == "slab":
        box = np.array([20.0, 24.0, 4.0])
        pos = (rng.random((300, 3)) - 0.5) * box
        rc = 5.0
    else:
        box = np.array([15.0, 15.0, 15.0])
        pos = rng.random((300, 3)) * box * 3.0 - box
        rc = 4.0
    dip = rng.standard_normal(pos.shape)
    for pz in ((False,) if shape == "slab" else (True, False)):
        if shape == "uncentred" and not pz:
            continue  # open z with particles outside the box: cell list and all-pairs differ in the reference too
        f0, e0 = fo.compute_forces(pos, dip, box, rc, periodic_z=pz)
        for cells in (True, False):
            f1, e1 = dfe.compute_forces(pos, dip, box, rc, periodic_z=pz, use_cell_list=cells)
            assert _rel(f1, f0) < 1e-12 and _rel(e1, e0) < 1e-12


@pytest.mark.gpu
def test_gpu_forces_vs_reference_kernel_large(dfe):
    """N = 40000 against the reference's own compiled kernel (cell list)."""
    ref = _ref_module()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    pos, dip, box = _random_config(40000, 0.3, 23)
    for pz in (True, False):
        f0, e0 = ref.compute_forces(pos, dip, box, 3.5, 1.0, periodic_z=pz, use_cell_list=True)
        f1, e1 = dfe.compute_forces(pos, dip, box, 3.5, 1.0, periodic_z=pz)
        assert _rel(f1, f0) < 1e-12 and _rel(e1, e0) < 1e-12


@pytest.mark.gpu
def test_gpu_forces_device_tensors_and_empty(dfe, gpu):
    import torch
    pos, dip, box = _random_config(200, 0.1, 8)
    f0, e0 = fo.compute_forces(pos, dip, box, 2.5)
    pt = torch.as_tensor(pos, device=gpu)
    dt = torch.as_tensor(dip, device=gpu)
    f1, e1 = dfe.compute_forces(pt, dt, box, 2.5)
    assert isinstance(f1, torch.Tensor) and f1.is_cuda
    assert _rel(f1.cpu().numpy(), f0) < 1e-12 and _rel(e1.cpu().numpy(), e0) < 1e-12
    f, e = dfe.compute_forces(np.zeros((0, 3)), np.zeros((0, 3)), box, 2.5)
    assert f.shape == (0, 3) and e.shape == (0,)


@pytest.mark.gpu
def test_gpu_forces_errors(dfe):
    pos = np.zeros((2, 3))
    with pytest.raises(RuntimeError, match="Invalid shapes"):
        dfe.compute_forces(np.zeros((2, 4)), np.zeros((2, 4)), BOX80, 1.0)
    with pytest.raises(RuntimeError, match="Invalid shapes"):
        dfe.compute_forces(pos, np.zeros((3, 3)), BOX80, 1.0)
    with pytest.raises(RuntimeError, match="Box lengths must be positive"):
        dfe.compute_forces(pos, pos, [1.0, -1.0, 1.0], 0.1)
    with pytest.raises(RuntimeError, match="r_cut must be > 0"):
        dfe.compute_forces(pos, pos, BOX80, 0.0)
    with pytest.raises(RuntimeError, match=r"Require r_cut < 0.5\*min\(Lx,Ly\)"):
        dfe.compute_forces(pos, pos, BOX80, 40.0)
    with pytest.raises(RuntimeError, match=r"Require r_cut < 0.5\*Lz"):
        dfe.compute_forces(pos, pos, [80.0, 80.0, 20.0], 12.0, periodic_z=True)
    with pytest.raises(ValueError):
        dfe.compute_forces(np.zeros(6), np.zeros(6), BOX80, 1.0)


# ----------------------------------------------------------------- HOOMD wrappers


class _FakeSystem:
    """The parts of a HOOMD v6 simulation that hoomd.md.force.Custom exposes
    to the wrappers: particle positions and tags (local snapshot), the box
    lengths, and the force arrays the wrapper writes."""

    def __init__(self, pos, tags, L):
        self.pos, self.tags, self.L = pos, tags, L
        n = pos.shape[0]
        self.arrays = types.SimpleNamespace(force=np.zeros((n, 3)), potential_energy=np.zeros(n),
                                            torque=np.ones((n, 3)), virial=np.ones((n, 6)))

    @property
    def cpu_local_snapshot(self):
        @contextlib.contextmanager
        def cm():
            yield types.SimpleNamespace(particles=types.SimpleNamespace(position=self.pos.copy(),
                                                                         tag=self.tags.copy()))
        return cm()


def _hoomd_force_module(monkeypatch, system):
    class Custom:
        def __init__(self):
            self._state = system
            self._simulation = types.SimpleNamespace(state=types.SimpleNamespace(
                box=types.SimpleNamespace(L=system.L)))

        @property
        def cpu_local_force_arrays(self):
            @contextlib.contextmanager
            def cm():
                yield system.arrays
            return cm()

    hoomd = types.ModuleType("hoomd")
    hoomd.md = types.ModuleType("hoomd.md")
    hoomd.md.force = types.SimpleNamespace(Custom=Custom)
    monkeypatch.setitem(sys.modules, "hoomd", hoomd)
    monkeypatch.setitem(sys.modules, "hoomd.md", hoomd.md)
    monkeypatch.delitem(sys.modules, "dipole_forces_ext.hoomd_force", raising=False)
    return importlib.import_module("dipole_forces_ext.hoomd_force")


@pytest.mark.gpu
def test_gpu_hoomd_fixed_dipole_force(dfe, monkeypatch):
    pos, _, box = _random_config(150, 0.1, 31)
    system = _FakeSystem(pos, np.arange(150), box)
    hf = _hoomd_force_module(monkeypatch, system)
    force = hf.FixedDipoleForce(Hhat=[0.0, 0.0, 1.0], mu=1.5, r_cut=3.0, prefactor=0.25)
    force.set_forces(0)
    dip = np.tile([0.0, 0.0, 1.5], (150, 1))
    f0, e0 = fo.compute_forces(pos, dip, box, 3.0, 0.25)
    assert _rel(system.arrays.force, f0) < 1e-12
    assert _rel(system.arrays.potential_energy, e0) < 1e-12
    assert np.all(system.arrays.torque == 0.0) and np.all(system.arrays.virial == 0.0)
    assert force.calls == 1


@pytest.mark.gpu
def test_gpu_hoomd_mutual_polarization_cached(dfe, monkeypatch):
    """Induced dipoles from the GPU solve, cached by tag: a permuted particle
    order between steps gives the permuted forces, and the solve is repeated
    only every pol_every_steps steps."""
    from helpers import rsa
    from oracle import capacitance_oracle as co

    x, box = rsa(40, 0.2, 9)
    x = x - 0.5 * box  # HOOMD-centred coordinates
    tags = np.arange(40)
    system = _FakeSystem(x, tags, box)
    hf = _hoomd_force_module(monkeypatch, system)
    force = hf.MutualPolarizationDipoleForceCached(
        eps_p=10.0, xi=0.9, errtol=1e-6, slab_factor=0.0, Hhat=[0.0, 0.0, 1.0], r_cut=3.5, prefactor=1.0,
        field_strength=2.0, pol_every_steps=3)
    force.set_forces(0)
    assert force.pol_solves == 1
    _, p_ref = co.run_capacitance(x, box, np.full(40, 10.0), 0.9, np.array([0.0, 0.0, 1.0]), errortol=1e-6,
                                  slab_factor=0.0)
    dip_ref = 2.0 * np.real(np.asarray(p_ref)[:, :, 0])
    assert _rel(force.last_dipoles, dip_ref) < 1e-6
    f0, e0 = fo.compute_forces(x, force.last_dipoles, box, 3.5, 1.0)
    assert _rel(system.arrays.force, f0) < 1e-12
    assert _rel(system.arrays.potential_energy, e0) < 1e-12
    # reorder particles (same configuration): cached dipoles follow the tags
    perm = np.random.default_rng(2).permutation(40)
    system.pos, system.tags = x[perm], tags[perm]
    f_before = system.arrays.force.copy()
    force.set_forces(1)
    assert force.pol_solves == 1
    assert _rel(system.arrays.force, f_before[perm]) < 1e-12
    force.set_forces(3)
    assert force.pol_solves == 2


@pytest.mark.gpu
def test_gpu_hoomd_warm_start_reuses_cached_dipoles(dfe, monkeypatch):
    """warm_start=True (not in the reference) starts each solve from the
    cached dipoles: after a small move it needs fewer operator applications
    than a cold start and gives the same dipoles to the solve tolerance."""
    from helpers import rsa

    x, box = rsa(60, 0.2, 13)
    tags = np.arange(60)
    kw = dict(eps_p=10.0, xi=0.9, errtol=1e-8, slab_factor=0.0, Hhat=[0.0, 0.0, 1.0], r_cut=3.5,
              prefactor=1.0, field_strength=2.0, pol_every_steps=1)
    moved = x + 0.01 * np.random.default_rng(4).standard_normal(x.shape)
    counts, dips = {}, {}
    for warm in (False, True):
        system = _FakeSystem(x.copy(), tags, box)
        hf = _hoomd_force_module(monkeypatch, system)
        force = hf.MutualPolarizationDipoleForceCached(warm_start=warm, **kw)
        force.set_forces(0)
        system.pos = moved.copy()
        force.set_forces(1)
        assert force.pol_solves == 2
        info = force.last_solve_info
        counts[warm] = info["matvecs_fp64"] + info["matvecs_fp32"]
        dips[warm] = force.last_dipoles.copy()
    assert counts[True] < counts[False]
    assert _rel(dips[True], dips[False]) < 1e-6
