"""dipole_forces_ext.hoomd_force: HOOMD v6 custom forces on the MI355X path.

Same classes, constructor arguments and per-step behaviour as
"CPU - HOOMD simulation test/dipole_forces_ext/hoomd_force.py":
  MutualPolarizationDipoleForceCached (:75-262) -- induced dipoles from the
      GPU GMRES solve (capacitance.run_capacitance), re-solved every
      pol_every_steps and cached by particle tag in between, then the GPU
      truncated dipole forces;
  FixedDipoleForce (:265-301) -- the same force kernel for dipoles mu * Hhat.
Torques and virials are written as zero (the kernel defines neither).

Extra keyword (not in the reference): precision="fp64" | "mixed" for the
dipole solve.  Importing this module without HOOMD raises ImportError.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

try:
    import hoomd
    import hoomd.md
except Exception as e:  # pragma: no cover - exercised only without HOOMD
    raise ImportError("dipole_forces_ext.hoomd_force requires HOOMD (v6.x); import it inside a "
                      "HOOMD-enabled environment") from e

from . import capacitance
from ._forces import compute_forces


# HOOMD's local force arrays: the energy attribute is `potential_energy`
# (`energy` on older builds); torques and virials of these forces are zero.
_ENERGY_ATTRS = ("potential_energy", "energy")
_ZEROED_ATTRS = ("torque", "virial")


def _write_force_energy_arrays(arrays, f: np.ndarray, e: np.ndarray) -> None:
    """Force and per-particle energy into HOOMD's local force arrays."""
    arrays.force[:] = f
    target = next((getattr(arrays, a) for a in _ENERGY_ATTRS if hasattr(arrays, a)), None)
    if target is not None:
        target[:] = e
    for a in _ZEROED_ATTRS:
        if hasattr(arrays, a):
            getattr(arrays, a)[:] = 0.0


def _evaluate(pos, dip, boxL, r_cut, prefactor, periodic_z, use_cell_list):
    f, e = compute_forces(pos, dip, boxL, r_cut, prefactor, periodic_z=periodic_z, use_cell_list=use_cell_list)
    f = np.asarray(f, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64).reshape(-1)
    n = pos.shape[0]
    if f.shape != (n, 3):
        raise RuntimeError(f"Force shape mismatch: got {f.shape}, expected {(n, 3)}")
    if e.shape != (n,):
        raise RuntimeError(f"Energy shape mismatch: got {e.shape}, expected {(n,)}")
    return f, e


class MutualPolarizationDipoleForceCached(hoomd.md.force.Custom):
    """Mutual polarization + truncated dipole forces, with the dipole solve
    cached for pol_every_steps steps (keyed by particle tag, so particle
    reordering between steps is harmless)."""

    def __init__(self, *, eps_p: float, xi: float, errtol: float, slab_factor: float, Hhat, r_cut: float,
                 prefactor: float, field_strength: float, pol_every_steps: int = 1, use_scipy_gmres: bool = True,
                 periodic_z: bool = True, use_cell_list: bool = True, precision: str = "fp64",
                 warm_start: bool = False):
        super().__init__()
        self.eps_p = float(eps_p)
        self.xi = float(xi)
        self.errtol = float(errtol)
        self.slab_factor = float(slab_factor)
        self.Hhat = np.asarray(Hhat, dtype=np.float64).reshape(3)
        self.r_cut = float(r_cut)
        self.prefactor = float(prefactor)
        self.field_strength = float(field_strength)
        self.use_scipy_gmres = bool(use_scipy_gmres)
        self.periodic_z = bool(periodic_z)
        self.use_cell_list = bool(use_cell_list)
        self.pol_every_steps = max(1, int(pol_every_steps))
        self.precision = precision
        # not in the reference: start each solve from the cached dipoles
        # (hoomd_force.py:183-188 there caches them but never reuses them)
        self.warm_start = bool(warm_start)
        self.last_solve_info: Optional[dict] = None
        self._solved: Optional[np.ndarray] = None
        self.calls = 0
        self.pol_solves = 0
        self.last_dipoles: Optional[np.ndarray] = None
        self.last_forces: Optional[np.ndarray] = None
        self.last_energy: Optional[np.ndarray] = None
        self._by_tag: Optional[np.ndarray] = None
        self._last_pol_step: Optional[int] = None

    def _grow_cache(self, tags: np.ndarray) -> None:
        need = (int(tags.max()) + 1) if tags.size else 0
        have = 0 if self._by_tag is None else self._by_tag.shape[0]
        if self._by_tag is None or need > have:
            grown = np.zeros((max(need, have), 3), dtype=np.float64)
            solved = np.zeros(max(need, have), dtype=bool)
            if self._by_tag is not None:
                grown[:have] = self._by_tag
                solved[:have] = self._solved
            self._by_tag = grown
            self._solved = solved

    def _solve_dipoles(self, pos: np.ndarray, boxL: np.ndarray, tags: np.ndarray) -> None:
        eps = np.full(pos.shape[0], self.eps_p, dtype=np.float64)
        guess = None
        if (self.warm_start and self.field_strength != 0.0 and self._solved is not None and tags.size
                and bool(self._solved[tags].all())):
            guess = self._by_tag[tags] / self.field_strength
        _, p = capacitance.run_capacitance(pos, boxL, eps, self.xi, self.Hhat.copy(), errortol=self.errtol,
                                           slab_factor=self.slab_factor, verbose=0,
                                           use_scipy_gmres=self.use_scipy_gmres, precision=self.precision,
                                           p_guess=guess)
        self.last_solve_info = dict(capacitance.magnetic_dipole.last_info or {})
        p = np.asarray(p)
        if p.ndim == 3:
            p0 = np.real(p[:, :, 0])
        elif p.ndim == 2:
            p0 = np.real(p)
        else:
            raise RuntimeError(f"Unexpected p shape from run_capacitance: {p.shape}")
        self._grow_cache(tags)
        self._by_tag[tags] = p0.astype(np.float64) * self.field_strength
        self._solved[tags] = True
        self.pol_solves += 1

    def set_forces(self, timestep):
        self.calls += 1
        with self._state.cpu_local_snapshot as snap:
            pos = np.array(snap.particles.position, dtype=np.float64)
            tags = np.array(snap.particles.tag, dtype=np.int64)
        boxL = np.array(self._simulation.state.box.L, dtype=np.float64).reshape(3)
        if pos.shape[0] == 0:
            return
        self._grow_cache(tags)
        if self._last_pol_step is None or int(timestep) - self._last_pol_step >= self.pol_every_steps:
            self._solve_dipoles(pos, boxL, tags)
            self._last_pol_step = int(timestep)
        dip = self._by_tag[tags]
        f, e = _evaluate(pos, dip, boxL, self.r_cut, self.prefactor, self.periodic_z, self.use_cell_list)
        if not (np.isfinite(f).all() and np.isfinite(e).all() and np.isfinite(dip).all()):
            raise RuntimeError("Non-finite values in dipoles/forces/energies")
        with self.cpu_local_force_arrays as arrays:
            _write_force_energy_arrays(arrays, f, e)
        self.last_dipoles = dip.copy()
        self.last_forces = f.copy()
        self.last_energy = e.copy()


class FixedDipoleForce(hoomd.md.force.Custom):
    """Truncated dipole forces for identical fixed dipoles mu * Hhat."""

    def __init__(self, *, Hhat, mu: float, r_cut: float, prefactor: float, periodic_z: bool = True,
                 use_cell_list: bool = True):
        super().__init__()
        self.Hhat = np.asarray(Hhat, dtype=np.float64).reshape(3)
        self.mu = float(mu)
        self.r_cut = float(r_cut)
        self.prefactor = float(prefactor)
        self.periodic_z = bool(periodic_z)
        self.use_cell_list = bool(use_cell_list)
        self.calls = 0
        self.last_forces: Optional[np.ndarray] = None
        self.last_energy: Optional[np.ndarray] = None

    def set_forces(self, timestep):
        self.calls += 1
        with self._state.cpu_local_snapshot as snap:
            pos = np.array(snap.particles.position, dtype=np.float64)
        boxL = np.array(self._simulation.state.box.L, dtype=np.float64).reshape(3)
        if pos.shape[0] == 0:
            return
        dip = np.broadcast_to(self.mu * self.Hhat, pos.shape).copy()
        f, e = _evaluate(pos, dip, boxL, self.r_cut, self.prefactor, self.periodic_z, self.use_cell_list)
        with self.cpu_local_force_arrays as arrays:
            _write_force_energy_arrays(arrays, f, e)
        self.last_forces = f.copy()
        self.last_energy = e.copy()
