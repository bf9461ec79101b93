"""capacitance.gmres_restarted (the reference's in-file restarted GMRES,
"CPU - Ewald test/capacitance.py":60-199, with a caller's matvec) on the GPU.

Against the oracle's restatement of the same algorithm (pinned by the
reference's own golden use_scipy_gmres=False solves in test_oracle_golden):
same iterates to rounding, same flag, same relative residual, on a dense
complex system -- converged, and stopped by maxit through the best-so-far
least-squares path -- with NumPy and with device-tensor matvecs; and on the
dipole operator against the golden solve the reference wrote.
"""
import numpy as np
import pytest

from oracle import capacitance_oracle as co

from conftest import golden

pytestmark = pytest.mark.gpu


def _system(n=60, seed=4):
    rng = np.random.default_rng(seed)
    A = (3.0 + 1.0j) * np.eye(n) + 0.4 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = 0.1 * rng.standard_normal(n)
    return A, b, x0


@pytest.mark.parametrize("tol,restart,maxit", [(1e-11, 10, 30), (1e-14, 6, 2), (1e-8, 60, 1)])
def test_gmres_restarted_matches_oracle(gpu, tol, restart, maxit):
    import torch

    import capacitance
    A, b, x0 = _system()
    xo, flago, relo, _ = co.gmres_restarted(lambda v: A @ v, b, x0, tol, restart, maxit)
    x, flag, rel = capacitance.gmres_restarted(lambda v: A @ v, b, x0, tol, restart, maxit)
    assert flag == flago
    assert np.max(np.abs(x - xo)) <= 1e-10 * np.max(np.abs(xo))
    assert abs(rel - relo) <= 1e-6 * relo + 1e-15
    At = torch.as_tensor(A, device=gpu)
    xd, flagd, reld = capacitance.gmres_restarted(lambda v: At @ v, b, x0, tol, restart, maxit, device_matvec=True)
    assert flagd == flago
    assert np.max(np.abs(xd - xo)) <= 1e-10 * np.max(np.abs(xo))


def test_gmres_restarted_on_the_operator_matches_golden(gpu):
    """The golden tight solve of the reference's own gmres_restarted path
    (magnetic_dipole(use_scipy_gmres=False) at tol 1e-12, N = 64, 21^3)."""
    import capacitance
    import ewald
    g = golden("solve_n64_g21_slab0_Hz.npz")
    op = golden("operator_n64_g21.npz")
    x, box, eps, H = g["x"], g["box"], g["eps"], g["H"]
    xi, P, rc, Ngrid = float(op["xi"]), int(op["P"]), float(op["rc"]), op["ngrid"]
    h = box / Ngrid
    eta = P * (h * xi) ** 2 / np.pi
    r = np.arange(0.001, 10.0 + 1e-12, 0.001)
    t = ewald.real_space_table(r, xi)
    perp, para, rv = t[2], t[3], np.concatenate(([0.0], r))
    n = x.shape[0]

    def mv(v):
        return ewald.magnetic_field(x, v.reshape(n, 3), eps, box, None, None, Ngrid, h, P, xi, eta, rc, None, None,
                                    perp, para, rv, slab_factor=0.0).reshape(-1)

    b = np.tile(H, n)
    p, flag, rel = capacitance.gmres_restarted(mv, b, g["guess"].reshape(-1), 1e-12, min(3 * n, 10),
                                               min(3 * n, 100))
    assert flag == 0 and rel <= 1e-12
    ref = g["p_python"].reshape(-1)
    assert np.max(np.abs(p - ref)) <= 1e-10 * np.max(np.abs(ref))
