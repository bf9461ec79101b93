"""Adjoint forces and dU/dbeta at full size (N = 1e5, phi = 0.4, 256^3, P = 10)
against central finite differences (h = 1e-4) of the stationary energy, on 8
random particles: max |F_adjoint - F_fd| <= 1e-5 max |F_fd|, and dU/dbeta to
1e-5 relative.  The real-space tables use a 1e-6 grid so that the linear
interpolation's slope jumps (which a 1e-4 step crosses) stay below the bar;
workloads/adjoint_check.py explains the construction."""
import pytest

pytestmark = pytest.mark.gpu


def test_adjoint_forces_at_full_size(gpu):
    from workloads import adjoint_check
    print("case n1e5_phi0.4, 8 particles, h 1e-4", flush=True)
    out, rows = adjoint_check.run(particles=8, h=1e-4, solve_tol=1e-12, seed=11, verbose=False)
    print(out, flush=True)
    assert out["components"] >= 18
    assert out["force_rel_err"] < 1e-5
    assert out["dU_dbeta_rel_err"] < 1e-5
