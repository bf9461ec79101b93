"""Full-size parity properties of the other workloads: the dipole +
quadrupole solve at N = 2e4 (BCT lattice, two permittivities, 160^3, P = 12)
in mixed precision against a tight FP64 solve (1e-6 relative in dipoles and
quadrupoles), and the FP64 solve's own residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_quadrupole_full_size_mixed_vs_fp64(gpu):
    import torch

    import quadrupole as quad
    from workloads.configs import bct_lattice
    print("case quadrupole N=2e4 160^3 P=12", flush=True)
    x, box = bct_lattice()
    n = x.shape[0]
    ng, P, tol = 160, 12, 1e-6
    s = np.sqrt(-np.log(tol))
    xi = (ng - 1) * np.pi / (2.0 * float(np.max(box)) * s)
    rc = s / xi
    ngrid = np.full(3, ng)
    eta = P * (box / ngrid * xi) ** 2 / np.pi
    r = np.concatenate(([0.0], np.arange(0.001, 10.0 + 1e-12, 0.001)))
    ctx = quad._context(x, box, ngrid, P, xi, eta, rc, r, quad._driver_tables(r, xi))
    eps = np.where(np.random.default_rng(17).random(n) < 0.5, 10.0, 0.5)
    ctx.set_permittivity(eps)
    ctx._eps_key = None
    b = np.zeros((1, n, 8))
    b[0, :, 2] = 1.0
    bt = torch.as_tensor(b, device=gpu)
    ref = ctx.quad_solve(bt, bt * 0.0, 1e-11, restart=10, maxit=100)
    assert ref.flag == 0
    mixed = ctx.quad_solve(bt, bt * 0.0, 1e-7, restart=10, maxit=100, mixed=True)
    assert mixed.flag == 0 and mixed.matvecs_fp32 > 0
    a, c = mixed.x.cpu().numpy(), ref.x.cpu().numpy()
    assert np.max(np.abs(a - c)) / np.max(np.abs(c)) < 1e-6
    # the tight solution's true residual in FP64
    res = ctx.quad_matvec(ref.x).cpu().numpy() - b
    assert np.linalg.norm(res) / np.linalg.norm(b) < 1e-10
