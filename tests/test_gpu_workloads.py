"""The benchmark's other workloads at their real sizes, on the GPU.

Three-field tensor (N = 1e4 unit spheres, phi = 0.3, beta = 0.8, 128^3,
P = 10, fields along x, y, z batched): one FP64 operator application against
the oracle; the tight FP64 solve's true residual; mixed and plain FP64 solves
at the tolerances the benchmark quotes within 1e-6 of it; the capacitance
tensor symmetric (the operator is).

Heterogeneous set (256 independent polydisperse configurations, N
log-uniform in [500, 8000], 16-point permittivity spectrum x 3 fields): one
configuration's FP64 operator and solve against the polydisperse oracle;
batched mixed solves at the benchmark's GMRES tolerance within 1e-6 of
per-configuration tight FP64 solves on a sample; and the whole set through
properties that hold for every configuration (convergence, zero dipoles at
eps = 1, a symmetric capacitance tensor, diagonal entries increasing with the
permittivity).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# GMRES tolerance of the batched mixed solves of the heterogeneous set: the
# loosest one whose dipoles are within 1e-6 of tight FP64 solves
HET_GMRES_TOL = 1e-8


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_three_field_tensor_n1e4(gpu):
    import torch

    import workloads
    from mpol_hip import FP64, Context, default_tables
    from oracle import ewald_oracle as eo
    print("case tensor N=1e4 phi=0.3 128^3 P=10", flush=True)
    cfg = workloads.configuration("n1e4_phi0.3")
    n = cfg["x"].shape[0]
    perp, para, rv = default_tables(cfg["xi"])
    ctx = Context(gpu)
    ctx.set_tables(perp, para, rv)
    ctx.set_geometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["eta"], cfg["rc"])
    ctx.set_permittivity(cfg["eps"])
    # one FP64 application against the oracle (real input: the real part)
    m = np.random.default_rng(5).standard_normal((n, 3))
    out = ctx.matvec(m[None], precision=FP64, slab_factor=0.0).cpu().numpy()[0]
    geo = eo.OperatorGeometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["rc"], perp, para, rv)
    ref_mv = np.real(eo.field(geo, m, cfg["eps"], 0.0))
    assert _rel(out, ref_mv) < 1e-12
    del geo
    phi = 4.0 * np.pi * n / (3.0 * np.prod(cfg["box"]))
    mg = 4.0 * np.pi * cfg["beta"] / (1.0 - cfg["beta"] * phi)
    b = torch.as_tensor(np.stack([np.tile(e, n).reshape(n, 3) for e in np.eye(3)]), device=gpu)
    x0 = mg * b
    ref = ctx.solve(b, x0, 1e-11, restart=10, maxit=100, slab_factor=0.0)
    assert ref.flag == 0
    res = ctx.matvec(ref.x, precision=FP64, slab_factor=0.0) - b
    for k in range(3):
        assert float(torch.linalg.vector_norm(res[k]) / torch.linalg.vector_norm(b[k])) < 1e-10
    p_ref = ref.x.cpu().numpy()
    for mixed, tol in ((True, 2e-7), (False, 1e-7)):
        r = ctx.solve(b, x0, tol, restart=10, maxit=100, mixed=mixed, slab_factor=0.0)
        assert r.flag == 0 and r.max_relres <= tol
        assert _rel(r.x.cpu().numpy(), p_ref) < 1e-6, (mixed, _rel(r.x.cpu().numpy(), p_ref))
        if mixed:
            assert r.matvecs_fp32 > 0
    C = p_ref.mean(axis=1)  # C[field, component]
    assert np.max(np.abs(C - C.T)) < 1e-8 * np.max(np.abs(C))
    assert np.all(np.diag(C) > 0.0)
    ctx.close()


@pytest.fixture(scope="module")
def het_set():
    import workloads
    return workloads.heterogeneous_set(seed=7)


def _sample(configs):
    ns = np.array([c["x"].shape[0] for c in configs])
    order = np.argsort(ns)
    return [int(order[0]), int(order[len(order) // 2]), int(order[-1])]


def test_heterogeneous_configuration_against_oracle(gpu, het_set):
    """The smallest configuration (N = 505): the batch's FP64 operator (closed-
    form pair couplings) and its solve at the hardest spectrum point
    (eps = 100, field along z) against the polydisperse oracle."""
    from mpol_hip import FP64, Context, batch
    from oracle import sizes_oracle as so
    configs, eps = het_set
    i = _sample(configs)[0]
    c = configs[i]
    print(f"case heterogeneous config {i} N={c['x'].shape[0]}", flush=True)
    xi, ng, P, eta, rc = batch.grid_candidates(c["box"], 1e-3)[0]
    geo = so.SizesGeometry(c["x"], c["a"], c["box"], xi, 1e-3, None, None, None, Ngrid=ng, P=P, exact=True)
    assert abs(geo.rc - rc) < 1e-12 and np.allclose(geo.eta, eta[0])
    n = c["x"].shape[0]
    ctx = Context(gpu)
    ctx.set_tables(np.zeros(2), np.zeros(2), np.array([0.0, 2.0 * rc]))
    ctx.set_geometry(c["x"], c["box"], ng, P, xi, eta, rc)
    ctx.set_sizes(c["a"])
    e = float(eps[-1])
    ctx.set_permittivity(np.full(n, e))
    m = np.random.default_rng(9).standard_normal((n, 3))
    out = ctx.sizes_matvec(m[None], precision=FP64).cpu().numpy()[0]
    assert _rel(out, np.real(so.field(geo, m, np.full(n, e)))) < 1e-12
    res = batch.solve_configurations([c], eps, gmres_tol=1e-12, precision="fp64", workers=1, tune=False, device=gpu)
    alpha = (e - 1.0) / (e + 2.0)
    phi = 4.0 * np.pi * float(np.sum(c["a"] ** 3)) / (3.0 * float(np.prod(c["box"])))
    guess = (4.0 * np.pi * c["a"] ** 3 * alpha / (1.0 - alpha * phi))[:, None] * np.array([0.0, 0.0, 1.0])
    p_or = np.real(so.solve(geo, np.full(n, e), np.array([0.0, 0.0, 1.0]), guess, 1e-12))
    assert _rel(res.p[0][-1, 2], p_or) < 1e-9
    ctx.close()


def test_heterogeneous_mixed_batch_vs_fp64(gpu, het_set):
    """Batched mixed solves (every spectrum point x field of a configuration
    in one solve) at the benchmark's GMRES tolerance against tight FP64
    solves of the same configurations: within the 1e-6 bar."""
    from mpol_hip import batch
    configs, eps = het_set
    idx = _sample(configs)
    sample = [configs[i] for i in idx]
    print(f"case heterogeneous sample {idx}", flush=True)
    grids = [32] * len(sample)
    r64 = batch.solve_configurations(sample, eps, gmres_tol=1e-11, precision="fp64", workers=3, tune=False,
                                     grids=grids, device=gpu)
    r32 = batch.solve_configurations(sample, eps, gmres_tol=HET_GMRES_TOL, precision="mixed", workers=3,
                                     tune=False, grids=grids, device=gpu)
    for k in range(len(sample)):
        assert _rel(r32.p[k], r64.p[k]) < 1e-6, (idx[k], _rel(r32.p[k], r64.p[k]))
        assert r32.matvecs[k][1] > 0


def test_heterogeneous_full_set_properties(gpu, het_set):
    """All 256 configurations x 16 permittivities x 3 fields, batched mixed."""
    from mpol_hip import batch
    configs, eps = het_set
    print("case heterogeneous full set (256 configurations)", flush=True)
    res = batch.solve_configurations(configs, eps, gmres_tol=HET_GMRES_TOL, precision="mixed", workers=8,
                                     device=gpu)
    live = np.abs(eps - 1.0) > 1e-12
    assert res.rhs == 256 * 16 * 3
    assert np.all(res.converged == 3 * int(live.sum()))
    for i in range(len(configs)):
        assert np.all(res.p[i][~live] == 0.0)
        C = res.C[i]  # (Nk, fields, components)
        for k in np.nonzero(live)[0]:
            Ck = C[k]
            assert np.max(np.abs(Ck - Ck.T)) < 1e-5 * np.max(np.abs(Ck)), (i, k)
            assert np.all(np.sign(np.diag(Ck)) == np.sign(eps[k] - 1.0)), (i, k)
        for f in range(3):
            assert np.all(np.diff(C[:, f, f]) > 0.0), i  # polarisability grows with the permittivity


@pytest.mark.parametrize("slab", [0.0, 1.0])
def test_config1_tensor_against_dense_oracle(gpu, slab):
    """The CPU-runnable workload (N = 64 unit spheres by RSA at phi = 0.2,
    eps = 4, xi = 0.5, 32^3, P = 8, rc = 8, GMRES tol 1e-8), three fields in
    one solve: against the dense solve of the oracle's assembled matrix --
    within 1e-6 at the workload's tolerance, 1e-10 at a tight one -- and a
    symmetric capacitance tensor."""
    import workloads
    from mpol_hip import Context
    from oracle import ewald_oracle as eo

    from helpers import tables
    cfg = workloads.configuration("n64_phi0.2")
    n = cfg["x"].shape[0]
    perp, para, rv = tables(cfg["xi"])
    geo = eo.OperatorGeometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["rc"], perp, para, rv)
    A = np.zeros((3 * n, 3 * n))
    for k in range(3 * n):
        e = np.zeros(3 * n)
        e[k] = 1.0
        A[:, k] = np.real(eo.field(geo, e.reshape(n, 3), cfg["eps"], slab)).reshape(-1)
    b = np.stack([np.tile(f, n).reshape(n, 3) for f in np.eye(3)])
    dense = np.stack([np.linalg.solve(A, b[k].reshape(-1)).reshape(n, 3) for k in range(3)])
    ctx = Context(gpu)
    ctx.set_tables(perp, para, rv)
    ctx.set_geometry(cfg["x"], cfg["box"], cfg["ngrid"], cfg["P"], cfg["xi"], cfg["eta"], cfg["rc"])
    ctx.set_permittivity(cfg["eps"])
    phi = 4.0 * np.pi * n / (3.0 * np.prod(cfg["box"]))
    x0 = 4.0 * np.pi * cfg["beta"] / (1.0 - cfg["beta"] * phi) * b
    r = ctx.solve(b, x0, cfg["tol"], restart=min(3 * n, 10), maxit=min(3 * n, 100), slab_factor=slab)
    assert r.flag == 0 and r.max_relres <= cfg["tol"]
    assert _rel(r.x.cpu().numpy(), dense) < 1e-6
    t = ctx.solve(b, x0, 1e-13, restart=min(3 * n, 10), maxit=min(3 * n, 100), slab_factor=slab)
    p = t.x.cpu().numpy()
    assert _rel(p, dense) < 1e-10
    C = p.mean(axis=1)
    assert np.max(np.abs(C - C.T)) < 1e-10 * np.max(np.abs(C))
    ctx.close()
