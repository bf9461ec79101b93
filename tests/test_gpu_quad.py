"""Dipole + quadrupole operator and solve on the GPU against the oracle.

FP64 bar: 1e-12 relative for the operator (both sides FP64; different
summation orders), 1e-10 for solves at tight tolerance; FP32 operator 2e-6;
mixed solves 1e-6.  On even grids the GPU computes the Hermitian part of the
reference operator (real grids), i.e. the real part of the oracle's complex
result for real moments; odd grids have no Nyquist bins and match exactly.
"""
import numpy as np
import pytest

from oracle import ewald_oracle as eo
from oracle import quad_oracle as qo

from helpers import rsa

pytestmark = pytest.mark.gpu

PI = np.pi


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


class Case:
    def __init__(self, gpu, n=24, phi=0.1, seed=3, xi=0.9, tol=1e-6, box=None, ngrid=None, P=None, lam=None):
        from mpol_hip import Context
        if box is None:
            x, box = rsa(n, phi, seed)
        else:
            box = np.asarray(box, dtype=np.float64)
            x, _ = rsa(n, phi, seed, L=float(np.min(box)))
            x = x * (box / np.min(box))[None, :]
        self.x, self.box = x, box
        rc = np.sqrt(-np.log(tol)) / xi
        kc = 2 * xi * np.sqrt(-np.log(tol))
        ng = np.ceil(1 + box * kc / PI).astype(np.int64) if ngrid is None else np.asarray(ngrid, dtype=np.int64)
        P = int(np.ceil(-2 * np.log(tol) / PI)) if P is None else int(P)
        self.ng, self.P, self.rc, self.xi = ng, P, rc, xi
        self.eta = P * (box / ng * xi) ** 2 / PI
        tabs, rv = qo.operator_tables(eo.default_rtable(), xi)
        self.geo = qo.QuadGeometry(x, box, ng, P, xi, rc, tabs, rv)
        self.ctx = Context(gpu)
        self.ctx.set_tables(tabs[0], tabs[1], rv)
        self.ctx.set_geometry(x, box, ng, P, xi, self.eta, rc)
        self.ctx.set_quad_tables(tabs[2:])
        rng = np.random.default_rng(seed + 100)
        self.lam = rng.uniform(2.0, 30.0, n) if lam is None else lam
        self.ctx.set_permittivity(self.lam)
        self.n = n


def _moments(n, seed, k=None, cplx=False):
    rng = np.random.default_rng(seed)
    shape = (n, 8) if k is None else (k, n, 8)
    m = rng.standard_normal(shape)
    if cplx:
        m = m + 1j * rng.standard_normal(shape)
    return m


def _oracle(case, m, lam=None):
    E, G = qo.field(case.geo, m[:, :3], m[:, 3:], case.lam if lam is None else lam)
    return np.concatenate([E, G], axis=1)


@pytest.mark.parametrize("odd", [True, False])
def test_quad_operator_fp64_matches_oracle(gpu, odd):
    ng = np.array([21, 21, 21]) if odd else np.array([24, 24, 24])
    c = Case(gpu, ngrid=ng, P=8)
    m = _moments(c.n, 1)
    out = c.ctx.quad_matvec(m[None]).cpu().numpy()[0]
    ref = _oracle(c, m)
    if odd:
        assert np.max(np.abs(ref.imag)) < 1e-13 * np.max(np.abs(ref))
    assert _rel(out, ref.real) < 1e-12
    # both blocks are exercised
    assert np.max(np.abs(out[:, 3:])) > 1e-2 * np.max(np.abs(out[:, :3]))


def test_quad_operator_components(gpu):
    c = Case(gpu, ngrid=[32, 32, 32], P=8)
    m = _moments(c.n, 2)
    hk, hr = c.ctx.quad_components(m[None])
    (Ek, Gk), (Er, Gr) = qo.field(c.geo, m[:, :3], m[:, 3:], c.lam, parts=True)
    assert _rel(hk.cpu().numpy()[0], np.concatenate([Ek, Gk], axis=1).real) < 1e-12
    assert _rel(hr.cpu().numpy()[0], np.concatenate([Er, Gr], axis=1).real) < 1e-13


@pytest.mark.parametrize("kind", ["noncubic", "prime", "generic"])
def test_quad_operator_grids_and_boxes(gpu, kind):
    """Non-cubic box (k-space coefficients computed per k), a prime grid
    (direct DFT) and a 2^a 3^b 5^c grid without a compile-time plan."""
    if kind == "noncubic":
        c = Case(gpu, n=20, box=[11.0, 12.5, 14.0], ngrid=[28, 30, 36], P=8)
    elif kind == "prime":
        c = Case(gpu, n=20, ngrid=[23, 23, 23], P=7)
    else:
        c = Case(gpu, n=20, ngrid=[36, 36, 36], P=8)
    m = _moments(c.n, 3)
    out = c.ctx.quad_matvec(m[None]).cpu().numpy()[0]
    assert _rel(out, _oracle(c, m).real) < 1e-12


def test_quad_operator_fp32(gpu):
    from mpol_hip import FP32
    c = Case(gpu, n=40, ngrid=[32, 32, 32], P=8)
    m = _moments(c.n, 4)
    o32 = c.ctx.quad_matvec(m[None], precision=FP32).cpu().numpy()[0]
    o64 = c.ctx.quad_matvec(m[None]).cpu().numpy()[0]
    assert _rel(o32, o64) < 2e-6


def test_quad_operator_batched_rhs_and_sets(gpu):
    c = Case(gpu, n=20, ngrid=[24, 24, 24], P=8)
    m = _moments(c.n, 5, k=3)
    lam3 = np.stack([c.lam, c.lam * 1.7, c.lam + 3.0])
    c.ctx.set_permittivity(lam3)
    out = c.ctx.quad_matvec(m).cpu().numpy()
    for k in range(3):
        assert _rel(out[k], _oracle(c, m[k], lam3[k]).real) < 1e-12


@pytest.mark.parametrize("ngrid", [[21, 21, 21], [24, 24, 24]], ids=["odd", "even"])
def test_quad_operator_complex_permittivity(gpu, ngrid):
    """Complex eps (a dielectric function): the diagonal couples real and
    imaginary parts.  On the even grid the reference's Nyquist planes give its
    wave-space operator an anti-Hermitian part that couples them too; the
    extra wave-space pass applies it (the oracle follows the reference)."""
    c = Case(gpu, n=18, ngrid=ngrid, P=8)
    rng = np.random.default_rng(7)
    lam = rng.uniform(2.0, 20.0, c.n) + 1j * rng.uniform(0.5, 10.0, c.n)
    c.ctx.set_permittivity(lam)
    m = _moments(c.n, 8, cplx=True)
    out = c.ctx.quad_matvec(m[None]).cpu().numpy()[0]
    ref = _oracle(c, m, lam)
    assert _rel(out, ref) < 1e-12
    # and the dipole operator in the same complex mode
    S = m[:, :3]
    Hd = c.ctx.matvec(S[None], slab_factor=0.0).cpu().numpy()[0]
    dgeo = eo.OperatorGeometry(c.x, c.box, c.ng, c.P, c.xi, c.rc, c.geo.tabs[0], c.geo.tabs[1], c.geo.rvals)
    assert _rel(Hd, eo.field(dgeo, S, lam, 0.0, complex_eps=True)) < 1e-12


def test_quad_solve_matches_dense_solve(gpu):
    c = Case(gpu, n=10, phi=0.06, ngrid=[21, 21, 21], P=8, lam=np.full(10, 12.0))
    b = np.zeros((1, c.n, 8))
    b[0, :, 2] = 1.0
    res = c.ctx.quad_solve(b, np.zeros_like(b), 1e-12)
    assert res.flag == 0
    A = qo.dense_operator(c.geo, c.lam).real
    perm = np.concatenate([np.arange(c.n)[:, None] * 3 + np.arange(3)[None, :],
                           3 * c.n + np.arange(c.n)[:, None] * 5 + np.arange(5)[None, :]], axis=1).reshape(-1)
    rhs = np.concatenate([np.tile([0.0, 0.0, 1.0], c.n), np.zeros(5 * c.n)])
    sol = np.linalg.solve(A, rhs)[perm].reshape(c.n, 8)
    assert _rel(res.x.cpu().numpy()[0], sol) < 1e-10


def test_quad_solve_fp64_and_mixed_vs_oracle_solve(gpu):
    c = Case(gpu, n=30, phi=0.15, ngrid=[24, 24, 24], P=8)
    b = np.zeros((3, c.n, 8))
    for d in range(3):
        b[d, :, d] = 1.0
    r64 = c.ctx.quad_solve(b, np.zeros_like(b), 1e-11)
    r32 = c.ctx.quad_solve(b, np.zeros_like(b), 1e-7, mixed=True)
    assert r64.flag == 0 and r32.flag == 0
    assert r32.matvecs_fp32 > 0
    x64 = r64.x.cpu().numpy()
    x32 = r32.x.cpu().numpy()
    for d in range(3):
        E0 = np.eye(3)[d]
        S, Q = qo.solve(c.geo, c.lam, E0, np.zeros((c.n, 3)), np.zeros((c.n, 5)), 1e-12)
        ref = np.concatenate([S, Q], axis=1).real
        assert _rel(x64[d], ref) < 1e-9
        assert _rel(x32[d], ref) < 1e-6


def test_quadrupole_module_drop_in(gpu):
    """quadrupole.py (the Capacitance_Quadrupole interface) end to end:
    MatrixVectorMultiply, Capacitance and CapacitanceTensor against the oracle."""
    import quadrupole as qd
    x, box = rsa(14, 0.02, 11)
    xi, tol = 0.9, 1e-6
    rc, _, Ngrid, h, P, eta = qd.ewald_parameters(box, xi, tol)
    r_table = np.concatenate(([0.0], np.arange(1, 100001) * 1e-4))
    tabs = dict(zip(qd.TABLE_NAMES, qd.RealSpaceTable(r_table, xi)))
    otabs = dict(zip(qo.TABLE_NAMES, qo.real_space_table(r_table, xi)))
    for name in qo.OPERATOR_TABLES:  # the product's tables (C++) against the oracle's transcription
        a, b = tabs[name], otabs[name]
        ok = np.isfinite(b)
        ok[:3000] = False  # r < 0.3: deep overlap, where both lose digits to cancellation
        assert np.max(np.abs(a[ok] - b[ok])) < 1e-9 * np.max(np.abs(b[ok]))
    T = [tabs[n] for n in qo.OPERATOR_TABLES]
    lam = np.full(14, 9.0)
    rng = np.random.default_rng(1)
    S = rng.standard_normal((14, 3))
    Q = rng.standard_normal((14, 5))
    E, G = qd.MatrixVectorMultiply(x, S, Q, lam, box, None, None, Ngrid, h, P, xi, eta, rc, None, None, r_table, *T)
    # the drivers' tables carry one extra leading entry; the product cuts them to len(r_table)
    otab = np.stack([otabs[n][:r_table.size] for n in qo.OPERATOR_TABLES])
    geo = qo.QuadGeometry(x, box, Ngrid, P, xi, rc, otab, r_table)
    Eo, Go = qo.field(geo, S, Q, lam)
    assert _rel(np.concatenate([E, G], 1), np.concatenate([Eo, Go], 1).real) < 1e-11
    C, Sz, Qz = qd.Capacitance(x, lam, box, np.zeros((14, 3)), np.zeros((14, 5)), xi, 1e-10, None, None,
                               r_table, *T)
    Ct, St, Qt = qd.CapacitanceTensor(x, lam, box, np.zeros((14, 3)), np.zeros((14, 5)), xi, 1e-10, None, None,
                                      r_table, *T)
    rc2, _, Ng2, h2, P2, eta2 = qd.ewald_parameters(box, xi, 1e-10)
    geo2 = qo.QuadGeometry(x, box, Ng2, P2, xi, rc2, otab, r_table)
    for d in range(3):
        Sd, Qd = qo.solve(geo2, lam, np.eye(3)[d], np.zeros((14, 3)), np.zeros((14, 5)), 1e-12)
        assert _rel(St[:, :, d], Sd.real) < 1e-8
        assert _rel(Qt[:, :, d], Qd.real) < 1e-8
        assert np.allclose(Ct[:, d], St[:, :, d].mean(axis=0))
    assert abs(C - Ct[2, 2]) < 1e-8 * abs(C)
    assert _rel(Sz, St[:, :, 2]) < 1e-8


def test_quad_spectrum_complex_eps(gpu):
    """CapacitanceTensorSpectrum with a complex dielectric function against
    per-point oracle solves (complex GMRES)."""
    import quadrupole as qd
    x, box = rsa(8, 0.08, 5)
    xi = 0.8
    eps = np.stack([np.full(8, 5.0 + 2.0j), np.full(8, 20.0 + 0.5j)], axis=1)
    C, p, Q = qd.CapacitanceTensorSpectrum(x, box, eps, xi)
    assert C.shape == (3, 3, 2) and p.shape == (8, 3, 3, 2, 1) and Q.shape == (8, 5, 3, 2, 1)
    rc, _, Ngrid, h, P, eta = qd.ewald_parameters(box, xi, 1e-3)
    r_table = np.concatenate(([0.0], 1.0 + np.arange(0, 90001) * 1e-4))
    otabs = dict(zip(qo.TABLE_NAMES, qo.real_space_table(r_table, xi)))
    otab = np.stack([otabs[n][:r_table.size] for n in qo.OPERATOR_TABLES])
    geo = qo.QuadGeometry(x, box, Ngrid, P, xi, rc, otab, r_table)
    for k in range(2):
        for d in range(3):
            Sd, Qd = qo.solve(geo, eps[:, k], np.eye(3)[d], np.zeros((8, 3)), np.zeros((8, 5)), 1e-12)
            # the solves stop at the reference's tolerance 1e-3; compare to that accuracy
            assert _rel(p[:, :, d, k, 0], Sd) < 3e-3
            assert _rel(Q[:, :, d, k, 0], Qd) < 3e-3
    assert np.max(np.abs(C.imag)) > 0.0


def test_dipole_tensor_spectrum_complex_eps(gpu):
    """capacitance.capacitance_tensor_spectrum with complex eps against
    dense complex solves of the dipole oracle."""
    import capacitance as cap
    x, box = rsa(10, 0.1, 2)
    xi = 1.4
    eps = np.stack([np.full(10, 4.0 + 1.0j), np.full(10, 30.0 + 5.0j)], axis=1)
    C, p = cap.capacitance_tensor_spectrum(x, box, eps, xi, errortol=1e-10)
    rc, _, ng, h, P, eta = cap.ewald_parameters(box, xi, 1e-10)
    r = eo.default_rtable()
    t = eo.real_space_table(r, xi)
    geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, t[2], t[3], np.concatenate(([0.0], r)))
    for k in range(2):
        A = np.zeros((30, 30), dtype=np.complex128)
        for j in range(30):
            e = np.zeros(30)
            e[j] = 1.0
            A[:, j] = eo.field(geo, e.reshape(10, 3), eps[:, k], 0.0, complex_eps=True).reshape(-1)
        for d in range(3):
            sol = np.linalg.solve(A, np.tile(np.eye(3)[d], 10)).reshape(10, 3)
            print("eps", eps[0, k], "field", d, "vs dense solve", _rel(p[:, :, d, k], sol), flush=True)
            assert _rel(p[:, :, d, k], sol) < 1e-10


@pytest.mark.parametrize("tile,chunk", [((2, 2), 0), ((4, 2), 32)])
def test_quad_spread_contract_engines_agree(gpu, tile, chunk):
    """Eight-component matrix-core spread/contract against the vector loops
    and the oracle (several chunks per tile with chunk 32)."""
    from mpol_hip import FP32
    c = Case(gpu, n=60, phi=0.05, ngrid=[64, 40, 36], P=9)
    m = _moments(c.n, 9, k=2)
    out = {}
    for eng in (1, 2):
        c.ctx.set_tuning(tile_ty=tile[0], tile_tz=tile[1], chunk=chunk if eng == 1 else 0, engine=eng)
        out[eng, 64] = c.ctx.quad_matvec(m).cpu().numpy()
        out[eng, 32] = c.ctx.quad_matvec(m, precision=FP32).cpu().numpy()
    assert _rel(out[1, 64], out[2, 64]) < 1e-13
    assert _rel(out[1, 32], out[2, 64]) < 2e-6
    assert _rel(out[1, 64][0], _oracle(c, m[0]).real) < 1e-12


def test_quad_operator_keeps_default_tile_for_dipole_only_shape(gpu):
    """16-row tiles exist for the dipole operator only: a context tuned to
    them (the launch tuner times the dipole operator) still applies the
    quadrupole operator on the matrix cores with its default tile shape."""
    c = Case(gpu, n=60, phi=0.05, ngrid=[64, 40, 36], P=9)
    m = _moments(c.n, 9, k=2)
    c.ctx.set_tuning(engine=1)
    ref = c.ctx.quad_matvec(m).cpu().numpy()
    c.ctx.set_tuning(tile_ty=4, tile_tz=4, chunk=64, engine=1)
    out = c.ctx.quad_matvec(m).cpu().numpy()
    assert _rel(out, ref) < 1e-13


def test_lone_sphere_quadrupole_clausius_mossotti_gpu(gpu):
    """mpol_quad_solve on a lone sphere under an imposed traceless field
    gradient: S = 0 and Q = g / D2, D2 = 3/(4 pi) (eps + 3/2)/(eps - 1) (the
    l = 2 Clausius-Mossotti response; pins the dielectric term
    15/(8 pi (eps-1)) with the self terms), four permittivities as four
    right-hand sides of one solve; and the oracle's dense solve to 1e-10."""
    from mpol_hip import Context

    from test_quad_oracle import _lone_quad_geometry, lone_quad_expected
    geo = _lone_quad_geometry()
    ctx = Context(gpu)
    ctx.set_tables(geo.tabs[0], geo.tabs[1], geo.rvals)
    ctx.set_geometry(geo.x, geo.box, geo.Ngrid, geo.P, geo.xi, geo.eta, geo.rc)
    ctx.set_quad_tables(geo.tabs[2:])
    epss = np.array([4.0, 0.3, 40.0, 1.5])
    ctx.set_permittivity(epss[:, None])
    g = np.array([1.0, 0.3, -0.2, -0.5, 0.4])
    b = np.zeros((4, 1, 8))
    b[:, 0, 3:] = g
    res = ctx.quad_solve(b, np.zeros_like(b), 1e-13, restart=8, maxit=20)
    out = res.x.cpu().numpy()[:, 0]
    for k, eps in enumerate(epss):
        S, Q = out[k, :3], out[k, 3:]
        ref = lone_quad_expected(eps) * g
        assert np.max(np.abs(S)) < 1e-10
        assert np.max(np.abs(Q - ref)) <= 1e-4 * np.max(np.abs(ref)), (eps, Q / g)
        dense = np.linalg.solve(qo.dense_operator(geo, np.array([eps])), np.concatenate([np.zeros(3), g])).real
        assert _rel(out[k], dense) < 1e-10, (eps, _rel(out[k], dense))
    ctx.close()


def test_quad_mixed_solve_with_a_coarse_correction_operator(gpu):
    """Two-level dipole + quadrupole solve: correction solves on a coarser
    operator (32^3, P = 6 at rule tolerance 1e-4, the library's quadrupole
    tables at its xi) while FP64 residuals stay on the full one (64^3, P = 9,
    the oracle's tables).  The two operators differ by the coarse one's
    Ewald error only (a wrong quadrupole table would put them O(1) apart);
    the two-level result meets tol for the full operator and agrees with its
    tight FP64 solve."""
    import torch

    from mpol_hip import FP32, FP64
    from mpol_hip.autotune import coarse_operator
    c = Case(gpu, n=800, phi=0.2, seed=5, xi=1.04, ngrid=[64, 64, 64], P=9)
    inner, (xi2, ng2, P2, _, rc2) = coarse_operator(c.x, c.box, c.lam, 1e-4, 32, device=gpu, quadrupole=True)
    print("coarse", xi2, ng2, P2, rc2, flush=True)
    m = torch.as_tensor(_moments(c.n, 9, k=1), device=gpu)
    full = c.ctx.quad_matvec(m, precision=FP64).cpu().numpy()
    coarse = inner.quad_matvec(m, precision=FP64).cpu().numpy()
    dist = _rel(coarse, full)
    print("coarse vs full operator", dist, flush=True)
    assert 1e-8 < dist < 1e-2
    b = np.zeros((1, c.n, 8))
    b[0, :, 2] = 1.0
    bt = torch.as_tensor(b, device=gpu)
    x0 = torch.zeros_like(bt)
    tol = 1e-8
    ref = c.ctx.quad_solve(bt, x0, 1e-12, restart=10, maxit=200).x
    c.ctx.set_inner_operator(inner, 1e-3)
    two = c.ctx.quad_solve(bt, x0, tol, restart=10, maxit=100, mixed=True)
    c.ctx.set_inner_operator(None)
    res = bt - c.ctx.quad_matvec(two.x, precision=FP64)
    rel = float(torch.linalg.norm(res) / torch.linalg.norm(bt))
    err = _rel(two.x.cpu().numpy(), ref.cpu().numpy())
    print("two levels", two.matvecs_inner, two.matvecs_fp32, two.matvecs_fp64, two.refinements, "true relres", rel,
          "err vs tight FP64", err, flush=True)
    assert two.converged == 1 and rel <= tol * 1.001 and err < 1e-6
    assert two.matvecs_inner > 0
    inner.close()
