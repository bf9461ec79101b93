"""Polydisperse operator and solve on the GPU against the oracle.

FP64 bar: 1e-12 relative for the operator in table mode (the reference's
interpolated pair tables) and in closed-form mode (each pair's couplings at
its own separation, against the oracle evaluated the same way); solves at
tight tolerance to 1e-10; FP32 operator 2e-6; mixed solves 1e-6.  The scalar
grid's scale depends on |q| only, so even grids need no Nyquist treatment.
"""
import numpy as np
import pytest

from oracle import sizes_oracle as so

from helpers import rsa

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _config(n, phi, seed, L=None, radii=(0.8, 1.0, 1.25), box=None):
    x, b = rsa(n, phi, seed, L=L)
    if box is not None:
        box = np.asarray(box, dtype=np.float64)
        x = x * (box / b)[None, :]
        b = box
    rng = np.random.default_rng(seed + 7)
    a = rng.choice(radii, size=n) if radii is not None else rng.lognormal(0.0, 0.2, n)
    return x, b, a


class Case:
    def __init__(self, gpu, x, box, a, xi=0.9, tol=1e-6, closed_form=False, eps=None):
        from mpol_hip import Context
        self.x, self.box, self.a, self.xi, self.tol = x, box, a, xi, tol
        n = x.shape[0]
        au, ind = so.uniquetol(a)
        self.r = np.concatenate(([0.0], np.arange(1e-3, 10.0 + 5e-4, 1e-3)))
        if closed_form:
            self.geo = so.SizesGeometry(x, a, box, xi, tol, None, None, None, exact=True)
        else:
            self.perp, self.para = so.real_space_table(self.r, au, xi)
            self.geo = so.SizesGeometry(x, a, box, xi, tol, self.r, self.perp, self.para)
        g = self.geo
        self.ctx = Context(gpu)
        if closed_form:
            self.ctx.set_tables(np.zeros(2), np.zeros(2), np.array([0.0, 2 * g.rc]))
        else:
            self.ctx.set_tables(self.perp[:, 0], self.para[:, 0], self.r)
        self.ctx.set_geometry(x, box, g.Ngrid, g.P, xi, np.full(3, g.eta), g.rc)
        if closed_form:
            self.ctx.set_sizes(a)
        else:
            self.ctx.set_sizes(a, ind, self.perp, self.para)
        rng = np.random.default_rng(11)
        self.eps = rng.uniform(2.0, 30.0, n) if eps is None else eps
        self.ctx.set_permittivity(self.eps)
        self.n = n


@pytest.mark.parametrize("closed_form", [False, True])
def test_sizes_operator_fp64(gpu, closed_form):
    x, box, a = _config(24, 0.1, 3)
    c = Case(gpu, x, box, a, closed_form=closed_form)
    p = np.random.default_rng(1).standard_normal((c.n, 3))
    out = c.ctx.sizes_matvec(p[None]).cpu().numpy()[0]
    ref = so.field(c.geo, p, c.eps)
    assert np.max(np.abs(ref.imag)) < 1e-12 * np.max(np.abs(ref))
    assert _rel(out, ref.real) < 1e-12


def test_sizes_operator_components_and_noncubic(gpu):
    x, box, a = _config(20, 0.1, 5, box=[10.0, 11.5, 12.5])
    c = Case(gpu, x, box, a)
    p = np.random.default_rng(2).standard_normal((c.n, 3))
    hk, hr = c.ctx.sizes_components(p[None])
    Eq, Er = so.field(c.geo, p, c.eps, parts=True)
    assert _rel(hk.cpu().numpy()[0], Eq.real) < 1e-12
    assert _rel(hr.cpu().numpy()[0], Er.real) < 1e-13


def test_sizes_operator_continuous_radii_fp32_and_batched(gpu):
    from mpol_hip import FP32
    x, box, a = _config(30, 0.1, 6, radii=None)
    c = Case(gpu, x, box, a, closed_form=True)
    p = np.random.default_rng(3).standard_normal((3, c.n, 3))
    eps3 = np.stack([c.eps, c.eps * 2.0, c.eps + 5.0])
    c.ctx.set_permittivity(eps3)
    o64 = c.ctx.sizes_matvec(p).cpu().numpy()
    o32 = c.ctx.sizes_matvec(p, precision=FP32).cpu().numpy()
    for k in range(3):
        assert _rel(o64[k], so.field(c.geo, p[k], eps3[k]).real) < 1e-12
    assert _rel(o32, o64) < 2e-6


def test_sizes_entry_kernels_match_particle_kernels_and_follow_geometry(gpu):
    """The operator from stored entries (default: per-node gather spread,
    particle-major contraction of the transposed grid) against the
    per-particle kernels (engine knob 2: atomic scatter, LDS-staged
    contraction) on 20 right-hand sides (a partial group of 16), then after
    moving the particles and changing the radii (the entries must be
    rebuilt)."""
    from mpol_hip import FP32, FP64
    x, box, a = _config(40, 0.15, 8, radii=None)
    c = Case(gpu, x, box, a, closed_form=True)
    p = np.random.default_rng(4).standard_normal((20, c.n, 3))
    c.ctx.set_permittivity(np.stack([c.eps * (1.0 + 0.1 * k) for k in range(20)]))
    for prec, bar in ((FP64, 1e-13), (FP32, 1e-6)):
        c.ctx.set_tuning(precision=prec, engine=0)
        gather = c.ctx.sizes_matvec(p, precision=prec).cpu().numpy()
        c.ctx.set_tuning(precision=prec, engine=2)
        scatter = c.ctx.sizes_matvec(p, precision=prec).cpu().numpy()
        c.ctx.set_tuning(precision=prec, engine=0)
        assert _rel(gather, scatter) < bar, prec
    rng = np.random.default_rng(9)
    x2 = np.mod(x + 0.05 * rng.standard_normal(x.shape), box)
    a2 = a * rng.uniform(0.9, 1.0, a.size)
    geo = so.SizesGeometry(x2, a2, box, c.xi, c.tol, None, None, None, exact=True)
    c.ctx.set_geometry(x2, box, geo.Ngrid, geo.P, c.xi, np.full(3, geo.eta), geo.rc)
    c.ctx.set_sizes(a2)
    c.ctx.set_permittivity(c.eps)
    out = c.ctx.sizes_matvec(p[:1]).cpu().numpy()[0]
    assert _rel(out, so.field(geo, p[0], c.eps).real) < 1e-12
    out32 = c.ctx.sizes_matvec(p[:1], precision=FP32).cpu().numpy()[0]
    assert _rel(out32, out) < 2e-6


@pytest.mark.parametrize("ngrid", [(32, 32, 32), (40, 32, 32)])
def test_sizes_operator_yz_plane_grids(gpu, ngrid):
    """ny = nz = 32 takes the one-kernel y/z passes (per y-z plane): FP64
    against the oracle on 18 right-hand sides (partial column groups: kx = 17
    and 21), FP32 against FP64, and the separate passes (engine knob 2) equal
    to it."""
    from mpol_hip import FP32, FP64
    x, box, a = _config(36, 0.12, 13, radii=None)
    xi, tol = 0.9, 1e-6
    geo = so.SizesGeometry(x, a, box, xi, tol, None, None, None, Ngrid=np.array(ngrid), P=8, exact=True)
    from mpol_hip import Context
    ctx = Context(gpu)
    ctx.set_tables(np.zeros(2), np.zeros(2), np.array([0.0, 2 * geo.rc]))
    ctx.set_geometry(x, box, geo.Ngrid, geo.P, xi, np.full(3, geo.eta), geo.rc)
    ctx.set_sizes(a)
    rng = np.random.default_rng(21)
    eps = np.stack([rng.uniform(2.0, 30.0, x.shape[0]) for _ in range(18)])
    ctx.set_permittivity(eps)
    p = rng.standard_normal((18, x.shape[0], 3))
    o64 = ctx.sizes_matvec(p, precision=FP64).cpu().numpy()
    for k in (0, 7, 17):
        assert _rel(o64[k], so.field(geo, p[k], eps[k]).real) < 1e-12, k
    o32 = ctx.sizes_matvec(p, precision=FP32).cpu().numpy()
    assert _rel(o32, o64) < 2e-6
    ctx.set_tuning(engine=2)
    sep = ctx.sizes_matvec(p, precision=FP64).cpu().numpy()
    assert _rel(sep, o64) < 1e-13


def test_sizes_operator_complex_permittivity(gpu):
    x, box, a = _config(18, 0.1, 8)
    rng = np.random.default_rng(4)
    eps = rng.uniform(2.0, 20.0, 18) + 1j * rng.uniform(0.5, 8.0, 18)
    c = Case(gpu, x, box, a, eps=eps)
    p = rng.standard_normal((18, 3)) + 1j * rng.standard_normal((18, 3))
    out = c.ctx.sizes_matvec(p[None]).cpu().numpy()[0]
    assert _rel(out, so.field(c.geo, p, eps)) < 1e-12


def test_sizes_solve_vs_dense_and_mixed(gpu):
    x, box, a = _config(12, 0.08, 9)
    c = Case(gpu, x, box, a, eps=np.full(12, 10.0))
    b = np.zeros((3, c.n, 3))
    for d in range(3):
        b[d, :, d] = 1.0
    r64 = c.ctx.sizes_solve(b, np.zeros_like(b), 1e-12)
    r32 = c.ctx.sizes_solve(b, np.zeros_like(b), 1e-7, mixed=True)
    assert r64.flag == 0 and r32.flag == 0 and r32.matvecs_fp32 > 0
    A = so.dense_operator(c.geo, c.eps).real
    for d in range(3):
        sol = np.linalg.solve(A, np.tile(np.eye(3)[d], c.n)).reshape(c.n, 3)
        assert _rel(r64.x.cpu().numpy()[d], sol) < 1e-10
        assert _rel(r32.x.cpu().numpy()[d], sol) < 1e-6


def test_sizes_module_drop_in(gpu):
    """sizes.py (the Capacitance_Sizes interface): ComputeField, Dipole and
    CapacitanceTensor against the oracle, tables from RealSpaceTable."""
    import sizes as sz
    x, box, a = _config(16, 0.02, 12)
    xi, tol = 1.0, 1e-8
    r_table = np.concatenate(([0.0], 1.0 + np.arange(0, 9001) * 1e-3))
    au, _ = sz.uniquetol(a)
    Ep_perp, Ep_para = sz.RealSpaceTable(r_table, au, xi)
    op, oq = so.real_space_table(r_table, au, xi)
    ok = r_table >= 1.0
    assert _rel(Ep_perp[ok], op[ok]) < 1e-12 and _rel(Ep_para[ok], oq[ok]) < 1e-12
    eps = np.full(16, 6.0)
    rng = np.random.default_rng(5)
    p = rng.standard_normal((16, 3))
    E = sz.ComputeField(x, a, p, eps, box, None, None, xi, r_table, Ep_perp, Ep_para, tol)
    geo = so.SizesGeometry(x, a, box, xi, tol, r_table, op, oq)
    assert _rel(E, so.field(geo, p, eps).real) < 1e-11
    C, pt = sz.CapacitanceTensor(x, a, eps, box, np.zeros((16, 3)), xi, tol, r_table, Ep_perp, Ep_para)
    assert C.shape == (9,) and pt.shape == (16, 9)
    for d in range(3):
        ref = so.solve(geo, eps, np.eye(3)[d], np.zeros((16, 3)), 1e-12)
        assert _rel(pt[:, 3 * d:3 * d + 3], ref.real) < 1e-7
    pave, pz = sz.Dipole(x, a, eps, box, [0.0, 0.0, 1.0], np.zeros((16, 3)), xi, tol, r_table, Ep_perp, Ep_para)
    assert _rel(pz, pt[:, 6:9]) < 1e-7 and np.allclose(pave, pz.mean(axis=0))


def test_sizes_spectrum_complex_and_closed_form(gpu):
    """CapacitanceTensorSpectrum with a complex dielectric function, in table
    mode against per-point oracle solves, and closed-form mode close to it
    (they differ by the table interpolation only)."""
    import sizes as sz
    x, box, a = _config(10, 0.05, 13)
    xi = 0.9
    eps = np.stack([np.full(10, 4.0 + 1.5j), np.full(10, 15.0 + 0.3j)], axis=1)
    C, p = sz.CapacitanceTensorSpectrum(x, box, a, eps, xi)
    Cc, pc = sz.CapacitanceTensorSpectrum(x, box, a, eps, xi, closed_form=True)
    assert C.shape == (2, 9) and p.shape == (10, 9, 2, 1)
    au, _ = sz.uniquetol(a)
    r_table = np.concatenate(([0.0], 1.0 + np.arange(0, 9001) * 1e-3))
    op, oq = so.real_space_table(r_table, au, xi)
    geo = so.SizesGeometry(x, a, box, xi, 1e-3, r_table, op, oq)
    for k in range(2):
        for d in range(3):
            ref = so.solve(geo, eps[:, k], np.eye(3)[d], np.zeros((10, 3)), 1e-12)
            assert _rel(p[:, 3 * d:3 * d + 3, k, 0], ref) < 3e-3  # the solves stop at tol 1e-3
    assert _rel(pc, p) < 1e-2
    assert np.max(np.abs(C.imag)) > 0


def test_lone_sphere_clausius_mossotti_gpu(gpu):
    """mpol_sizes_solve on a lone sphere (one per cubic cell) for several
    radii and permittivities, all in one batched solve: p = 4 pi a^3 beta /
    (1 - beta phi) E0 (the Clausius-Mossotti closed form, which pins the
    dielectric term 3p/(4 pi a^3 (eps-1)) and the self terms at each radius),
    and the oracle's dense solve to 1e-10."""
    from mpol_hip import Context

    from test_sizes_oracle import LONE_CASES, lone_sphere_expected
    L, xi, tol = 20.0, 0.6, 1e-10
    x = np.array([[3.1, 7.4, 11.9]])
    box = np.array([L, L, L])
    for a, eps in LONE_CASES:
        geo = so.SizesGeometry(x, np.array([a]), box, xi, tol, None, None, None, Ngrid=np.array([40, 40, 40]),
                               exact=True)
        ctx = Context(gpu)
        ctx.set_tables(np.zeros(2), np.zeros(2), np.array([0.0, 2 * geo.rc]))
        ctx.set_geometry(x, box, geo.Ngrid, geo.P, xi, np.full(3, geo.eta), geo.rc)
        ctx.set_sizes(np.array([a]))
        ctx.set_permittivity(np.array([eps]))
        b = np.array([[[0.0, 0.0, 1.0]]])
        res = ctx.sizes_solve(b, np.zeros_like(b), 1e-13, restart=3, maxit=20)
        p = res.x.cpu().numpy()[0, 0]
        ref = lone_sphere_expected(a, eps, L)
        assert abs(p[2] - ref) <= 2e-6 * abs(ref), (a, eps, p, ref)
        dense = np.linalg.solve(so.dense_operator(geo, np.array([eps])), np.array([0.0, 0.0, 1.0])).real
        assert _rel(p, dense) < 1e-10, (a, eps, _rel(p, dense))
        ctx.close()
