"""Parity of the HIP operator (libmpol.so through the C ABI) with the FP64
oracle and the reference's golden vectors.

Tolerances: FP64 operator 1e-12 relative to max |field| (the target is 1e-10);
FP32 operator 2e-6 relative (mixed-mode target 1e-6 on the solution).
"""
import glob
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, golden
from helpers import oracle_geometry, relerr, rsa, tables
from oracle import ewald_oracle as eo

pytestmark = pytest.mark.gpu

OPERATOR_FILES = sorted(glob.glob(os.path.join(GOLDEN, "operator_*.npz")))


def ctx_for(d):
    from mpol_hip import Context
    xi = float(d["xi"])
    perp, para, rv = tables(xi)
    c = Context()
    c.set_tables(perp, para, rv)
    h = d["box"] / d["ngrid"]
    eta = int(d["P"]) * (h * xi) ** 2 / np.pi
    c.set_geometry(d["x"], d["box"], d["ngrid"], int(d["P"]), xi, eta, float(d["rc"]))
    c.set_permittivity(d["eps"])
    return c


@pytest.mark.parametrize("path", OPERATOR_FILES, ids=os.path.basename)
def test_operator_fp64_matches_reference(gpu, path):
    d = np.load(path)
    print("case", os.path.basename(path), flush=True)
    c = ctx_for(d)
    info = c.geometry_info()
    assert info["npairs"] == int(d["npairs"])
    m = d["m"]
    for slab in (0, 1):
        hk, hr, hs = c.components(np.stack((m.real, m.imag)), slab_factor=float(slab))
        hk = hk.cpu().numpy()
        hr = hr.cpu().numpy()
        hs = hs.cpu().numpy()
        Hk = hk[0] + 1j * hk[1]
        Hr = hr[0] + 1j * hr[1]
        Hs = hs[0] + 1j * hs[1]
        # the reference's wave part for complex input equals ours to the size of
        # its Nyquist-plane anti-Hermitian part (< 1e-9 of |Hk|)
        assert relerr(Hk, d[f"Hk_slab{slab}"]) < 2e-9
        assert relerr(Hr, d[f"Hr_slab{slab}"]) < 1e-13
        np.testing.assert_allclose(Hs, d[f"Hslab_slab{slab}"], rtol=1e-13, atol=1e-15)
        out = c.matvec(np.stack((m.real, m.imag)), slab_factor=float(slab)).cpu().numpy()
        full = d[f"Hk_slab{slab}"] + d[f"Hr_slab{slab}"] - d[f"Hslab_slab{slab}"]
        assert relerr(out[0] + 1j * out[1], full) < 5e-11


@pytest.mark.parametrize("path", OPERATOR_FILES, ids=os.path.basename)
def test_operator_real_input_matches_oracle_hermitian_part(gpu, path):
    """For real dipoles the GPU result equals the real part of the reference
    operator to rounding (FP64)."""
    d = np.load(path)
    print("case", os.path.basename(path), flush=True)
    c = ctx_for(d)
    geo = oracle_geometry(d["x"], d["box"], d["ngrid"], int(d["P"]), float(d["xi"]), float(d["rc"]))
    m = np.real(d["m"])
    ref = np.real(eo.field(geo, m, d["eps"], 1.0))
    out = c.matvec(m[None], slab_factor=1.0).cpu().numpy()[0]
    assert relerr(out, ref) < 1e-12


@pytest.mark.parametrize("path", OPERATOR_FILES, ids=os.path.basename)
def test_operator_fp32(gpu, path):
    from mpol_hip import FP32
    d = np.load(path)
    print("case", os.path.basename(path), flush=True)
    c = ctx_for(d)
    m = np.real(d["m"])
    ref = c.matvec(m[None], slab_factor=1.0).cpu().numpy()[0]
    out = c.matvec(m[None], precision=FP32, slab_factor=1.0).cpu().numpy()[0]
    assert relerr(out, ref) < 2e-6


def test_batched_rhs_and_permittivity_sets(gpu):
    d = golden("operator_n512_g42.npz")
    c = ctx_for(d)
    rng = np.random.default_rng(7)
    m = rng.standard_normal((5, d["x"].shape[0], 3))
    eps = 1.5 + 20 * rng.random((5, d["x"].shape[0]))
    c.set_permittivity(eps)
    batched = c.matvec(m, slab_factor=1.0).cpu().numpy()
    for k in range(5):
        c.set_permittivity(eps[k])
        single = c.matvec(m[k:k + 1], slab_factor=1.0).cpu().numpy()[0]
        assert relerr(batched[k], single) < 1e-13


def test_operator_is_symmetric_and_linear(gpu):
    """<u, A v> = <A u, v> (every block of A is symmetric for real eps) and linearity."""
    x, box = rsa(400, 0.3, 11)
    xi, P, ng = 0.9, 8, np.array([40, 36, 30])
    rc = np.sqrt(-np.log(1e-6)) / xi
    from mpol_hip import Context
    perp, para, rv = tables(xi)
    c = Context()
    c.set_tables(perp, para, rv)
    h = box / ng
    c.set_geometry(x, box, ng, P, xi, P * (h * xi) ** 2 / np.pi, rc)
    c.set_permittivity(np.full(400, 5.0))
    rng = np.random.default_rng(3)
    u = rng.standard_normal((400, 3))
    v = rng.standard_normal((400, 3))
    Au, Av, Auv = (c.matvec(np.stack((u, v, 2 * u - 3 * v)), slab_factor=1.0).cpu().numpy())
    assert abs(np.sum(u * Av) - np.sum(Au * v)) / abs(np.sum(u * Av)) < 1e-12
    assert relerr(Auv, 2 * Au - 3 * Av) < 1e-12
    # anisotropic grid against the oracle
    geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp, para, rv)
    ref = np.real(eo.field(geo, u, np.full(400, 5.0), 1.0))
    assert relerr(Au, ref) < 1e-12


def test_geometry_errors(gpu):
    from mpol_hip import Context
    d = golden("operator_n64_g32.npz")
    perp, para, rv = tables(float(d["xi"]))
    c = Context()
    c.set_tables(perp[:3000], para[:3000], rv[:3000])  # table ends at r = 2.999 < rc
    h = d["box"] / d["ngrid"]
    with pytest.raises(ValueError):
        c.set_geometry(d["x"], d["box"], d["ngrid"], int(d["P"]), float(d["xi"]),
                       int(d["P"]) * (h * float(d["xi"])) ** 2 / np.pi, float(d["rc"]))
    with pytest.raises(Exception):
        c.matvec(np.zeros((1, 64, 3)))


def test_operator_noncubic_box(gpu):
    """Orthorhombic box (per-k coefficient path) against the oracle."""
    from mpol_hip import Context
    rng = np.random.default_rng(21)
    box = np.array([14.0, 12.5, 16.0])
    pts = []
    while len(pts) < 60:
        c = rng.random(3) * box
        if all(np.sum(((c - p) - box * np.rint((c - p) / box)) ** 2) >= 4.0 for p in pts):
            pts.append(c)
    x = np.array(pts)
    xi, P, ng = 0.9, 8, np.array([32, 30, 36])
    rc = 5.0
    perp, para, rv = tables(xi)
    eps = 2.0 + 10 * rng.random(60)
    c = Context()
    c.set_tables(perp, para, rv)
    h = box / ng
    c.set_geometry(x, box, ng, P, xi, P * (h * xi) ** 2 / np.pi, rc)
    c.set_permittivity(eps)
    m = rng.standard_normal((60, 3))
    out = c.matvec(m[None], slab_factor=1.0).cpu().numpy()[0]
    geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp, para, rv)
    ref = np.real(eo.field(geo, m, eps, 1.0))
    assert relerr(out, ref) < 1e-12


@pytest.mark.parametrize("ng,P,tile,chunk", [((32, 32, 32), 8, (0, 0), 0), ((64, 30, 36), 10, (4, 2), 32),
                                              ((96, 40, 24), 9, (2, 2), 64), ((64, 64, 64), 12, (2, 2), 256),
                                              ((64, 44, 38), 10, (4, 4), 128), ((128, 36, 52), 9, (4, 4), 32)])
def test_spread_contract_engines_agree(gpu, ng, P, tile, chunk):
    """The matrix-core spread/contract (per-tile candidate images, several
    chunks per tile, partial edge tiles) against the vector-loop kernels and
    the FP64 oracle, both precisions."""
    from mpol_hip import FP32, Context
    print("case", ng, P, tile, chunk, flush=True)
    rng = np.random.default_rng(sum(ng) + P)
    box = np.array(ng, dtype=float) * 0.41
    n = 700
    x = rng.random((n, 3)) * box
    xi = 1.0
    rc = 3.2
    perp, para, rv = tables(xi)
    eps = 2.0 + 10 * rng.random(n)
    c = Context()
    c.set_tables(perp, para, rv)
    h = box / np.array(ng)
    eta = P * (h * xi) ** 2 / np.pi
    c.set_geometry(x, box, np.array(ng), P, xi, eta, rc)
    c.set_permittivity(eps)
    m = rng.standard_normal((2, n, 3))
    out = {}
    for eng in (1, 2):
        tt = tile if eng == 1 or tile != (4, 4) else (0, 0)  # 16-row tiles: matrix-core engine only
        c.set_tuning(tile_ty=tt[0], tile_tz=tt[1], chunk=chunk if eng == 1 else 0, engine=eng)
        out[eng, 64] = c.matvec(m, slab_factor=1.0).cpu().numpy()
        out[eng, 32] = c.matvec(m, precision=FP32, slab_factor=1.0).cpu().numpy()
    assert relerr(out[1, 64], out[2, 64]) < 1e-13
    assert relerr(out[1, 32], out[2, 64]) < 2e-6
    geo = eo.OperatorGeometry(x, box, np.array(ng), P, xi, rc, perp, para, rv)
    ref = np.real(eo.field(geo, m[1], eps, 1.0))
    assert relerr(out[1, 64][1], ref) < 1e-12


def test_register_z_pass_matches_lds_z_pass(gpu):
    """nz = 256: the register-resident z transforms + scale against the
    LDS-resident z pass (forced by a z column width) and the FP64 oracle,
    including the Nyquist planes of an even anisotropic grid."""
    from mpol_hip import FP32, Context
    rng = np.random.default_rng(41)
    ng = np.array([32, 48, 256])
    box = ng * 0.45
    n = 500
    x = rng.random((n, 3)) * box
    xi, P, rc = 1.0, 8, 3.5
    perp, para, rv = tables(xi)
    eps = 2.0 + 10 * rng.random(n)
    c = Context()
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    c.set_permittivity(eps)
    m = rng.standard_normal((2, n, 3))
    reg64 = c.matvec(m, slab_factor=1.0).cpu().numpy()
    reg32 = c.matvec(m, precision=FP32, slab_factor=1.0).cpu().numpy()
    c.set_tuning(z_cols=8)
    lds64 = c.matvec(m, slab_factor=1.0).cpu().numpy()
    lds32 = c.matvec(m, precision=FP32, slab_factor=1.0).cpu().numpy()
    geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp, para, rv)
    for k in range(2):
        ref = np.real(eo.field(geo, m[k], eps, 1.0))
        print("rhs", k, "register z vs oracle", relerr(reg64[k], ref), "lds z vs oracle", relerr(lds64[k], ref),
              flush=True)
        assert relerr(lds64[k], ref) < 1e-12
        assert relerr(reg64[k], ref) < 1e-12
    assert relerr(reg64, lds64) < 1e-13
    assert relerr(reg32, lds32) < 2e-6


def test_sparse_tiles_repeatable(gpu):
    """Tiles without candidate particles (sparse system on a large grid) must
    give exactly zero spread, on every call: repeated multi-RHS matvecs are
    bitwise identical and equal the vector-loop engine."""
    from mpol_hip import Context
    rng = np.random.default_rng(43)
    ng = np.array([64, 48, 256])
    box = ng * 0.45
    n = 500
    x = rng.random((n, 3)) * box
    xi, P, rc = 1.0, 8, 3.5
    perp, para, rv = tables(xi)
    c = Context()
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    c.set_permittivity(np.full(n, 7.0))
    m = rng.standard_normal((2, n, 3))
    c.set_tuning(engine=2)
    ref = c.matvec(m, slab_factor=1.0).cpu().numpy()
    c.set_tuning(engine=1)
    outs = [c.matvec(m, slab_factor=1.0).cpu().numpy() for _ in range(6)]
    for o in outs:
        assert np.array_equal(o, outs[0])
    assert relerr(outs[0], ref) < 1e-13


def test_rank_one_z_scheme_matches_oracle(gpu):
    """Cubic box, nz = 256: the rank-one scheme (x and y components as one
    grid between the y passes, four z transforms per regular line, the last
    column and the row ky = ny/2 in the full form) against the FP64 oracle and
    against the full three-component passes (z_rank1 = 2), in both precisions,
    with two right-hand sides, then with complex permittivities."""
    from mpol_hip import FP32, Context
    rng = np.random.default_rng(47)
    ng = np.array([64, 96, 256])
    L = 40.0
    box = np.array([L, L, L])
    n = 400
    x = rng.random((n, 3)) * box
    xi, P, rc = 1.2, 8, 3.5
    perp, para, rv = tables(xi)
    eps = 2.0 + 10 * rng.random(n)
    c = Context()
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    c.set_permittivity(eps)
    m = rng.standard_normal((2, n, 3))
    out = {}
    for mode in (0, 2):
        c.set_tuning(z_rank1=mode)
        out[mode, 64] = c.matvec(m, slab_factor=1.0).cpu().numpy()
        out[mode, 32] = c.matvec(m, precision=FP32, slab_factor=1.0).cpu().numpy()
    geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp, para, rv)
    for k in range(2):
        ref = np.real(eo.field(geo, m[k], eps, 1.0))
        print("rhs", k, "rank-one vs oracle", relerr(out[0, 64][k], ref), "full vs oracle",
              relerr(out[2, 64][k], ref), flush=True)
        assert relerr(out[0, 64][k], ref) < 1e-12
        assert relerr(out[0, 32][k], ref) < 2e-6
    print("rank-one vs full", relerr(out[0, 64], out[2, 64]), relerr(out[0, 32], out[2, 32]), flush=True)
    assert relerr(out[0, 64], out[2, 64]) < 1e-13
    assert relerr(out[0, 32], out[2, 32]) < 2e-6
    # the two schemes round differently: equal outputs would mean one path ran twice
    assert not np.array_equal(out[0, 64], out[2, 64])
    # complex permittivities (each system a (re, im) pair of right-hand sides):
    # against the full passes and the oracle, which follows the reference's
    # complex FFTs (their Nyquist planes' anti-Hermitian part is applied by
    # an extra wave-space pass; without it, cplx_nyquist = 2, the result is
    # off by that part)
    epsc = eps + 1j * (0.5 + rng.random(n))
    mc = rng.standard_normal((2, n, 3)) + 1j * rng.standard_normal((2, n, 3))
    c.set_permittivity(epsc)
    outc = {}
    for mode in (0, 2):
        c.set_tuning(z_rank1=mode)
        outc[mode] = c.matvec(mc, slab_factor=0.0).cpu().numpy()
    c.set_tuning(cplx_nyquist=2)
    herm = c.matvec(mc, slab_factor=0.0).cpu().numpy()
    c.set_tuning(precision=FP32)
    out32 = c.matvec(mc, precision=FP32, slab_factor=0.0).cpu().numpy()
    for k in range(2):
        refc = eo.field(geo, mc[k], epsc, 0.0, complex_eps=True)
        print("complex rhs", k, "vs oracle", relerr(outc[0][k], refc), "FP32", relerr(out32[k], refc),
              "Hermitian part only", relerr(herm[k], refc), flush=True)
        assert relerr(outc[0][k], refc) < 1e-12
        assert relerr(out32[k], refc) < 2e-6
        assert relerr(herm[k], refc) > 1e-11
    print("complex rank-one vs full", relerr(outc[0], outc[2]), flush=True)
    assert relerr(outc[0], outc[2]) < 1e-13
    assert not np.array_equal(outc[0], outc[2])


@pytest.mark.parametrize("ng", [(64, 96, 256), (64, 64, 64)])
def test_pass_order_and_spread_group_keep_the_operator(gpu, ng):
    """Several systems: the y, z and inverse y passes system by system
    (fft_order 1, the automatic order) and pass by pass over all systems
    (fft_order 2) give bitwise the same outputs, with the FP32 spread at one
    right-hand side per workgroup (automatic) or in groups of three, for
    real permittivities (three systems) and complex ones (two systems of two
    right-hand sides), each output also against the oracle.  Both grids take
    the matrix-core tile kernels (fixed summation orders; nz = 256 the
    register z pass, nz = 64 the LDS one); off that path the spread sums
    with atomics and repeats agree to rounding only."""
    from mpol_hip import FP32, FP64, Context
    rng = np.random.default_rng(53)
    ng = np.array(ng)
    box = np.array([40.0, 40.0, 40.0])
    n = 500
    x = rng.random((n, 3)) * box
    xi, P, rc = 1.2, 8, 3.5
    perp, para, rv = tables(xi)
    eps = 2.0 + 10 * rng.random(n)
    c = Context()
    c.set_tables(perp, para, rv)
    c.set_geometry(x, box, ng, P, xi, P * (box / ng * xi) ** 2 / np.pi, rc)
    geo = eo.OperatorGeometry(x, box, ng, P, xi, rc, perp, para, rv)
    cases = (("real", eps, rng.standard_normal((3, n, 3))),
             ("complex", eps + 1j * (0.5 + rng.random(n)),
              rng.standard_normal((2, n, 3)) + 1j * rng.standard_normal((2, n, 3))))
    for name, e, m in cases:
        c.set_permittivity(e)
        for prec, tol in ((FP64, 1e-12), (FP32, 2e-6)):
            outs = {}
            for fo, rg in ((2, 3), (1, 3), (0, 0), (2, 0)):
                c.set_tuning(precision=prec, fft_order=fo, rhs_group=rg)
                outs[fo, rg] = c.matvec(m, precision=prec, slab_factor=0.0).cpu().numpy()
            c.set_tuning()
            base = outs[2, 3]
            for key, o in outs.items():
                assert np.array_equal(o, base), (name, prec, key, relerr(o, base))
            for k in range(m.shape[0]):
                ref = eo.field(geo, m[k], e, 0.0, complex_eps=(name == "complex"))
                if name == "real":
                    ref = np.real(ref)
                err = relerr(base[k], ref)
                print(name, "precision", int(prec), "rhs", k, "vs oracle", err, flush=True)
                assert err < tol
