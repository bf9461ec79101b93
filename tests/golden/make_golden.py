"""Write golden input/output vectors from the reference Python solver.

Run in a container where the reference tree is mounted (it is never needed at
test time):

    python tests/golden/make_golden.py /root/reference

It imports ``CPU - Ewald test/ewald.py`` and ``capacitance.py`` from that tree
and records, as small .npz fixtures next to this script:

  * tables_xi*.npz     -- real_space_table (all six couplings) on a subsample
                          of the reference r grid, for xi = 0.5 and 1.0;
  * operator_*.npz     -- magnetic_field_components (Hk, Hr, H_slab) for seeded
                          positions, dipoles and permittivities at explicit
                          grids (32^3 power of two, 21^3 = 3*7, 47^3 prime,
                          42^3 with Ncell > 3), slab_factor 0 and 1;
  * solve_*.npz        -- magnetic_dipole at tol 1e-12 through SciPy GMRES and
                          the in-file gmres_restarted;
  * run_*.npz          -- run_capacitance end to end at derived parameters;
  * dropin_n24.npz     -- the step-by-step functions of the fast path
                          (prepare_grid_kernel, spread_fast, spread,
                          contract_fast, contract, scale, precompute_kspace,
                          scale_fast, prepare_real_space_pairs,
                          real_space_fast, real_space) on a small
                          anisotropic case with random dipoles and grids.

    python tests/golden/make_golden.py /root/reference dropin   # only that one

Only inputs and outputs are stored; no reference source is copied.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def rsa(n, phi, seed, radius=1.0):
    """Seeded random sequential addition of n unit spheres in a cube at volume fraction phi."""
    L = (n * 4.0 / 3.0 * np.pi * radius ** 3 / phi) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pts = np.zeros((0, 3))
    while pts.shape[0] < n:
        c = rng.random(3) * L
        if pts.shape[0]:
            d = pts - c
            d -= L * np.rint(d / L)
            if np.min(np.sum(d * d, axis=1)) < (2.0 * radius) ** 2:
                continue
        pts = np.vstack((pts, c))
    return pts, np.array([L, L, L])


def dropin_case(ewald):
    """The reference's step-by-step functions on 24 spheres in a non-cubic box
    (8 x 10 x 7 grid, P = 5, coordinates partly negative), with random complex
    dipoles, grids and permittivities as inputs."""
    n, xi, P = 24, 0.7, 5
    x, box = rsa(n, 0.15, 21)
    box = box * np.array([1.0, 1.15, 0.9])
    x = x * np.array([1.0, 1.15, 0.9]) - 0.3 * box
    Ngrid = np.array([8, 10, 7])
    h = box / Ngrid
    eta = P * (h * xi) ** 2 / np.pi
    rc = np.sqrt(-np.log(1e-4)) / xi
    rng = np.random.default_rng(77)
    m = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    lam = 1.0 + 20.0 * rng.random(n)
    G = tuple(int(v) for v in Ngrid) + (3,)
    Htilde = rng.standard_normal(G) + 1j * rng.standard_normal(G)
    fH = rng.standard_normal(G) + 1j * rng.standard_normal(G)
    r_full = np.arange(0.001, 10.0 + 1e-12, 0.001, dtype=np.float64)
    r_table = np.concatenate(([0.0], r_full))
    _, _, perp, para, _, _ = ewald.real_space_table(r_full, xi)
    offset, offsetxyz = ewald.pre_calculations(P, h)
    linnode, w_spread, w_contract = ewald.prepare_grid_kernel(x, Ngrid, h, xi, eta, P, offset, offsetxyz)
    H_fast = ewald.spread_fast(m, Ngrid, linnode, w_spread)
    Hx, Hy, Hz = ewald.spread(x, m, Ngrid, h, xi, eta, P, offset, offsetxyz)
    Hk_fast = ewald.contract_fast(Htilde, Ngrid, linnode, w_contract)
    Hk = ewald.contract(x, Ngrid, h, xi, eta, P, Htilde, offset, offsetxyz)
    ks = [np.arange(-int(np.ceil((N - 1) / 2.0)), int(np.floor((N - 1) / 2.0)) + 1) * (2.0 * np.pi / L)
          for N, L in zip(Ngrid, box)]
    k = np.stack(np.meshgrid(*ks, indexing="ij"), axis=3)
    fHt = ewald.scale(fH, k, Ngrid, xi, eta)
    khat, coeff = ewald.precompute_kspace(Ngrid, box, xi, eta)
    fHt_fast = ewald.scale_fast(fH, khat, coeff)
    cell, ncell = ewald.cell_list(x, box, rc)
    p1, p2 = ewald.neighbor_list(cell, ncell)
    p1f, p2f, rhat, A, B = ewald.prepare_real_space_pairs(x, box, p1, p2, rc, perp, para, r_table)
    Hr_fast = ewald.real_space_fast(m, lam, p1f, p2f, rhat, A, B, float(perp[0]))
    Hr = ewald.real_space(x, m, lam, box, p1, p2, rc, perp, para, r_table)
    np.savez_compressed(os.path.join(HERE, "dropin_n24.npz"), x=x, box=box, Ngrid=Ngrid, h=h, xi=xi, P=P, eta=eta,
                        rc=rc, m=m, lam=lam, Htilde=Htilde, fH=fH, k=k, offset=offset, offsetxyz=offsetxyz,
                        linnode=linnode, w_spread=w_spread, w_contract=w_contract, H_fast=H_fast,
                        H_loop=np.stack((Hx, Hy, Hz), axis=3), Hk_fast=Hk_fast, Hk_loop=Hk, fHt=fHt, khat=khat,
                        coeff=coeff, fHt_fast=fHt_fast, p1=p1, p2=p2, p1f=p1f, p2f=p2f, rhat=rhat, A=A, B=B,
                        Hr_fast=Hr_fast, Hr_loop=Hr)


def main(ref_root, only=None):
    sys.path.insert(0, os.path.join(ref_root, "CPU - Ewald test"))
    import ewald  # noqa: E402  (reference module, read-only tree)
    import capacitance  # noqa: E402

    if only == "dropin":
        dropin_case(ewald)
        print("golden fixture dropin_n24.npz written to", HERE)
        return
    dropin_case(ewald)

    r_full = np.arange(0.001, 10.0 + 1e-12, 0.001, dtype=np.float64)
    sub = np.unique(np.concatenate((np.arange(0, 60), np.arange(1985, 2016), np.arange(0, r_full.size, 7),
                                    [r_full.size - 1])))
    for xi in (0.5, 1.0):
        tabs = ewald.real_space_table(r_full, xi)
        # index 0 of every table is the self term; r_full[k] sits at index k+1
        np.savez_compressed(os.path.join(HERE, f"tables_xi{xi:g}.npz"), xi=xi, r=r_full[sub],
                            self_terms=np.array([t[0] for t in tabs]),
                            tables=np.stack([t[1:][sub] for t in tabs]))

    r_table = np.concatenate(([0.0], r_full))

    def operator_case(name, n, phi, seed, xi, ngrid, P, rc, shift=0.0):
        x, box = rsa(n, phi, seed)
        x = x + shift
        rng = np.random.default_rng(seed + 1000)
        eps = 1.0 + 30.0 * rng.random(n)
        m = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
        Ngrid = np.array([ngrid] * 3)
        h = box / Ngrid
        eta = P * (h * xi) ** 2 / np.pi
        _, _, f1, f2, _, _ = ewald.real_space_table(r_full, xi)
        perp = np.asarray(f1)
        para = np.asarray(f2)
        cell, ncell = ewald.cell_list(x, box, rc)
        p1, p2 = ewald.neighbor_list(cell, ncell)
        offset, offsetxyz = ewald.pre_calculations(P, h)
        real = ewald.prepare_real_space_pairs(x, box, p1, p2, rc, perp, para, r_table)
        grid = ewald.prepare_grid_kernel(x, Ngrid, h, xi, eta, P, offset, offsetxyz)
        ks = ewald.precompute_kspace(Ngrid, box, xi, eta)
        out = {}
        for slab in (0.0, 1.0):
            Hk, Hr, Hs = ewald.magnetic_field_components(
                x, m, eps, box, p1, p2, Ngrid, h, P, xi, eta, rc, offset, offsetxyz, perp, para,
                r_table, real_precomp=real, grid_precomp=grid, k_precomp=ks, slab_factor=slab)
            out[f"Hk_slab{int(slab)}"] = Hk
            out[f"Hr_slab{int(slab)}"] = Hr
            out[f"Hslab_slab{int(slab)}"] = Hs
        np.savez_compressed(os.path.join(HERE, f"operator_{name}.npz"), x=x, box=box, eps=eps, m=m, xi=xi,
                            ngrid=Ngrid, P=P, rc=rc, npairs=real[0].size, **out)
        return x, box, eps, Ngrid, h, P, xi, eta, rc, perp, para, p1, p2, offset, offsetxyz, real, grid, ks

    operator_case("n64_g32", 64, 0.2, 1, 0.5, 32, 8, 8.0)
    c2 = operator_case("n64_g21", 64, 0.2, 2, 0.8, 21, 8, np.sqrt(-np.log(1e-5)) / 0.8)
    operator_case("n64_g21_centred", 64, 0.2, 3, 0.8, 21, 8, np.sqrt(-np.log(1e-5)) / 0.8, shift=-5.513)
    operator_case("n256_g47", 256, 0.25, 4, 1.0, 47, 9, np.sqrt(-np.log(1e-6)) / 1.0)
    operator_case("n512_g42", 512, 0.3, 5, 0.9, 42, 9, np.sqrt(-np.log(1e-6)) / 0.9)

    # tight-tolerance solves on the 21^3 case through both GMRES paths
    x, box, eps, Ngrid, h, P, xi, eta, rc, perp, para, p1, p2, offset, offsetxyz, real, grid, ks = c2
    n = x.shape[0]
    phi = 4.0 * np.pi * n / (3.0 * np.prod(box))
    for slab in (0.0, 1.0):
        for H in ((0.0, 0.0, 1.0), (1.0, 0.0, 0.0)):
            H = np.array(H)
            beta = (eps - 1.0) / (eps + 2.0)
            guess = (4.0 * np.pi * beta / (1.0 - beta * phi))[:, None] * H[None, :]
            sols = {}
            for use_scipy in (True, False):
                S = capacitance.magnetic_dipole(
                    x, eps, H, box, p1, p2, Ngrid, h, P, xi, eta, rc, guess, offset, offsetxyz, perp, para,
                    r_table, 1e-12, real_precomp=real, grid_precomp=grid, k_precomp=ks, slab_factor=slab,
                    use_scipy_gmres=use_scipy, verbose=0)
                sols["scipy" if use_scipy else "python"] = S
            tag = f"slab{int(slab)}_H{'xyz'[int(np.argmax(H))]}"
            np.savez_compressed(os.path.join(HERE, f"solve_n64_g21_{tag}.npz"), x=x, box=box, eps=eps, H=H,
                                guess=guess, tol=1e-12, p_scipy=sols["scipy"], p_python=sols["python"])

    # end-to-end run_capacitance at derived parameters (errortol 1e-5, xi 0.8 -> 21^3, P=8)
    x, box = rsa(64, 0.2, 6)
    eps = np.full(64, 4.0)
    for slab in (0.0, 1.0):
        C, p = capacitance.run_capacitance(x, box, eps, 0.8, np.array([0.0, 0.0, 1.0]), errortol=1e-5,
                                           slab_factor=slab, use_scipy_gmres=True, verbose=0)
        np.savez_compressed(os.path.join(HERE, f"run_n64_slab{int(slab)}.npz"), x=x, box=box, eps=eps,
                            xi=0.8, errortol=1e-5, H0=np.array([0.0, 0.0, 1.0]), C=C, p=p)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", sys.argv[2] if len(sys.argv) > 2 else None)
