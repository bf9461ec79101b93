// Vector-Jacobian product of the dipole operator with respect to positions
// and permittivities (gfx950), for energy gradients (forces) through an
// adjoint solve.
//
// For right-hand sides p and adjoint vectors lam, g = d/dx sum_rhs lam^T A(x) p.
// A's position dependence is the real-space pair sum and the Gaussian
// spread/contract weights (the slab and diagonal terms do not depend on x):
//   real space:  for every pair, grad_r [lam_i^T T(r) p_j + lam_j^T T(r) p_i]
//                with T = A(d) I + (B(d) - A(d)) rhat rhat^T, A, B the table
//                interpolants and A', B' their segment slopes (the exact
//                derivative of the interpolated operator);
//   wave space:  h^3 sum_nodes dW_k/dx [lam_k . Phi_p + p_k . Phi_lam], with
//                Phi_v the real-space field grid of v (spread, FFT, scale,
//                inverse FFT, before contraction) and
//                dW/dx_d = W 4 xi^2 r_d / eta_d for r = node - x.
#pragma once

#include "mpol_kernels.h"

namespace mpol {

// Real space: one thread per particle i, summing over its CSR neighbours j
// (both orientations of each pair, so no atomics).  Recomputes the pair
// separation with the FP64 operations of k_pairs.
static __global__ void k_real_space_vjp(PairParams p, const double* __restrict__ x, const int* __restrict__ off,
                                        const int* __restrict__ pj, const double* __restrict__ rv,
                                        const double* __restrict__ perp, const double* __restrict__ para,
                                        const double* __restrict__ m, const double* __restrict__ lam, int nrhs,
                                        double* __restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  for (int w = off[i]; w < off[i + 1]; ++w) {
    const int j = pj[w];
    double r0, r1, r2, d;
    {
#pragma clang fp contract(off)
      r0 = xi0 - x[3 * j];
      r1 = xi1 - x[3 * j + 1];
      r2 = xi2 - x[3 * j + 2];
      r0 = r0 - p.box[0] * trunc(2.0 * r0 / p.box[0]);
      r1 = r1 - p.box[1] * trunc(2.0 * r1 / p.box[1]);
      r2 = r2 - p.box[2] * trunc(2.0 * r2 / p.box[2]);
      d = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
    }
    int idx = table_lower_bound(rv, p.ntab, d);
    if (idx <= 0) idx = 1;
    if (idx >= p.ntab) idx = p.ntab - 1;
    const double lo = rv[idx - 1], hi = rv[idx];
    const double t = (d - lo) / (hi - lo);
    const double A = perp[idx - 1] * (1.0 - t) + perp[idx] * t;
    const double B = para[idx - 1] * (1.0 - t) + para[idx] * t;
    const double dA = (perp[idx] - perp[idx - 1]) / (hi - lo);
    const double dB = (para[idx] - para[idx - 1]) / (hi - lo);
    const double C = B - A, dC = dB - dA;
    const double u0 = r0 / d, u1 = r1 / d, u2 = r2 / d;
    for (int rhs = 0; rhs < nrhs; ++rhs) {
      const double* mr = m + (size_t)rhs * p.n * 3;
      const double* lr = lam + (size_t)rhs * p.n * 3;
      // both orientations: (lam_i, p_j) and (lam_j, p_i)
      for (int o = 0; o < 2; ++o) {
        const double* a = o == 0 ? lr + 3 * (size_t)i : lr + 3 * (size_t)j;
        const double* b = o == 0 ? mr + 3 * (size_t)j : mr + 3 * (size_t)i;
        const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        const double ar = a[0] * u0 + a[1] * u1 + a[2] * u2;
        const double br = b[0] * u0 + b[1] * u1 + b[2] * u2;
        // grad_r [A (a.b) + C (a.u)(b.u)]
        const double radial = dA * ab + dC * ar * br - 2.0 * C * ar * br / d;
        const double cd = C / d;
        g0 += radial * u0 + cd * (br * a[0] + ar * b[0]);
        g1 += radial * u1 + cd * (br * a[1] + ar * b[1]);
        g2 += radial * u2 + cd * (br * a[2] + ar * b[2]);
      }
    }
  }
  grad[3 * i] = g0;
  grad[3 * i + 1] = g1;
  grad[3 * i + 2] = g2;
}

// Real-space part of the forces of the Ewald-summed polarisation energy
// U = 1/2 p^T A p - b^T p at solved dipoles p (F = -dU/dx = -1/2 d(p^T A p)/dx),
// from the tabulated pair-force couplings f1 = dA/dr = (B - A)/r and
// f2 = d(B - A)/dr (field_dip_force_1/2 of "CPU - Ewald test/ewald.py":798-916,
// interpolated on the table grid like the couplings): with u = r_ij/d,
// a = p_i, b = p_j, the pair energy e = A (a.b) + (B - A)(a.u)(b.u) has
//   grad_r e = f1 [(a.b) u + (b.u) a + (a.u) b - 2 (a.u)(b.u) u] + f2 (a.u)(b.u) u
// and F_i = -sum_j grad_r e (each pair enters p^T A p twice).  Owner computes
// over the CSR pair list; summed over the right-hand sides.
static __global__ void k_ewald_forces_real(PairParams p, const double* __restrict__ x, const int* __restrict__ off,
                                           const int* __restrict__ pj, const double* __restrict__ rv,
                                           const double* __restrict__ f1t, const double* __restrict__ f2t,
                                           const double* __restrict__ m, int nrhs, double* __restrict__ force) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  for (int w = off[i]; w < off[i + 1]; ++w) {
    const int j = pj[w];
    double r0, r1, r2, d;
    {
#pragma clang fp contract(off)
      r0 = xi0 - x[3 * j];
      r1 = xi1 - x[3 * j + 1];
      r2 = xi2 - x[3 * j + 2];
      r0 = r0 - p.box[0] * trunc(2.0 * r0 / p.box[0]);
      r1 = r1 - p.box[1] * trunc(2.0 * r1 / p.box[1]);
      r2 = r2 - p.box[2] * trunc(2.0 * r2 / p.box[2]);
      d = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
    }
    int idx = table_lower_bound(rv, p.ntab, d);
    if (idx <= 0) idx = 1;
    if (idx >= p.ntab) idx = p.ntab - 1;
    const double lo = rv[idx - 1], hi = rv[idx];
    const double t = (d - lo) / (hi - lo);
    const double F1 = f1t[idx - 1] * (1.0 - t) + f1t[idx] * t;
    const double F2 = f2t[idx - 1] * (1.0 - t) + f2t[idx] * t;
    const double u0 = r0 / d, u1 = r1 / d, u2 = r2 / d;
    for (int rhs = 0; rhs < nrhs; ++rhs) {
      const double* a = m + 3 * ((size_t)rhs * p.n + i);
      const double* b = m + 3 * ((size_t)rhs * p.n + j);
      const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
      const double au = a[0] * u0 + a[1] * u1 + a[2] * u2;
      const double bu = b[0] * u0 + b[1] * u1 + b[2] * u2;
      const double radial = F1 * (ab - 2.0 * au * bu) + F2 * au * bu;
      g0 += radial * u0 + F1 * (bu * a[0] + au * b[0]);
      g1 += radial * u1 + F1 * (bu * a[1] + au * b[1]);
      g2 += radial * u2 + F1 * (bu * a[2] + au * b[2]);
    }
  }
  force[3 * i] = -g0;
  force[3 * i + 1] = -g1;
  force[3 * i + 2] = -g2;
}

struct WaveVjpArgs {
  int nx, ny, nz;
  int P, half;
  int n, nrhs;
  double h[3];
  double cexp[3];     // 2 xi^2 / eta_d
  double norm;        // (2 xi^2/pi)^{3/2} / sqrt(prod eta)
  double hvol;        // prod(h)
  const int4* rec;    // sorted: first stencil node x, y, z; caller's index
  const double4* base;  // sorted: node0*h - x per axis
};

// Wave space: one workgroup per particle (sorted order).  grid holds the real
// field grids of the 2*nrhs inputs [p_0..p_{R-1}, lam_0..lam_{R-1}] as
// grid[(v*3 + c)][z][y][x].  Adds scale * h^3 sum dW/dx (lam_k . Phi_p +
// p_k . Phi_lam) to grad (caller's order).
static __global__ void __launch_bounds__(256) k_wave_vjp(WaveVjpArgs a, const double* __restrict__ grid,
                                                         const double* __restrict__ m,
                                                         const double* __restrict__ lam,
                                                         double* __restrict__ grad, double scale) {
  __shared__ double red[3][4];
  const int s = blockIdx.x;
  if (s >= a.n) return;
  const int4 rc = a.rec[s];
  const double4 b = a.base[s];
  const int k = rc.w;
  const int P = a.P, P2 = P * P, P3 = P2 * P;
  const size_t G = (size_t)a.nx * a.ny * a.nz;
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
  for (int o = threadIdx.x; o < P3; o += blockDim.x) {
    const int ox = o % P, oy = (o / P) % P, oz = o / P2;
    const double r0 = b.x + (double)(ox - a.half) * a.h[0];
    const double r1 = b.y + (double)(oy - a.half) * a.h[1];
    const double r2 = b.z + (double)(oz - a.half) * a.h[2];
    const double W = a.norm * exp(-a.cexp[0] * r0 * r0) * exp(-a.cexp[1] * r1 * r1) * exp(-a.cexp[2] * r2 * r2);
    int gx = rc.x + ox, gy = rc.y + oy, gz = rc.z + oz;
    while (gx >= a.nx) gx -= a.nx;
    while (gy >= a.ny) gy -= a.ny;
    while (gz >= a.nz) gz -= a.nz;
    const size_t node = ((size_t)gz * a.ny + gy) * a.nx + gx;
    double s_val = 0.0;
    for (int rhs = 0; rhs < a.nrhs; ++rhs) {
      const double* pk = m + 3 * ((size_t)rhs * a.n + k);
      const double* lk = lam + 3 * ((size_t)rhs * a.n + k);
      const double* Fp = grid + (size_t)(rhs * 3) * G + node;
      const double* Fl = grid + (size_t)((a.nrhs + rhs) * 3) * G + node;
      s_val += lk[0] * Fp[0] + lk[1] * Fp[G] + lk[2] * Fp[2 * G] + pk[0] * Fl[0] + pk[1] * Fl[G] +
               pk[2] * Fl[2 * G];
    }
    const double f = W * s_val;
    // dW/dx_d = W * 2 cexp_d * r_d (r = node - x)
    acc0 += f * 2.0 * a.cexp[0] * r0;
    acc1 += f * 2.0 * a.cexp[1] * r1;
    acc2 += f * 2.0 * a.cexp[2] * r2;
  }
  for (int offs = 32; offs > 0; offs >>= 1) {
    acc0 += __shfl_xor(acc0, offs);
    acc1 += __shfl_xor(acc1, offs);
    acc2 += __shfl_xor(acc2, offs);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = acc0;
    red[1][wave] = acc1;
    red[2][wave] = acc2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[threadIdx.x][w];
    grad[3 * (size_t)k + threadIdx.x] += scale * (v * a.hvol);
  }
}

// Permittivity part: d/d eps_i of lam^T diag p with diag = -3/(4 pi (1 - eps))
// + self, per permittivity set: out[set][i] = sum over its rhs of
// lam_i . p_i * (-3/(4 pi (1 - eps)^2)).
static __global__ void k_eps_vjp(const double* __restrict__ eps, int n, int nsets, int nrhs,
                                 const double* __restrict__ m, const double* __restrict__ lam,
                                 double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)nsets * n) return;
  const int set = (int)(k / n), i = (int)(k % n);
  const double e = eps[k];
  const double dd = -3.0 / (4.0 * kPi * (1.0 - e) * (1.0 - e));
  double s = 0.0;
  for (int rhs = 0; rhs < nrhs; ++rhs) {
    if (nsets != 1 && rhs != set) continue;
    const double* pk = m + 3 * ((size_t)rhs * n + i);
    const double* lk = lam + 3 * ((size_t)rhs * n + i);
    s += lk[0] * pk[0] + lk[1] * pk[1] + lk[2] * pk[2];
  }
  out[k] = s * dd;
}

}  // namespace mpol
