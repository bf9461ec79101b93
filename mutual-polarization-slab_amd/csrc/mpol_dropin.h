// Kernels of the drop-in precompute and fast-path functions of
// "CPU - Ewald test/ewald.py" that take and return the reference's own
// intermediate arrays (linnode / w_spread / w_contract, complex grids in
// (Nx, Ny, Nz, 3) C order, khat / Htildecoeff, filtered pair lists).  The
// operator itself never materialises these arrays (mpol_kernels.h,
// mpol_mfma.h); these kernels exist so that a caller of the reference's
// step-by-step API gets the same arrays from the GPU.
//
// Arithmetic follows the reference expression by expression (same operand
// order, no fused multiply-add), so integer decisions (nearest node, cutoff,
// table interval) match it exactly and floating-point results match to
// rounding; sums the reference forms with bincount / add.at are formed here
// with atomics, so their order (and last bits) can differ.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mpol {

constexpr double kDropPi = 3.14159265358979323846;

// MATLAB round (ties away from zero): sign(x) * floor(|x| + 0.5) ([E]ewald.py:71-78)
__device__ inline double drop_round(double v) {
  const double s = v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0);
  return s * floor(fabs(v) + 0.5);
}

__device__ inline int64_t drop_mod(int64_t a, int64_t n) {
  int64_t r = a % n;
  return r < 0 ? r + n : r;
}

// first index k with rvals[k] >= d (numpy searchsorted, side "left")
__device__ inline int drop_lower_bound(const double* __restrict__ rvals, int ntab, double d) {
  int lo = 0, hi = ntab;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rvals[mid] < d) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ inline void drop_atomic_add(double* a, double v) { atomicAdd(a, v); }

// numpy.nan_to_num: NaN -> 0, +-inf -> +-DBL_MAX
__device__ inline double drop_finite(double v) {
  if (v != v) return 0.0;
  if (v == INFINITY) return 1.7976931348623157e308;
  if (v == -INFINITY) return -1.7976931348623157e308;
  return v;
}

struct DropGridParams {
  int64_t n, P3;
  int nx, ny, nz;
  double h[3], inv_eta[3];
  double m2xi2;  // -2 xi^2
  double cnst;   // (2 xi^2/pi)^(3/2) sqrt(1/prod(eta))
  double prodh;
};

// prepare_grid_kernel ([E]ewald.py:233-296): one thread per (particle,
// stencil point).  linnode = ix + Nx iy + Nx Ny iz of the node
// mod(node0 + offset - 1, Ngrid), w_spread = const exp(-2 xi^2 sum r^2/eta)
// with r = (node0 h - x) + offsetxyz, w_contract = w_spread prod(h).
static __global__ void k_drop_grid_kernel(DropGridParams p, const double* __restrict__ x,
                                          const int64_t* __restrict__ offset, const double* __restrict__ offxyz,
                                          int64_t* __restrict__ linnode, double* __restrict__ wsp,
                                          double* __restrict__ wct) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= p.n * p.P3) return;
  const int64_t i = idx / p.P3, k = idx - i * p.P3;
  const int64_t ng[3] = {p.nx, p.ny, p.nz};
  int64_t node[3];
  double s = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double xd = x[3 * i + d];
    const double n0 = drop_round(xd / p.h[d]);
    node[d] = drop_mod((int64_t)n0 + offset[3 * k + d] - 1, ng[d]);
    const double r = (n0 * p.h[d] - xd) + offxyz[3 * k + d];
    const double t = (r * r) * p.inv_eta[d];
    s = (d == 0) ? t : s + t;
  }
  linnode[idx] = node[0] + ng[0] * node[1] + (ng[0] * ng[1]) * node[2];
  const double w = p.cnst * exp(p.m2xi2 * s);
  wsp[idx] = w;
  wct[idx] = w * p.prodh;
}

__device__ inline int64_t drop_grid_addr(int64_t l, int64_t nx, int64_t ny, int64_t nz) {
  const int64_t ix = l % nx, iy = (l / nx) % ny, iz = l / (nx * ny);
  return (ix * ny + iy) * nz + iz;  // (Nx, Ny, Nz) C order
}

// spread_fast ([E]ewald.py:299-331): H[linnode, c] += w m_c, complex grid
// (Nx, Ny, Nz, 3) C order as (re, im) pairs; linnode outside the grid sets err.
static __global__ void k_drop_spread(int64_t n, int64_t P3, int nx, int ny, int nz, const int64_t* __restrict__ linnode,
                                     const double* __restrict__ w, const double2* __restrict__ m,
                                     double2* __restrict__ H, int* __restrict__ err) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * P3) return;
  const int64_t i = idx / P3;
  const int64_t l = linnode[idx];
  const int64_t G = (int64_t)nx * ny * nz;
  if (l < 0 || l >= G) {
    atomicOr(err, 1);
    return;
  }
  const double wk = w[idx];
  const int64_t a = drop_grid_addr(l, nx, ny, nz) * 3;
  for (int c = 0; c < 3; ++c) {
    const double2 mc = m[3 * i + c];
    drop_atomic_add(&H[a + c].x, wk * mc.x);
    drop_atomic_add(&H[a + c].y, wk * mc.y);
  }
}

// contract_fast ([E]ewald.py:334-368): Hk[n, c] = sum_k w[n, k] H[linnode[n, k], c];
// one 64-lane wave per particle, a fixed butterfly over the lanes.
static __global__ void k_drop_contract(int64_t n, int64_t P3, int nx, int ny, int nz,
                                       const int64_t* __restrict__ linnode, const double* __restrict__ w,
                                       const double2* __restrict__ H, double2* __restrict__ Hk, int* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int64_t G = (int64_t)nx * ny * nz;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t k = lane; k < P3; k += 64) {
    const int64_t l = linnode[i * P3 + k];
    if (l < 0 || l >= G) {
      atomicOr(err, 1);
      continue;
    }
    const double wk = w[i * P3 + k];
    const int64_t a = drop_grid_addr(l, nx, ny, nz) * 3;
    for (int c = 0; c < 3; ++c) {
      const double2 v = H[a + c];
      acc[2 * c] += wk * v.x;
      acc[2 * c + 1] += wk * v.y;
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int q = 0; q < 6; ++q) acc[q] += __shfl_xor(acc[q], off, 64);
  if (lane == 0)
    for (int c = 0; c < 3; ++c) Hk[3 * i + c] = make_double2(acc[2 * c], acc[2 * c + 1]);
}

// J_{3/2}(x) = sqrt(2/(pi x)) (sin x/x - cos x), 0 at x = 0 ([E]ewald.py:51-65)
__device__ inline double drop_j32(double x) {
  if (x == 0.0) return 0.0;
  return sqrt(2.0 / (kDropPi * x)) * (sin(x) / x - cos(x));
}

// Htildecoeff of Scale.m for wave vector k ([E]ewald.py:500-507, 571-576):
// 9 pi/(2 |k|) J_{3/2}(|k|)^2 exp(-sum (1 - eta_i) k_i^2/(4 xi^2)) / |k|^2
__device__ inline double drop_coeff(const double k[3], double k2, const double one_m_eta[3], double four_xi2) {
  const double etak2 = (one_m_eta[0] * (k[0] * k[0]) + one_m_eta[1] * (k[1] * k[1])) + one_m_eta[2] * (k[2] * k[2]);
  const double sq = sqrt(k2);
  const double J = drop_j32(sq);
  return (9.0 * kDropPi) / (2.0 * sq) * (J * J) * exp(-etak2 / four_xi2) / k2;
}

struct DropKParams {
  int nx, ny, nz;
  int k0x, k0y, k0z;     // the k = 0 entry in fftshift order: ceil((N-1)/2)
  double one_m_eta[3];   // 1 - eta
  double four_xi2;       // 4 xi^2
  double kunit[3];       // 2 pi / box
};

// scale ([E]ewald.py:473-516) for a given k grid (Nx, Ny, Nz, 3): the same
// khat and coefficient as precompute_kspace, then fHtilde = coeff (khat . fH)
// khat, zero at the k = 0 entry, nan_to_num.
static __global__ void k_drop_scale_k(DropKParams p, const double* __restrict__ kv, const double2* __restrict__ fH,
                                      double2* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t G = (int64_t)p.nx * p.ny * p.nz;
  if (g >= G) return;
  const int iz = (int)(g % p.nz), iy = (int)((g / p.nz) % p.ny), ix = (int)(g / ((int64_t)p.nz * p.ny));
  const bool zero = (ix == p.k0x && iy == p.k0y && iz == p.k0z);
  const double k[3] = {kv[3 * g], kv[3 * g + 1], kv[3 * g + 2]};
  const double k2 = (k[0] * k[0] + k[1] * k[1]) + k[2] * k[2];
  const double sq = sqrt(k2);
  const double den = sq == 0.0 ? 1.0 : sq;
  double kh[3];
  for (int c = 0; c < 3; ++c) kh[c] = zero ? 0.0 : k[c] / den;
  const double coeff = zero ? 0.0 : drop_coeff(k, k2, p.one_m_eta, p.four_xi2);
  double dre = 0.0, dim = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double2 f = fH[3 * g + c];
    dre = (c == 0) ? kh[c] * f.x : dre + kh[c] * f.x;
    dim = (c == 0) ? kh[c] * f.y : dim + kh[c] * f.y;
  }
  const double are = coeff * dre, aim = coeff * dim;
  for (int c = 0; c < 3; ++c) {
    double2 o = make_double2(are * kh[c], aim * kh[c]);
    if (zero) o = make_double2(0.0, 0.0);
    out[3 * g + c] = make_double2(drop_finite(o.x), drop_finite(o.y));
  }
}

// precompute_kspace ([E]ewald.py:523-582): k on the fftshift-ordered MATLAB
// grid, khat = k/|k| (0 at k = 0, complex with zero imaginary part) and the
// coefficient (0 at k = 0, nan_to_num).
static __global__ void k_drop_kspace(DropKParams p, double2* __restrict__ khat, double* __restrict__ coeff) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t G = (int64_t)p.nx * p.ny * p.nz;
  if (g >= G) return;
  const int iz = (int)(g % p.nz), iy = (int)((g / p.nz) % p.ny), ix = (int)(g / ((int64_t)p.nz * p.ny));
  const bool zero = (ix == p.k0x && iy == p.k0y && iz == p.k0z);
  const double k[3] = {(double)(ix - p.k0x) * p.kunit[0], (double)(iy - p.k0y) * p.kunit[1],
                       (double)(iz - p.k0z) * p.kunit[2]};
  const double k2 = (k[0] * k[0] + k[1] * k[1]) + k[2] * k[2];
  const double sq = sqrt(k2);
  const double den = sq == 0.0 ? 1.0 : sq;
  for (int c = 0; c < 3; ++c) khat[3 * g + c] = make_double2(zero ? 0.0 : k[c] / den, 0.0);
  coeff[g] = zero ? 0.0 : drop_finite(drop_coeff(k, k2, p.one_m_eta, p.four_xi2));
}

// scale_fast ([E]ewald.py:585-599): fHtilde_c = (coeff (khat . fH)) khat_c in
// complex arithmetic.
static __global__ void k_drop_scale_fast(int64_t G, const double2* __restrict__ fH, const double2* __restrict__ khat,
                                         const double* __restrict__ coeff, double2* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  double dre = 0.0, dim = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double2 a = khat[3 * g + c], f = fH[3 * g + c];
    const double pre = a.x * f.x - a.y * f.y, pim = a.x * f.y + a.y * f.x;
    dre = (c == 0) ? pre : dre + pre;
    dim = (c == 0) ? pim : dim + pim;
  }
  const double cf = coeff[g];
  const double sre = cf * dre, sim = cf * dim;
  for (int c = 0; c < 3; ++c) {
    const double2 a = khat[3 * g + c];
    out[3 * g + c] = make_double2(sre * a.x - sim * a.y, sre * a.y + sim * a.x);
  }
}

struct DropPairParams {
  int64_t m;   // candidate pairs
  int ntab;
  double box[3];
  double rc;
};

// minimum image r = x[p1] - x[p2] - box trunc(2 r/box) and d = |r|
__device__ inline double drop_pair_r(const DropPairParams& p, const double* __restrict__ x, int64_t i1, int64_t i2,
                                     double r[3]) {
  for (int c = 0; c < 3; ++c) {
    const double v = x[3 * i1 + c] - x[3 * i2 + c];
    r[c] = v - p.box[c] * trunc((2.0 * v) / p.box[c]);
  }
  return sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
}

// prepare_real_space_pairs ([E]ewald.py:939-1016), pass 1: keep[q] = d < rc.
static __global__ void k_drop_pairs_keep(DropPairParams p, const double* __restrict__ x, const int64_t* __restrict__ p1,
                                         const int64_t* __restrict__ p2, int* __restrict__ keep) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.m) return;
  double r[3];
  keep[q] = drop_pair_r(p, x, p1[q], p2[q], r) < p.rc ? 1 : 0;
}

// pass 2: kept pairs in their input order, r hat = r/d, the table interval
// (first rvals >= d) and the interpolated couplings A, B with
// t = (d - rlow)/(rhigh - rlow); an interval outside the table sets err.
static __global__ void k_drop_pairs_fill(DropPairParams p, const double* __restrict__ x,
                                         const int64_t* __restrict__ p1, const int64_t* __restrict__ p2,
                                         const int* __restrict__ keep, const int* __restrict__ pos,
                                         const double* __restrict__ perp, const double* __restrict__ para,
                                         const double* __restrict__ rvals, int64_t* __restrict__ p1f,
                                         int64_t* __restrict__ p2f, double* __restrict__ rhat, double* __restrict__ A,
                                         double* __restrict__ B, int* __restrict__ err) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.m || !keep[q]) return;
  double r[3];
  const double d = drop_pair_r(p, x, p1[q], p2[q], r);
  const int64_t o = pos[q];
  p1f[o] = p1[q];
  p2f[o] = p2[q];
  for (int c = 0; c < 3; ++c) rhat[3 * o + c] = r[c] / d;
  const int k = drop_lower_bound(rvals, p.ntab, d);
  if (k <= 0 || k >= p.ntab) {
    atomicOr(err, 1);
    A[o] = 0.0;
    B[o] = 0.0;
    return;
  }
  const double rlow = rvals[k - 1], rhigh = rvals[k];
  const double t = (d - rlow) / (rhigh - rlow);
  A[o] = perp[k - 1] * (1.0 - t) + perp[k] * t;
  B[o] = para[k - 1] * (1.0 - t) + para[k] * t;
}

// The diagonal of RealSpace.m: Hr = -3/(4 pi (1 - lambda)) m + m self
// ([E]ewald.py:1061-1062, 1095-1096).
static __global__ void k_drop_rs_diag(int64_t n, const double2* __restrict__ m, const double* __restrict__ lam,
                                      double self_perp, double2* __restrict__ Hr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c = -3.0 / ((4.0 * kDropPi) * (1.0 - lam[i]));
  for (int k = 0; k < 3; ++k) {
    const double2 v = m[3 * i + k];
    Hr[3 * i + k] = make_double2(c * v.x + v.x * self_perp, c * v.y + v.y * self_perp);
  }
}

// real_space_fast ([E]ewald.py:1019-1072): Hr[p1] += A m_j + (B - A)(m_j . r hat) r hat
static __global__ void k_drop_rs_fast(int64_t m_pairs, const double2* __restrict__ m, const int64_t* __restrict__ p1,
                                      const int64_t* __restrict__ p2, const double* __restrict__ rhat,
                                      const double* __restrict__ A, const double* __restrict__ B,
                                      double2* __restrict__ Hr) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m_pairs) return;
  const int64_t i1 = p1[q], i2 = p2[q];
  const double rh[3] = {rhat[3 * q], rhat[3 * q + 1], rhat[3 * q + 2]};
  double dre = 0.0, dim = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double2 v = m[3 * i2 + c];
    dre = (c == 0) ? v.x * rh[c] : dre + v.x * rh[c];
    dim = (c == 0) ? v.y * rh[c] : dim + v.y * rh[c];
  }
  const double a = A[q], ba = B[q] - a;
  const double bre = ba * dre, bim = ba * dim;
  for (int c = 0; c < 3; ++c) {
    const double2 v = m[3 * i2 + c];
    drop_atomic_add(&Hr[3 * i1 + c].x, a * v.x + bre * rh[c]);
    drop_atomic_add(&Hr[3 * i1 + c].y, a * v.y + bim * rh[c]);
  }
}

// real_space ([E]ewald.py:1076-1139), the pair loop: d = 0 skipped, d < rc
// kept, the table interval from the first rvals >= d (interval 0 skipped,
// past the table sets err), A and B interpolated as
// perp[k-1] (rhigh - d)/(rhigh - rlow) + perp[k] (d - rlow)/(rhigh - rlow),
// Hr[p1] += A (m_j - (m_j . r hat) r hat) + B (m_j . r hat) r hat.
static __global__ void k_drop_rs_direct(DropPairParams p, const double* __restrict__ x, const double2* __restrict__ m,
                                        const int64_t* __restrict__ p1, const int64_t* __restrict__ p2,
                                        const double* __restrict__ perp, const double* __restrict__ para,
                                        const double* __restrict__ rvals, double2* __restrict__ Hr,
                                        int* __restrict__ err) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.m) return;
  const int64_t i1 = p1[q], i2 = p2[q];
  double r[3];
  const double d = drop_pair_r(p, x, i1, i2, r);
  if (d == 0.0 || !(d < p.rc)) return;
  double rh[3];
  for (int c = 0; c < 3; ++c) rh[c] = r[c] / d;
  const int k = drop_lower_bound(rvals, p.ntab, d);
  if (k <= 0) return;
  if (k >= p.ntab) {
    atomicOr(err, 1);
    return;
  }
  const double rhigh = rvals[k], rlow = rvals[k - 1];
  const double A = perp[k - 1] * (rhigh - d) / (rhigh - rlow) + perp[k] * (d - rlow) / (rhigh - rlow);
  const double B = para[k - 1] * (rhigh - d) / (rhigh - rlow) + para[k] * (d - rlow) / (rhigh - rlow);
  double dre = 0.0, dim = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double2 v = m[3 * i2 + c];
    dre = (c == 0) ? v.x * rh[c] : dre + v.x * rh[c];
    dim = (c == 0) ? v.y * rh[c] : dim + v.y * rh[c];
  }
  for (int c = 0; c < 3; ++c) {
    const double2 v = m[3 * i2 + c];
    const double pre = dre * rh[c], pim = dim * rh[c];
    drop_atomic_add(&Hr[3 * i1 + c].x, A * (v.x - pre) + B * pre);
    drop_atomic_add(&Hr[3 * i1 + c].y, A * (v.y - pim) + B * pim);
  }
}

}  // namespace mpol
