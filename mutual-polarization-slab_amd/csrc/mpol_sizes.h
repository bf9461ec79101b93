// Polydisperse dipole operator (spheres of different radii), gfx950.
//
// Method of Capacitance_Sizes/ (ComputeField.m:22-57): the dipoles are spread
// "as charges" onto ONE scalar grid per right-hand side through the
// radius-dependent kernel
//   K(d; a) = 3 / (8 sqrt2 pi^{3/2} a^3 sqrt(eta) xi d^2)
//             [ (eta + 4 a xi^2 d) e^{-2 (d+a)^2 xi^2/eta} - (eta - 4 a xi^2 d) e^{-2 (d-a)^2 xi^2/eta} ]
// with grid value K (r.p)/d for r = node - x (Spread.m:54-61), the spectrum is
// scaled by exp(-(1-eta) q^2/(4 xi^2))/q^2 (Scale.m:17-21) and contracted with
// the vector kernel K r/d times prod(h) (Contract.m:54-71).  The real-space
// pair couplings (I - rr and rr components) depend on both radii
// (RealSpaceTable.m:34-135); here they are either interpolated from
// per-radius-pair tables (the reference's way) or evaluated in closed form
// per pair (any radius distribution).
#pragma once

#include "mpol_kernels.h"

namespace mpol {

// ---------------------------------------------------------------------------
// closed-form real-space couplings
// ---------------------------------------------------------------------------
struct SizesPoly {
  // ascending coefficients of r^0 .. r^6 (unused ones zero)
  double c[7];
  __host__ __device__ double eval(double r) const {
    double v = c[6];
    for (int k = 5; k >= 0; --k) v = v * r + c[k];
    return v;
  }
};

// (perp, para) at separation r > 0 for radii ai, aj.  Each component is
//   sum over the four shifts u in {a_i+a_j, -(a_i+a_j), a_j-a_i, a_i-a_j} of
//   E_u(r) exp(-(r+u)^2 xi^2) + F_u(r) erfc((r+u) xi),   plus the overlap
// regularisation, with E_u = poly/(r^3 xi^5), F_u = poly/(r^3 xi^6) scaled.
__host__ __device__ inline void sizes_coupling(double r, double ai, double aj, double x, double& perp,
                                               double& para, bool regularize = true) {
  const double pi = kPi, p32 = pi * sqrt(pi);
  const double s = ai + aj, d = ai - aj;
  const double x2 = x * x, x4 = x2 * x2, x6 = x4 * x2;
  const double ai2 = ai * ai, aj2 = aj * aj, ai3 = ai2 * ai, aj3 = aj2 * aj;
  const double c3p = 4 * ai3 - 3 * ai2 * aj - 3 * ai * aj2 + 4 * aj3;
  const double c3m = 4 * ai3 + 3 * ai2 * aj - 3 * ai * aj2 - 4 * aj3;
  const double b3p = ai3 - 3 * ai2 * aj - 3 * ai * aj2 + aj3;
  const double b3m = ai3 + 3 * ai2 * aj - 3 * ai * aj2 - aj3;
  const double qp = ai2 - 4 * ai * aj + aj2, qm = ai2 + 4 * ai * aj + aj2;
  const double sq = ai2 + aj2, dq = ai2 - aj2;
  const double r3 = r * r * r;
  const double ex[4] = {exp(-(r + s) * (r + s) * x2), exp(-(r - s) * (r - s) * x2), exp(-(r - d) * (r - d) * x2),
                        exp(-(r + d) * (r + d) * x2)};
  const double ef[4] = {erfc((r + s) * x), erfc((r - s) * x), erfc((r - d) * x), erfc((r + d) * x)};
  const double den = ai3 * aj3 * r3;
  // ---- I - rr ----
  {
    const double lin_p = 3 - 12 * (ai2 - ai * aj + aj2) * x2 - 4 * s * s * qp * x4;
    const double lin_m = -3 + 12 * (ai2 + ai * aj + aj2) * x2 + 4 * d * d * qm * x4;
    const double cube_p = 16 * x2 + 8 * (-4 * ai2 + ai * aj - 4 * aj2) * x4;
    const double cube_m = -16 * x2 + 8 * (4 * ai2 + ai * aj + 4 * aj2) * x4;
    const double k0p = 3 * s + 4 * c3p * x2 + 4 * s * s * s * qp * x4;
    const double k0m = 3 * d + 4 * c3m * x2 + 4 * d * d * d * qm * x4;
    const SizesPoly e1 = {{k0p, lin_p, -12 * s * x2 - 8 * c3p * x4, cube_p, -4 * s * x4, 4 * x4, 0}};
    const SizesPoly e2 = {{-k0p, lin_p, 12 * s * x2 + 8 * c3p * x4, cube_p, 4 * s * x4, 4 * x4, 0}};
    const SizesPoly e3 = {{k0m, lin_m, -12 * d * x2 - 8 * c3m * x4, cube_m, -4 * d * x4, -4 * x4, 0}};
    const SizesPoly e4 = {{-k0m, lin_m, 12 * d * x2 + 8 * c3m * x4, cube_m, 4 * d * x4, -4 * x4, 0}};
    const double r4c = 36 * x4 * (-1 + 2 * sq * x2);
    const double r2c = -18 * x2 + 72 * sq * x4 + 72 * dq * dq * x6;
    const double k0 = 3 - 18 * sq * x2 - 36 * dq * dq * x4;
    const SizesPoly f1 = {{k0 - 8 * s * s * s * s * qp * x6, 0, r2c, 128 * (ai3 + aj3) * x6, r4c, 0, -8 * x6}};
    const SizesPoly f2 = {{k0 - 8 * s * s * s * s * qp * x6, 0, r2c, -128 * (ai3 + aj3) * x6, r4c, 0, -8 * x6}};
    const SizesPoly f3 = {{-k0 + 8 * d * d * d * d * qm * x6, 0, -r2c, 128 * (ai3 - aj3) * x6, -r4c, 0, 8 * x6}};
    const SizesPoly f4 = {{-k0 + 8 * d * d * d * d * qm * x6, 0, -r2c, -128 * (ai3 - aj3) * x6, -r4c, 0, 8 * x6}};
    const double ce = 1.0 / (1024 * p32 * den * x4 * x), cf = 1.0 / (2048 * pi * den * x6);
    double v = ce * (e1.eval(r) * ex[0] + e2.eval(r) * ex[1] + e3.eval(r) * ex[2] + e4.eval(r) * ex[3]) +
               cf * (f1.eval(r) * ef[0] + f2.eval(r) * ef[1] + f3.eval(r) * ef[2] + f4.eval(r) * ef[3]);
    if (!regularize) {
      // field-map tables: no overlap terms (Compute_Field_Sizes/RealSpaceTable.m)
    } else if (r < s && r >= d && r >= -d) {
      const double t = s - r;
      v += t * t * t / (128 * pi * den) * (-r3 - 3 * s * r * r + 3 * qp * r + ai3 - 3 * ai2 * aj - 3 * ai * aj2 + aj3);
    } else if (r < -d && r > d) {
      v += (r3 - aj3) / (4 * pi * aj3 * r3);
    } else if (r < d && r > -d) {
      v += (r3 - ai3) / (4 * pi * ai3 * r3);
    }
    perp = v;
  }
  // ---- rr ----
  {
    const double lin_p = -3 + 12 * (ai2 - ai * aj + aj2) * x2 + 4 * s * s * qp * x4;
    const double lin_m = 3 - 12 * (ai2 + ai * aj + aj2) * x2 - 4 * d * d * qm * x4;
    const double cube_p = 14 * x2 - 4 * (7 * ai2 - 4 * ai * aj + 7 * aj2) * x4;
    const double cube_m = -14 * x2 + 4 * (7 * ai2 + 4 * ai * aj + 7 * aj2) * x4;
    const double k0p = 3 * s + 4 * c3p * x2 + 4 * s * s * s * qp * x4;
    const double k0m = 3 * d + 4 * c3m * x2 + 4 * d * d * d * qm * x4;
    const SizesPoly e1 = {{-k0p, lin_p, -6 * s * x2 - 4 * b3p * x4, cube_p, -8 * s * x4, 8 * x4, 0}};
    const SizesPoly e2 = {{k0p, lin_p, 6 * s * x2 + 4 * b3p * x4, cube_p, 8 * s * x4, 8 * x4, 0}};
    const SizesPoly e3 = {{-k0m, lin_m, -6 * d * x2 - 4 * b3m * x4, cube_m, -8 * d * x4, -8 * x4, 0}};
    const SizesPoly e4 = {{k0m, lin_m, 6 * d * x2 + 4 * b3m * x4, cube_m, 8 * d * x4, -8 * x4, 0}};
    const double r4c = 36 * x4 * (-1 + 2 * sq * x2);
    const double k0 = -3 + 18 * sq * x2 + 36 * dq * dq * x4;
    const SizesPoly f1 = {{k0 + 8 * s * s * s * s * qp * x6, 0, 0, 64 * (ai3 + aj3) * x6, r4c, 0, -16 * x6}};
    const SizesPoly f2 = {{k0 + 8 * s * s * s * s * qp * x6, 0, 0, -64 * (ai3 + aj3) * x6, r4c, 0, -16 * x6}};
    const SizesPoly f3 = {{-k0 - 8 * d * d * d * d * qm * x6, 0, 0, 64 * (ai3 - aj3) * x6, -r4c, 0, 16 * x6}};
    const SizesPoly f4 = {{-k0 - 8 * d * d * d * d * qm * x6, 0, 0, -64 * (ai3 - aj3) * x6, -r4c, 0, 16 * x6}};
    const double ce = 1.0 / (512 * p32 * den * x4 * x), cf = 1.0 / (1024 * pi * den * x6);
    double v = ce * (e1.eval(r) * ex[0] + e2.eval(r) * ex[1] + e3.eval(r) * ex[2] + e4.eval(r) * ex[3]) +
               cf * (f1.eval(r) * ef[0] + f2.eval(r) * ef[1] + f3.eval(r) * ef[2] + f4.eval(r) * ef[3]);
    if (!regularize) {
    } else if (r < s && r >= d && r >= -d) {
      const double r2 = r * r;
      v += -1.0 / (64 * pi * den) * (s * s * s * s * qp - 8 * (ai3 + aj3) * r3 + 9 * sq * r2 * r2 - 2 * r3 * r3);
    } else if (r < -d && r > d) {
      v += (r3 + 2 * aj3) / (4 * pi * aj3 * r3);
    } else if (r < d && r > -d) {
      v += (r3 + 2 * ai3) / (4 * pi * ai3 * r3);
    }
    para = v;
  }
}

// The r -> 0 value of both components (RealSpaceTable.m:96-99).
__host__ __device__ inline double sizes_self(double ai, double aj, double x) {
  const double pi = kPi, p32 = pi * sqrt(pi);
  const double x2 = x * x, ai3 = ai * ai * ai, aj3 = aj * aj * aj;
  return 1.0 / (16 * p32 * ai3 * aj3 * x2 * x) *
             ((1 - 2 * ai * ai * x2 + 2 * ai * aj * x2 - 2 * aj * aj * x2) * exp(-(ai + aj) * (ai + aj) * x2) +
              (-1 + 2 * ai * ai * x2 + 2 * ai * aj * x2 + 2 * aj * aj * x2) * exp(-(ai - aj) * (ai - aj) * x2)) +
         1.0 / (8 * pi * ai3 * aj3) * ((ai3 - aj3) * erf((ai - aj) * x) - (ai3 + aj3) * erf((ai + aj) * x)) +
         fmin(ai3, aj3) / (4 * pi * ai3 * aj3);
}

// ---------------------------------------------------------------------------
// pair coefficients and diagonal
// ---------------------------------------------------------------------------
// Per CSR pair: A = perp, B - A = para - perp, from the tables of the radius
// types (ntypes > 0; column col[ti*M+tj] of perp/para, each ntab long on rv)
// or from the closed form with the particles' own radii (ntypes == 0).
// Same distance decisions as k_pairs (FP64, no contraction).
static __global__ void k_pairs_sizes(PairParams p, const double* __restrict__ x, const int* __restrict__ off,
                              const int* __restrict__ pj, const double* __restrict__ rv,
                              const double* __restrict__ radius, const int* __restrict__ type, int ntypes,
                              const int* __restrict__ col, const double* __restrict__ tperp,
                              const double* __restrict__ tpara, double xi, double* __restrict__ pA,
                              double* __restrict__ pBA, float* __restrict__ pA32, float* __restrict__ pBA32) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  for (int w = off[i]; w < off[i + 1]; ++w) {
    const int j = pj[w];
    double r0 = xi0 - x[3 * j], r1 = xi1 - x[3 * j + 1], r2 = xi2 - x[3 * j + 2];
    r0 = r0 - p.box[0] * trunc(2.0 * r0 / p.box[0]);
    r1 = r1 - p.box[1] * trunc(2.0 * r1 / p.box[1]);
    r2 = r2 - p.box[2] * trunc(2.0 * r2 / p.box[2]);
    const double d = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
    double A, B;
    if (ntypes > 0) {
      int idx = table_lower_bound(rv, p.ntab, d);
      if (idx <= 0) idx = 1;
      if (idx >= p.ntab) idx = p.ntab - 1;
      const double lo = rv[idx - 1], hi = rv[idx];
      const double wl = (hi - d) / (hi - lo), wh = (d - lo) / (hi - lo);
      const int c = col[type[i] * ntypes + type[j]];
      const double* tp = tperp + (size_t)c * p.ntab;
      const double* tq = tpara + (size_t)c * p.ntab;
      A = tp[idx - 1] * wl + tp[idx] * wh;
      B = tq[idx - 1] * wl + tq[idx] * wh;
    } else {
      sizes_coupling(d, radius[i], radius[j], xi, A, B);
    }
    pA[w] = A;
    pBA[w] = B - A;
    pA32[w] = (float)A;
    pBA32[w] = (float)(B - A);
  }
}

// diag = 3/(4 pi a^3 (eps - 1)) + self coupling (RealSpace.m:36-40); complex
// eps gives the imaginary part too.
static __global__ void k_diag_sizes(const double* __restrict__ er, const double* __restrict__ ei, int n, int64_t total,
                             const double* __restrict__ radius, const int* __restrict__ type, int ntypes,
                             const int* __restrict__ col, const double* __restrict__ tperp, int ntab, double xi,
                             double* __restrict__ dre, double* __restrict__ dim) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const int i = (int)(k % n);
  const double a = radius[i];
  const double self = ntypes > 0 ? tperp[(size_t)col[type[i] * ntypes + type[i]] * ntab] : sizes_self(a, a, xi);
  const double coef = 3.0 / (4.0 * kPi * a * a * a);
  const double aa = 1.0 - er[k], bb = ei ? ei[k] : 0.0;
  if (ei) {
    const double inv = 1.0 / (aa * aa + bb * bb);
    dre[k] = -coef * aa * inv + self;
    dim[k] = -coef * bb * inv;
  } else {
    dre[k] = -coef / aa + self;
  }
}

static __global__ void k_permute_scalar(const double* __restrict__ v, const int* __restrict__ perm, int n,
                                 double* __restrict__ vs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) vs[s] = v[perm[s]];
}

// ---------------------------------------------------------------------------
// spread / contract with the radius-dependent kernel
// ---------------------------------------------------------------------------
struct SizesGridArgs {
  int nx, ny, nz;
  int P, half;
  int n, nrhs;
  double h[3];
  double xi, eta;
  double hvol;        // prod(h)
  const int4* rec;    // sorted: first stencil node x, y, z; caller's index
  const double4* base;  // sorted: node0*h - x per axis
  const double* rad;  // sorted radii
  const int* gate;    // optional per-system state (rhs_idle): idle right-hand sides are skipped
  int gdiv;
};

// Weight w with grid value w (r.p) for r = node - x: K(d)/d, or the
// reference's small-d form below d = 1e-6 (written with exp(-2 a xi^2/eta)).
template <typename T>
__device__ inline T sizes_weight(T d, T a, T xi, T eta) {
  const T pi = (T)kPi;
  if (d < (T)1e-6) {
    return (T)8 * sqrt((T)2) * xi * xi * xi * xi * xi * exp(-(T)2 * a * xi * xi / eta) /
           (pi * sqrt(pi) * eta * eta * sqrt(eta));
  }
  const T c = (T)2 * xi * xi / eta;
  const T K = (T)3 / ((T)8 * sqrt((T)2) * pi * sqrt(pi) * a * a * a * sqrt(eta) * xi * d * d) *
              ((eta + (T)4 * a * xi * xi * d) * exp(-c * (d + a) * (d + a)) -
               (eta - (T)4 * a * xi * xi * d) * exp(-c * (d - a) * (d - a)));
  return K / d;
}

// One workgroup per particle (sorted order): every node of its P^3 stencil
// gets w (r.p) for each right-hand side (global atomics into the real grid
// grid[rhs][z][y][x]).
template <typename T>
__global__ void __launch_bounds__(256) k_sizes_spread(SizesGridArgs a, const double* __restrict__ ms,
                                                      T* __restrict__ grid) {
  const int s = blockIdx.x;
  if (s >= a.n) return;
  const int4 rc = a.rec[s];
  const double4 b = a.base[s];
  const T rad = (T)a.rad[s];
  const int P = a.P, P2 = P * P, P3 = P2 * P;
  const size_t plane = (size_t)a.nx * a.ny * a.nz;
  for (int o = threadIdx.x; o < P3; o += blockDim.x) {
    const int ox = o % P, oy = (o / P) % P, oz = o / P2;
    const T r0 = (T)(b.x + (double)(ox - a.half) * a.h[0]);
    const T r1 = (T)(b.y + (double)(oy - a.half) * a.h[1]);
    const T r2 = (T)(b.z + (double)(oz - a.half) * a.h[2]);
    const T d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    const T w = sizes_weight<T>(d, rad, (T)a.xi, (T)a.eta);
    int gx = rc.x + ox, gy = rc.y + oy, gz = rc.z + oz;
    gx -= (gx >= a.nx) ? a.nx : 0;  // P <= n on each axis: one wrap at most
    gy -= (gy >= a.ny) ? a.ny : 0;
    gz -= (gz >= a.nz) ? a.nz : 0;
    const size_t node = ((size_t)gz * a.ny + gy) * a.nx + gx;
    for (int rhs = 0; rhs < a.nrhs; ++rhs) {
      if (rhs_idle(a.gate, a.gdiv, rhs)) continue;
      const double* p = ms + 3 * ((size_t)rhs * a.n + s);
      atomicAdd(grid + rhs * plane + node, w * (r0 * (T)p[0] + r1 * (T)p[1] + r2 * (T)p[2]));
    }
  }
}

// ---------------------------------------------------------------------------
// node-major spread (a gather: each node written once, no atomics)
// ---------------------------------------------------------------------------
// Built once per geometry and radii: for every grid node the sorted particles
// whose stencil covers it, in ascending order (a fixed summation order), each
// with w r of that node as one 4-vector {w r_x, w r_y, w r_z, particle bits}.
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
  __device__ static float bits(int s) { return __int_as_float(s); }
  __device__ static int index(float v) { return __float_as_int(v); }
};
template <>
struct Vec4<double> {
  using type = double4;
  __device__ static double bits(int s) { return __longlong_as_double((long long)s); }
  __device__ static int index(double v) { return (int)__double_as_longlong(v); }
};

// Stencil point o of sorted particle s -> node index (z, y, x order) and the
// per-axis stencil offsets.
__device__ inline int sizes_node(const SizesGridArgs& a, int4 rc, int o, int& ox, int& oy, int& oz) {
  const int P = a.P;
  ox = o % P;
  oy = (o / P) % P;
  oz = o / (P * P);
  int gx = rc.x + ox, gy = rc.y + oy, gz = rc.z + oz;
  gx -= (gx >= a.nx) ? a.nx : 0;  // P <= n on each axis: one wrap at most
  gy -= (gy >= a.ny) ? a.ny : 0;
  gz -= (gz >= a.nz) ? a.nz : 0;
  return (gz * a.ny + gy) * a.nx + gx;
}

// Pass 1: stencil hits per node (one workgroup per sorted particle).
static __global__ void __launch_bounds__(256) k_sizes_node_count(SizesGridArgs a, int* __restrict__ cnt) {
  const int s = blockIdx.x;
  if (s >= a.n) return;
  const int4 rc = a.rec[s];
  const int P3 = a.P * a.P * a.P;
  int ox, oy, oz;
  for (int o = threadIdx.x; o < P3; o += blockDim.x) atomicAdd(&cnt[sizes_node(a, rc, o, ox, oy, oz)], 1);
}

// Pass 2: each hit's particle into its node's segment (arbitrary order).
static __global__ void __launch_bounds__(256) k_sizes_node_fill(SizesGridArgs a, const int* __restrict__ off,
                                                                int* __restrict__ cursor, int* __restrict__ pidx) {
  const int s = blockIdx.x;
  if (s >= a.n) return;
  const int4 rc = a.rec[s];
  const int P3 = a.P * a.P * a.P;
  int ox, oy, oz;
  for (int o = threadIdx.x; o < P3; o += blockDim.x) {
    const int node = sizes_node(a, rc, o, ox, oy, oz);
    pidx[off[node] + atomicAdd(&cursor[node], 1)] = s;
  }
}

// Pass 3: sort every node's segment (insertion sort: a few tens of entries).
static __global__ void k_sizes_node_sort(int nodes, const int* __restrict__ off, int* __restrict__ pidx) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nodes) return;
  const int lo = off[node], hi = off[node + 1];
  for (int k = lo + 1; k < hi; ++k) {
    const int v = pidx[k];
    int j = k - 1;
    while (j >= lo && pidx[j] > v) {
      pidx[j + 1] = pidx[j];
      --j;
    }
    pidx[j + 1] = v;
  }
}

// Pass 4: w r per entry in precision T, with the same expressions as the
// per-particle kernels (r = node0 h - x + (o - half) h on each axis).
template <typename T>
__global__ void k_sizes_node_weights(SizesGridArgs a, const int* __restrict__ off, const int* __restrict__ pidx,
                                     typename Vec4<T>::type* __restrict__ ent) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= a.nx * a.ny * a.nz) return;
  const int gx = node % a.nx, gy = (node / a.nx) % a.ny, gz = node / (a.nx * a.ny);
  for (int k = off[node]; k < off[node + 1]; ++k) {
    const int s = pidx[k];
    const int4 rc = a.rec[s];
    const double4 b = a.base[s];
    int ox = gx - rc.x, oy = gy - rc.y, oz = gz - rc.z;
    ox += (ox < 0) ? a.nx : 0;
    oy += (oy < 0) ? a.ny : 0;
    oz += (oz < 0) ? a.nz : 0;
    const T r0 = (T)(b.x + (double)(ox - a.half) * a.h[0]);
    const T r1 = (T)(b.y + (double)(oy - a.half) * a.h[1]);
    const T r2 = (T)(b.z + (double)(oz - a.half) * a.h[2]);
    const T d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    const T w = sizes_weight<T>(d, (T)a.rad[s], (T)a.xi, (T)a.eta);
    typename Vec4<T>::type e;
    e.x = w * r0;
    e.y = w * r1;
    e.z = w * r2;
    e.w = Vec4<T>::bits(s);
    ent[k] = e;
  }
}

// Dipoles of the right-hand sides, particle-major in sorted order:
// pm[s][rhs] = {p_x, p_y, p_z, 0} (m is [rhs][i][3] in the caller's order).
template <typename T>
__global__ void k_sizes_permute_pm(const double* __restrict__ m, const int* __restrict__ perm, int n, int nrhs,
                                   typename Vec4<T>::type* __restrict__ pm, const int* __restrict__ gate,
                                   int gdiv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * nrhs) return;
  const int s = (int)(t / nrhs), rhs = (int)(t - (int64_t)s * nrhs);
  if (rhs_idle(gate, gdiv, rhs)) return;
  const double* q = m + 3 * ((size_t)rhs * n + perm[s]);
  typename Vec4<T>::type v;
  v.x = (T)q[0];
  v.y = (T)q[1];
  v.z = (T)q[2];
  v.w = (T)0;
  pm[t] = v;
}

// grid[rhs][node] = sum over the node's entries of (w r) . p.  W lanes per
// node with RL consecutive right-hand sides each (blockIdx.y picks the group
// of W RL), 64 / W consecutive nodes per wave: an entry's dipoles of a lane
// are RL consecutive 4-vectors of pm, the nodes of a wave share most
// particles, and an entry load serves W RL right-hand sides.  Idle
// right-hand sides are not written.
template <typename T, int W, int RL>
__global__ void __launch_bounds__(256) k_sizes_spread_nodes(SizesGridArgs a, const int* __restrict__ off,
                                                            const typename Vec4<T>::type* __restrict__ ent,
                                                            const typename Vec4<T>::type* __restrict__ pm,
                                                            T* __restrict__ grid) {
  using V = typename Vec4<T>::type;
  const int nodes = a.nx * a.ny * a.nz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int node = (int)(t / W);
  const int r0 = (blockIdx.y * W + (int)(t % W)) * RL;
  if (node >= nodes || r0 >= a.nrhs) return;
  bool live[RL];
  bool any = false;
#pragma unroll
  for (int u = 0; u < RL; ++u) {
    live[u] = r0 + u < a.nrhs && !rhs_idle(a.gate, a.gdiv, r0 + u);
    any |= live[u];
  }
  if (!any) return;
  T acc[RL];
#pragma unroll
  for (int u = 0; u < RL; ++u) acc[u] = (T)0;
  const int hi = off[node + 1];
#pragma unroll 2
  for (int k = off[node]; k < hi; ++k) {
    const V e = ent[k];
    const V* q = pm + (size_t)Vec4<T>::index(e.w) * a.nrhs + r0;
#pragma unroll
    for (int u = 0; u < RL; ++u) {
      if (r0 + u < a.nrhs) {  // a particle's row ends at nrhs
        const V v = q[u];
        acc[u] += e.x * v.x + e.y * v.y + e.z * v.z;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < RL; ++u)
    if (live[u]) grid[(size_t)(r0 + u) * nodes + node] = acc[u];
}

// The same gather with 16 right-hand sides per node, one DPP row per node
// (lane r = right-hand side r of the group): the row's lanes load 16
// consecutive entries of the node at once, one each, and broadcast them in
// turn (row_newbcast), so an entry is read once per row instead of once per
// lane.  Every lane runs the wave-uniform entry loop (the broadcasts need
// the whole row); the sum per (node, right-hand side) is over the node's
// entries in ascending order, as in k_sizes_spread_nodes.
template <typename T>
__global__ void __launch_bounds__(256) k_sizes_spread_rows(SizesGridArgs a, const int* __restrict__ off,
                                                           const typename Vec4<T>::type* __restrict__ ent,
                                                           const typename Vec4<T>::type* __restrict__ pm,
                                                           T* __restrict__ grid) {
  using V = typename Vec4<T>::type;
  const int nodes = a.nx * a.ny * a.nz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int node = (int)(t >> 4), r = (int)(t & 15);
  const int rhs = blockIdx.y * 16 + r;
  const bool in = node < nodes;
  const int lo = in ? off[node] : 0;
  const int cnt = in ? off[node + 1] - lo : 0;
  const bool on = in && rhs < a.nrhs && !rhs_idle(a.gate, a.gdiv, rhs);
  const V* q = pm + (on ? rhs : 0);
  T acc = (T)0;
  for (int j0 = 0; __any(j0 < cnt); j0 += 16) {
    V mine;
    mine.x = mine.y = mine.z = mine.w = (T)0;
    if (j0 + r < cnt) mine = ent[lo + j0 + r];
#define SPREAD_ROW_STEP(J)                                                   \
  {                                                                          \
    const T ex = dpp_mov<kDppRowBcast + (J)>(mine.x);                        \
    const T ey = dpp_mov<kDppRowBcast + (J)>(mine.y);                        \
    const T ez = dpp_mov<kDppRowBcast + (J)>(mine.z);                        \
    const T ew = dpp_mov<kDppRowBcast + (J)>(mine.w);                        \
    if (on && j0 + (J) < cnt) {                                              \
      const V v = q[(size_t)Vec4<T>::index(ew) * a.nrhs];                    \
      acc += ex * v.x + ey * v.y + ez * v.z;                                 \
    }                                                                        \
  }
    SPREAD_ROW_STEP(0) SPREAD_ROW_STEP(1) SPREAD_ROW_STEP(2) SPREAD_ROW_STEP(3)
    SPREAD_ROW_STEP(4) SPREAD_ROW_STEP(5) SPREAD_ROW_STEP(6) SPREAD_ROW_STEP(7)
    SPREAD_ROW_STEP(8) SPREAD_ROW_STEP(9) SPREAD_ROW_STEP(10) SPREAD_ROW_STEP(11)
    SPREAD_ROW_STEP(12) SPREAD_ROW_STEP(13) SPREAD_ROW_STEP(14) SPREAD_ROW_STEP(15)
#undef SPREAD_ROW_STEP
  }
  if (on) grid[(size_t)rhs * nodes + node] = acc;
}

// Particle-major entries for the contraction: pent[s][o] = {w r_x, w r_y,
// w r_z, node bits} for stencil point o of sorted particle s (one workgroup
// per particle; the same expressions as the per-particle kernels).
template <typename T>
__global__ void __launch_bounds__(256) k_sizes_particle_weights(SizesGridArgs a,
                                                                typename Vec4<T>::type* __restrict__ pent) {
  const int s = blockIdx.x;
  if (s >= a.n) return;
  const int4 rc = a.rec[s];
  const double4 b = a.base[s];
  const T rad = (T)a.rad[s];
  const int P3 = a.P * a.P * a.P;
  for (int o = threadIdx.x; o < P3; o += blockDim.x) {
    int ox, oy, oz;
    const int node = sizes_node(a, rc, o, ox, oy, oz);
    const T r0 = (T)(b.x + (double)(ox - a.half) * a.h[0]);
    const T r1 = (T)(b.y + (double)(oy - a.half) * a.h[1]);
    const T r2 = (T)(b.z + (double)(oz - a.half) * a.h[2]);
    const T d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    const T w = sizes_weight<T>(d, rad, (T)a.xi, (T)a.eta);
    typename Vec4<T>::type e;
    e.x = w * r0;
    e.y = w * r1;
    e.z = w * r2;
    e.w = Vec4<T>::bits(node);
    pent[(size_t)s * P3 + o] = e;
  }
}

// gT[node][rhs] = grid[rhs][node] (row stride ldg >= nrhs, a multiple of 4):
// tiles of 64 nodes x 16 right-hand sides through LDS (reads along nodes,
// writes along right-hand sides).
template <typename T>
__global__ void __launch_bounds__(256) k_sizes_grid_transpose(const T* __restrict__ grid, int nodes, int nrhs,
                                                              int ldg, T* __restrict__ gT) {
  __shared__ T tile[16][65];
  const int n0 = blockIdx.x * 64, r0 = blockIdx.y * 16;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int i = ty; i < 16; i += 4) {
    const int r = r0 + i, node = n0 + tx;
    if (r < nrhs && node < nodes) tile[i][tx] = grid[(size_t)r * nodes + node];
  }
  __syncthreads();
  const int ri = threadIdx.x & 15, nj = threadIdx.x >> 4;
  for (int j = nj; j < 64; j += 16) {
    const int r = r0 + ri, node = n0 + j;
    if (r < nrhs && node < nodes) gT[(size_t)node * ldg + r] = tile[ri][j];
  }
}

// E_i += prod(h) sum over particle i's stencil of (w r) g(node) with the
// transposed grid.  A wave holds 64 / (K W) particles; for each, K stencil
// splits (points o = k, k + K, ...) x W lanes of RL consecutive right-hand
// sides (blockIdx.y picks the group of W RL), so a lane reads a node's RL
// values as one aligned 4-vector and an entry load serves W RL right-hand
// sides.  The splits are summed by a fixed butterfly, then in FP64.
// out[rhs][i][3] in the caller's order.
template <typename T, int W, int K>
__global__ void __launch_bounds__(256) k_sizes_contract_nodes(SizesGridArgs a,
                                                              const typename Vec4<T>::type* __restrict__ pent,
                                                              const T* __restrict__ gT, int ldg,
                                                              double* __restrict__ out) {
  using V = typename Vec4<T>::type;
  constexpr int RL = 4, PPW = 64 / (K * W);
  static_assert(PPW * K * W == 64, "lanes");
  // XCD-contiguous blocks: the grid nodes a block reads stay in one L2
  const int lb = xcd_block(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const int bx = lb % gridDim.x, by = lb / gridDim.x;
  const int lane = threadIdx.x & 63;
  const int64_t wv = ((int64_t)bx * blockDim.x + threadIdx.x) >> 6;
  const int p = lane / (K * W), k = (lane / W) % K, rl = lane % W;
  const int s = (int)(wv * PPW) + p;
  const int r0 = (by * W + rl) * RL;
  const bool on = s < a.n && r0 < a.nrhs;  // every lane takes part in the butterfly
  const int P3 = a.P * a.P * a.P;
  T acc[RL][3];
#pragma unroll
  for (int u = 0; u < RL; ++u) acc[u][0] = acc[u][1] = acc[u][2] = (T)0;
  if (on) {
    const V* e = pent + (size_t)s * P3;
    const T* g = gT + r0;
#pragma unroll 2
    for (int o = k; o < P3; o += K) {
      const V w = e[o];
      const V v = *reinterpret_cast<const V*>(g + (size_t)Vec4<T>::index(w.w) * ldg);
      const T vv[RL] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int u = 0; u < RL; ++u) {
        acc[u][0] += w.x * vv[u];
        acc[u][1] += w.y * vv[u];
        acc[u][2] += w.z * vv[u];
      }
    }
  }
#pragma unroll
  for (int off = W; off < K * W; off <<= 1)
#pragma unroll
    for (int u = 0; u < RL; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[u][c] += __shfl_xor(acc[u][c], off);
  if (!on || k != 0) return;
  const int i = a.rec[s].w;
#pragma unroll
  for (int u = 0; u < RL; ++u) {
    const int rhs = r0 + u;
    if (rhs >= a.nrhs || rhs_idle(a.gate, a.gdiv, rhs)) continue;
    double* o3 = out + ((size_t)rhs * a.n + i) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) o3[c] += (double)acc[u][c] * a.hvol;
  }
}

// One workgroup per particle: E_i += prod(h) sum over its stencil of
// w r E~(node) for each right-hand side (out[rhs][i][3], caller's order).
template <typename T>
__global__ void __launch_bounds__(256) k_sizes_contract(SizesGridArgs a, const T* __restrict__ grid,
                                                        double* __restrict__ out) {
  __shared__ double red[3][256 / 64];
  const int s = blockIdx.x;
  if (s >= a.n) return;
  const int4 rc = a.rec[s];
  const double4 b = a.base[s];
  const T rad = (T)a.rad[s];
  const int P = a.P, P2 = P * P, P3 = P2 * P;
  const size_t plane = (size_t)a.nx * a.ny * a.nz;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  for (int rhs = 0; rhs < a.nrhs; ++rhs) {
    if (rhs_idle(a.gate, a.gdiv, rhs)) continue;  // uniform over the block
    T acc0 = 0, acc1 = 0, acc2 = 0;
    const T* g = grid + rhs * plane;
    for (int o = threadIdx.x; o < P3; o += blockDim.x) {
      const int ox = o % P, oy = (o / P) % P, oz = o / P2;
      const T r0 = (T)(b.x + (double)(ox - a.half) * a.h[0]);
      const T r1 = (T)(b.y + (double)(oy - a.half) * a.h[1]);
      const T r2 = (T)(b.z + (double)(oz - a.half) * a.h[2]);
      const T d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
      const T w = sizes_weight<T>(d, rad, (T)a.xi, (T)a.eta);
      int gx = rc.x + ox, gy = rc.y + oy, gz = rc.z + oz;
      gx -= (gx >= a.nx) ? a.nx : 0;
      gy -= (gy >= a.ny) ? a.ny : 0;
      gz -= (gz >= a.nz) ? a.nz : 0;
      const T v = w * g[((size_t)gz * a.ny + gy) * a.nx + gx];
      acc0 += v * r0;
      acc1 += v * r1;
      acc2 += v * r2;
    }
    for (int off = 32; off > 0; off >>= 1) {
      acc0 += __shfl_xor(acc0, off);
      acc1 += __shfl_xor(acc1, off);
      acc2 += __shfl_xor(acc2, off);
    }
    if (lane == 0) {
      red[0][wave] = (double)acc0;
      red[1][wave] = (double)acc1;
      red[2][wave] = (double)acc2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double v = 0.0;
      for (int k = 0; k < nwave; ++k) v += red[threadIdx.x][k];
      out[((size_t)rhs * a.n + rc.w) * 3 + threadIdx.x] += v * a.hvol;
    }
    __syncthreads();
  }
}

// Same contraction with the stencil staged once: phase 1 stores w r (3 T) and
// the node index of every stencil point in LDS (P^3 (3 sizeof(T) + 4) bytes of
// dynamic LDS); phase 2 gives each wave whole right-hand sides (rhs = wave,
// wave + 4, ...), lanes over the stencil, a fixed butterfly per wave.
template <typename T>
__global__ void __launch_bounds__(256) k_sizes_contract_lds(SizesGridArgs a, const T* __restrict__ grid,
                                                            double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int s = blockIdx.x;
  if (s >= a.n) return;
  const int P = a.P, P2 = P * P, P3 = P2 * P;
  T* wr = reinterpret_cast<T*>(smem_raw);  // [3][P3]
  int* node = reinterpret_cast<int*>(wr + 3 * P3);
  const int4 rc = a.rec[s];
  const double4 b = a.base[s];
  const T rad = (T)a.rad[s];
  for (int o = threadIdx.x; o < P3; o += blockDim.x) {
    const int ox = o % P, oy = (o / P) % P, oz = o / P2;
    const T r0 = (T)(b.x + (double)(ox - a.half) * a.h[0]);
    const T r1 = (T)(b.y + (double)(oy - a.half) * a.h[1]);
    const T r2 = (T)(b.z + (double)(oz - a.half) * a.h[2]);
    const T d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    const T w = sizes_weight<T>(d, rad, (T)a.xi, (T)a.eta);
    int gx = rc.x + ox, gy = rc.y + oy, gz = rc.z + oz;
    gx -= (gx >= a.nx) ? a.nx : 0;
    gy -= (gy >= a.ny) ? a.ny : 0;
    gz -= (gz >= a.nz) ? a.nz : 0;
    wr[o] = w * r0;
    wr[P3 + o] = w * r1;
    wr[2 * P3 + o] = w * r2;
    node[o] = (gz * a.ny + gy) * a.nx + gx;
  }
  __syncthreads();
  const size_t plane = (size_t)a.nx * a.ny * a.nz;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  for (int rhs = wave; rhs < a.nrhs; rhs += nwave) {
    if (rhs_idle(a.gate, a.gdiv, rhs)) continue;
    const T* g = grid + rhs * plane;
    T acc0 = 0, acc1 = 0, acc2 = 0;
    for (int o = lane; o < P3; o += 64) {
      const T v = g[node[o]];
      acc0 += v * wr[o];
      acc1 += v * wr[P3 + o];
      acc2 += v * wr[2 * P3 + o];
    }
    for (int off = 32; off > 0; off >>= 1) {
      acc0 += __shfl_xor(acc0, off);
      acc1 += __shfl_xor(acc1, off);
      acc2 += __shfl_xor(acc2, off);
    }
    if (lane == 0) {
      double* o3 = out + ((size_t)rhs * a.n + rc.w) * 3;
      o3[0] += (double)acc0 * a.hvol;
      o3[1] += (double)acc1 * a.hvol;
      o3[2] += (double)acc2 * a.hvol;
    }
  }
}

// ---------------------------------------------------------------------------
// y and z passes of a small cubic-section grid in one kernel
// ---------------------------------------------------------------------------
// For ny == nz == N (compile-time) a workgroup holds the whole y-z plane of
// TX kx columns of one right-hand side in LDS and does forward y, forward z,
// the scalar scale exp(-(1-eta) q^2/(4 xi^2))/q^2 (times 1/G), inverse z and
// inverse y, transposing the plane in LDS between the axes: the half
// spectrum is read and written once instead of three times.  The same line
// transforms and scale expressions as the separate y and z passes.
// A[(c N + z) ls + pidx(y)] holds y-lines, B[(c N + y) ls + pidx(z)] z-lines.
template <typename T, int N, int TX>
__global__ void __launch_bounds__(256) k_sizes_yz_plane(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ls = padded_len(N);
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* A = tw + N;
  cplx<T>* B = A + TX * N * ls;
  // XCD-contiguous blocks: the column groups of one plane share the spectrum
  // rows' cache lines in one L2 (grid (kx groups, 1, right-hand sides))
  const int lb = xcd_block(blockIdx.x + gridDim.x * blockIdx.z, gridDim.x * gridDim.z);
  const int kx0 = (lb % gridDim.x) * TX, rhs = lb / gridDim.x;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;  // uniform over the block
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  cplx<T>* base = spec + (size_t)rhs * N * N * a.kxs;
  constexpr int TOT = TX * N * N;
  // consecutive threads on consecutive kx of a row
  for (int t = threadIdx.x; t < TOT; t += blockDim.x) {
    const int c = t % TX, y = (t / TX) % N, z = t / (TX * N);
    const int kx = kx0 + c;
    A[(c * N + z) * ls + pidx(y)] = kx < a.kx ? base[((size_t)z * N + y) * a.kxs + kx] : cmk<T>(0, 0);
  }
  __syncthreads();
  fft_ct<T, N>(A, ls, TX * N, tw, false);
  for (int t = threadIdx.x; t < TOT; t += blockDim.x) {
    const int z = t % N, y = (t / N) % N, c = t / (N * N);
    B[(c * N + y) * ls + pidx(z)] = A[(c * N + z) * ls + pidx(y)];
  }
  __syncthreads();
  fft_ct<T, N>(B, ls, TX * N, tw, false);
  for (int t = threadIdx.x; t < TOT; t += blockDim.x) {
    const int z = t % N, y = (t / N) % N, c = t / (N * N);
    const int kx = kx0 + c;
    if (kx >= a.kx) continue;
    const int mx = fft_freq(kx, a.nx), my = fft_freq(y, N), mz = fft_freq(z, N);
    const double k0 = mx * sp.twopi_over_l[0], k1 = my * sp.twopi_over_l[1], k2 = mz * sp.twopi_over_l[2];
    const double kk = (k0 * k0 + k1 * k1) + k2 * k2;
    cplx<T>* f = B + (c * N + y) * ls + pidx(z);
    if (kk == 0.0) {
      *f = cmk<T>(0, 0);
      continue;
    }
    const double cf = (sp.ex[kx] * sp.ey[y] * sp.ez[z]) / kk * sp.inv_g;
    const cplx<T> F0 = *f;
    *f = cmk<T>((T)cf * F0.x, (T)cf * F0.y);
  }
  __syncthreads();
  fft_ct<T, N>(B, ls, TX * N, tw, true);
  for (int t = threadIdx.x; t < TOT; t += blockDim.x) {
    const int y = t % N, z = (t / N) % N, c = t / (N * N);
    A[(c * N + z) * ls + pidx(y)] = B[(c * N + y) * ls + pidx(z)];
  }
  __syncthreads();
  fft_ct<T, N>(A, ls, TX * N, tw, true);
  for (int t = threadIdx.x; t < TOT; t += blockDim.x) {
    const int c = t % TX, y = (t / TX) % N, z = t / (TX * N);
    const int kx = kx0 + c;
    if (kx < a.kx) base[((size_t)z * N + y) * a.kxs + kx] = A[(c * N + z) * ls + pidx(y)];
  }
}

// ---------------------------------------------------------------------------
// x-axis transforms of real grids (scalar field, tiles of ty*tz rows)
// ---------------------------------------------------------------------------
// grid[rhs][z][y][x] rows -> half spectrum spec[rhs][z][y][kx]
template <typename T, int E, int N>
__global__ void __launch_bounds__(256) k_rows_r2c(TileArgs a, const T* __restrict__ grid, cplx<T>* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.nx;
  cplx<T>* scratch = buf + a.ncl * a.lstride;
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.nx; k += blockDim.x) tw[k] = twg[k];
  const T* g = grid + (size_t)rhs * a.nx * a.ny * a.nz;
  const int tot = a.rows * a.nx;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int row = idx / a.nx, xx = idx - row * a.nx;
    const int y = ty0 + row % a.ty, z = tz0 + row / a.ty;
    const T v = (y < a.ny && z < a.nz) ? g[((size_t)z * a.ny + y) * a.nx + xx] : (T)0;
    reinterpret_cast<T*>(buf + (row >> 1) * a.lstride)[2 * pidx(xx) + (row & 1)] = v;
  }
  if (a.rows & 1) {  // odd row count: the last line's second half is zero
    for (int xx = threadIdx.x; xx < a.nx; xx += blockDim.x)
      reinterpret_cast<T*>(buf + (a.rows >> 1) * a.lstride)[2 * pidx(xx) + 1] = (T)0;
  }
  __syncthreads();
  fft_any<T, E, N>(buf, a.lstride, a.ncl, a.plan, tw, false, scratch);
  store_half_spectrum<T, T, 1>(a, buf, ty0, tz0, rhs, spec);
}

// half spectrum -> real rows of grid (inverse transform, unnormalised)
template <typename T, int E, int N>
__global__ void __launch_bounds__(256) k_rows_c2r(TileArgs a, const cplx<T>* __restrict__ spec, T* __restrict__ grid) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.nx;
  cplx<T>* scratch = buf + a.ncl * a.lstride;
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.nx; k += blockDim.x) tw[k] = twg[k];
  const int tot = a.ncl * a.kx;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int q = idx / a.kx, k = idx - q * a.kx;
    cplx<T> A = cmk<T>(0, 0), B = cmk<T>(0, 0);
    const int la = 2 * q, lb = 2 * q + 1;
    {
      const int y = ty0 + la % a.ty, z = tz0 + la / a.ty;
      if (y < a.ny && z < a.nz) A = spec[(((size_t)rhs * a.nz + z) * a.ny + y) * a.kxs + k];
    }
    if (lb < a.rows) {
      const int y = ty0 + lb % a.ty, z = tz0 + lb / a.ty;
      if (y < a.ny && z < a.nz) B = spec[(((size_t)rhs * a.nz + z) * a.ny + y) * a.kxs + k];
    }
    cplx<T>* line = buf + q * a.lstride;
    const bool selfmirror = (k == 0) || (2 * k == a.nx);
    if (selfmirror) {
      line[pidx(k)] = cmk<T>(A.x, B.x);
    } else {
      line[pidx(k)] = cmk<T>(A.x - B.y, A.y + B.x);
      line[pidx(a.nx - k)] = cmk<T>(A.x + B.y, B.x - A.y);
    }
  }
  __syncthreads();
  fft_any<T, E, N>(buf, a.lstride, a.ncl, a.plan, tw, true, scratch);
  T* g = grid + (size_t)rhs * a.nx * a.ny * a.nz;
  const int tot2 = a.rows * a.nx;
  for (int idx = threadIdx.x; idx < tot2; idx += blockDim.x) {
    const int row = idx / a.nx, xx = idx - row * a.nx;
    const int y = ty0 + row % a.ty, z = tz0 + row / a.ty;
    if (y < a.ny && z < a.nz)
      g[((size_t)z * a.ny + y) * a.nx + xx] = reinterpret_cast<const T*>(buf + (row >> 1) * a.lstride)[2 * pidx(xx) + (row & 1)];
  }
}

}  // namespace mpol
