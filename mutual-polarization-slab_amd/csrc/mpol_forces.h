// Truncated point-dipole forces and energies for fixed dipoles (gfx950).
//
// Same physics, cutoff and conventions as compute_forces of
// "CPU - HOOMD simulation test/dipole_forces_ext/_dipole_forces_ext.cpp":
//   U_ij = pref [(mi.mj)/r^3 - 3 (mi.r)(mj.r)/r^5]
//   F_i  = 3 pref/r^5 [(mi.mj) r + (mj.r) mi + (mi.r) mj - 5 (mi.r)(mj.r) r/r^2]
// with r = r_i - r_j (minimum image by rint; z optional), pairs with
// r^2 > r_cut^2 or r^2 < 1e-24 skipped, energy split 1/2 per particle.
// Owner computes: particle i sums over all its neighbours j != i, so no
// atomics are needed (the reference visits j > i and scatters to j).
#pragma once

#include <hip/hip_runtime.h>

namespace mpol {

struct ForceParams {
  int n;
  double box[3];
  double rc2;
  double pref;
  int periodic_z;
  int use_cells;
  int nc[3];
};

// cell of a particle in HOOMD-centred coordinates [-L/2, L/2) (or any
// coordinates: they are wrapped), cell size L/Nc, clamped to [0, Nc).
__device__ inline int force_cell_1d(double xc, double L, int nc) {
  double x = xc + 0.5 * L;
  x -= L * floor(x / L);
  const double h = L / (double)nc;
  int ix = (int)floor(x / h);
  if (ix < 0) ix = 0;
  if (ix >= nc) ix = nc - 1;
  return ix;
}

static __global__ void k_force_cells(ForceParams p, const double* __restrict__ pos, int* __restrict__ cell,
                              int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const int ix = force_cell_1d(pos[3 * i], p.box[0], p.nc[0]);
  const int iy = force_cell_1d(pos[3 * i + 1], p.box[1], p.nc[1]);
  const int iz = force_cell_1d(pos[3 * i + 2], p.box[2], p.nc[2]);
  const int cid = (ix * p.nc[1] + iy) * p.nc[2] + iz;
  cell[i] = cid;
  atomicAdd(&count[cid], 1);
}

static __global__ void k_force_scatter(int n, const int* __restrict__ cell, const int* __restrict__ start,
                                int* __restrict__ cursor, int* __restrict__ members) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell[i];
  members[start[c] + atomicAdd(&cursor[c], 1)] = i;
}

static __global__ void k_fill_int(int* __restrict__ a, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__device__ inline double wrap_rint(double dx, double L) { return dx - L * rint(dx / L); }

// distinct neighbour cells of c along one axis (1..3 of them)
__device__ inline int neighbor_cells(int c, int nc, bool periodic, int* out) {
  int cnt = 0;
  for (int d = -1; d <= 1; ++d) {
    int cc = c + d;
    if (periodic) {
      cc %= nc;
      if (cc < 0) cc += nc;
    } else if (cc < 0 || cc >= nc) {
      continue;
    }
    bool dup = false;
    for (int k = 0; k < cnt; ++k) dup |= (out[k] == cc);
    if (!dup) out[cnt++] = cc;
  }
  return cnt;
}

__device__ inline void pair_force(const ForceParams& p, const double* __restrict__ pos, const double* __restrict__ dip,
                                  int i, int j, double& fx, double& fy, double& fz, double& e) {
  const double dx = wrap_rint(pos[3 * i] - pos[3 * j], p.box[0]);
  const double dy = wrap_rint(pos[3 * i + 1] - pos[3 * j + 1], p.box[1]);
  double dz = pos[3 * i + 2] - pos[3 * j + 2];
  if (p.periodic_z) dz = wrap_rint(dz, p.box[2]);
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 > p.rc2 || r2 < 1e-24) return;
  const double inv_r2 = 1.0 / r2;
  const double inv_r = sqrt(inv_r2);
  const double inv_r3 = inv_r2 * inv_r;
  const double inv_r5 = inv_r3 * inv_r2;
  const double mix = dip[3 * i], miy = dip[3 * i + 1], miz = dip[3 * i + 2];
  const double mjx = dip[3 * j], mjy = dip[3 * j + 1], mjz = dip[3 * j + 2];
  const double mimj = mix * mjx + miy * mjy + miz * mjz;
  const double mir = mix * dx + miy * dy + miz * dz;
  const double mjr = mjx * dx + mjy * dy + mjz * dz;
  e += 0.5 * p.pref * (mimj * inv_r3 - 3.0 * (mir * mjr) * inv_r5);
  const double coef = p.pref * 3.0 * inv_r5;
  const double five = 5.0 * (mir * mjr) * inv_r2;
  fx += coef * (mimj * dx + mjr * mix + mir * mjx - five * dx);
  fy += coef * (mimj * dy + mjr * miy + mir * mjy - five * dy);
  fz += coef * (mimj * dz + mjr * miz + mir * mjz - five * dz);
}

static __global__ void k_forces(ForceParams p, const double* __restrict__ pos, const double* __restrict__ dip,
                         const int* __restrict__ cell, const int* __restrict__ start, const int* __restrict__ members,
                         double* __restrict__ f, double* __restrict__ pe) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  if (p.use_cells) {
    const int cid = cell[i];
    const int iz = cid % p.nc[2], iy = (cid / p.nc[2]) % p.nc[1], ix = cid / (p.nc[1] * p.nc[2]);
    int nx[3], ny[3], nz[3];
    const int cx = neighbor_cells(ix, p.nc[0], true, nx);
    const int cy = neighbor_cells(iy, p.nc[1], true, ny);
    const int cz = neighbor_cells(iz, p.nc[2], p.periodic_z != 0, nz);
    for (int a = 0; a < cx; ++a)
      for (int b = 0; b < cy; ++b)
        for (int c = 0; c < cz; ++c) {
          const int nb = (nx[a] * p.nc[1] + ny[b]) * p.nc[2] + nz[c];
          for (int q = start[nb]; q < start[nb + 1]; ++q) {
            const int j = members[q];
            if (j != i) pair_force(p, pos, dip, i, j, fx, fy, fz, e);
          }
        }
  } else {
    for (int j = 0; j < p.n; ++j)
      if (j != i) pair_force(p, pos, dip, i, j, fx, fy, fz, e);
  }
  f[3 * i] = fx;
  f[3 * i + 1] = fy;
  f[3 * i + 2] = fz;
  pe[i] = e;
}

}  // namespace mpol
