// Device kernels of the spectral-Ewald dipole operator (gfx950).
//
// Grid spectrum layout in HBM (one buffer per context, reused every matvec):
//   spec[b][z][y][kx],  b = rhs*3 + component,  kx in [0, nx/2] (half spectrum
//   along x: the grids are real), complex<T>; rows are kxs >= nx/2 + 1 long
//   (rounded up to 16 elements so that column segments start on 128-byte
//   boundaries).
// Particles are kept in "column-bin" order: key = g0z*ny + g0y where g0 is the
// first grid index of the particle's P-wide stencil on each axis (wrapped).
#pragma once

#include "mpol_common.h"
#include "mpol_fft.h"

namespace mpol {

constexpr int kMaxRanges = 160;  // candidate particle ranges per tile (2 per z row)
constexpr int kPUnroll = 16;    // stencil widths up to this are unrolled in registers
constexpr double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__host__ __device__ inline long long imod(long long a, long long n) {
  long long r = a % n;
  return r < 0 ? r + n : r;
}

// numpy's float modulus (result carries the divisor's sign), used by cell_list.
__device__ inline double np_mod(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0) {
    if ((b < 0.0) != (m < 0.0)) m += b;
  } else {
    m = copysign(0.0, b);
  }
  return m;
}

__device__ inline double round_half_away(double v) {
  return v >= 0.0 ? floor(v + 0.5) : -floor(-v + 0.5);
}

// ---------------------------------------------------------------------------
// geometry: per-particle stencil data and cell keys
// ---------------------------------------------------------------------------
struct GeoParams {
  int n;
  int nx, ny, nz;
  int P, half;
  double box[3];
  double h[3];
  double eta[3];
  double xi;
  double norm;   // (2 xi^2/pi)^{3/2} / sqrt(prod eta)
  int nc[3];     // real-space cells per axis
};

// Pass 1: bin keys and histograms.
static __global__ void k_geo_keys(GeoParams g, const double* __restrict__ x, int* __restrict__ bin_key,
                           int* __restrict__ bin_count, int* __restrict__ cell_key, int* __restrict__ cell_count,
                           double* __restrict__ base, int* __restrict__ g0_unsorted) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int ng[3] = {g.nx, g.ny, g.nz};
  int g0[3];
  int c[3];
  for (int d = 0; d < 3; ++d) {
    const double xv = x[3 * i + d];
    const double node0 = round_half_away(xv / g.h[d]);
    base[3 * i + d] = node0 * g.h[d] - xv;
    g0[d] = (int)imod((long long)node0 - g.half - 1, ng[d]);
    g0_unsorted[3 * i + d] = g0[d];
    const double xw = np_mod(xv, g.box[d]);
    int cc = (int)floor((xw / g.box[d]) * (double)g.nc[d]);
    if (cc < 0) cc = 0;
    if (cc >= g.nc[d]) cc = g.nc[d] - 1;
    c[d] = cc;
  }
  const int key = g0[2] * g.ny + g0[1];
  bin_key[i] = key;
  atomicAdd(&bin_count[key], 1);
  const int ck = c[0] + g.nc[0] * (c[1] + g.nc[1] * c[2]);
  cell_key[i] = ck;
  atomicAdd(&cell_count[ck], 1);
}

// Single-workgroup exclusive scan of n ints (out has n+1 entries; out[n] = total).
static __global__ void k_scan(const int* __restrict__ in, int* __restrict__ out, int n) {
  __shared__ int part[1024];
  const int t = threadIdx.x, nt = blockDim.x;
  const int chunk = (n + nt - 1) / nt;
  const int lo = min(n, t * chunk), hi = min(n, lo + chunk);
  int s = 0;
  for (int k = lo; k < hi; ++k) s += in[k];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int k = lo; k < hi; ++k) {
    out[k] = run;
    run += in[k];
  }
  if (t == nt - 1) out[n] = part[t];
}

// Pass 2: scatter the caller's indices into bin order (perm) and cell order
// (cell_members); the places inside a bin or cell come from atomics, so the
// segments are sorted next (k_sort_segments) for an order that does not
// depend on timing.
static __global__ void k_geo_scatter(GeoParams g, const int* __restrict__ bin_key, const int* __restrict__ bin_start,
                              int* __restrict__ bin_cursor, const int* __restrict__ cell_key,
                              const int* __restrict__ cell_start, int* __restrict__ cell_cursor,
                              int* __restrict__ perm, int* __restrict__ cell_members) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int key = bin_key[i];
  perm[bin_start[key] + atomicAdd(&bin_cursor[key], 1)] = i;
  const int ck = cell_key[i];
  cell_members[cell_start[ck] + atomicAdd(&cell_cursor[ck], 1)] = i;
}

// Ascending order inside every segment [start[k], start[k+1]) of vals.
// Segments of at most kSortSmall entries: one thread each, insertion sort;
// longer ones are left to k_sort_segments_big.
constexpr int kSortSmall = 32;
static __global__ void k_sort_segments(int nseg, const int* __restrict__ start, int* __restrict__ vals) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nseg) return;
  const int lo = start[k], hi = start[k + 1];
  if (hi - lo > kSortSmall) return;
  for (int a = lo + 1; a < hi; ++a) {
    const int v = vals[a];
    int b = a - 1;
    while (b >= lo && vals[b] > v) {
      vals[b + 1] = vals[b];
      --b;
    }
    vals[b + 1] = v;
  }
}

// Segments longer than kSortSmall: one workgroup per segment (grid-stride
// over segments), an in-place bitonic network in global memory.  The
// network's first step of every merge compares i with the mirror entry of
// its block, so every comparison is ascending and the entries past the end
// (padding to a power of two) act as +infinity: a comparison whose upper
// index is past the end is skipped.  O(m log^2 m) work per segment.
static __global__ void k_sort_segments_big(int nseg, const int* __restrict__ start, int* __restrict__ vals) {
  for (int k = blockIdx.x; k < nseg; k += gridDim.x) {
    const int lo = start[k], m = start[k + 1] - lo;
    if (m <= kSortSmall) continue;
    int* v = vals + lo;
    int M = 1;
    while (M < m) M <<= 1;
    for (int blk = 2; blk <= M; blk <<= 1) {
      for (int j = blk >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < (M >> 1); t += blockDim.x) {
          const int i = (t / j) * 2 * j + (t % j);
          // first step of a merge: mirror partner inside the block of size blk
          const int l = (j == (blk >> 1)) ? ((i / blk) * blk + blk - 1 - (i % blk)) : (i + j);
          if (l < m) {
            const int a = v[i], b = v[l];
            if (a > b) {
              v[i] = b;
              v[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
  }
}

// Pass 3: per sorted particle s: rec[s] = (first stencil node on x, y, z,
// caller's index) and base[s] = node0*h - x on each axis (FP64 and FP32
// copies), from which the kernels recompute the Gaussian factors
// exp(-2 xi^2 (base_d + (o - half) h_d)^2 / eta_d).
static __global__ void k_geo_records(int n, const int* __restrict__ perm, const double* __restrict__ base,
                                     const int* __restrict__ g0_unsorted, int4* __restrict__ rec,
                                     double4* __restrict__ base64, float4* __restrict__ base32) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  rec[s] = make_int4(g0_unsorted[3 * i], g0_unsorted[3 * i + 1], g0_unsorted[3 * i + 2], i);
  const double b0 = base[3 * i], b1 = base[3 * i + 1], b2 = base[3 * i + 2];
  base64[s] = make_double4(b0, b1, b2, 0.0);
  base32[s] = make_float4((float)b0, (float)b1, (float)b2, 0.0f);
}

// ms[rhs][s] = m[rhs][perm[s]]: dipoles in stencil-bin order, once per matvec.
// Moments of NC components per particle (3 dipole, or 3 dipole + 5
// quadrupole) into sorted order.
template <int NC>
__global__ void k_permute_m(const double* __restrict__ m, const int* __restrict__ perm, int n,
                            double* __restrict__ ms, const int* __restrict__ gate, int gdiv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = blockIdx.y;
  if (s >= n || rhs_idle(gate, gdiv, rhs)) return;
  const size_t o = (size_t)rhs * n;
  const int i = perm[s];
#pragma unroll
  for (int c = 0; c < NC; ++c) ms[NC * (o + s) + c] = m[NC * (o + i) + c];
}

// ---------------------------------------------------------------------------
// real-space pair list (CSR by particle, owner computes)
// ---------------------------------------------------------------------------
struct PairParams {
  int n;
  double box[3];
  double rc;
  int nc[3];
  int ntab;
};

__device__ inline int table_lower_bound(const double* __restrict__ rv, int ntab, double d) {
  int lo = 0, hi = ntab;  // first index with rv[idx] >= d
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rv[mid] < d) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// FILL=false: count pairs per particle.  FILL=true: write j, A, B-A, rhat.
template <bool FILL>
__global__ void k_pairs(PairParams p, const double* __restrict__ x, const int* __restrict__ cell_key,
                        const int* __restrict__ cell_start, const int* __restrict__ cell_members,
                        const double* __restrict__ rv, const double* __restrict__ perp,
                        const double* __restrict__ para, int* __restrict__ count, const int* __restrict__ off,
                        int* __restrict__ pj, double* __restrict__ pA, double* __restrict__ pBA,
                        double* __restrict__ prh, float* __restrict__ pA32, float* __restrict__ pBA32,
                        float* __restrict__ prh32, int* __restrict__ err) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  const int ck = cell_key[i];
  const int cx = ck % p.nc[0], cy = (ck / p.nc[0]) % p.nc[1], cz = ck / (p.nc[0] * p.nc[1]);
  int cnt = 0;
  int w = FILL ? off[i] : 0;
  for (int dz = -1; dz <= 1; ++dz) {
    const int nz = (int)imod(cz + dz, p.nc[2]);
    for (int dy = -1; dy <= 1; ++dy) {
      const int ny = (int)imod(cy + dy, p.nc[1]);
      for (int dx = -1; dx <= 1; ++dx) {
        const int nx = (int)imod(cx + dx, p.nc[0]);
        const int nb = nx + p.nc[0] * (ny + p.nc[1] * nz);
        for (int q = cell_start[nb]; q < cell_start[nb + 1]; ++q) {
          const int j = cell_members[q];
          if (j == i) continue;
          double r0 = xi0 - x[3 * j], r1 = xi1 - x[3 * j + 1], r2 = xi2 - x[3 * j + 2];
          r0 = r0 - p.box[0] * trunc(2.0 * r0 / p.box[0]);
          r1 = r1 - p.box[1] * trunc(2.0 * r1 / p.box[1]);
          r2 = r2 - p.box[2] * trunc(2.0 * r2 / p.box[2]);
          const double d = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
          if (!(d < p.rc)) continue;
          if (!FILL) {
            ++cnt;
            if (d <= rv[0] || d > rv[p.ntab - 1]) atomicExch(err, 1);
          } else {
            int idx = table_lower_bound(rv, p.ntab, d);
            if (idx <= 0) idx = 1;
            if (idx >= p.ntab) idx = p.ntab - 1;
            const double lo = rv[idx - 1], hi = rv[idx];
            const double t = (d - lo) / (hi - lo);
            const double A = perp[idx - 1] * (1.0 - t) + perp[idx] * t;
            const double B = para[idx - 1] * (1.0 - t) + para[idx] * t;
            pj[w] = j;
            pA[w] = A;
            pBA[w] = B - A;
            prh[3 * (size_t)w] = r0 / d;
            prh[3 * (size_t)w + 1] = r1 / d;
            prh[3 * (size_t)w + 2] = r2 / d;
            pA32[w] = (float)A;
            pBA32[w] = (float)(B - A);
            prh32[3 * (size_t)w] = (float)(r0 / d);
            prh32[3 * (size_t)w + 1] = (float)(r1 / d);
            prh32[3 * (size_t)w + 2] = (float)(r2 / d);
            ++w;
          }
        }
      }
    }
  }
  if (!FILL) count[i] = cnt;
}

// ---------------------------------------------------------------------------
// per-matvec kernels
// ---------------------------------------------------------------------------
// Partial sums of the z components per right-hand side (slab term);
// partial[rhs*gridDim.x + block], deterministic.
static __global__ void k_sum_z(const double* __restrict__ m, int n, double* __restrict__ out) {
  __shared__ double red[256];
  const int rhs = blockIdx.y;
  const double* mr = m + (size_t)rhs * n * 3;
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += mr[3 * i + 2];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(size_t)rhs * gridDim.x + blockIdx.x] = red[0];
}

static __global__ void k_sum_finish(const double* __restrict__ partial, int nblk, int nrhs, double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrhs) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)r * nblk + b];
  out[r] = s;
}

// Real-space sum + dielectric/self diagonal + slab term.  Writes out (the
// wave-space contraction is accumulated on top afterwards).
// Complex permittivities (diag_im != null): right-hand sides come in pairs
// (real part, imaginary part) of one complex system, and the diagonal
// couples them: out_re += d_re m_re - d_im m_im, out_im += d_re m_im + d_im m_re.
// Permittivity sets are per right-hand side (real) or per pair (complex).
// Sixteen lanes per particle: lane g of the group takes pairs off[i] + g,
// off[i] + g + 16, ... (consecutive lanes read consecutive pair entries), then
// a fixed lane-sum; lane 0 adds the diagonal and writes.
template <typename T>
__global__ void __launch_bounds__(256) k_real_space(int n, int nrhs, int nsets, const double* __restrict__ m,
                                                    double* __restrict__ out, const double* __restrict__ diag,
                                                    const double* __restrict__ diag_im, const int* __restrict__ off,
                                                    const int* __restrict__ pj, const T* __restrict__ pA,
                                                    const T* __restrict__ pBA, const T* __restrict__ prh,
                                                    const double* __restrict__ zsum, double slab_over_v,
                                                    int write_diag_only_parts, double* __restrict__ hr_out,
                                                    double* __restrict__ slab_out, const int* __restrict__ gate,
                                                    int gdiv) {
  const int g = threadIdx.x & 15;
  const int i = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 4);
  const int rhs = blockIdx.y;
  if (rhs_idle(gate, gdiv, rhs)) return;  // uniform over the block
  const bool valid = i < n && rhs < nrhs;  // every lane takes part in the lane sum
  const int ii = valid ? i : 0;
  const double* mr = m + (size_t)rhs * n * 3;
  T a0 = 0, a1 = 0, a2 = 0;
  const int q0 = off[ii], q1 = valid ? off[ii + 1] : q0;
  // the particle's own terms load with its pair range (lane 0 uses them)
  const int set = (nsets == 1) ? 0 : (diag_im ? rhs >> 1 : rhs);
  const double dg = diag[(size_t)set * n + ii];
  const double mi0 = mr[3 * ii], mi1 = mr[3 * ii + 1], mi2 = mr[3 * ii + 2];
  for (int q = q0 + g; q < q1; q += 16) {
    const int j = pj[q];
    const T m0 = (T)mr[3 * j], m1 = (T)mr[3 * j + 1], m2 = (T)mr[3 * j + 2];
    const T r0 = prh[3 * (size_t)q], r1 = prh[3 * (size_t)q + 1], r2 = prh[3 * (size_t)q + 2];
    const T A = pA[q];
    const T c = pBA[q] * (m0 * r0 + m1 * r1 + m2 * r2);
    a0 += A * m0 + c * r0;
    a1 += A * m1 + c * r1;
    a2 += A * m2 + c * r2;
  }
  a0 = group_sum<T>(a0, 16);
  a1 = group_sum<T>(a1, 16);
  a2 = group_sum<T>(a2, 16);
  if (!valid || g != 0) return;
  double h0 = dg * mi0 + (double)a0;
  double h1 = dg * mi1 + (double)a1;
  double h2 = dg * mi2 + (double)a2;
  if (diag_im) {
    const double di = (rhs & 1) ? diag_im[(size_t)set * n + i] : -diag_im[(size_t)set * n + i];
    const double* mp = m + (size_t)(rhs ^ 1) * n * 3;
    h0 += di * mp[3 * i];
    h1 += di * mp[3 * i + 1];
    h2 += di * mp[3 * i + 2];
  }
  const double sz = slab_over_v != 0.0 ? slab_over_v * zsum[rhs] : 0.0;  // zsum is not formed without a slab term
  double* o = out + ((size_t)rhs * n + i) * 3;
  if (write_diag_only_parts) {
    // component mode: out <- 0 (wave part accumulates), hr_out <- Hr, slab_out <- H_slab
    o[0] = 0.0; o[1] = 0.0; o[2] = 0.0;
    double* hr = hr_out + ((size_t)rhs * n + i) * 3;
    hr[0] = h0; hr[1] = h1; hr[2] = h2;
    double* sl = slab_out + ((size_t)rhs * n + i) * 3;
    sl[0] = 0.0; sl[1] = 0.0; sl[2] = -sz;
  } else {
    o[0] = h0;
    o[1] = h1;
    o[2] = h2 + sz;
  }
}

// ---------------------------------------------------------------------------
// dipole + quadrupole real space (8 moments per particle: S_x, S_y, S_z and
// the traceless Q_xx, Q_xy, Q_xz, Q_yy, Q_yz; outputs E (3) and G (5))
// ---------------------------------------------------------------------------
constexpr int kQuadTabs = 7;  // field_quad_1..3, grad_quad_1..4 (the field/dipole pair is A, B-A)

// Coefficients C..I of every CSR pair from the seven quadrupole tables, with
// the distance and interval decisions of k_pairs (same FP64 operations).
static __global__ void k_pairs_quad(PairParams p, const double* __restrict__ x, const int* __restrict__ off,
                             const int* __restrict__ pj, const double* __restrict__ rv,
                             const double* __restrict__ qtab, double* __restrict__ pq,
                             float* __restrict__ pq32) {
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
    int idx = table_lower_bound(rv, p.ntab, d);
    if (idx <= 0) idx = 1;
    if (idx >= p.ntab) idx = p.ntab - 1;
    const double lo = rv[idx - 1], hi = rv[idx];
    const double t = (d - lo) / (hi - lo);
    for (int k = 0; k < kQuadTabs; ++k) {
      const double* tb = qtab + (size_t)k * p.ntab;
      const double v = tb[idx - 1] * (1.0 - t) + tb[idx] * t;
      pq[(size_t)w * kQuadTabs + k] = v;
      pq32[(size_t)w * kQuadTabs + k] = (float)v;
    }
  }
}

// E_i, G_i from the moments of the neighbours (RealSpace.m:55-96 restated
// owner-computes) plus the diagonal: dielectric 3/(4 pi (eps-1)) S and
// 15/(8 pi (eps-1)) Q with the self terms field_dip_1(0) S and
// grad_quad_2(0) Q / 2 (folded into dg_s, dg_q).  Sixteen lanes per
// particle over its pair range, as in k_real_space.
template <typename T>
__global__ void __launch_bounds__(256) k_real_space_quad(int n, int nrhs, int nsets, const double* __restrict__ m, double* __restrict__ out,
                                  const double* __restrict__ dg_s, const double* __restrict__ dg_q,
                                  const double* __restrict__ dgi_s, const double* __restrict__ dgi_q,
                                  const int* __restrict__ off, const int* __restrict__ pj,
                                  const T* __restrict__ pA, const T* __restrict__ pBA, const T* __restrict__ prh,
                                  const T* __restrict__ pq, int parts, double* __restrict__ hr_out,
                                  const int* __restrict__ gate, int gdiv) {
  const int lane = threadIdx.x & 15;
  const int i = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 4);
  const int rhs = blockIdx.y;
  if (rhs_idle(gate, gdiv, rhs)) return;  // uniform over the block
  const bool valid = i < n && rhs < nrhs;  // every lane takes part in the lane sum
  const int ii = valid ? i : 0;
  const double* mr = m + (size_t)rhs * n * 8;
  T e[3] = {0, 0, 0}, g[5] = {0, 0, 0, 0, 0};
  const T half = (T)0.5;
  const int q0 = off[ii], q1e = valid ? off[ii + 1] : q0;
  for (int q = q0 + lane; q < q1e; q += 16) {
    const int j = pj[q];
    const double* mj = mr + 8 * (size_t)j;
    const T S0 = (T)mj[0], S1 = (T)mj[1], S2 = (T)mj[2];
    const T q1 = (T)mj[3], q2 = (T)mj[4], q3 = (T)mj[5], q4 = (T)mj[6], q5 = (T)mj[7];
    const T r0 = prh[3 * (size_t)q], r1 = prh[3 * (size_t)q + 1], r2 = prh[3 * (size_t)q + 2];
    const T A = pA[q], BA = pBA[q];
    const T* cq = pq + (size_t)q * kQuadTabs;
    const T C = cq[0], D = cq[1], E = cq[2], F = cq[3], G = cq[4], H = cq[5], I = cq[6];
    const T Sr = S0 * r0 + S1 * r1 + S2 * r2;
    // Q.r (third row uses Q_zz = -(Q_xx + Q_yy)) and Q:rr
    const T Qr0 = q1 * r0 + q2 * r1 + q3 * r2;
    const T Qr1 = q2 * r0 + q4 * r1 + q5 * r2;
    const T Qr2 = q3 * r0 + q5 * r1 - (q1 + q4) * r2;
    const T qdd = (r0 * r0 - r2 * r2) * q1 + 2 * r0 * r1 * q2 + 2 * r0 * r2 * q3 + (r1 * r1 - r2 * r2) * q4 +
                  2 * r1 * r2 * q5;
    // field: A S + (B - A)(S.r) r - (C (Q:rr) r + 2 D Q.r)/2
    const T ce = BA * Sr - half * C * qdd;
    e[0] += A * S0 + ce * r0 - D * Qr0;
    e[1] += A * S1 + ce * r1 - D * Qr1;
    e[2] += A * S2 + ce * r2 - D * Qr2;
    // gradient, components (xx, xy, xz, yy, yz) of
    //   C (S.r) rr + D (S r + r S) + (D + E)(S.r) I
    //   + (F (Q:rr) I + G Q + H ((Q:rr) I + 2 Q.r r + 2 r Q.r) + I (Q:rr) rr)/2
    const T rr[5] = {r0 * r0, r0 * r1, r0 * r2, r1 * r1, r1 * r2};
    const T Sv[5] = {S0 * r0 + r0 * S0, S0 * r1 + r0 * S1, S0 * r2 + r0 * S2, S1 * r1 + r1 * S1,
                     S1 * r2 + r1 * S2};
    const T Qv[5] = {Qr0 * r0 + r0 * Qr0, Qr0 * r1 + r0 * Qr1, Qr0 * r2 + r0 * Qr2, Qr1 * r1 + r1 * Qr1,
                     Qr1 * r2 + r1 * Qr2};
    const T Qj[5] = {q1, q2, q3, q4, q5};
    const T cdiag = (D + E) * Sr + half * (F + H) * qdd;  // coefficient of I (xx and yy slots)
    const T crr = C * Sr + half * I * qdd;
#pragma unroll
    for (int p = 0; p < 5; ++p) {
      T v = crr * rr[p] + D * Sv[p] + half * (G * Qj[p] + 2 * H * Qv[p]);
      if (p == 0 || p == 3) v += cdiag;
      g[p] += v;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) e[c] = group_sum<T>(e[c], 16);
#pragma unroll
  for (int c = 0; c < 5; ++c) g[c] = group_sum<T>(g[c], 16);
  if (!valid || lane != 0) return;
  const int set = (nsets == 1) ? 0 : (dgi_s ? rhs >> 1 : rhs);
  const double ds = dg_s[(size_t)set * n + i], dq = dg_q[(size_t)set * n + i];
  const double* mi = mr + 8 * (size_t)i;
  double h[8];
  for (int c = 0; c < 3; ++c) h[c] = ds * mi[c] + (double)e[c];
  for (int c = 0; c < 5; ++c) h[3 + c] = dq * mi[3 + c] + (double)g[c];
  if (dgi_s) {
    const double sgn = (rhs & 1) ? 1.0 : -1.0;
    const double is = sgn * dgi_s[(size_t)set * n + i], iq = sgn * dgi_q[(size_t)set * n + i];
    const double* mp = m + (size_t)(rhs ^ 1) * n * 8 + 8 * (size_t)i;
    for (int c = 0; c < 3; ++c) h[c] += is * mp[c];
    for (int c = 0; c < 5; ++c) h[3 + c] += iq * mp[3 + c];
  }
  double* o = out + ((size_t)rhs * n + i) * 8;
  if (parts) {
    double* hr = hr_out + ((size_t)rhs * n + i) * 8;
    for (int c = 0; c < 8; ++c) {
      hr[c] = h[c];
      o[c] = 0.0;
    }
  } else {
    for (int c = 0; c < 8; ++c) o[c] = h[c];
  }
}

// ---------------------------------------------------------------------------
// spread + forward x-FFT (fused), and inverse x-FFT + contract (fused)
//
// A workgroup owns a tile of ty*tz grid rows along x (all nx nodes of each
// row, 3 components), held in LDS as complex lines that pack two real rows
// (row 2q in .x, row 2q+1 in .y).  The particles whose stencils reach the tile
// are staged in chunks of `ch`: phase 1 loads, with one thread per particle
// and no dependent loads between particles, the stencil start, the dipole, the
// x factors and the weight of every tile row (product of the y and z
// factors); phase 2 walks (particle, row) items.
// ---------------------------------------------------------------------------
struct TileArgs {
  int nx, ny, nz, kx;
  int kxs;           // spectrum row stride (>= kx)
  int ty, tz;        // grid rows (y, z) per tile
  int rows;          // ty*tz, a power of two <= 16
  int ncl;           // complex LDS lines = ceil(3*rows/2)
  int lstride;       // complex elements per LDS line (padded)
  int ch;            // particles staged per chunk
  int P;
  int half;          // (P-1)/2
  int n;             // particles
  int nrhs;
  int rg;            // right-hand sides per workgroup (matrix-core spread/contract)
  int alias;         // matrix-core spread: the staged chunk shares LDS with the tile lines
  const int4* rec;   // sorted: first stencil node x, y, z; caller's index
  const void* base;  // sorted: node0*h - x per axis (double4 or float4)
  const double* ms;  // sorted dipoles [nrhs][n][3]
  const int* bin_start;
  double h[3];
  double cexp[3];    // 2 xi^2 / eta_d
  double norm;       // (2 xi^2/pi)^{3/2} / sqrt(prod eta)
  const void* tw;    // nx twiddles
  FftPlan plan;
  double hvol;       // prod(h), trapezoid factor of the contraction
  const int* gate;   // optional per-system state (rhs_idle): tiles of idle systems exit
  int gdiv;          // right-hand sides per system
};

// Build the candidate particle ranges of a tile into shared arrays.
__device__ inline void tile_ranges(const TileArgs& a, int ty0, int tz0, int* s_start, int* s_cum, int* s_nr) {
  int ylo = ty0 - a.P + 1, ycount = a.ty + a.P - 1;
  if (ycount >= a.ny) { ylo = 0; ycount = a.ny; }
  int zlo = tz0 - a.P + 1, zcount = a.tz + a.P - 1;
  if (zcount >= a.nz) { zlo = 0; zcount = a.nz; }
  const int y0 = (int)imod(ylo, a.ny);
  const int len1 = min(ycount, a.ny - y0);
  const int len2 = ycount - len1;
  // two slots per z row; lengths first
  for (int zi = threadIdx.x; zi < zcount; zi += blockDim.x) {
    const int gz = (int)imod(zlo + zi, a.nz);
    const int k0 = gz * a.ny + y0;
    s_start[2 * zi] = a.bin_start[k0];
    s_cum[2 * zi] = a.bin_start[k0 + len1] - a.bin_start[k0];
    if (len2 > 0) {
      const int k1 = gz * a.ny;
      s_start[2 * zi + 1] = a.bin_start[k1];
      s_cum[2 * zi + 1] = a.bin_start[k1 + len2] - a.bin_start[k1];
    } else {
      s_start[2 * zi + 1] = 0;
      s_cum[2 * zi + 1] = 0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nr = 2 * zcount;
    int run = 0;
    for (int r = 0; r < nr; ++r) {
      const int c = s_cum[r];
      s_cum[r] = run;
      run += c;
    }
    s_cum[nr] = run;
    *s_nr = nr;
  }
  __syncthreads();
}

template <typename T>
struct Stage {
  int* gx;     // [ch] first x node of the stencil
  int* orig;   // [ch] caller's particle index
  T* m;        // [ch][NC] moments (spread)
  T* fx;       // [ch][P] x factors
  T* w;        // [rows][ch] y*z factor of each tile row
};

template <typename T, int NC>
__device__ inline Stage<T> stage_layout(unsigned char* base, int ch, int P, int rows) {
  Stage<T> st;
  T* tb = reinterpret_cast<T*>(base);
  st.m = tb;
  st.fx = st.m + NC * ch;
  st.w = st.fx + (size_t)P * ch;
  st.gx = reinterpret_cast<int*>(st.w + (size_t)rows * ch);
  st.orig = st.gx + ch;
  return st;
}

template <typename T>
struct Vec4Of;
template <>
struct Vec4Of<double> { using type = double4; };
template <>
struct Vec4Of<float> { using type = float4; };

// Gaussian factor of stencil offset o on axis d (same expression as the
// precompute of the reference weights).
template <typename T>
__device__ inline T stencil_factor(const TileArgs& a, int d, T b, int o) {
  const T r = b + (T)(o - a.half) * (T)a.h[d];
  return exp(-(T)a.cexp[d] * (r * r));
}

template <typename T>
__device__ inline T axis_weight_fn(const TileArgs& a, int d, T b, int g, int g0, int n) {
  int o = g - g0;
  if (o < 0) o += n;
  T w = 0;
  for (; o < a.P; o += n) w += stencil_factor<T>(a, d, b, o);
  return w;
}

// Locate candidate q of a tile: the sorted particle index.
__device__ inline int candidate_particle(const int* s_start, const int* s_cum, int nr, int q) {
  int lo = 0, hi = nr;  // largest r with s_cum[r] <= q
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_cum[mid] <= q) lo = mid;
    else hi = mid;
  }
  return s_start[lo] + (q - s_cum[lo]);
}

// Write the staged data of sorted particle s (record rc, node offsets b) into
// stage slot t.
template <typename T, bool WITH_M, int NC, typename V4>
__device__ inline void stage_at(const TileArgs& a, const Stage<T>& st, int t, int s, int4 rc, V4 b, int ty0, int tz0,
                                int rhs) {
  st.gx[t] = rc.x;
  st.orig[t] = rc.w;
  if (WITH_M) {
    const double* mp = a.ms + NC * ((size_t)rhs * a.n + s);
#pragma unroll
    for (int c = 0; c < NC; ++c) st.m[NC * t + c] = (T)mp[c];
  }
  for (int o = 0; o < a.P; ++o) st.fx[t * a.P + o] = (T)a.norm * stencil_factor<T>(a, 0, b.x, o);
  T wy[4], wz[4];
  for (int l = 0; l < a.ty; ++l) {
    const int y = ty0 + l;
    wy[l] = (y < a.ny) ? axis_weight_fn<T>(a, 1, b.y, y, rc.y, a.ny) : (T)0;
  }
  for (int l = 0; l < a.tz; ++l) {
    const int z = tz0 + l;
    wz[l] = (z < a.nz) ? axis_weight_fn<T>(a, 2, b.z, z, rc.z, a.nz) : (T)0;
  }
  for (int r = 0; r < a.rows; ++r) st.w[r * a.ch + t] = wy[r % a.ty] * wz[r / a.ty];
}

// Phase 1: thread t stages candidate base+t (if any): one 16-byte record, the
// node offsets and (spread) the sorted dipole -- independent loads -- then
// the P x factors and the weight of every tile row from the Gaussians.
template <typename T, bool WITH_M, int NC>
__device__ inline void stage_candidates(const TileArgs& a, const Stage<T>& st, int base, int total, const int* s_start,
                                        const int* s_cum, int nr, int ty0, int tz0, int rhs) {
  const int t = threadIdx.x;
  if (t >= a.ch) return;
  const int q = base + t;
  if (q >= total) {
    for (int r = 0; r < a.rows; ++r) st.w[r * a.ch + t] = (T)0;
    st.orig[t] = -1;
    return;
  }
  int lo = 0, hi = nr;  // largest r with s_cum[r] <= q
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_cum[mid] <= q) lo = mid;
    else hi = mid;
  }
  const int s = s_start[lo] + (q - s_cum[lo]);
  const int4 rc = a.rec[s];
  const typename Vec4Of<T>::type b = reinterpret_cast<const typename Vec4Of<T>::type*>(a.base)[s];
  st.gx[t] = rc.x;
  st.orig[t] = rc.w;
  if (WITH_M) {
    const double* mp = a.ms + NC * ((size_t)rhs * a.n + s);
#pragma unroll
    for (int c = 0; c < NC; ++c) st.m[NC * t + c] = (T)mp[c];
  }
  for (int o = 0; o < a.P; ++o) st.fx[t * a.P + o] = (T)a.norm * stencil_factor<T>(a, 0, b.x, o);
  T wy[4], wz[4];
  for (int l = 0; l < a.ty; ++l) {
    const int y = ty0 + l;
    wy[l] = (y < a.ny) ? axis_weight_fn<T>(a, 1, b.y, y, rc.y, a.ny) : (T)0;
  }
  for (int l = 0; l < a.tz; ++l) {
    const int z = tz0 + l;
    wz[l] = (z < a.nz) ? axis_weight_fn<T>(a, 2, b.z, z, rc.z, a.nz) : (T)0;
  }
  for (int r = 0; r < a.rows; ++r) st.w[r * a.ch + t] = wy[r % a.ty] * wz[r / a.ty];
}

// Write the NC*rows real rows of a tile (packed two per complex LDS line,
// forward-transformed along x) into the half spectrum.
template <typename T, typename TS, int NC>
__device__ inline void store_half_spectrum(const TileArgs& a, const cplx<T>* buf, int ty0, int tz0, int rhs,
                                           cplx<TS>* __restrict__ spec) {
  const int nreal = NC * a.rows;
  const int tot = nreal * a.kx;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int l = idx / a.kx, k = idx - l * a.kx;
    const int c = l / a.rows, row = l - c * a.rows;
    const int y = ty0 + row % a.ty, z = tz0 + row / a.ty;
    if (y >= a.ny || z >= a.nz) continue;
    const cplx<T>* line = buf + (l >> 1) * a.lstride;
    const cplx<T> Z = line[pidx(k)];
    const cplx<T> Zc = cconj(line[pidx(k == 0 ? 0 : a.nx - k)]);
    cplx<TS> v;
    if ((l & 1) == 0) v = cmk<TS>((TS)((Z.x + Zc.x) * (T)0.5), (TS)((Z.y + Zc.y) * (T)0.5));
    else v = cmk<TS>((TS)((Z.y - Zc.y) * (T)0.5), (TS)((Zc.x - Z.x) * (T)0.5));
    spec[(((size_t)(rhs * NC + c) * a.nz + z) * a.ny + y) * a.kxs + k] = v;
  }
}

template <typename T, int E, int N, typename TS, int NC>
__global__ void __launch_bounds__(256) k_spread_xfft(TileArgs a, const double* __restrict__ m,
                                                     cplx<TS>* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_start[kMaxRanges], s_cum[kMaxRanges + 1], s_nr;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.nx;
  cplx<T>* scratch = buf + a.ncl * a.lstride;
  unsigned char* stage_base = reinterpret_cast<unsigned char*>(buf + a.ncl * a.lstride * (E == 0 && N == 0 ? 2 : 1));
  const Stage<T> st = stage_layout<T, NC>(stage_base, a.ch, a.P, a.rows);
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.nx; k += blockDim.x) tw[k] = twg[k];
  for (int k = threadIdx.x; k < a.ncl * a.lstride; k += blockDim.x) buf[k] = cmk<T>(0, 0);
  tile_ranges(a, ty0, tz0, s_start, s_cum, &s_nr);
  const int total = s_cum[s_nr];
  const int P = a.P;
  for (int base = 0; base < total; base += a.ch) {
    stage_candidates<T, true, NC>(a, st, base, total, s_start, s_cum, s_nr, ty0, tz0, rhs);
    __syncthreads();
    const int nitems = min(a.ch, total - base) * a.rows;
    for (int item = threadIdx.x; item < nitems; item += blockDim.x) {
      const int cnd = item / a.rows, row = item - cnd * a.rows;
      const T wyz = st.w[row * a.ch + cnd];
      if (wyz == (T)0) continue;
      T mv[NC];
      T* pl[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        mv[c] = st.m[NC * cnd + c];
        const int l = c * a.rows + row;
        pl[c] = reinterpret_cast<T*>(buf + (l >> 1) * a.lstride) + (l & 1);
      }
      const T* fx = st.fx + cnd * P;
      int x = st.gx[cnd];
      if (P <= kPUnroll) {
        // factors into registers first so the LDS atomics issue back to back
        T f[kPUnroll];
#pragma unroll
        for (int o = 0; o < kPUnroll; ++o) f[o] = (o < P) ? fx[o] : (T)0;
#pragma unroll
        for (int o = 0; o < kPUnroll; ++o) {
          if (o < P) {
            const T w = wyz * f[o];
            const int px = 2 * pidx(x);
#pragma unroll
            for (int c = 0; c < NC; ++c) atomicAdd(pl[c] + px, w * mv[c]);
            if (++x == a.nx) x = 0;
          }
        }
      } else {
        for (int o = 0; o < P; ++o) {
          const T w = wyz * fx[o];
          const int px = 2 * pidx(x);
#pragma unroll
          for (int c = 0; c < NC; ++c) atomicAdd(pl[c] + px, w * mv[c]);
          if (++x == a.nx) x = 0;
        }
      }
    }
    __syncthreads();
  }
  fft_any<T, E, N>(buf, a.lstride, a.ncl, a.plan, tw, false, scratch);
  store_half_spectrum<T, TS, NC>(a, buf, ty0, tz0, rhs, spec);
}

// Exclusive prefix sum of a[0..n) in place (whole workgroup; part has
// blockDim entries).
__device__ inline void block_exclusive_scan(int* a, int n, int* part) {
  const int t = threadIdx.x, nt = blockDim.x;
  const int chunk = (n + nt - 1) / nt;
  const int lo = min(n, t * chunk), hi = min(n, lo + chunk);
  int s = 0;
  for (int k = lo; k < hi; ++k) s += a[k];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int k = lo; k < hi; ++k) {
    const int c = a[k];
    a[k] = run;
    run += c;
  }
  __syncthreads();
}

// Spread + forward x-FFT without atomics: the staged particles of each chunk
// are counting-sorted by stencil start gx, then every thread owns XS
// consecutive nodes of one tile row and sums the particles whose x window
// [gx, gx+P) covers them (a contiguous bucket range of the sorted order).
// Needs P + XS - 1 <= nx.
template <typename T, int E, int N, int XS, int NC>
__global__ void __launch_bounds__(256) k_spread_gather_xfft(TileArgs a, const double* __restrict__ m,
                                                            cplx<T>* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_start[kMaxRanges], s_cum[kMaxRanges + 1], s_nr;
  __shared__ int s_part[256];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.nx;
  cplx<T>* scratch = buf + a.ncl * a.lstride;
  unsigned char* stage_base = reinterpret_cast<unsigned char*>(buf + a.ncl * a.lstride * (E == 0 && N == 0 ? 2 : 1));
  const Stage<T> st = stage_layout<T, NC>(stage_base, a.ch, a.P, a.rows);
  int* cnt = st.orig + a.ch;     // nx + 1 bucket starts
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.nx; k += blockDim.x) tw[k] = twg[k];
  for (int k = threadIdx.x; k < a.ncl * a.lstride; k += blockDim.x) buf[k] = cmk<T>(0, 0);
  tile_ranges(a, ty0, tz0, s_start, s_cum, &s_nr);
  const int total = s_cum[s_nr];
  const int P = a.P, nx = a.nx;
  const int nseg = nx / XS;
  const int nitems = a.rows * nseg;
  using V4 = typename Vec4Of<T>::type;
  for (int base = 0; base < total; base += a.ch) {
    const int nc = min(a.ch, total - base);
    // 1a: records into registers, bucket counts by stencil start gx
    const int t = threadIdx.x;
    int s = 0, rk = 0;
    int4 rc = make_int4(0, 0, 0, 0);
    V4 bv;
    for (int k = threadIdx.x; k <= nx; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    if (t < nc) {
      s = candidate_particle(s_start, s_cum, s_nr, base + t);
      rc = a.rec[s];
      bv = reinterpret_cast<const V4*>(a.base)[s];
      rk = atomicAdd(&cnt[rc.x], 1);
    }
    __syncthreads();
    block_exclusive_scan(cnt, nx + 1, s_part);
    // 1b: stage straight into the sorted slot
    if (t < nc) stage_at<T, true, NC>(a, st, cnt[rc.x] + rk, s, rc, bv, ty0, tz0, rhs);
    __syncthreads();
    for (int item = threadIdx.x; item < nitems; item += blockDim.x) {
      const int row = item / nseg, x0 = (item - row * nseg) * XS;
      T acc[XS][NC];
#pragma unroll
      for (int k = 0; k < XS; ++k)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[k][c] = (T)0;
      const int lo = x0 - P + 1, hi = x0 + XS;  // buckets [lo, hi)
      int q0[2], q1[2];
      if (lo >= 0) {
        q0[0] = cnt[lo]; q1[0] = cnt[hi]; q0[1] = 0; q1[1] = 0;
      } else {
        q0[0] = cnt[lo + nx]; q1[0] = cnt[nx]; q0[1] = cnt[0]; q1[1] = cnt[hi];
      }
      for (int part = 0; part < 2; ++part) {
        for (int cc = q0[part]; cc < q1[part]; ++cc) {
          const T wv = st.w[row * a.ch + cc];
          if (wv == (T)0) continue;
          int o0 = x0 - st.gx[cc];
          if (o0 < 0) o0 += nx;
          T mv[NC];
#pragma unroll
          for (int c = 0; c < NC; ++c) mv[c] = st.m[NC * cc + c];
          const T* fx = st.fx + cc * P;
#pragma unroll
          for (int k = 0; k < XS; ++k) {
            int o = o0 + k;
            if (o >= nx) o -= nx;
            if (o < P) {
              const T f = wv * fx[o];
#pragma unroll
              for (int c = 0; c < NC; ++c) acc[k][c] += f * mv[c];
            }
          }
        }
      }
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
        const int l = comp * a.rows + row;
        T* pl = reinterpret_cast<T*>(buf + (l >> 1) * a.lstride) + (l & 1);
#pragma unroll
        for (int k = 0; k < XS; ++k) pl[2 * pidx(x0 + k)] += acc[k][comp];
      }
    }
    __syncthreads();
  }
  fft_any<T, E, N>(buf, a.lstride, a.ncl, a.plan, tw, false, scratch);
  store_half_spectrum<T, T, NC>(a, buf, ty0, tz0, rhs, spec);
}

template <typename T, int E, int N, int NC>
__global__ void __launch_bounds__(256) k_xifft_contract(TileArgs a, const cplx<T>* __restrict__ spec,
                                                        double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_start[kMaxRanges], s_cum[kMaxRanges + 1], s_nr;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.nx;
  cplx<T>* scratch = buf + a.ncl * a.lstride;
  unsigned char* stage_base = reinterpret_cast<unsigned char*>(buf + a.ncl * a.lstride * (E == 0 && N == 0 ? 2 : 1));
  const Stage<T> st = stage_layout<T, NC>(stage_base, a.ch, a.P, a.rows);
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.nx; k += blockDim.x) tw[k] = twg[k];
  // pack two half spectra per complex line: Z = A + iB (Hermitian extension)
  const int nreal = NC * a.rows;
  const int tot = a.ncl * a.kx;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int q = idx / a.kx, k = idx - q * a.kx;
    cplx<T> A = cmk<T>(0, 0), B = cmk<T>(0, 0);
    const int la = 2 * q, lb = 2 * q + 1;
    {
      const int c = la / a.rows, row = la - c * a.rows;
      const int y = ty0 + row % a.ty, z = tz0 + row / a.ty;
      if (y < a.ny && z < a.nz) A = spec[(((size_t)(rhs * NC + c) * a.nz + z) * a.ny + y) * a.kxs + k];
    }
    if (lb < nreal) {
      const int c = lb / a.rows, row = lb - c * a.rows;
      const int y = ty0 + row % a.ty, z = tz0 + row / a.ty;
      if (y < a.ny && z < a.nz) B = spec[(((size_t)(rhs * NC + c) * a.nz + z) * a.ny + y) * a.kxs + k];
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
  tile_ranges(a, ty0, tz0, s_start, s_cum, &s_nr);  // contains __syncthreads
  fft_any<T, E, N>(buf, a.lstride, a.ncl, a.plan, tw, true, scratch);
  const int total = s_cum[s_nr];
  const int P = a.P;
  double* outr = out + (size_t)rhs * NC * (size_t)a.n;
  for (int base = 0; base < total; base += a.ch) {
    stage_candidates<T, false, NC>(a, st, base, total, s_start, s_cum, s_nr, ty0, tz0, rhs);
    __syncthreads();
    const int nitems = min(a.ch, total - base) * a.rows;
    const int iters = (nitems + (int)blockDim.x - 1) / (int)blockDim.x;
    for (int it = 0; it < iters; ++it) {
      const int item = it * blockDim.x + threadIdx.x;
      T acc[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = (T)0;
      const int cnd = item / a.rows, row = item - cnd * a.rows;
      if (item < nitems) {
        const T wyz = st.w[row * a.ch + cnd];
        if (wyz != (T)0) {
          const T* fx = st.fx + cnd * P;
          const T* pl[NC];
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const int l = c * a.rows + row;
            pl[c] = reinterpret_cast<const T*>(buf + (l >> 1) * a.lstride) + (l & 1);
          }
          int x = st.gx[cnd];
          if (P <= kPUnroll) {
#pragma unroll
            for (int o = 0; o < kPUnroll; ++o) {
              if (o < P) {
                const T w = fx[o];
                const int px = 2 * pidx(x);
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[c] += w * pl[c][px];
                if (++x == a.nx) x = 0;
              }
            }
          } else {
            for (int o = 0; o < P; ++o) {
              const T w = fx[o];
              const int px = 2 * pidx(x);
#pragma unroll
              for (int c = 0; c < NC; ++c) acc[c] += w * pl[c][px];
              if (++x == a.nx) x = 0;
            }
          }
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] *= wyz;
        }
      }
      // the rows of one particle are consecutive lanes (rows | 64): reduce them
      for (int o = a.rows / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += __shfl_xor(acc[c], o, a.rows);
      }
      if (item < nitems && row == 0) {
        const int orig = st.orig[cnd];
        double* po = outr + NC * (size_t)orig;
#pragma unroll
        for (int c = 0; c < NC; ++c) atomicAdd(po + c, (double)acc[c] * a.hvol);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// y-axis FFT (both directions) and the fused z-axis FFT + k-space scale
// ---------------------------------------------------------------------------
struct ColArgs {
  int nx, ny, nz, kx;
  int kxs;          // spectrum row stride (>= kx)
  int tx;           // kx columns per workgroup
  int lstride;
  FftPlan plan;
  const void* tw;   // twiddles for the transformed axis
  const int* gate;  // optional per-system state (rhs_idle) of blockIdx.z / gdiv
  int gdiv;
  // register z passes (mpol_zreg.h): tail > 0 when kx = tail*tx + 1
  int tail;
};

template <typename T, int E, int N>
__global__ void __launch_bounds__(256) k_fft_y(ColArgs a, cplx<T>* __restrict__ spec, int inverse) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.ny;
  cplx<T>* scratch = buf + a.tx * a.lstride;
  const int kx0 = blockIdx.x * a.tx, z = blockIdx.y, b = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, b)) return;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.ny; k += blockDim.x) tw[k] = twg[k];
  cplx<T>* base = spec + ((size_t)b * a.nz + z) * a.ny * a.kxs;
  // thread -> (column c, first row); rows step by the threads per column set
  const int c = threadIdx.x % a.tx, ystep = blockDim.x / a.tx;
  const int kx = kx0 + c;
  const bool act = (int)threadIdx.x < ystep * a.tx;
  if (act) {
    // batches of 8 loads in flight before the LDS writes
    for (int y0 = threadIdx.x / a.tx; y0 < a.ny; y0 += 8 * ystep) {
      cplx<T> v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int y = y0 + u * ystep;
        v[u] = (kx < a.kx && y < a.ny) ? base[(size_t)y * a.kxs + kx] : cmk<T>(0, 0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int y = y0 + u * ystep;
        if (y < a.ny) buf[c * a.lstride + pidx(y)] = v[u];
      }
    }
  }
  __syncthreads();
  fft_any<T, E, N>(buf, a.lstride, a.tx, a.plan, tw, inverse != 0, scratch);
  if (act && kx < a.kx) {
    for (int y = threadIdx.x / a.tx; y < a.ny; y += ystep) base[(size_t)y * a.kxs + kx] = buf[c * a.lstride + pidx(y)];
  }
}

struct ScaleParams {
  double twopi_over_l[3];
  double one_minus_eta[3];
  double inv_4xi2;
  double inv_g;     // 1/(nx ny nz): inverse-FFT normalisation folded into the scale
  // Separable tables (natural FFT index): e_d[i] = exp(-(1-eta_d) k_d^2/(4 xi^2)).
  const double* ex;
  const double* ey;
  const double* ez;
  // Cubic boxes: jt[mx^2+my^2+mz^2] = 9 pi/(2|k|) J_{3/2}(|k|)^2/|k|^2 (0 at k=0);
  // null for other boxes (then computed per k).
  const double* jt;
  // Cubic boxes, dipole + quadrupole operator: jq[3 m^2 + (0,1,2)] =
  // (j1^2, j1 j2, j2^2)/|k|^2 with the spherical Bessel functions j1, j2 of
  // |k| (0 at k=0); null otherwise.
  const double* jq;
  // Cubic boxes: jk[m^2] = jt[m^2] / |k|^2 (0 at k = 0); null otherwise.
  const double* jk;
  // FP32 copies of jk and ez (cubic boxes; the FP32 rank-one z pass)
  const float* jk32;
  const float* ez32;
  // k_fft_z_scale, dipole and quadrupole blocks: 0 the Hermitian part of the
  // reference's operator (real grids stay real); 1 its anti-Hermitian part
  // times -i, (M(k) - M(k'))/2 F * (-i) with k' the wave vector with its
  // Nyquist components' signs flipped (zero off the Nyquist planes): the
  // reference applied to a real grid gives the Hermitian part's result plus
  // i times this one's
  int odd;
};

// Spherical Bessel functions j1(k), j2(k): power series below k = 0.6 (the
// closed forms lose digits to cancellation there), closed forms above.
__host__ __device__ inline void sph_j12(double k, double& j1, double& j2) {
  if (k < 0.6) {
    const double q = k * k;
    j1 = k / 3.0 *
         (1.0 - q / 10.0 * (1.0 - q / 28.0 * (1.0 - q / 54.0 * (1.0 - q / 88.0 * (1.0 - q / 130.0 * (1.0 - q / 180.0))))));
    j2 = q / 15.0 *
         (1.0 - q / 14.0 * (1.0 - q / 36.0 * (1.0 - q / 66.0 * (1.0 - q / 104.0 * (1.0 - q / 150.0 * (1.0 - q / 204.0))))));
  } else {
    const double s = sin(k), c = cos(k), ik = 1.0 / k;
    j1 = (s * ik - c) * ik;
    j2 = ((3.0 * ik * ik - 1.0) * s - 3.0 * c * ik) * ik;
  }
}

// The four k-space blocks of the dipole + quadrupole operator at one wave
// vector with unit vector u: with sS = u.F_S and sQ = (u u^T - u_z^2 ..):F_Q
// in the 5-component traceless form,
//   E   = (ces sS + i ceq sQ) u,    G_p = (i cgs sS + cgq sQ) (u u^T - I/3)_p.
template <typename T>
__device__ inline void quad_scale_point(T u0, T u1, T u2, T ces, T ceq, T cgs, T cgq, const cplx<T>* F,
                                        cplx<T>* G) {
  const cplx<T> sS = cmk<T>(u0 * F[0].x + u1 * F[1].x + u2 * F[2].x, u0 * F[0].y + u1 * F[1].y + u2 * F[2].y);
  const T d0 = u0 * u0 - u2 * u2, d1 = 2 * u0 * u1, d2 = 2 * u0 * u2, d3 = u1 * u1 - u2 * u2, d4 = 2 * u1 * u2;
  const cplx<T> sQ = cmk<T>(d0 * F[3].x + d1 * F[4].x + d2 * F[5].x + d3 * F[6].x + d4 * F[7].x,
                            d0 * F[3].y + d1 * F[4].y + d2 * F[5].y + d3 * F[6].y + d4 * F[7].y);
  const cplx<T> e = cmk<T>(ces * sS.x - ceq * sQ.y, ces * sS.y + ceq * sQ.x);
  const cplx<T> g = cmk<T>(cgq * sQ.x - cgs * sS.y, cgq * sQ.y + cgs * sS.x);
  const T third = (T)(1.0 / 3.0);
  const T f[5] = {u0 * u0 - third, u0 * u1, u0 * u2, u1 * u1 - third, u1 * u2};
  G[0] = cmk<T>(e.x * u0, e.y * u0);
  G[1] = cmk<T>(e.x * u1, e.y * u1);
  G[2] = cmk<T>(e.x * u2, e.y * u2);
#pragma unroll
  for (int p = 0; p < 5; ++p) G[3 + p] = cmk<T>(g.x * f[p], g.y * f[p]);
}

// c(k) = 9 pi/(2|k|) J_{3/2}(|k|)^2 exp(-sum (1-eta_d) k_d^2 / (4 xi^2)) / |k|^2
__device__ inline double kspace_coef(const ScaleParams& s, double k0, double k1, double k2) {
  const double kk = (k0 * k0 + k1 * k1) + k2 * k2;
  if (kk == 0.0) return 0.0;
  const double kn = sqrt(kk);
  const double j = sqrt(2.0 / (kPi * kn)) * (sin(kn) / kn - cos(kn));
  const double ek = s.one_minus_eta[0] * k0 * k0 + s.one_minus_eta[1] * k1 * k1 + s.one_minus_eta[2] * k2 * k2;
  return 9.0 * kPi / (2.0 * kn) * (j * j) * exp(-ek * s.inv_4xi2) / kk;
}

// a[i] = v, i < n.
static __global__ void k_fill_double(double* __restrict__ a, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Complex systems as (real, imaginary) right-hand side pairs: system s =
// blockIdx.y adds the anti-Hermitian wave-space part B (odd[rhs], from the
// pass with ScaleParams::odd) as out[2s] -= odd[2s + 1], out[2s + 1] +=
// odd[2s]; len values per right-hand side.
static __global__ void k_add_odd_part(int64_t len, const double* __restrict__ odd, double* __restrict__ out,
                                      const int* __restrict__ gate) {
  const int s = blockIdx.y;
  if (rhs_idle(gate, 1, s)) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const double* oa = odd + (size_t)(2 * s) * len;
  double* re = out + (size_t)(2 * s) * len;
  const double ob = oa[len + i], oa_i = oa[i];
  re[i] -= ob;
  re[len + i] += oa_i;
}

template <typename T, int E, int N, int NC>
__global__ void __launch_bounds__(1024) k_fft_z_scale(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.nz;
  cplx<T>* scratch = buf + NC * a.tx * a.lstride;
  const int kx0 = blockIdx.x * a.tx, y = blockIdx.y, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.nz; k += blockDim.x) tw[k] = twg[k];
  const size_t plane = (size_t)a.ny * a.kxs;
  // thread -> (column cc, first z); z steps by the threads per column set
  const int cc_t = threadIdx.x % a.tx, zstep = blockDim.x / a.tx, z_t = threadIdx.x / a.tx;
  const int kx_t = kx0 + cc_t;
  const bool act = (int)threadIdx.x < zstep * a.tx;
  if (act) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const cplx<T>* src = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx_t;
      cplx<T>* dst = buf + (c * a.tx + cc_t) * a.lstride;
      // batches of 8 loads in flight before the LDS writes
      for (int z0 = z_t; z0 < a.nz; z0 += 8 * zstep) {
        cplx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int z = z0 + u * zstep;
          v[u] = (kx_t < a.kx && z < a.nz) ? src[(size_t)z * plane] : cmk<T>(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int z = z0 + u * zstep;
          if (z < a.nz) dst[pidx(z)] = v[u];
        }
      }
    }
  }
  __syncthreads();
  fft_any<T, E, N>(buf, a.lstride, NC * a.tx, a.plan, tw, false, scratch);
  // k-space scale.  Even grids: the Hermitian part of the reference's
  // operator, whose Nyquist bins carry k = -N/2 on both sides of the
  // spectrum: the matrix averaged over the sign of each Nyquist component
  // (for the dipole block that removes the cross terms with a Nyquist axis),
  // so real grids stay real.
  const int my = fft_freq(y, a.ny);
  const bool nyq_y = (a.ny % 2 == 0) && (2 * my == -a.ny);
  for (int z = z_t; act && kx_t < a.kx && z < a.nz; z += zstep) {
    const int cc = cc_t, kx = kx_t;
    const int mx = fft_freq(kx, a.nx), mz = fft_freq(z, a.nz);
    const bool nyq_x = (a.nx % 2 == 0) && (2 * mx == -a.nx);
    const bool nyq_z = (a.nz % 2 == 0) && (2 * mz == -a.nz);
    const double k0 = mx * sp.twopi_over_l[0], k1 = my * sp.twopi_over_l[1], k2 = mz * sp.twopi_over_l[2];
    cplx<T>* f[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = buf + (c * a.tx + cc) * a.lstride + pidx(z);
    const double kk = (k0 * k0 + k1 * k1) + k2 * k2;
    if (kk == 0.0) {
#pragma unroll
      for (int c = 0; c < NC; ++c) *f[c] = cmk<T>(0, 0);
      continue;
    }
    const double ikn = 1.0 / sqrt(kk);
    const T u0 = (T)(k0 * ikn), u1 = (T)(k1 * ikn), u2 = (T)(k2 * ikn);
    if constexpr (NC == 1) {
      // polydisperse scalar grid: exp(-(1-eta) q^2/(4 xi^2))/q^2
      const double cf = (sp.ex[kx] * sp.ey[y] * sp.ez[z]) / kk * sp.inv_g;
      const cplx<T> F0 = *f[0];
      *f[0] = cmk<T>((T)cf * F0.x, (T)cf * F0.y);
    } else if constexpr (NC == 3) {
      double cf;
      if (sp.jt) cf = sp.jt[mx * mx + my * my + mz * mz] * (sp.ex[kx] * sp.ey[y] * sp.ez[z]) * sp.inv_g;
      else cf = kspace_coef(sp, k0, k1, k2) * sp.inv_g;
      const T c = (T)cf;
      // u_a u_b changes sign under the flip iff exactly one of a, b is a Nyquist axis
      const T x01 = (nyq_x != nyq_y) ? (T)0 : u0 * u1;
      const T x02 = (nyq_x != nyq_z) ? (T)0 : u0 * u2;
      const T x12 = (nyq_y != nyq_z) ? (T)0 : u1 * u2;
      const cplx<T> F0 = *f[0], F1 = *f[1], F2 = *f[2];
      if (sp.odd) {
        // the cross terms the Hermitian part drops, times -i
        const T o01 = (nyq_x != nyq_y) ? u0 * u1 : (T)0;
        const T o02 = (nyq_x != nyq_z) ? u0 * u2 : (T)0;
        const T o12 = (nyq_y != nyq_z) ? u1 * u2 : (T)0;
        const cplx<T> G0 = cmk<T>(c * (o01 * F1.x + o02 * F2.x), c * (o01 * F1.y + o02 * F2.y));
        const cplx<T> G1 = cmk<T>(c * (o01 * F0.x + o12 * F2.x), c * (o01 * F0.y + o12 * F2.y));
        const cplx<T> G2 = cmk<T>(c * (o02 * F0.x + o12 * F1.x), c * (o02 * F0.y + o12 * F1.y));
        *f[0] = cmk<T>(G0.y, -G0.x);
        *f[1] = cmk<T>(G1.y, -G1.x);
        *f[2] = cmk<T>(G2.y, -G2.x);
        continue;
      }
      const T d00 = u0 * u0, d11 = u1 * u1, d22 = u2 * u2;
      *f[0] = cmk<T>(c * (d00 * F0.x + x01 * F1.x + x02 * F2.x), c * (d00 * F0.y + x01 * F1.y + x02 * F2.y));
      *f[1] = cmk<T>(c * (x01 * F0.x + d11 * F1.x + x12 * F2.x), c * (x01 * F0.y + d11 * F1.y + x12 * F2.y));
      *f[2] = cmk<T>(c * (x02 * F0.x + x12 * F1.x + d22 * F2.x), c * (x02 * F0.y + x12 * F1.y + d22 * F2.y));
    } else {
      double a11, a12, a22, ex;
      if (sp.jq) {
        const double* q = sp.jq + 3 * (mx * mx + my * my + mz * mz);
        a11 = q[0]; a12 = q[1]; a22 = q[2];
        ex = sp.ex[kx] * sp.ey[y] * sp.ez[z];
      } else {
        double j1, j2;
        sph_j12(kk * ikn, j1, j2);
        a11 = j1 * j1 / kk; a12 = j1 * j2 / kk; a22 = j2 * j2 / kk;
        const double ek =
            sp.one_minus_eta[0] * k0 * k0 + sp.one_minus_eta[1] * k1 * k1 + sp.one_minus_eta[2] * k2 * k2;
        ex = exp(-ek * sp.inv_4xi2);
      }
      const double w = ex * sp.inv_g;
      const T ces = (T)(9.0 * a11 * w), ceq = (T)(-22.5 * a12 * w), cgs = (T)(45.0 * a12 * w),
              cgq = (T)(112.5 * a22 * w);
      cplx<T> F[NC], G[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) F[c] = *f[c];
      quad_scale_point<T>(u0, u1, u2, ces, ceq, cgs, cgq, F, G);
      if (nyq_x || nyq_y || nyq_z) {
        cplx<T> G2[NC];
        quad_scale_point<T>(nyq_x ? -u0 : u0, nyq_y ? -u1 : u1, nyq_z ? -u2 : u2, ces, ceq, cgs, cgq, F, G2);
        if (sp.odd) {
          // (G - G2)/2 times -i
#pragma unroll
          for (int c = 0; c < NC; ++c) G[c] = cmk<T>((G[c].y - G2[c].y) * (T)0.5, (G2[c].x - G[c].x) * (T)0.5);
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c) G[c] = cmk<T>((G[c].x + G2[c].x) * (T)0.5, (G[c].y + G2[c].y) * (T)0.5);
        }
      } else if (sp.odd) {
#pragma unroll
        for (int c = 0; c < NC; ++c) G[c] = cmk<T>(0, 0);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) *f[c] = G[c];
    }
  }
  __syncthreads();
  fft_any<T, E, N>(buf, a.lstride, NC * a.tx, a.plan, tw, true, scratch);
  if (act && kx_t < a.kx) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cplx<T>* dst = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx_t;
      const cplx<T>* src = buf + (c * a.tx + cc_t) * a.lstride;
      for (int z = z_t; z < a.nz; z += zstep) dst[(size_t)z * plane] = src[pidx(z)];
    }
  }
}

}  // namespace mpol
