// Spread and contract on the matrix cores of gfx950 (v_mfma_*_16x16x4).
//
// Along x the Gaussian stencil of a particle is a band: particle q adds
// fx_q[x - gx_q] (x - gx_q in [0, P)) to node x of every tile row, weighted
// by the row weight w_q[row] (product of its y and z factors) and its moment
// m_q[c].  For one tile (R rows, NC components, all nx nodes of each row) the
// spread is therefore a matrix product
//
//     G[x][(c, row)] = sum_q  A[x][q] * B[q][(c, row)],
//     A[x][q] = fx_q[x - gx_q],   B[q][(c, row)] = w_q[row] * m_q[c],
//
// and the contraction its transpose followed by a row-weighted sum:
//
//     out_q[c] = sum_row w_q[row] * sum_x A[x][q] * G[x][(c, row)].
//
// With the candidates of a tile sorted by gx, the candidates that reach a
// 16-node x block are a contiguous range (about P + 15 of them), and the 16
// consecutive candidates of a block span about 16 + P nodes, so both products
// are short bands of 16 x 16 x 4 MFMA steps instead of per-candidate VALU
// loops.
//
// The per-tile candidate data (first x node, sorted particle index, P x
// factors times the Gaussian norm, R row weights) depends only on the
// geometry, so it is built once per geometry into a "tile image" in HBM,
// sorted by gx within chunks of ch candidates; the per-matvec kernels copy a
// chunk into LDS with straight coalesced loads.
#pragma once

#include "mpol_kernels.h"

namespace mpol {

constexpr int kMaxMfmaNx = 512;  // largest x grid of the compile-time FFT plans

typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x4_t __attribute__((ext_vector_type(4)));

// One 16x16x4 step D += A B.  Lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; its four results are D[row(l, r)][l % 16], r = 0..3
// (the two precisions order the rows differently).
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  using acc_t = f32x4_t;
  static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct Mfma<double> {
  using acc_t = f64x4_t;
  static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

struct ImgArgs {
  const int* off;    // [ntiles + 1] first entry of each tile
  const int2* gs;    // [total] {first stencil node along x, sorted particle index}
  const void* fw;    // [total][P + rows] T: norm * x factors, then the row weights
  const int* perm;   // sorted particle index -> caller's index
  long long total;   // image entries (slots)
  void* part;        // contraction partials [nrhs][total][NC] T
};

// Per-particle slot lists of an image (CSR by sorted particle index), so the
// contraction can sum each particle's tile partials without atomics.
static __global__ void k_img_pcount(const int2* __restrict__ gs, int64_t total, int* __restrict__ cnt) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < total) atomicAdd(&cnt[gs[q].y], 1);
}
static __global__ void k_img_pscatter(const int2* __restrict__ gs, int64_t total, const int* __restrict__ poff,
                                      int* __restrict__ cursor, int* __restrict__ pslot) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int s = gs[q].y;
  pslot[poff[s] + atomicAdd(&cursor[s], 1)] = (int)q;
}
// Sort each particle's slots ascending (insertion sort, a few tens of entries)
// so that the contraction's summation order is fixed.
static __global__ void k_img_psort(int n, const int* __restrict__ poff, int* __restrict__ pslot) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int* a = pslot + poff[s];
  const int m = poff[s + 1] - poff[s];
  for (int i = 1; i < m; ++i) {
    const int v = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = v;
  }
}

// out[rhs][perm[s]][c] += hvol * sum over the slots of particle s of the
// contraction partials (FP64 sums in a fixed order: sixteen lanes per
// particle, lane g taking slots g, g + 16, ... of the sorted slot list, then a
// fixed lane-sum).
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_contract_reduce(int n, long long total, const int* __restrict__ poff,
                                                         const int* __restrict__ pslot, const int* __restrict__ perm,
                                                         const T* __restrict__ part, double hvol,
                                                         double* __restrict__ out, const int* __restrict__ gate,
                                                         int gdiv) {
  const int g = threadIdx.x & 15;
  // XCD-contiguous blocks: the partials of neighbouring particles stay in one L2
  const int lb = xcd_block(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const int s = (int)(((size_t)(lb % gridDim.x) * blockDim.x + threadIdx.x) >> 4);
  const int rhs = lb / gridDim.x;
  if (rhs_idle(gate, gdiv, rhs)) return;  // uniform over the block
  const bool valid = s < n;  // every lane takes part in the lane sum
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  const T* pr = part + (size_t)rhs * total * NC;
  const int k0 = valid ? poff[s] : 0, k1 = valid ? poff[s + 1] : 0;
  // the output the sums add to loads with the slot range (lane 0 uses it)
  double* o = out + ((size_t)rhs * n + perm[valid ? s : 0]) * NC;
  double ov[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) ov[c] = o[c];
  for (int k = k0 + g; k < k1; k += 16) {
    const T* p = pr + (size_t)pslot[k] * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += (double)p[c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = group_sum<double>(acc[c], 16);
  if (!valid || g != 0) return;
#pragma unroll
  for (int c = 0; c < NC; ++c) o[c] = ov[c] + acc[c] * hvol;
}

// Candidates of tile (ty0, tz0): the total of the column-bin ranges that
// tile_ranges enumerates (one thread per tile).
static __global__ void k_img_count(TileArgs a, int ntx, int ntiles, int* __restrict__ cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int ty0 = (t % ntx) * a.ty, tz0 = (t / ntx) * a.tz;
  int ylo = ty0 - a.P + 1, ycount = a.ty + a.P - 1;
  if (ycount >= a.ny) { ylo = 0; ycount = a.ny; }
  int zlo = tz0 - a.P + 1, zcount = a.tz + a.P - 1;
  if (zcount >= a.nz) { zlo = 0; zcount = a.nz; }
  const int y0 = (int)imod(ylo, a.ny);
  const int len1 = min(ycount, a.ny - y0);
  const int len2 = ycount - len1;
  int total = 0;
  for (int zi = 0; zi < zcount; ++zi) {
    const int gz = (int)imod(zlo + zi, a.nz);
    const int k0 = gz * a.ny + y0;
    total += a.bin_start[k0 + len1] - a.bin_start[k0];
    if (len2 > 0) total += a.bin_start[gz * a.ny + len2] - a.bin_start[gz * a.ny];
  }
  cnt[t] = total;
}

// Tile image: one workgroup per tile.  Candidates come in chunks of a.ch in
// the enumeration order of tile_ranges; within a chunk each candidate's slot
// is its rank by (gx, enumeration index), a deterministic stable sort.
template <typename T>
__global__ void __launch_bounds__(256) k_img_fill(TileArgs a, const int* __restrict__ off, int2* __restrict__ gs,
                                                  T* __restrict__ fw) {
  __shared__ int s_start[kMaxRanges], s_cum[kMaxRanges + 1], s_nr;
  __shared__ int s_gx[256];
  const int tile = blockIdx.x + blockIdx.y * gridDim.x;
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz;
  tile_ranges(a, ty0, tz0, s_start, s_cum, &s_nr);
  const int total = s_cum[s_nr];
  const int q0 = off[tile];
  const int PW = a.P + a.rows;
  using V4 = typename Vec4Of<T>::type;
  for (int base = 0; base < total; base += a.ch) {
    const int nc = min(a.ch, total - base);
    const int t = threadIdx.x;
    int s = 0;
    int4 rc = make_int4(0, 0, 0, 0);
    if (t < nc) {
      s = candidate_particle(s_start, s_cum, s_nr, base + t);
      rc = a.rec[s];
      s_gx[t] = rc.x;
    }
    __syncthreads();
    if (t < nc) {
      int rank = 0;
      for (int u = 0; u < nc; ++u) {
        const int g = s_gx[u];
        rank += (g < rc.x) || (g == rc.x && u < t);
      }
      const size_t slot = (size_t)q0 + base + rank;
      gs[slot] = make_int2(rc.x, s);
      const V4 b = reinterpret_cast<const V4*>(a.base)[s];
      T* f = fw + slot * PW;
      for (int o = 0; o < a.P; ++o) f[o] = (T)a.norm * stencil_factor<T>(a, 0, b.x, o);
      T wy[4], wz[4];
      for (int l = 0; l < a.ty; ++l) {
        const int y = ty0 + l;
        wy[l] = (y < a.ny) ? axis_weight_fn<T>(a, 1, b.y, y, rc.y, a.ny) : (T)0;
      }
      for (int l = 0; l < a.tz; ++l) {
        const int z = tz0 + l;
        wz[l] = (z < a.nz) ? axis_weight_fn<T>(a, 2, b.z, z, rc.z, a.nz) : (T)0;
      }
      for (int r = 0; r < a.rows; ++r) f[a.P + r] = wy[r % a.ty] * wz[r / a.ty];
    }
    __syncthreads();
  }
}

// Spectrum row of real line l = c*rows + row of tile (ty0, tz0) for
// right-hand side rhs, or -1 when the row lies outside the grid (tiles at the
// upper y/z edges).  rows, ty and tz are powers of two.
__device__ __forceinline__ long long tile_line_row(const TileArgs& a, int l, int ty0, int tz0, int rhs, int nc_comp) {
  const int lr = __builtin_ctz(a.rows), lty = __builtin_ctz(a.ty);
  const int c = l >> lr, row = l & (a.rows - 1);
  const int y = ty0 + (row & (a.ty - 1)), z = tz0 + (row >> lty);
  if (y >= a.ny || z >= a.nz) return -1;
  return ((long long)(rhs * nc_comp + c) * a.nz + z) * a.ny + y;
}

// Candidate entries past a chunk's end that the stage fills with zeros: the
// spread's MFMA steps take four candidates at a time, read up to three
// entries past the last one without a range check, and load the next step's
// four candidates ahead (up to seven entries past the last one).
constexpr int kStagePad = 8;

// Moments per staged candidate: NC, zero-padded to the components of the
// NB = ceil(NC*R/16) column blocks, so that a column past the last line
// reads a zero moment.
__host__ __device__ constexpr int mfma_ncp(int nc, int rows) { return ((nc * rows + 15) / 16) * 16 / rows; }

// LDS layout of one staged chunk (after the twiddles and the tile lines).
template <typename T>
struct MfmaStage {
  T* fw;     // [ch + kStagePad][P + rows]
  T* m;      // [ch + kStagePad][ncp] (spread)
  int* gx;   // [ch + kStagePad]
};

template <typename T>
__device__ inline MfmaStage<T> mfma_stage_layout(unsigned char* base, int ch, int PW, int ncp) {
  MfmaStage<T> st;
  st.fw = reinterpret_cast<T*>(base);
  st.m = st.fw + (size_t)(ch + kStagePad) * PW;
  st.gx = reinterpret_cast<int*>(st.m + (size_t)(ch + kStagePad) * ncp);
  return st;
}

// LDS bytes of the MFMA tile kernels: twiddles, tile lines, staged chunk.
template <int NC>
inline size_t mfma_tile_lds(int nx, size_t csize, int P, int ncl, int lstride, int ch, int rows) {
  return ((size_t)nx + (size_t)ncl * lstride) * csize +
         (size_t)(ch + kStagePad) * ((P + rows + mfma_ncp(NC, rows)) * (csize / 2) + sizeof(int));
}

// LDS bytes of the matrix-core spread when the staged chunk lies over the
// tile lines (TileArgs::alias): twiddles, then the larger of the two.
template <int NC>
inline size_t mfma_tile_lds_alias(int nx, size_t csize, int P, int ncl, int lstride, int ch, int rows) {
  const size_t lines = (size_t)ncl * lstride * csize;
  const size_t stage = (size_t)(ch + kStagePad) * ((P + rows + mfma_ncp(NC, rows)) * (csize / 2) + sizeof(int));
  return (size_t)nx * csize + (lines > stage ? lines : stage);
}

// Whether k_spread_mfma_xfft keeps a wave's accumulators in registers across
// chunks (its kRegAcc): the x blocks of a wave (four waves) times the column
// blocks, four values each, in at most 64 registers.
inline bool mfma_reg_acc(int nx, int nc, int rows, size_t tbytes) {
  const int xpw = (nx / 16 + 3) / 4, nb = (nc * rows + 15) / 16;
  return (size_t)xpw * nb * 4 * tbytes <= 256;
}

// Candidate ranges of every 16-node x block of a staged chunk (candidates
// sorted by first node gx): block xb needs the candidates with gx in
// [16 xb - P + 1, 16 xb + 15] (mod N), one range r[0..1], or two when the
// window wraps below node 0: [r[0], nc) for gx >= 16 xb - P + 1 + N, then
// [r[2] = 0, r[3]).  Every bound is lower_bound(v) = #{u : gx[u] < v}, which
// is the u with gx[u - 1] < v <= gx[u] (gx[-1] = -1, gx[nc] = N): thread u
// writes the bounds whose v falls in its interval, so each bound is written
// once, without a search.  gs: the chunk's image records.
template <int N>
__device__ inline void mfma_chunk_ranges(const int2* __restrict__ gs, int nc, int P, int* s_xr) {
  constexpr int NXB = N / 16;
  const int nwrap = (P - 1 + 15) >> 4;  // blocks with 16 xb - P + 1 < 0
  for (int xb = threadIdx.x; xb < NXB; xb += blockDim.x) {
    int* r = s_xr + 4 * xb;
    r[2] = 0;
    if (xb < nwrap) r[1] = nc;
    else r[3] = 0;
  }
  for (int u = threadIdx.x; u <= nc; u += blockDim.x) {
    const int lo = u > 0 ? gs[u - 1].x : -1;  // this thread's bounds: v in (lo, hi]
    const int hi = u < nc ? gs[u].x : N;
    // upper bounds v = 16 (xb + 1)
    for (int xb = max(lo >> 4, 0); xb <= min((hi >> 4) - 1, NXB - 1); ++xb) s_xr[4 * xb + (xb < nwrap ? 3 : 1)] = u;
    // lower bounds v = 16 xb - P + 1 (blocks xb >= nwrap)
    for (int xb = max(((lo + P - 1) >> 4) + 1, nwrap); xb <= min((hi + P - 1) >> 4, NXB - 1); ++xb)
      s_xr[4 * xb] = u;
    // wrapped lower bounds v = 16 xb - P + 1 + N (blocks xb < nwrap)
    for (int xb = max(((lo + P - 1 - N) >> 4) + 1, 0); xb <= min((hi + P - 1 - N) >> 4, nwrap - 1); ++xb)
      s_xr[4 * xb] = u;
  }
}

// Copy chunk [q, q + nc) of the tile image into LDS (IMAGE); the spread also
// gathers the moments of right-hand side rhs and zero-fills kStagePad entries
// past the chunk (WITH_M).  Without IMAGE only the moments are staged (the
// image chunk is in LDS from an earlier right-hand side).
template <typename T, bool WITH_M, int NC, int NCP>
__device__ inline void mfma_stage_chunk(const TileArgs& a, const ImgArgs& im, const MfmaStage<T>& st, size_t q,
                                        int nc, int rhs, bool image = true) {
  const int PW = a.P + a.rows;
  const T* src = reinterpret_cast<const T*>(im.fw) + q * PW;
  const int nt = blockDim.x, t = threadIdx.x;
  const int tot = image ? nc * PW : 0;
  // every load of the chunk issued before the first LDS write: this thread's
  // slot record, the image rows (up to kBatch per thread in one batch: 256
  // FP32 candidates of P + rows = 18 entries), then the moments the record
  // points to (issued once the record is in, while the rows are in flight)
  constexpr int kBatch = 18;
  int2 g = make_int2(0, 0);
  if (t < nc) g = im.gs[q + t];
  int k0 = 0, u0 = t;
  if (tot <= kBatch * nt) {
    T v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int k = u * nt + t;
      v[u] = k < tot ? src[k] : (T)0;
    }
    double mv[NC];
    if (WITH_M) {
      const double* mp = a.ms + NC * ((size_t)rhs * a.n + g.y);
#pragma unroll
      for (int c = 0; c < NC; ++c) mv[c] = t < nc ? mp[c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int k = u * nt + t;
      if (k < tot) st.fw[k] = v[u];
    }
    if (t < nc) {
      if (image) st.gx[t] = g.x;
      if (WITH_M) {
#pragma unroll
        for (int c = 0; c < NCP; ++c) st.m[NCP * t + c] = c < NC ? (T)mv[c] : (T)0;
      }
    }
    k0 = tot;
    u0 = t + nt;
  }
  for (; k0 < tot; k0 += 8 * nt) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u * nt + t;
      v[u] = k < tot ? src[k] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u * nt + t;
      if (k < tot) st.fw[k] = v[u];
    }
  }
  for (int u = u0; u < nc; u += nt) {
    if (u != t) g = im.gs[q + u];
    if (image) st.gx[u] = g.x;
    if (WITH_M) {
      const double* mp = a.ms + NC * ((size_t)rhs * a.n + g.y);
#pragma unroll
      for (int c = 0; c < NCP; ++c) st.m[NCP * u + c] = c < NC ? (T)mp[c] : (T)0;
    }
  }
  if (WITH_M) {
    if (image) {
      for (int k = t; k < kStagePad * PW; k += nt) st.fw[nc * PW + k] = (T)0;
      if (t < kStagePad) st.gx[nc + t] = 0;
    }
    for (int k = t; k < kStagePad * NCP; k += nt) st.m[nc * NCP + k] = (T)0;
  }
}

// First right-hand side at or after r0, below r1, whose system is not idle
// (r1 when there is none).
__device__ __forceinline__ int next_live_rhs(const TileArgs& a, int r0, int r1) {
  int r = r0;
  while (r < r1 && rhs_idle(a.gate, a.gdiv, r)) ++r;
  return r;
}

// Workgroups per CU that k_spread_mfma_xfft's registers are budgeted for:
// four for FP32 lines up to 256 nodes and at most 24 tile lines (the LDS of an
// aliased 256-candidate chunk allows four), three for longer FP32 lines or more
// lines (whose accumulators and line FFTs would spill at four), two for FP64.
__host__ __device__ constexpr int spread_mfma_min_blocks(size_t tbytes, int n, int ncol) {
  return tbytes == 4 ? ((n <= 256 && ncol <= 24) ? 4 : 3) : 2;
}

// Spread + forward x-FFT.  Wave w handles x blocks w, w + 4, ...; lane l of a
// block holds node x0 + l % 16 (A rows) and columns (c, row) = nb*16 + l % 16
// of the tile lines (column index = line index c*R + row).  R (tile rows) is
// a template parameter so that the column blocks, the row of a column and
// the lane reductions are compile-time.  Needs nx % 16 == 0, P + 16 <= nx,
// 256 threads (four waves) and lines of padded_len(N) entries.
//
// Per MFMA step (four candidates) a lane reads the candidate's first node,
// one x factor (clamped to the stencil and selected to 0 outside it), its row
// weight and NB moments: the steps read up to three staged entries past the
// candidate range, which are either candidates whose stencil misses the x
// block (their x factor selects to 0) or the zero-filled pad past the chunk.
// When a wave's accumulators for all its x blocks fit in 64 registers they
// stay in registers across chunks and are stored into the tile lines once;
// otherwise each chunk adds its accumulators into the (cleared) lines.  With
// a.alias (register accumulators, set by the host) the staged chunk lies over
// the tile lines, which are written only after the last chunk's steps, and
// every right-hand side restages the image.  Registers are budgeted by
// spread_mfma_min_blocks.
template <typename T, int N, int NC, int R>
__global__ void __launch_bounds__(256, spread_mfma_min_blocks(sizeof(T), N, NC * R))
    k_spread_mfma_xfft(TileArgs a, ImgArgs im, cplx<T>* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  static_assert(R == 4 || R == 8 || R == 16, "tile rows");
  constexpr int NCOL = NC * R, NB = (NCOL + 15) / 16, NXB = N / 16, NW = 4;
  constexpr int NCP = mfma_ncp(NC, R), XPW = (NXB + NW - 1) / NW, LS = padded_len(N);
  constexpr bool kRegAcc = XPW * NB * 4 * (int)sizeof(T) <= 256;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + N;
  const int P = a.P, PW = P + R;
  const MfmaStage<T> st =
      mfma_stage_layout<T>(reinterpret_cast<unsigned char*>(a.alias ? buf : buf + a.ncl * LS), a.ch, PW, NCP);
  __shared__ int s_xr[4 * NXB];
  const int tile = blockIdx.x + blockIdx.y * gridDim.x;
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz;
  // right-hand sides [r0, r1) in turn; the first live one is staged with
  // the image, later ones only restage their moments when the tile's
  // candidates fit in one chunk
  const int r0 = blockIdx.z * a.rg, r1 = min(r0 + a.rg, a.nrhs);
  const int rfirst = next_live_rhs(a, r0, r1);
  if (rfirst >= r1) return;
  const size_t qt = (size_t)im.off[tile];
  const int total = im.off[tile + 1] - (int)qt;
  const bool one_chunk = total <= a.ch;
  // the first chunk of the tile image goes to LDS with the twiddles
  if (total > 0) {
    mfma_stage_chunk<T, true, NC, NCP>(a, im, st, qt, min(a.ch, total), rfirst);
    mfma_chunk_ranges<N>(im.gs + qt, min(a.ch, total), P, s_xr);
  }
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int i16 = lane & 15, k4 = lane >> 4;
  const int colrow = i16 % R;  // row of column nb*16 + i16 for every nb (R | 16)
  const int colm = i16 / R;    // component of column i16; block nb adds nb*16/R
  // one x block's accumulators go to its tile lines (ADD: onto what is there)
  // pidx(x0 + row) = x0 + x0/16 + row for the block's rows (row < 16,
  // x0 % 16 == 0): one lane base per column block, the rest uniform or
  // immediate, so the addresses need no registers across the kernel
  auto store_block = [&](int x0, const acc_t* acc, bool add) {
    const int xo = x0 + (x0 >> 4);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int l = nb * 16 + i16;
      if (l < NCOL) {
        T* pl = reinterpret_cast<T*>(buf + (l >> 1) * LS) + (l & 1) + 2 * xo;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T* d = pl + 2 * MF::row(lane, r);
          *d = add ? *d + acc[nb][r] : acc[nb][r];
        }
      }
    }
  };
  for (int rhs = rfirst; rhs < r1; rhs = next_live_rhs(a, rhs + 1, r1)) {
    const bool first = rhs == rfirst;
    if (!kRegAcc)
      for (int k = threadIdx.x; k < a.ncl * LS; k += blockDim.x) buf[k] = cmk<T>(0, 0);
    // the staged chunk, the cleared lines and the twiddles must be complete
    // before any other thread uses them (a tile without candidates goes
    // straight to the FFT)
    __syncthreads();
    acc_t acc[XPW][NB];
#pragma unroll
    for (int i = 0; i < XPW; ++i)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[i][nb] = acc_t{0, 0, 0, 0};
    for (int base = 0; base < total; base += a.ch) {
      const int nc = min(a.ch, total - base);
      if (!(first && base == 0)) {
        // the candidate ranges depend only on the image, so a restage of the
        // moments alone keeps them
        const bool image = !one_chunk || a.alias;
        mfma_stage_chunk<T, true, NC, NCP>(a, im, st, qt + base, nc, rhs, image);
        if (image) mfma_chunk_ranges<N>(im.gs + qt + base, nc, P, s_xr);
        __syncthreads();
      }
#pragma unroll
      for (int i = 0; i < XPW; ++i) {
        const int xb = wave + NW * i;
        if (xb < NXB) {
          const int x0 = xb * 16;
          if (!kRegAcc)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[i][nb] = acc_t{0, 0, 0, 0};
#pragma unroll
          for (int sg = 0; sg < 2; ++sg) {
            const int qa = __builtin_amdgcn_readfirstlane(s_xr[4 * xb + 2 * sg]);
            const int qb = __builtin_amdgcn_readfirstlane(s_xr[4 * xb + 2 * sg + 1]);
            // node offset o = x - gx, plus N for the candidates that wrap below node 0
            const int xo = x0 + i16 + ((sg == 0 && x0 - P + 1 < 0) ? N : 0);
            if (qa < qb) {
              // the first node, row weight and moments of each step's
              // candidate are loaded one step ahead, so a step waits only for
              // its x factor (the stage is zero-padded for the look-ahead)
              int cn = qa + k4;
              int gn = st.gx[cn];
              T wn = st.fw[__umul24(cn, PW) + P + colrow];
              T mn[NB];
#pragma unroll
              for (int nb = 0; nb < NB; ++nb) mn[nb] = st.m[cn * NCP + colm + nb * 16 / R];
              for (int q = qa; q < qb; q += 4) {
                const int cand = cn;
                const int o = xo - gn;
                const T wv = wn;
                T mq[NB];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) mq[nb] = mn[nb];
                cn = cand + 4;
                gn = st.gx[cn];
                wn = st.fw[__umul24(cn, PW) + P + colrow];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) mn[nb] = st.m[cn * NCP + colm + nb * 16 / R];
                // o outside [0, P) (negative ones wrap to large unsigned
                // values) reads the stencil's last factor and selects 0
                const T fv = st.fw[__umul24(cand, PW) + min((unsigned)o, (unsigned)(P - 1))];
                const T av = (unsigned)o < (unsigned)P ? fv : (T)0;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[i][nb] = MF::mac(av, wv * mq[nb], acc[i][nb]);
              }
            }
          }
          if (!kRegAcc) store_block(x0, acc[i], true);
        }
      }
      __syncthreads();
    }
    if (kRegAcc) {
#pragma unroll
      for (int i = 0; i < XPW; ++i) {
        const int xb = wave + NW * i;
        if (xb < NXB) store_block(xb * 16, acc[i], false);
      }
      __syncthreads();
    }
    fft_any<T, 0, N>(buf, LS, a.ncl, a.plan, tw, false, nullptr);
    // half spectrum of the NC*R real lines, packed two per complex line
    // (line pair p): the even line from the Hermitian part, the odd one from
    // the anti-Hermitian part of Z[k] and conj(Z[N - k]).  One wave per pair,
    // lanes along kx < N/2, then the kx = N/2 entries of every line.
    constexpr int NPAIR = NCOL / 2, HALF = N / 2;
    for (int pr = wave; pr < NPAIR; pr += NW) {
      const long long re = tile_line_row(a, 2 * pr, ty0, tz0, rhs, NC);
      const long long ro = tile_line_row(a, 2 * pr + 1, ty0, tz0, rhs, NC);
      const cplx<T>* line = buf + pr * LS;
      for (int k = lane; k < HALF; k += 64) {
        const cplx<T> Z = line[pidx(k)];
        const cplx<T> Zc = cconj(line[pidx(k == 0 ? 0 : N - k)]);
        if (re >= 0) spec[re * a.kxs + k] = cmk<T>((Z.x + Zc.x) * (T)0.5, (Z.y + Zc.y) * (T)0.5);
        if (ro >= 0) spec[ro * a.kxs + k] = cmk<T>((Z.y - Zc.y) * (T)0.5, (Zc.x - Z.x) * (T)0.5);
      }
    }
    for (int l = threadIdx.x; l < NCOL; l += blockDim.x) {
      const long long row = tile_line_row(a, l, ty0, tz0, rhs, NC);
      if (row < 0) continue;
      const cplx<T> Z = buf[(l >> 1) * LS + pidx(HALF)];
      const cplx<T> Zc = cconj(Z);
      spec[row * a.kxs + HALF] = (l & 1) == 0 ? cmk<T>((Z.x + Zc.x) * (T)0.5, (Z.y + Zc.y) * (T)0.5)
                                              : cmk<T>((Z.y - Zc.y) * (T)0.5, (Zc.x - Z.x) * (T)0.5);
    }
    // the lines are read until here; the next right-hand side writes them
    __syncthreads();
  }
}

// Sum over groups of R lanes (R = 4 or 8, aligned) of V values per lane (V a
// power of two, V >= R), transposed: on return lane g (= its index in the
// group) holds in v[0 .. V/R) the group sums of its values g V/R .. (g + 1)
// V/R - 1.  Stages pair lanes g and R-1-g (half-mirror, R = 8), then g^2, g^1
// (quad swaps); the lane in the upper half of a pair keeps the upper half of
// its current values.  Fixed order, so the sums are deterministic.
template <typename T, int CNT, int CTRL>
__device__ __forceinline__ void transpose_sum_stage(T* v, bool upper) {
  constexpr int H = CNT / 2;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const T keep = upper ? v[H + k] : v[k];
    const T send = upper ? v[k] : v[H + k];
    v[k] = keep + dpp_mov<CTRL>(send);
  }
}

template <typename T, int V, int R>
__device__ __forceinline__ void transpose_sum(T* v, int g) {
  static_assert((R == 4 || R == 8) && V % R == 0, "lane groups of 4 or 8");
  constexpr int V2 = R == 8 ? V / 2 : V;
  if constexpr (R == 8) transpose_sum_stage<T, V, kDppRowHalfMirror>(v, (g & 4) != 0);
  transpose_sum_stage<T, V2, kDppSwap2>(v, (g & 2) != 0);
  transpose_sum_stage<T, V2 / 2, kDppSwap1>(v, (g & 1) != 0);
}

// Inverse x-FFT + contraction.  Wave w handles blocks of 16 consecutive
// sorted candidates w, w + 4, ...: A[q][x] over the x span of the block (in
// unwrapped coordinates, the grid read at x mod nx), then the row weights and
// a reduction over the R lanes of each component, stored as the partial of
// image slot q (k_contract_reduce sums each particle's partials).  Needs 256
// threads (four waves) and lines of padded_len(N) entries.
template <typename T, int N, int NC, int R>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 1 : 2)
    k_xifft_contract_mfma(TileArgs a, ImgArgs im, const cplx<T>* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  static_assert(R == 4 || R == 8 || R == 16, "tile rows");
  constexpr int NCOL = NC * R, NB = (NCOL + 15) / 16, NW = 4, LS = padded_len(N);
  constexpr int NCP = mfma_ncp(NC, R), NPAIR = NCOL / 2, HALF = N / 2, KPL = (HALF + 63) / 64;
  // the transposing lane sum needs 4 NB (values per lane) a power of two >= R
  constexpr bool kTransposeSum = (R == 4 || R == 8) && (4 * NB) % R == 0 && ((4 * NB) & (4 * NB - 1)) == 0;
  // line pairs whose spectrum loads a wave keeps in flight together (32 VGPRs)
  constexpr int GP = (8 / (KPL * (int)sizeof(T) / 4)) > 0 ? 8 / (KPL * (int)sizeof(T) / 4) : 1;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + N;
  const int P = a.P, PW = P + R;
  const MfmaStage<T> st = mfma_stage_layout<T>(reinterpret_cast<unsigned char*>(buf + a.ncl * LS), a.ch, PW, NCP);
  const int tile = blockIdx.x + blockIdx.y * gridDim.x;
  const int ty0 = blockIdx.x * a.ty, tz0 = blockIdx.y * a.tz;
  // right-hand sides [r0, r1) in turn; the staged image chunk serves all of
  // them when the tile's candidates fit in one chunk
  const int r0 = blockIdx.z * a.rg, r1 = min(r0 + a.rg, a.nrhs);
  const int rfirst = next_live_rhs(a, r0, r1);
  if (rfirst >= r1) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const size_t qt = (size_t)im.off[tile];
  const int total = im.off[tile + 1] - (int)qt;
  const bool one_chunk = total <= a.ch;
  // the first chunk of the tile image goes to LDS while the spectrum loads
  mfma_stage_chunk<T, false, NC, NCP>(a, im, st, qt, min(a.ch, total), rfirst);
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  const int i16 = lane & 15, k4 = lane >> 4;
  const int colrow = i16 % R;
  int colc[NB];
  const T* colp[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int j = nb * 16 + i16;
    colc[nb] = j / R;
    // columns past the last line read the last line; their outputs are dropped
    const int jl = min(j, NCOL - 1);
    colp[nb] = reinterpret_cast<const T*>(buf + (jl >> 1) * LS) + (jl & 1);
  }
  for (int rhs = rfirst; rhs < r1; rhs = next_live_rhs(a, rhs + 1, r1)) {
    const bool first = rhs == rfirst;
    // pack the half spectra of line pair pr (real lines 2pr, 2pr + 1) into
    // one complex line Z = A + iB (Hermitian extension): one wave per pair,
    // lanes along kx < N/2, GP pairs' loads in flight before their LDS
    // writes; then the kx = N/2 entries
    for (int p0 = wave; p0 < NPAIR; p0 += NW * GP) {
      cplx<T> va[GP][KPL], vb[GP][KPL];
#pragma unroll
      for (int g = 0; g < GP; ++g) {
        const int pr = p0 + NW * g;
        const long long ra = pr < NPAIR ? tile_line_row(a, 2 * pr, ty0, tz0, rhs, NC) : -1;
        const long long rb = pr < NPAIR ? tile_line_row(a, 2 * pr + 1, ty0, tz0, rhs, NC) : -1;
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
          const int k = lane + 64 * j;
          va[g][j] = cmk<T>(0, 0);
          vb[g][j] = cmk<T>(0, 0);
          if (k < HALF) {
            if (ra >= 0) va[g][j] = spec[ra * a.kxs + k];
            if (rb >= 0) vb[g][j] = spec[rb * a.kxs + k];
          }
        }
      }
#pragma unroll
      for (int g = 0; g < GP; ++g) {
        const int pr = p0 + NW * g;
        if (pr < NPAIR) {
          cplx<T>* line = buf + pr * LS;
#pragma unroll
          for (int j = 0; j < KPL; ++j) {
            const int k = lane + 64 * j;
            if (k < HALF) {
              const cplx<T> A = va[g][j], B = vb[g][j];
              line[pidx(k)] = k == 0 ? cmk<T>(A.x, B.x) : cmk<T>(A.x - B.y, A.y + B.x);
              if (k != 0) line[pidx(N - k)] = cmk<T>(A.x + B.y, B.x - A.y);
            }
          }
        }
      }
    }
    for (int pr = threadIdx.x; pr < NPAIR; pr += blockDim.x) {
      const long long ra = tile_line_row(a, 2 * pr, ty0, tz0, rhs, NC);
      const long long rb = tile_line_row(a, 2 * pr + 1, ty0, tz0, rhs, NC);
      const T ax = ra >= 0 ? spec[ra * a.kxs + HALF].x : (T)0;
      const T bx = rb >= 0 ? spec[rb * a.kxs + HALF].x : (T)0;
      buf[pr * LS + pidx(HALF)] = cmk<T>(ax, bx);
    }
    __syncthreads();
    fft_any<T, 0, N>(buf, LS, a.ncl, a.plan, tw, true, nullptr);
    T* partr = reinterpret_cast<T*>(im.part) + (size_t)rhs * im.total * NC;
    for (int base = 0; base < total; base += a.ch) {
      const int nc = min(a.ch, total - base);
      if (base > 0 || (!first && !one_chunk)) {
        mfma_stage_chunk<T, false, NC, NCP>(a, im, st, qt + base, nc, rhs);
        __syncthreads();
      }
      T* pb = partr + (qt + base) * NC;
      const int nblk = (nc + 15) >> 4;
      for (int cb = wave; cb < nblk; cb += NW) {
        const int c0 = cb * 16;
        const int cand = c0 + i16;
        const bool ok = cand < nc;
        const int cq = ok ? cand : c0;
        // a lane past the chunk's last candidate gets a first node that puts
        // every node of the span outside its stencil
        const int gxc = ok ? st.gx[cq] : (1 << 29);
        const T* f = st.fw + __umul24(cq, PW);
        // x span [xs, xe) of the block's stencils, widened to multiples of 4
        // so that no step of four nodes straddles node N (N % 16 == 0); the
        // added nodes lie outside every stencil of the block
        const int xs = __builtin_amdgcn_readfirstlane(st.gx[c0]) & ~3;
        const int xe = (__builtin_amdgcn_readfirstlane(st.gx[min(c0 + 15, nc - 1)]) + P + 3) & ~3;
        acc_t acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = acc_t{0, 0, 0, 0};
        // nodes below N, then the ones that wrap to x - N (xe - N < N)
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
          const int lo = sg == 0 ? xs : N;
          const int hi = sg == 0 ? min(xe, N) : xe;
          const int wrap = sg == 0 ? 0 : N;
          for (int xx = lo; xx < hi; xx += 4) {
            const int xk = xx + k4;
            // o outside [0, P) (negative ones wrap to large unsigned values)
            // reads the stencil's last factor and selects 0
            const int o = xk - gxc;
            const T fv = f[min((unsigned)o, (unsigned)(P - 1))];
            const T av = (unsigned)o < (unsigned)P ? fv : (T)0;
            // pidx(xw + k4) = xw + xw/16 + k4 (xw % 4 == 0, k4 < 4): the
            // lane adds 2 k4 to a uniform offset
            const int xw = xx - wrap;
            const int px = 2 * (xw + (xw >> 4)) + 2 * k4;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[nb] = MF::mac(av, colp[nb][px], acc[nb]);
          }
        }
        // out[q][c] = sum_row w_q[row] D[q][(c, row)]: weight, then sum the R
        // lanes of each component
        if constexpr (kTransposeSum) {
          // the 4 NB weighted values of a lane, v[nb*4 + r], summed over the
          // R lanes of each component by a transposing butterfly: every stage
          // pairs lanes (mirror, then swaps), each keeps half of its values
          // and adds the partner's copy of them, so lane g of the R ends with
          // the sums of values g*VL .. g*VL + VL - 1 and writes them itself
          T v[4 * NB];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int cr = c0 + MF::row(lane, r);
            const T wq = st.fw[__umul24(cr < nc ? cr : c0, PW) + P + colrow];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) v[nb * 4 + r] = acc[nb][r] * wq;
          }
          transpose_sum<T, 4 * NB, R>(v, colrow);
          constexpr int VL = 4 * NB / R;
#pragma unroll
          for (int k = 0; k < VL; ++k) {
            const int idx = colrow * VL + k, nb = idx >> 2, r = idx & 3;
            const int cr = c0 + MF::row(lane, r);
            const int cc = nb * 16 / R + i16 / R;
            if (cc < NC && cr < nc) pb[cr * NC + cc] = v[k];
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int cr = c0 + MF::row(lane, r);
            const bool okr = cr < nc;
            const int crq = okr ? cr : c0;
            const T wq = st.fw[__umul24(crq, PW) + P + colrow];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
              const T v = group_sum<T>(acc[nb][r] * wq, R);
              if (colc[nb] < NC && colrow == 0 && okr) pb[cr * NC + colc[nb]] = v;
            }
          }
        }
      }
      // the lines and the staged chunk are read until here
      __syncthreads();
    }
    // the next right-hand side writes the lines (a tile without candidates
    // leaves the chunk loop out)
    if (total == 0) __syncthreads();
  }
}

}  // namespace mpol
