// Device kernels of the batched restarted GMRES (FP64, real arithmetic).
//
// Krylov basis layout: V[i][rhs][k] (basis index outermost) so that V_j for all
// right-hand sides is one contiguous (nrhs, n) block, which is exactly the
// operator's input layout.  Hessenberg H[rhs][i][j], (restart+1) x restart.
#pragma once

#include <hip/hip_runtime.h>

#include "mpol_common.h"

namespace mpol {

constexpr int kDotBlocks = 256;  // partial-sum blocks per right-hand side (one per CU)

// Vectors whose per-wave sums are staged in LDS between two barriers, and
// vectors read together in one sweep over a block's chunk.
constexpr int kDotGroup = 64;
constexpr int kDotSweep = 8;

// Elements of a thread whose loads a sweep issues together (KU per vector).
constexpr int kDotUnroll = 4;

// Partial dot products of one block's chunk [lo, hi) of right-hand side rhs:
// partial[(rhs*cnt + i)*gridDim.x + blockIdx.x] = sum over the chunk of
// a_i . w, a_i = A + i*stride_i + rhs*n (kDotSweep vectors per sweep, w read
// once per sweep).  Thread t handles elements lo + t, lo + t + blockDim.x, ...
// in increasing order, kDotUnroll of them per batch: every load of a batch is
// unconditional (indices clamped into the chunk and to the last vector, the
// products past the chunk multiplied by a zero w), so a batch waits for memory
// once.  Each wave's sums go to LDS (red[vector][wave]) for up to kDotGroup
// vectors between two barriers, so the sweeps of a group follow each other
// without waiting; the four wave sums are added in a fixed order.
__device__ __forceinline__ void block_dots(const double* __restrict__ A, size_t stride_i, int cnt,
                                           const double* __restrict__ wr, size_t roff, int64_t lo, int64_t hi,
                                           int rhs, double* __restrict__ partial, double (*red)[4]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t step = blockDim.x;
  for (int g0 = 0; g0 < cnt; g0 += kDotGroup) {
    const int gc = min(kDotGroup, cnt - g0);
    for (int i0 = 0; i0 < gc; i0 += kDotSweep) {
      double s[kDotSweep];
#pragma unroll
      for (int u = 0; u < kDotSweep; ++u) s[u] = 0.0;
      const double* av[kDotSweep];
#pragma unroll
      for (int u = 0; u < kDotSweep; ++u) av[u] = A + (size_t)(g0 + min(i0 + u, gc - 1)) * stride_i + roff;
      for (int64_t kb = lo + threadIdx.x; kb < hi; kb += kDotUnroll * step) {
        double wv[kDotUnroll], a[kDotUnroll][kDotSweep];
#pragma unroll
        for (int e = 0; e < kDotUnroll; ++e) {
          const int64_t k = kb + e * step;
          const int64_t kc = k < hi ? k : kb;
          const double w0 = wr[kc];
          wv[e] = k < hi ? w0 : 0.0;
#pragma unroll
          for (int u = 0; u < kDotSweep; ++u) a[e][u] = av[u][kc];
        }
#pragma unroll
        for (int e = 0; e < kDotUnroll; ++e)
#pragma unroll
          for (int u = 0; u < kDotSweep; ++u) s[u] += a[e][u] * wv[e];
      }
#pragma unroll
      for (int u = 0; u < kDotSweep; ++u)
        for (int o = 32; o > 0; o >>= 1) s[u] += __shfl_xor(s[u], o);
      if (lane == 0) {
#pragma unroll
        for (int u = 0; u < kDotSweep; ++u)
          if (i0 + u < gc) red[i0 + u][wave] = s[u];
      }
    }
    __syncthreads();
    for (int u = threadIdx.x; u < gc; u += blockDim.x)
      partial[((size_t)rhs * cnt + g0 + u) * gridDim.x + blockIdx.x] = (red[u][0] + red[u][1]) + (red[u][2] + red[u][3]);
    __syncthreads();
  }
}

// Chunk of block blockIdx.x when n elements are split over gridDim.x blocks.
__device__ __forceinline__ void block_chunk(int64_t n, int64_t& lo, int64_t& hi) {
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  lo = (int64_t)blockIdx.x * chunk;
  hi = lo + chunk < n ? lo + chunk : n;
}

// partial dot products of a_i . w (i < cnt) for every rhs; grid (blocks, nrhs)
static __global__ void __launch_bounds__(256) k_dots(const double* __restrict__ A, size_t stride_i, int cnt,
                                              const double* __restrict__ w, int64_t n, int nrhs,
                                              double* __restrict__ partial) {
  __shared__ double red[kDotGroup][4];  // [vector of the group][wave]
  const int rhs = blockIdx.y;
  int64_t lo, hi;
  block_chunk(n, lo, hi);
  block_dots(A, stride_i, cnt, w + (size_t)rhs * n, (size_t)rhs * n, lo, hi, rhs, partial, red);
}

// Sum of the nblk partials at pt in a fixed order, by one wave: lane l adds
// partials l, l + 64, ... (four loads in flight), then a fixed butterfly;
// every lane returns the sum.
__device__ __forceinline__ double finish_sum(const double* __restrict__ pt, int nblk, int lane) {
  double s = 0.0;
  for (int b0 = lane; b0 < nblk; b0 += 4 * 64) {
    double p[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) p[e] = pt[min(b0 + 64 * e, nblk - 1)];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (b0 + 64 * e < nblk) s += p[e];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  return s;
}

// out[t] = sum of the nblk partials of output t (finish_sum, one wave per output).
static __global__ void k_dots_finish(const double* __restrict__ partial, int nblk, int total, double* __restrict__ out) {
  const int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= total) return;
  const double s = finish_sum(partial + (size_t)t * nblk, nblk, lane);
  if (lane == 0) out[t] = s;
}

// Largest Krylov dimension per cycle (the C ABI's restart bound).
constexpr int kMaxKrylov = 200;

// One Gram-Schmidt pass fused with the reductions on either side, per rhs in
// kIter (grid (blocks, nrhs), chunks as in k_dots; hpart holds the partial
// sums of the pass's coefficients from a launch of the same grid):
//   h[rhs*cnt + i] = sum of the partials of coefficient i (finish_sum, the
//                    order of k_dots_finish; block 0 also stores them to h)
//   w[rhs] -= sum_{i<cnt} h[rhs*cnt + i] V_i[rhs]
// then the partial sums of V_i . w (i < cnt; next_norm = 0, the second pass's
// coefficients) or of w . w (next_norm = 1, ||w||^2) over the block's chunk
// into partial (not hpart: other blocks still read it).  A thread re-reads
// only the elements of w that it wrote itself.
static __global__ void __launch_bounds__(256) k_sub_dots(const double* __restrict__ V, size_t stride_i, int cnt,
                                                  const double* __restrict__ hpart, double* __restrict__ h,
                                                  double* __restrict__ w, int64_t n,
                                                  const int* __restrict__ state, int next_norm,
                                                  double* __restrict__ partial) {
  __shared__ double red[kDotGroup][4];
  __shared__ double hs[kMaxKrylov + 1];
  const int rhs = blockIdx.y;
  if (state[rhs] != kIter) return;  // uniform over the block
  {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int t = wave; t < cnt; t += blockDim.x >> 6) {
      const double v = finish_sum(hpart + ((size_t)rhs * cnt + t) * gridDim.x, gridDim.x, lane);
      if (lane == 0) {
        hs[t] = v;
        if (blockIdx.x == 0) h[(size_t)rhs * cnt + t] = v;
      }
    }
    __syncthreads();
  }
  int64_t lo, hi;
  block_chunk(n, lo, hi);
  const size_t roff = (size_t)rhs * n;
  double* wr = w + roff;
  const double* hr = hs;
  // w_k -= h_i V_i,k in increasing i, kDotSweep vectors' loads of kDotUnroll
  // elements in flight together (clamped indices; the padded terms subtract
  // a zero coefficient)
  const int64_t step = blockDim.x;
  for (int64_t kb = lo + threadIdx.x; kb < hi; kb += kDotUnroll * step) {
    double acc[kDotUnroll];
    int64_t kc[kDotUnroll];
#pragma unroll
    for (int e = 0; e < kDotUnroll; ++e) {
      const int64_t k = kb + e * step;
      kc[e] = k < hi ? k : kb;
      acc[e] = wr[kc[e]];
    }
    for (int i0 = 0; i0 < cnt; i0 += kDotSweep) {
      double v[kDotUnroll][kDotSweep], hc[kDotSweep];
#pragma unroll
      for (int u = 0; u < kDotSweep; ++u) {
        const int i = min(i0 + u, cnt - 1);
        hc[u] = i0 + u < cnt ? hr[i] : 0.0;
        const double* vi = V + (size_t)i * stride_i + roff;
#pragma unroll
        for (int e = 0; e < kDotUnroll; ++e) v[e][u] = vi[kc[e]];
      }
#pragma unroll
      for (int e = 0; e < kDotUnroll; ++e)
#pragma unroll
        for (int u = 0; u < kDotSweep; ++u) acc[e] -= hc[u] * v[e][u];
    }
#pragma unroll
    for (int e = 0; e < kDotUnroll; ++e)
      if (kb + e * step < hi) wr[kb + e * step] = acc[e];
  }
  if (next_norm) block_dots(wr - roff, 0, 1, wr, roff, lo, hi, rhs, partial, red);
  else block_dots(V, stride_i, cnt, wr, roff, lo, hi, rhs, partial, red);
}

// r = b - Ax
static __global__ void k_residual(const double* __restrict__ b, const double* __restrict__ ax, double* __restrict__ r,
                           int64_t total) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < total) r[k] = b[k] - ax[k];
}

static __global__ void k_copy(const double* __restrict__ src, double* __restrict__ dst, int64_t total) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < total) dst[k] = src[k];
}

static __global__ void k_axpy_all(double a, const double* __restrict__ x, double* __restrict__ y, int64_t total) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < total) y[k] += a * x[k];
}

// to = from, then from = 0 (the residual becomes the right-hand side of a
// correction solve whose correction starts at zero in its place).
static __global__ void k_take(double* __restrict__ from, double* __restrict__ to, int64_t total) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  to[k] = from[k];
  from[k] = 0.0;
}

// ||b|| from ||b||^2, and every system's first state: kWait, or kDone for b = 0.
static __global__ void k_norms_init(int nrhs, const double* __restrict__ bnorm2, double* __restrict__ bnorm,
                                    int* __restrict__ state) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs >= nrhs) return;
  const double bn = sqrt(bnorm2[rhs]);
  bnorm[rhs] = bn;
  state[rhs] = bn > 0.0 ? kWait : kDone;
}

// Start of a restart cycle: rnorm2[rhs] = ||b - A x||^2.
// kWait/kIter -> kDone if converged, else kIter with g = [beta, 0, ...].
// The new state and relative residual also go to post_state / post_rel
// (page-locked host memory the host reads after an event).
static __global__ void k_cycle_start(int nrhs, int restart, const double* __restrict__ rnorm2,
                                     const double* __restrict__ bnorm, double tol, int* __restrict__ state,
                                     double* __restrict__ g, double* __restrict__ beta_out,
                                     double* __restrict__ relres, int* __restrict__ post_state,
                                     double* __restrict__ post_rel) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs >= nrhs) return;
  const double beta = sqrt(rnorm2[rhs]);
  const double rel = bnorm[rhs] > 0.0 ? beta / bnorm[rhs] : 0.0;
  relres[rhs] = rel;
  beta_out[rhs] = beta;
  post_rel[rhs] = rel;
  int s = state[rhs];
  if (s != kDone) {
    if (rel <= tol || beta == 0.0) {
      s = kDone;
    } else {
      s = kIter;
      double* gr = g + (size_t)rhs * (restart + 1);
      gr[0] = beta;
      for (int i = 1; i <= restart; ++i) gr[i] = 0.0;
    }
    state[rhs] = s;
  }
  post_state[rhs] = s;
}

// V_0[rhs] = r[rhs] / beta[rhs] for rhs in kIter.
static __global__ void k_scale_into(const double* __restrict__ r, const double* __restrict__ beta,
                             double* __restrict__ v, int64_t n, const int* __restrict__ state) {
  const int rhs = blockIdx.y;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double bt = beta[rhs];
  v[(size_t)rhs * n + k] = (state[rhs] == kIter && bt != 0.0) ? r[(size_t)rhs * n + k] / bt : 0.0;
}

// V_{j+1} = w / ||w|| (hnorm2 = ||w||^2) for rhs in kIter.
static __global__ void k_normalize_next(const double* __restrict__ w, double* __restrict__ vnext,
                                        const double* __restrict__ hnorm2, int64_t n, const int* __restrict__ state) {
  const int rhs = blockIdx.y;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || state[rhs] != kIter) return;
  const double hn = sqrt(hnorm2[rhs]);
  vnext[(size_t)rhs * n + k] = hn != 0.0 ? w[(size_t)rhs * n + k] / hn : 0.0;
}

// One Arnoldi step of the small problem, per right-hand side in kIter:
// column j of H from the two Gram-Schmidt passes (h1 + h2) and ||w||, the
// Givens rotations on that column, the decision whether the cycle ends
// (estimated relative residual |g_{j+1}| / ||b|| in relest), and for a cycle
// that ends, the back substitution of the rotated H y = g.  Every system's
// state and residual estimate after the step also go to post_state /
// post_rel (page-locked host memory the host reads after an event).
// hn2 = ||w||^2 of this system.
__device__ inline void arnoldi_step_rhs(int rhs, int restart, int j, const double* __restrict__ h1,
                                        const double* __restrict__ h2, double hn2, double* __restrict__ H,
                                        double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ g,
                                        const double* __restrict__ bnorm, double tol, int* __restrict__ state,
                                        int* __restrict__ kdim, double* __restrict__ relest, double* __restrict__ y,
                                        int* __restrict__ post_state, double* __restrict__ post_rel) {
  if (state[rhs] != kIter) {
    post_state[rhs] = state[rhs];
    post_rel[rhs] = relest[rhs];
    return;
  }
  double* Hr = H + (size_t)rhs * (restart + 1) * restart;
  double* c = cs + (size_t)rhs * restart;
  double* s = sn + (size_t)rhs * restart;
  double* gr = g + (size_t)rhs * (restart + 1);
  const double* h1r = h1 + (size_t)rhs * (j + 1);
  const double* h2r = h2 + (size_t)rhs * (j + 1);
  // the earlier rotations on the new column h1 + h2: rotation i mixes entries
  // i and i + 1, so entry i + 1 enters unrotated and the running entry stays
  // in a register (no store-to-load chain through H)
  double a = h1r[0] + h2r[0];
  for (int i = 0; i < j; ++i) {
    const double b = h1r[i + 1] + h2r[i + 1];
    Hr[i * restart + j] = c[i] * a + s[i] * b;
    a = -s[i] * a + c[i] * b;
  }
  const double b = sqrt(hn2);
  const bool breakdown = (b == 0.0);
  const double den = sqrt(a * a + b * b);
  const double cj = den == 0.0 ? 1.0 : a / den, sj = den == 0.0 ? 0.0 : b / den;
  c[j] = cj;
  s[j] = sj;
  Hr[j * restart + j] = cj * a + sj * b;
  Hr[(j + 1) * restart + j] = 0.0;
  const double gj = gr[j];
  gr[j] = cj * gj;
  gr[j + 1] = -sj * gj;
  const double rel = fabs(gr[j + 1]) / bnorm[rhs];
  relest[rhs] = rel;
  post_rel[rhs] = rel;
  if (!(rel <= tol || j == restart - 1 || breakdown)) {
    post_state[rhs] = kIter;
    return;
  }
  state[rhs] = kSolve;
  post_state[rhs] = kSolve;
  const int k = j + 1;
  kdim[rhs] = k;
  double* yr = y + (size_t)rhs * restart;
  for (int i = k - 1; i >= 0; --i) {
    double sum = gr[i];
    for (int l = i + 1; l < k; ++l) sum -= Hr[i * restart + l] * yr[l];
    const double d = Hr[i * restart + i];
    yr[i] = d != 0.0 ? sum / d : 0.0;
  }
}

// ||w||^2 of every system from the partials of the second Gram-Schmidt pass
// (finish_sum; one 64-thread block per system) into hnorm2, then the Arnoldi
// step (arnoldi_step_rhs) of the system.
static __global__ void k_finish_arnoldi(const double* __restrict__ partial, int nblk, int restart, int j,
                                        const double* __restrict__ h1, const double* __restrict__ h2,
                                        double* __restrict__ hnorm2, double* __restrict__ H, double* __restrict__ cs,
                                        double* __restrict__ sn, double* __restrict__ g,
                                        const double* __restrict__ bnorm, double tol, int* __restrict__ state,
                                        int* __restrict__ kdim, double* __restrict__ relest, double* __restrict__ y,
                                        int* __restrict__ post_state, double* __restrict__ post_rel) {
  const int rhs = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const double hn2 = finish_sum(partial + (size_t)rhs * nblk, nblk, lane);
  if (threadIdx.x != 0) return;
  hnorm2[rhs] = hn2;
  arnoldi_step_rhs(rhs, restart, j, h1, h2, hn2, H, cs, sn, g, bnorm, tol, state, kdim, relest, y, post_state,
                   post_rel);
}

// x[rhs] += sum_i y_i V_i[rhs] for rhs in kSolve.
static __global__ void k_update_x(const double* __restrict__ V, size_t stride_i, const double* __restrict__ y,
                           const int* __restrict__ kdim, const int* __restrict__ state, int restart,
                           double* __restrict__ x, int64_t n) {
  const int rhs = blockIdx.y;
  if (state[rhs] != kSolve) return;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int kd = kdim[rhs];
  const size_t o = (size_t)rhs * n + k;
  const double* yr = y + (size_t)rhs * restart;
  double acc = x[o];
  // increasing i, kDotSweep loads in flight (clamped; padded terms add 0)
  for (int i0 = 0; i0 < kd; i0 += kDotSweep) {
    double v[kDotSweep], yc[kDotSweep];
#pragma unroll
    for (int u = 0; u < kDotSweep; ++u) {
      const int i = min(i0 + u, kd - 1);
      yc[u] = i0 + u < kd ? yr[i] : 0.0;
      v[u] = V[i * stride_i + o];
    }
#pragma unroll
    for (int u = 0; u < kDotSweep; ++u) acc += yc[u] * v[u];
  }
  x[o] = acc;
}

static __global__ void k_solve_to_wait(int nrhs, int* __restrict__ state) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs < nrhs && state[rhs] == kSolve) state[rhs] = kWait;
}

}  // namespace mpol
