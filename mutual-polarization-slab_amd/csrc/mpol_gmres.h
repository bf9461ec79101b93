// Device kernels of the batched restarted GMRES (FP64, real arithmetic).
//
// Krylov basis layout: V[i][rhs][k] (basis index outermost) so that V_j for all
// right-hand sides is one contiguous (nrhs, n) block, which is exactly the
// operator's input layout.  Hessenberg H[rhs][i][j], (restart+1) x restart.
#pragma once

#include <hip/hip_runtime.h>

namespace mpol {

enum RhsState : int { kDone = 0, kIter = 1, kSolve = 2, kWait = 3 };

constexpr int kDotBlocks = 256;  // partial-sum blocks per right-hand side (one per CU)

// partial[(rhs*cnt + i)*kDotBlocks + blk] = sum over the block's chunk of a_i . w
// with a_i = A + i*stride_i (rhs offset rhs*n).  Deterministic.
static __global__ void __launch_bounds__(256) k_dots(const double* __restrict__ A, size_t stride_i, int cnt,
                                              const double* __restrict__ w, int64_t n, int nrhs,
                                              double* __restrict__ partial) {
  __shared__ double red[4][4];  // [vector of the group][wave]
  const int rhs = blockIdx.y;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  const double* wr = w + (size_t)rhs * n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // four dot products per sweep over the chunk (w read once per group)
  for (int i0 = 0; i0 < cnt; i0 += 4) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t k = lo + threadIdx.x; k < hi; k += blockDim.x) {
      const double wv = wr[k];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i0 + u < cnt) s[u] += A[(size_t)(i0 + u) * stride_i + (size_t)rhs * n + k] * wv;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      for (int o = 32; o > 0; o >>= 1) s[u] += __shfl_xor(s[u], o);
    if (lane == 0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) red[u][wave] = s[u];
    }
    __syncthreads();
    if ((int)threadIdx.x < 4 && i0 + (int)threadIdx.x < cnt) {
      const int u = threadIdx.x;
      partial[((size_t)rhs * cnt + i0 + u) * gridDim.x + blockIdx.x] = (red[u][0] + red[u][1]) + (red[u][2] + red[u][3]);
    }
    __syncthreads();
  }
}

// out[rhs*cnt + i] = sum of the partials (fixed order): one wave per output,
// lanes load the partials in parallel, then a fixed butterfly.
static __global__ void k_dots_finish(const double* __restrict__ partial, int nblk, int total, double* __restrict__ out) {
  const int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= total) return;
  double s = 0.0;
  for (int b = lane; b < nblk; b += 64) s += partial[(size_t)t * nblk + b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[t] = s;
}

// w[rhs] -= sum_i h[rhs*cnt + i] V_i[rhs]   (only rhs in state kIter)
static __global__ void k_sub_basis(const double* __restrict__ V, size_t stride_i, int cnt, const double* __restrict__ h,
                            double* __restrict__ w, int64_t n, const int* __restrict__ state) {
  const int rhs = blockIdx.y;
  if (state[rhs] != kIter) return;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double acc = w[(size_t)rhs * n + k];
  for (int i = 0; i < cnt; ++i) acc -= h[rhs * cnt + i] * V[i * stride_i + (size_t)rhs * n + k];
  w[(size_t)rhs * n + k] = acc;
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

// Start of a restart cycle: rnorm2[rhs] = ||b - A x||^2.
// kWait/kIter -> kDone if converged, else kIter with g = [beta, 0, ...].
static __global__ void k_cycle_start(int nrhs, int restart, const double* __restrict__ rnorm2,
                              const double* __restrict__ bnorm, double tol, int* __restrict__ state,
                              double* __restrict__ g, double* __restrict__ beta_out, double* __restrict__ relres) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs >= nrhs) return;
  const double beta = sqrt(rnorm2[rhs]);
  const double rel = bnorm[rhs] > 0.0 ? beta / bnorm[rhs] : 0.0;
  relres[rhs] = rel;
  beta_out[rhs] = beta;
  if (state[rhs] == kDone) return;
  if (rel <= tol || beta == 0.0) {
    state[rhs] = kDone;
    return;
  }
  state[rhs] = kIter;
  double* gr = g + (size_t)rhs * (restart + 1);
  gr[0] = beta;
  for (int i = 1; i <= restart; ++i) gr[i] = 0.0;
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

// Store h (+ h2) into column j of H and normalise w into V_{j+1}; hnorm2 = ||w||^2.
static __global__ void k_arnoldi_finish(int nrhs, int restart, int j, const double* __restrict__ h1,
                                 const double* __restrict__ h2, const double* __restrict__ hnorm2,
                                 double* __restrict__ H, const int* __restrict__ state) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs >= nrhs || state[rhs] != kIter) return;
  double* Hr = H + (size_t)rhs * (restart + 1) * restart;
  for (int i = 0; i <= j; ++i) Hr[i * restart + j] = h1[rhs * (j + 1) + i] + h2[rhs * (j + 1) + i];
  Hr[(j + 1) * restart + j] = sqrt(hnorm2[rhs]);
}

static __global__ void k_normalize_next(const double* __restrict__ w, double* __restrict__ vnext, const double* __restrict__ H,
                                 int restart, int j, int64_t n, const int* __restrict__ state) {
  const int rhs = blockIdx.y;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || state[rhs] != kIter) return;
  const double hn = H[(size_t)rhs * (restart + 1) * restart + (size_t)(j + 1) * restart + j];
  vnext[(size_t)rhs * n + k] = hn != 0.0 ? w[(size_t)rhs * n + k] / hn : 0.0;
}

// Givens rotations on column j; decide whether the cycle ends for each rhs;
// relest[rhs] = the estimated relative residual |g_{j+1}| / ||b||.
static __global__ void k_givens(int nrhs, int restart, int j, double* __restrict__ H, double* __restrict__ cs,
                         double* __restrict__ sn, double* __restrict__ g, const double* __restrict__ bnorm,
                         double tol, int* __restrict__ state, int* __restrict__ kdim, double* __restrict__ relest) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs >= nrhs || state[rhs] != kIter) return;
  double* Hr = H + (size_t)rhs * (restart + 1) * restart;
  double* c = cs + (size_t)rhs * restart;
  double* s = sn + (size_t)rhs * restart;
  double* gr = g + (size_t)rhs * (restart + 1);
  for (int i = 0; i < j; ++i) {
    const double a = Hr[i * restart + j], b = Hr[(i + 1) * restart + j];
    Hr[i * restart + j] = c[i] * a + s[i] * b;
    Hr[(i + 1) * restart + j] = -s[i] * a + c[i] * b;
  }
  const double a = Hr[j * restart + j], b = Hr[(j + 1) * restart + j];
  const bool breakdown = (b == 0.0);
  const double den = sqrt(a * a + b * b);
  if (den == 0.0) {
    c[j] = 1.0;
    s[j] = 0.0;
  } else {
    c[j] = a / den;
    s[j] = b / den;
  }
  Hr[j * restart + j] = c[j] * a + s[j] * b;
  Hr[(j + 1) * restart + j] = 0.0;
  const double gj = gr[j];
  gr[j] = c[j] * gj;
  gr[j + 1] = -s[j] * gj;
  const double rel = fabs(gr[j + 1]) / bnorm[rhs];
  relest[rhs] = rel;
  if (rel <= tol || j == restart - 1 || breakdown) {
    state[rhs] = kSolve;
    kdim[rhs] = j + 1;
  }
}

// Back substitution of the rotated (upper triangular) H y = g for rhs in kSolve.
static __global__ void k_backsolve(int nrhs, int restart, const double* __restrict__ H, const double* __restrict__ g,
                            const int* __restrict__ kdim, const int* __restrict__ state, double* __restrict__ y) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs >= nrhs || state[rhs] != kSolve) return;
  const double* Hr = H + (size_t)rhs * (restart + 1) * restart;
  const double* gr = g + (size_t)rhs * (restart + 1);
  double* yr = y + (size_t)rhs * restart;
  const int k = kdim[rhs];
  for (int i = k - 1; i >= 0; --i) {
    double s = gr[i];
    for (int l = i + 1; l < k; ++l) s -= Hr[i * restart + l] * yr[l];
    const double d = Hr[i * restart + i];
    yr[i] = d != 0.0 ? s / d : 0.0;
  }
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
  double acc = x[(size_t)rhs * n + k];
  for (int i = 0; i < kd; ++i) acc += y[(size_t)rhs * restart + i] * V[i * stride_i + (size_t)rhs * n + k];
  x[(size_t)rhs * n + k] = acc;
}

static __global__ void k_solve_to_wait(int nrhs, int* __restrict__ state) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs < nrhs && state[rhs] == kSolve) state[rhs] = kWait;
}

}  // namespace mpol
