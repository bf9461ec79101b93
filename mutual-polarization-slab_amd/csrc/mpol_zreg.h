// z-axis FFT + dipole k-space scale + inverse z-FFT for nz = 256 with the
// transforms in registers (gfx950).
//
// Four-step 16 x 16 decomposition, z = n1 + 16 n2, k = k2 + 16 k1:
//   Y[n1][k2] = W_256^(n1 k2) sum_n2 x[n1 + 16 n2] W_16^(n2 k2)   (thread (col, n1))
//   X[k2 + 16 k1] = sum_n1 Y[n1][k2] W_16^(n1 k1)                   (thread (col, k2))
// with one LDS transpose between the two 16-point DFTs; the inverse runs the
// same steps backwards, so a thread ends holding the z values it loaded.
// A thread keeps the three dipole components of its 16 values, so the scale
// G = c(k) M(k^) F needs no exchange, and one TX x 256 transpose buffer is
// reused by the components: with TX = 16 (FP32) / 8 (FP64) kx columns per
// workgroup every load and store is a 128-byte segment, at about a third of
// the LDS of the LDS-resident z pass.
#pragma once

#include "mpol_kernels.h"

namespace mpol {

template <typename T>
__device__ __forceinline__ void idft16(cplx<T>* v) {
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = cconj(v[i]);
  dft16<T>(v);
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = cconj(v[i]);
}

// Dipole k-space scale at (kx, y, kz): the same coefficient and Nyquist rule
// as k_fft_z_scale (cross terms with exactly one Nyquist axis vanish).
template <typename T>
__device__ __forceinline__ void dipole_scale_point(const ColArgs& a, const ScaleParams& sp, int kx, int y, int kz,
                                                   cplx<T>& F0, cplx<T>& F1, cplx<T>& F2) {
  const int mx = fft_freq(kx, a.nx), my = fft_freq(y, a.ny), mz = fft_freq(kz, a.nz);
  if (mx == 0 && my == 0 && mz == 0) {
    F0 = F1 = F2 = cmk<T>(0, 0);
    return;
  }
  const bool nyq_x = (a.nx % 2 == 0) && (2 * mx == -a.nx);
  const bool nyq_y = (a.ny % 2 == 0) && (2 * my == -a.ny);
  const bool nyq_z = (a.nz % 2 == 0) && (2 * mz == -a.nz);
  const double k0 = mx * sp.twopi_over_l[0], k1 = my * sp.twopi_over_l[1], k2 = mz * sp.twopi_over_l[2];
  const double kk = (k0 * k0 + k1 * k1) + k2 * k2;
  T u0, u1, u2;
  if constexpr (sizeof(T) == 4) {
    // FP32 operator: hardware reciprocal square root (about 1 ulp)
    const float ikn = __builtin_amdgcn_rsqf((float)kk);
    u0 = (float)k0 * ikn;
    u1 = (float)k1 * ikn;
    u2 = (float)k2 * ikn;
  } else {
    const double ikn = 1.0 / sqrt(kk);
    u0 = k0 * ikn;
    u1 = k1 * ikn;
    u2 = k2 * ikn;
  }
  double cf;
  if (sp.jt) cf = sp.jt[mx * mx + my * my + mz * mz] * (sp.ex[kx] * sp.ey[y] * sp.ez[kz]) * sp.inv_g;
  else cf = kspace_coef(sp, k0, k1, k2) * sp.inv_g;
  const T c = (T)cf;
  const T x01 = (nyq_x != nyq_y) ? (T)0 : u0 * u1;
  const T x02 = (nyq_x != nyq_z) ? (T)0 : u0 * u2;
  const T x12 = (nyq_y != nyq_z) ? (T)0 : u1 * u2;
  const T d00 = u0 * u0, d11 = u1 * u1, d22 = u2 * u2;
  const cplx<T> G0 = cmk<T>(c * (d00 * F0.x + x01 * F1.x + x02 * F2.x), c * (d00 * F0.y + x01 * F1.y + x02 * F2.y));
  const cplx<T> G1 = cmk<T>(c * (x01 * F0.x + d11 * F1.x + x12 * F2.x), c * (x01 * F0.y + d11 * F1.y + x12 * F2.y));
  const cplx<T> G2 = cmk<T>(c * (x02 * F0.x + x12 * F1.x + d22 * F2.x), c * (x02 * F0.y + x12 * F1.y + d22 * F2.y));
  F0 = G0;
  F1 = G1;
  F2 = G2;
}

// Thread (col, j), col = threadIdx.x % TX, j = threadIdx.x / TX (16 values).
template <typename T, int TX>
__global__ void __launch_bounds__(TX * 16) k_fft_z_scale_reg256(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  constexpr int N = 256, NC = 3, LS = 273;  // LS: padded line of the transpose buffer
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> xb[TX * LS];
  const int col = threadIdx.x % TX, j = threadIdx.x / TX;
  const int kx = blockIdx.x * TX + col, y = blockIdx.y, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const bool valid = kx < a.kx;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  const size_t plane = (size_t)a.ny * a.kxs;
  cplx<T> v[NC][16];
  // load z = j + 16 n2 of the three components (all loads in flight together)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const cplx<T>* src = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[c][n2] = valid ? src[(size_t)(j + 16 * n2) * plane] : cmk<T>(0, 0);
  }
  cplx<T>* line = xb + col * LS;
  __syncthreads();  // twiddles
  // forward: DFT over n2, twiddle, transpose, DFT over n1
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    dft16<T>(v[c]);
#pragma unroll
    for (int k2 = 1; k2 < 16; ++k2) v[c][k2] = cmul(v[c][k2], tw[(j * k2) & (N - 1)]);
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) line[pidx(k2 * 16 + j)] = v[c][k2];
    __syncthreads();
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[c][n1] = line[pidx(j * 16 + n1)];
    __syncthreads();
    dft16<T>(v[c]);  // thread (col, k2 = j) now holds X[j + 16 k1]
  }
  if (valid) {
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) dipole_scale_point<T>(a, sp, kx, y, j + 16 * k1, v[0][k1], v[1][k1], v[2][k1]);
  }
  // inverse: DFT over k1, conjugate twiddle, transpose, DFT over k2
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    idft16<T>(v[c]);  // index n1
#pragma unroll
    for (int n1 = 1; n1 < 16; ++n1) v[c][n1] = cmul(v[c][n1], cconj(tw[(j * n1) & (N - 1)]));
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) line[pidx(n1 * 16 + j)] = v[c][n1];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) v[c][k2] = line[pidx(j * 16 + k2)];
    __syncthreads();
    idft16<T>(v[c]);  // thread (col, n1 = j) now holds x[j + 16 n2]
  }
  if (valid) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cplx<T>* dst = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx;
#pragma unroll
      for (int n2 = 0; n2 < 16; ++n2) dst[(size_t)(j + 16 * n2) * plane] = v[c][n2];
    }
  }
}

}  // namespace mpol
