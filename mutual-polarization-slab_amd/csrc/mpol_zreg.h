// z-axis FFT + dipole k-space scale + inverse z-FFT for nz = 256 with the
// transforms in registers (gfx950).
//
// Four-step 16 x 16 decomposition, z = n1 + 16 n2, k = k2 + 16 k1:
//   Y[n1][k2] = W_256^(n1 k2) sum_n2 x[n1 + 16 n2] W_16^(n2 k2)   (lane pair (col, n1))
//   X[k2 + 16 k1] = sum_n1 Y[n1][k2] W_16^(n1 k1)                   (lane pair (col, k2))
// with one LDS transpose between the two 16-point DFTs; the inverse runs the
// same steps backwards.  A lane (k_fft_z_scale_reg256) or a lane pair
// (k_fft_z_scale_reg256_pair) keeps the three dipole components of its 16
// points, so the scale G = c(k) M(k^) F needs no exchange, and one TX x 256
// transpose buffer is reused by the components: with TX = 16 (FP32) / 8
// (FP64) kx columns per workgroup every load and store is a 128-byte
// segment.  In the pair kernel each 16-point DFT is split over neighbouring
// lanes: lane h holds inputs 8h .. 8h + 7, the pair exchanges halves with a
// DPP swap, the lanes finish the radix-2 step of their half and an 8-point
// DFT, and lane h ends with outputs 2k + h (the arithmetic of dft16, so both
// kernels give the same results).  Eight points per lane halve the
// registers: FP64 went from one wave per SIMD (344 registers) to two, 0.30 ->
// 0.25 ms at 256^3; FP32, at two waves with sixteen points per lane, stays
// with one lane per DFT (the pair kernel's exchanges and selects cost more
// issue than its occupancy returns: 0.139 against 0.128 ms).
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

// The z line (kx, y) of thread column col in the full register passes.
// Without a tail (a.tail == 0) workgroup (bx, by) holds columns bx*TX + col of
// row by.  With one (kx = tail*TX + 1) the last column kx - 1 is held by the
// workgroups bx == tail, TX rows each (y = by*TX + col); of the other
// workgroups, all rows take part (a.zsel == 0) or only the row y = ny/2 of an
// even ny, at by == 0 (a.zsel == 1: k_fft_z_scale_reg256_rank1 takes the
// rest).  Returns false when the workgroup holds no line.
__device__ __forceinline__ bool zreg_column(const ColArgs& a, int col, int TX, int& kx, int& y) {
  if (a.tail > 0 && (int)blockIdx.x == a.tail) {
    kx = a.kx - 1;
    y = blockIdx.y * TX + col;
    return (int)blockIdx.y * TX < a.ny;
  }
  kx = blockIdx.x * TX + col;
  if (a.zsel == 0) {
    y = blockIdx.y;
    return true;
  }
  y = a.ny / 2;
  return blockIdx.y == 0 && a.ny % 2 == 0;
}

// Thread (col, j), col = threadIdx.x % TX, j = threadIdx.x / TX (16 values).
template <typename T, int TX>
__global__ void __launch_bounds__(TX * 16) k_fft_z_scale_reg256(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  constexpr int N = 256, NC = 3, LS = 273;  // LS: padded line of the transpose buffer
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> xb[TX * LS];
  const int col = threadIdx.x % TX, j = threadIdx.x / TX;
  const int rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  int kx, yr;
  if (!zreg_column(a, col, TX, kx, yr)) return;  // uniform over the workgroup
  const bool valid = kx < a.kx && yr < a.ny;
  const int y = valid ? yr : 0;
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

// The value of the neighbouring lane (lane ^ 1).
template <typename T>
__device__ __forceinline__ cplx<T> lane_swap1(cplx<T> v) {
  return cmk<T>(dpp_mov<kDppSwap1>(v.x), dpp_mov<kDppSwap1>(v.y));
}

// 16-point DFT over a lane pair (h = lane & 1) with the arithmetic of dft16:
// lane h holds inputs 8h .. 8h + 7 in v[0..7] and returns outputs 2k + h in
// v[k].  Lane 0 forms the sums of dft16's radix-2 step (its even outputs),
// lane 1 the twiddled differences (its odd outputs).
template <typename T>
__device__ __forceinline__ void dft16_pair(cplx<T>* v, int h) {
  const T c1 = (T)FftConst<T>::c161, s1 = (T)FftConst<T>::s161, r2 = (T)FftConst<T>::r2;
  cplx<T> o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) o[k] = lane_swap1(v[k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = h == 0 ? cadd(v[k], o[k]) : csub(o[k], v[k]);
  // w16^k = cos(pi k/8) - i sin(pi k/8) on the differences
  const cplx<T> d1 = cmul(v[1], cmk(c1, -s1));
  const cplx<T> d2 = cmul(v[2], cmk(r2, -r2));
  const cplx<T> d3 = cmul(v[3], cmk(s1, -c1));
  const cplx<T> d4 = cmul_mi(v[4]);
  const cplx<T> d5 = cmul(v[5], cmk(-s1, -c1));
  const cplx<T> d6 = cmul(v[6], cmk(-r2, -r2));
  const cplx<T> d7 = cmul(v[7], cmk(-c1, -s1));
  if (h) {
    v[1] = d1;
    v[2] = d2;
    v[3] = d3;
    v[4] = d4;
    v[5] = d5;
    v[6] = d6;
    v[7] = d7;
  }
  dft8(v);
}

template <typename T>
__device__ __forceinline__ void idft16_pair(cplx<T>* v, int h) {
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = cconj(v[i]);
  dft16_pair<T>(v, h);
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = cconj(v[i]);
}

// Lane pair reorder: lane h holds x[2k + h] in v[k]; on return x[8h + k].
// Lane 0 keeps x[0, 2, 4, 6] and takes x[1, 3, 5, 7] from lane 1; lane 1
// keeps x[9, 11, 13, 15] and takes x[8, 10, 12, 14] from lane 0.  Both
// halves are swapped and the lane selects among values, never between two
// entries of v (a select of array entries would become a computed index and
// put the array in scratch memory).
template <typename T>
__device__ __forceinline__ void pair_to_halves(cplx<T>* v, int h) {
  cplx<T> lo[4], hi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    lo[i] = lane_swap1(v[i]);      // partner's x[2i + 1] (read by lane 0)
    hi[i] = lane_swap1(v[4 + i]);  // partner's x[8 + 2i] (read by lane 1)
  }
  cplx<T> w[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const cplx<T> own_lo = v[i], own_hi = v[4 + i];
    w[2 * i] = h == 0 ? own_lo : hi[i];
    w[2 * i + 1] = h == 0 ? lo[i] : own_hi;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = w[k];
}

// Lane (col, j, h): col = (threadIdx.x / 2) % TX, j = (threadIdx.x / 2) / TX,
// h = threadIdx.x % 2; the pair (col, j) does the work of one 16-point lane
// of k_fft_z_scale_reg256, with the same results.
template <typename T, int TX>
__global__ void __launch_bounds__(TX * 32) k_fft_z_scale_reg256_pair(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  constexpr int N = 256, NC = 3, LS = 273;  // LS: padded line of the transpose buffer
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> xb[TX * LS];
  const int h = threadIdx.x & 1, col = (threadIdx.x >> 1) % TX, j = (threadIdx.x >> 1) / TX;
  const int rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  int kx, yr;
  if (!zreg_column(a, col, TX, kx, yr)) return;  // uniform over the workgroup
  const bool valid = kx < a.kx && yr < a.ny;
  const int y = valid ? yr : 0;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  const size_t plane = (size_t)a.ny * a.kxs;
  cplx<T> v[NC][8];
  // load z = j + 16 n2, n2 = 8h + t, of the three components (all loads in flight together)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const cplx<T>* src = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[c][t] = valid ? src[(size_t)(j + 16 * (8 * h + t)) * plane] : cmk<T>(0, 0);
  }
  cplx<T>* line = xb + col * LS;
  __syncthreads();  // twiddles
  // forward: DFT over n2, twiddle, transpose, DFT over n1
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    dft16_pair<T>(v[c], h);  // v[k] = Y[n1 = j][k2 = 2k + h] before the twiddle
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int k2 = 2 * k + h;
      const cplx<T> t = cmul(v[c][k], tw[(j * k2) & (N - 1)]);
      if (k2 != 0) v[c][k] = t;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) line[pidx((2 * k + h) * 16 + j)] = v[c][k];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; ++t) v[c][t] = line[pidx(j * 16 + 8 * h + t)];
    __syncthreads();
    dft16_pair<T>(v[c], h);  // lane pair (col, k2 = j) now holds X[j + 16 k1], k1 = 2k + h
  }
  if (valid) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      dipole_scale_point<T>(a, sp, kx, y, j + 16 * (2 * k + h), v[0][k], v[1][k], v[2][k]);
  }
  // inverse: DFT over k1, conjugate twiddle, transpose, DFT over k2
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    pair_to_halves<T>(v[c], h);  // k1 = 8h + k
    idft16_pair<T>(v[c], h);     // index n1 = 2k + h
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int n1 = 2 * k + h;
      const cplx<T> t = cmul(v[c][k], cconj(tw[(j * n1) & (N - 1)]));
      if (n1 != 0) v[c][k] = t;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) line[pidx((2 * k + h) * 16 + j)] = v[c][k];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; ++t) v[c][t] = line[pidx(j * 16 + 8 * h + t)];
    __syncthreads();
    idft16_pair<T>(v[c], h);  // lane pair (col, n1 = j) now holds x[j + 16 n2], n2 = 2k + h
  }
  if (valid) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cplx<T>* dst = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx;
#pragma unroll
      for (int k = 0; k < 8; ++k) dst[(size_t)(j + 16 * (2 * k + h)) * plane] = v[c][k];
    }
  }
}

// Dipole z pass over the lines whose x and y wave numbers are not Nyquist
// frequencies (every column but the last, every row but y = ny/2 of an even
// ny; k_fft_z_scale_reg256 with a.zsel = 1 takes those).  There the scale is
// rank one, G = c k (k.F)/|k|^2, except at kz = nz/2, where the Hermitian part
// (cross terms with exactly one Nyquist axis vanish) keeps the (x, y) block
// and the z entry apart: G_xy = c k_xy (k_xy.F_xy)/|k|^2, G_z = c k_z^2
// F_z/|k|^2.  k_x and k_y are constant along a z line, so the forward pass
// transforms two lines, h = k_x f_x + k_y f_y and f_z (k.F = DFT(h) +
// k_z DFT(f_z)), and the inverse pass two, psi and k_z psi_z with psi =
// c (k.F)/|k|^2 (psi = c DFT(h)/|k|^2 and psi_z = c k_z DFT(f_z)/|k|^2 at
// kz = nz/2): g_x = k_x IDFT(psi), g_y = k_y IDFT(psi), g_z = IDFT(k_z psi_z).
// Four 256-point transforms per line instead of six, two lines in registers
// instead of three.  c(k)/|k|^2 = jk[m^2] e_x e_y e_z / G from the cubic-box
// tables (sp.jk; FP32: float copies sp.jk32, sp.ez32, whose eight loads per
// thread in flight stay within three waves' registers).  Thread (col, j) as in
// k_fft_z_scale_reg256; grid (tail, ny, nrhs).
template <typename T, int TX>
__global__ void __launch_bounds__(TX * 16) k_fft_z_scale_reg256_rank1(ColArgs a, ScaleParams sp,
                                                                      cplx<T>* __restrict__ spec) {
  constexpr int N = 256, LS = 273;  // LS: padded line of the transpose buffer
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> xb[TX * LS];
  const int col = threadIdx.x % TX, j = threadIdx.x / TX;
  const int kx = blockIdx.x * TX + col, y = blockIdx.y, rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  if (a.ny % 2 == 0 && 2 * y == a.ny) return;  // the Nyquist row (uniform)
  const bool valid = kx < a.kx;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  const size_t plane = (size_t)a.ny * a.kxs;
  const int mx = fft_freq(kx, a.nx), my = fft_freq(y, a.ny);
  const T k0t = (T)(mx * sp.twopi_over_l[0]), k1t = (T)(my * sp.twopi_over_l[1]);
  cplx<T> v[2][16];  // h, then f_z
  {
    const cplx<T>* s0 = spec + (size_t)(rhs * 3) * a.nz * plane + (size_t)y * a.kxs + kx;
    const cplx<T>* s1 = s0 + (size_t)a.nz * plane;
    const cplx<T>* s2 = s1 + (size_t)a.nz * plane;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const size_t o = (size_t)(j + 16 * n2) * plane;
      const cplx<T> a0 = valid ? s0[o] : cmk<T>(0, 0);
      const cplx<T> a1 = valid ? s1[o] : cmk<T>(0, 0);
      v[1][n2] = valid ? s2[o] : cmk<T>(0, 0);
      v[0][n2] = cmk<T>(k0t * a0.x + k1t * a1.x, k0t * a0.y + k1t * a1.y);
    }
  }
  cplx<T>* line = xb + col * LS;
  __syncthreads();  // twiddles
  // forward: DFT over n2, twiddle, transpose, DFT over n1
#pragma unroll
  for (int c = 0; c < 2; ++c) {
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
    const int mxy2 = mx * mx + my * my;
    const T tz = (T)sp.twopi_over_l[2];
    T exy;
    if constexpr (sizeof(T) == 4) exy = (T)(sp.ex[kx] * sp.ey[y] * sp.inv_g);
    else exy = sp.ex[kx] * sp.ey[y] * sp.inv_g;
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
      const int kz = j + 16 * k1;
      const int mz = fft_freq(kz, a.nz);
      T q;  // c(k)/|k|^2 (0 at k = 0)
      if constexpr (sizeof(T) == 4) q = sp.jk32[mxy2 + mz * mz] * sp.ez32[kz] * exy;
      else q = sp.jk[mxy2 + mz * mz] * sp.ez[kz] * exy;
      const T k2 = (T)mz * tz;
      const cplx<T> txy = v[0][k1];
      const cplx<T> tzv = cmk<T>(k2 * v[1][k1].x, k2 * v[1][k1].y);
      const bool split = a.nz % 2 == 0 && 2 * mz == -a.nz;  // kz = nz/2
      const cplx<T> s = split ? txy : cmk<T>(txy.x + tzv.x, txy.y + tzv.y);
      const cplx<T> sz = split ? tzv : s;
      v[0][k1] = cmk<T>(q * s.x, q * s.y);
      v[1][k1] = cmk<T>(q * k2 * sz.x, q * k2 * sz.y);
    }
  }
  // inverse: DFT over k1, conjugate twiddle, transpose, DFT over k2
#pragma unroll
  for (int c = 0; c < 2; ++c) {
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
    cplx<T>* d0 = spec + (size_t)(rhs * 3) * a.nz * plane + (size_t)y * a.kxs + kx;
    cplx<T>* d1 = d0 + (size_t)a.nz * plane;
    cplx<T>* d2 = d1 + (size_t)a.nz * plane;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const size_t o = (size_t)(j + 16 * n2) * plane;
      d0[o] = cmk<T>(k0t * v[0][n2].x, k0t * v[0][n2].y);
      d1[o] = cmk<T>(k1t * v[0][n2].x, k1t * v[0][n2].y);
      d2[o] = v[1][n2];
    }
  }
}

}  // namespace mpol
