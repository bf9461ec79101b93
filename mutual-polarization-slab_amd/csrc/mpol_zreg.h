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

// Line bodies of the register z passes.  Each transforms one z line
// (kx, y) of right-hand side rhs per lane (16 points, j = lane's z residue)
// or per lane pair (8 points each, h = lane in the pair), with one padded
// 256-entry LDS line per line of the workgroup (`line`) and the twiddles in
// LDS; all lanes of the workgroup call them (they hold barriers); lanes with
// valid = false transform zeros and store nothing.

// Three components, one line per lane.
template <typename T>
__device__ __forceinline__ void zline16_full(const ColArgs& a, const ScaleParams& sp, cplx<T>* __restrict__ spec,
                                             int rhs, int kx, int y, bool valid, int j, cplx<T>* line,
                                             const cplx<T>* tw) {
  constexpr int N = 256, NC = 3;
  const size_t plane = (size_t)a.ny * a.kxs;
  cplx<T> v[NC][16];
  // load z = j + 16 n2 of the three components (all loads in flight together)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const cplx<T>* src = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[c][n2] = valid ? src[(size_t)(j + 16 * n2) * plane] : cmk<T>(0, 0);
  }
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
    dft16<T>(v[c]);  // lane (col, k2 = j) now holds X[j + 16 k1]
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
    idft16<T>(v[c]);  // lane (col, n1 = j) now holds x[j + 16 n2]
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

// Three components, one line per lane pair, with the results of zline16_full.
template <typename T>
__device__ __forceinline__ void zline_pair_full(const ColArgs& a, const ScaleParams& sp, cplx<T>* __restrict__ spec,
                                                int rhs, int kx, int y, bool valid, int j, int h, cplx<T>* line,
                                                const cplx<T>* tw) {
  constexpr int N = 256, NC = 3;
  const size_t plane = (size_t)a.ny * a.kxs;
  cplx<T> v[NC][8];
  // load z = j + 16 n2, n2 = 8h + t, of the three components (all loads in flight together)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const cplx<T>* src = spec + (size_t)(rhs * NC + c) * a.nz * plane + (size_t)y * a.kxs + kx;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[c][t] = valid ? src[(size_t)(j + 16 * (8 * h + t)) * plane] : cmk<T>(0, 0);
  }
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

// Rank-one dipole scheme.  On the lines whose x and y wave numbers are not
// Nyquist frequencies (every column but the last, every row but ky = ny/2 of
// an even ny: the "regular" lines) the scale is rank one, G = c k (k.F)/|k|^2,
// except at kz = nz/2, where the Hermitian part (cross terms with exactly one
// Nyquist axis vanish) keeps the (x, y) block and the z entry apart:
// G_xy = c k_xy (k_xy.F_xy)/|k|^2, G_z = c k_z^2 F_z/|k|^2.  k_x and k_y are
// constant along a z line, so the forward z transforms take two lines, h =
// k_x f_x + k_y f_y and f_z (k.F = DFT(h) + k_z DFT(f_z)), and the inverse
// two, psi and k_z psi_z with psi = c (k.F)/|k|^2 (psi = c DFT(h)/|k|^2 and
// psi_z = c k_z DFT(f_z)/|k|^2 at kz = nz/2): g_x = k_x IDFT(psi), g_y = k_y
// IDFT(psi), g_z = IDFT(k_z psi_z).  k_fft_y_xy forms h after the forward y
// transform and k_x psi', k_y psi' (psi' = IDFT_z(psi)) before the inverse
// one, so between the y passes the x and y components travel as one grid
// (slot 0; slot 1 holds data only on the full-form lines, the last column and
// the row ky = ny/2, which keep the three-component form): four 256-point z
// transforms per regular line instead of six, and two grids read and written
// by the z pass instead of three.  c(k)/|k|^2 = jk[m^2] e_x e_y e_z / G from
// the cubic-box tables (sp.jk; FP32: float copies sp.jk32, sp.ez32).

// Scale of one point of a regular line: txy = DFT(h), f = DFT(f_z) at kz
// (mz = its signed wave number) become psi and k_z psi_z.
template <typename T>
__device__ __forceinline__ void rank1_scale_point(const ColArgs& a, const ScaleParams& sp, int mxy2, T exy, T tz,
                                                  int kz, cplx<T>& txy, cplx<T>& f) {
  const int mz = fft_freq(kz, a.nz);
  T q;  // c(k)/|k|^2 (0 at k = 0)
  if constexpr (sizeof(T) == 4) q = sp.jk32[mxy2 + mz * mz] * sp.ez32[kz] * exy;
  else q = sp.jk[mxy2 + mz * mz] * sp.ez[kz] * exy;
  const T k2 = (T)mz * tz;
  const cplx<T> tzv = cmk<T>(k2 * f.x, k2 * f.y);
  const bool split = a.nz % 2 == 0 && 2 * mz == -a.nz;  // kz = nz/2
  const cplx<T> s = split ? txy : cmk<T>(txy.x + tzv.x, txy.y + tzv.y);
  const cplx<T> sz = split ? tzv : s;
  txy = cmk<T>(q * s.x, q * s.y);
  f = cmk<T>(q * k2 * sz.x, q * k2 * sz.y);
}

// Per-line constants of rank1_scale_point: m_x^2 + m_y^2, e_x e_y / G and
// 2 pi / L_z.
template <typename T>
__device__ __forceinline__ void rank1_line_consts(const ColArgs& a, const ScaleParams& sp, int kx, int y, int& mxy2,
                                                  T& exy, T& tz) {
  const int mx = fft_freq(kx, a.nx), my = fft_freq(y, a.ny);
  mxy2 = mx * mx + my * my;
  exy = (T)(sp.ex[kx] * sp.ey[y] * sp.inv_g);
  tz = (T)sp.twopi_over_l[2];
}

// A regular line, one line per lane: slot 0 (h) and slot 2 (f_z) in, psi'
// and IDFT(k_z psi_z) out.
template <typename T>
__device__ __forceinline__ void zline16_rank1(const ColArgs& a, const ScaleParams& sp, cplx<T>* __restrict__ spec,
                                              int rhs, int kx, int y, bool valid, int j, cplx<T>* line,
                                              const cplx<T>* tw) {
  constexpr int N = 256;
  const size_t plane = (size_t)a.ny * a.kxs;
  cplx<T>* s0 = spec + (size_t)(rhs * 3) * a.nz * plane + (size_t)y * a.kxs + kx;
  cplx<T>* s2 = s0 + 2 * (size_t)a.nz * plane;
  cplx<T> v[2][16];  // h, then f_z
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) {
    const size_t o = (size_t)(j + 16 * n2) * plane;
    v[0][n2] = valid ? s0[o] : cmk<T>(0, 0);
    v[1][n2] = valid ? s2[o] : cmk<T>(0, 0);
  }
  __syncthreads();  // twiddles
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
    dft16<T>(v[c]);  // lane (col, k2 = j) now holds X[j + 16 k1]
  }
  if (valid) {
    int mxy2;
    T exy, tz;
    rank1_line_consts<T>(a, sp, kx, y, mxy2, exy, tz);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) rank1_scale_point<T>(a, sp, mxy2, exy, tz, j + 16 * k1, v[0][k1], v[1][k1]);
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    idft16<T>(v[c]);
#pragma unroll
    for (int n1 = 1; n1 < 16; ++n1) v[c][n1] = cmul(v[c][n1], cconj(tw[(j * n1) & (N - 1)]));
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) line[pidx(n1 * 16 + j)] = v[c][n1];
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) v[c][k2] = line[pidx(j * 16 + k2)];
    __syncthreads();
    idft16<T>(v[c]);  // lane (col, n1 = j) now holds x[j + 16 n2]
  }
  if (valid) {
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const size_t o = (size_t)(j + 16 * n2) * plane;
      s0[o] = v[0][n2];
      s2[o] = v[1][n2];
    }
  }
}

// A regular line, one line per lane pair (zline16_rank1's results).
template <typename T>
__device__ __forceinline__ void zline_pair_rank1(const ColArgs& a, const ScaleParams& sp, cplx<T>* __restrict__ spec,
                                                 int rhs, int kx, int y, bool valid, int j, int h, cplx<T>* line,
                                                 const cplx<T>* tw) {
  constexpr int N = 256;
  const size_t plane = (size_t)a.ny * a.kxs;
  cplx<T>* s0 = spec + (size_t)(rhs * 3) * a.nz * plane + (size_t)y * a.kxs + kx;
  cplx<T>* s2 = s0 + 2 * (size_t)a.nz * plane;
  cplx<T> v[2][8];  // h, then f_z
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const size_t o = (size_t)(j + 16 * (8 * h + t)) * plane;
    v[0][t] = valid ? s0[o] : cmk<T>(0, 0);
    v[1][t] = valid ? s2[o] : cmk<T>(0, 0);
  }
  __syncthreads();  // twiddles
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    dft16_pair<T>(v[c], h);
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
    int mxy2;
    T exy, tz;
    rank1_line_consts<T>(a, sp, kx, y, mxy2, exy, tz);
#pragma unroll
    for (int k = 0; k < 8; ++k) rank1_scale_point<T>(a, sp, mxy2, exy, tz, j + 16 * (2 * k + h), v[0][k], v[1][k]);
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
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
    for (int k = 0; k < 8; ++k) {
      const size_t o = (size_t)(j + 16 * (2 * k + h)) * plane;
      s0[o] = v[0][k];
      s2[o] = v[1][k];
    }
  }
}

// The z line (kx, y) of thread column col in the full register passes.
// Without a tail (a.tail == 0) workgroup (bx, by) holds columns bx*TX + col of
// row by; with one (kx = tail*TX + 1) the last column kx - 1 is held by the
// workgroups bx == tail, TX rows each (y = by*TX + col).  Returns false when
// the workgroup holds no line.
__device__ __forceinline__ bool zreg_column(const ColArgs& a, int col, int TX, int& kx, int& y) {
  if (a.tail > 0 && (int)blockIdx.x == a.tail) {
    kx = a.kx - 1;
    y = blockIdx.y * TX + col;
    return (int)blockIdx.y * TX < a.ny;
  }
  kx = blockIdx.x * TX + col;
  y = blockIdx.y;
  return true;
}

// Full three-component z pass, lane (col, j): col = threadIdx.x % TX, j =
// threadIdx.x / TX.
template <typename T, int TX>
__global__ void __launch_bounds__(TX * 16) k_fft_z_scale_reg256(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  constexpr int N = 256, LS = 273;  // LS: padded line of the transpose buffer
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> xb[TX * LS];
  const int col = threadIdx.x % TX, j = threadIdx.x / TX;
  if (rhs_idle(a.gate, a.gdiv, blockIdx.z)) return;
  int kx, yr;
  if (!zreg_column(a, col, TX, kx, yr)) return;  // uniform over the workgroup
  const bool valid = kx < a.kx && yr < a.ny;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  zline16_full<T>(a, sp, spec, blockIdx.z, kx, valid ? yr : 0, valid, j, xb + col * LS, tw);
}

// Full three-component z pass, lane (col, j, h): col = (threadIdx.x / 2) %
// TX, j = (threadIdx.x / 2) / TX, h = threadIdx.x % 2; the pair (col, j) does
// the work of one 16-point lane of k_fft_z_scale_reg256, with the same results.
template <typename T, int TX>
__global__ void __launch_bounds__(TX * 32) k_fft_z_scale_reg256_pair(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  constexpr int N = 256, LS = 273;
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> xb[TX * LS];
  const int h = threadIdx.x & 1, col = (threadIdx.x >> 1) % TX, j = (threadIdx.x >> 1) / TX;
  if (rhs_idle(a.gate, a.gdiv, blockIdx.z)) return;
  int kx, yr;
  if (!zreg_column(a, col, TX, kx, yr)) return;  // uniform over the workgroup
  const bool valid = kx < a.kx && yr < a.ny;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  zline_pair_full<T>(a, sp, spec, blockIdx.z, kx, valid ? yr : 0, valid, j, h, xb + col * LS, tw);
}

// Rank-one z pass: 256 threads per workgroup.  Workgroups bx < a.tail take
// regular lines, TX columns of row by (PAIR = false: one line per lane, TX =
// 16; PAIR = true: one line per lane pair, TX = 8); the row ky = ny/2 of an
// even ny returns at once.  Workgroups bx == a.tail take the full-form lines
// (the row ky = ny/2 over the columns but the last, then the last column over
// every row), eight per workgroup (lane pairs), by = their group.  Grid
// (tail + 1, max(ny, full-form groups), nrhs).
template <typename T, int TX, bool PAIR>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 3 : 2)
    k_fft_z_scale_reg256_rank1(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  static_assert(TX * (PAIR ? 32 : 16) == 256, "256 threads");
  constexpr int N = 256, LS = 273;
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> xb[16 * LS];
  const int rhs = blockIdx.z;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const int nrow = a.ny % 2 == 0 ? a.kx - 1 : 0;  // full-form lines of the row ky = ny/2
  if ((int)blockIdx.x == a.tail) {
    const int h = threadIdx.x & 1, col = (threadIdx.x >> 1) & 7, j = threadIdx.x >> 4;
    if ((int)blockIdx.y * 8 >= nrow + a.ny) return;  // uniform
    const int l = blockIdx.y * 8 + col;
    const bool valid = l < nrow + a.ny;
    const int kx = l < nrow ? l : a.kx - 1;
    const int y = l < nrow ? a.ny / 2 : (valid ? l - nrow : 0);
    const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
    for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
    zline_pair_full<T>(a, sp, spec, rhs, kx, y, valid, j, h, xb + col * LS, tw);
    return;
  }
  const int y = blockIdx.y;
  if (y >= a.ny || (a.ny % 2 == 0 && 2 * y == a.ny)) return;  // uniform
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  if constexpr (PAIR) {
    const int h = threadIdx.x & 1, col = (threadIdx.x >> 1) % TX, j = (threadIdx.x >> 1) / TX;
    const int kx = blockIdx.x * TX + col;
    zline_pair_rank1<T>(a, sp, spec, rhs, kx, y, kx < a.kx - 1, j, h, xb + col * LS, tw);
  } else {
    const int col = threadIdx.x % TX, j = threadIdx.x / TX;
    const int kx = blockIdx.x * TX + col;
    zline16_rank1<T>(a, sp, spec, rhs, kx, y, kx < a.kx - 1, j, xb + col * LS, tw);
  }
}

// The y transforms of the rank-one scheme, workgroups of two roles (grid
// (column groups, nz, 2 nrhs), blockIdx.z = 2 rhs + role).  Role 0, the x and
// y components: forward (inverse = 0), F_x, F_y (slots 0, 1) of tx columns
// of plane z are y-transformed together and combined into h = k_x F_x +
// k_y F_y (slot 0), except on the full-form lines (the last column; ky = ny/2
// of an even ny), where F_x and F_y stay in slots 0 and 1; inverse, slot 0
// holds psi' (or G_x, with G_y in slot 1, on the full-form lines) and
// k_x psi', k_y psi' (or G_x, G_y) are inverse-transformed into slots 0 and
// 1.  Role 1, the z component (slot 2): 2 tx columns per workgroup, the
// column groups past kx return at once.  Up to 512 threads (2 tx lines in
// one pass of the line FFTs); 2 tx lines in LDS.
template <typename T, int E, int N, bool INV>
__global__ void __launch_bounds__(512) k_fft_y_xy(ColArgs a, ScaleParams sp, cplx<T>* __restrict__ spec) {
  constexpr bool inverse = INV;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + a.ny;
  cplx<T>* scratch = buf + 2 * a.tx * a.lstride;
  const int z = blockIdx.y, rhs = blockIdx.z >> 1, role = blockIdx.z & 1;
  if (rhs_idle(a.gate, a.gdiv, rhs)) return;
  const size_t cgrid = (size_t)a.nz * a.ny * a.kxs;  // one component
  if (role == 1) {
    const int tz = 2 * a.tx, kz0 = blockIdx.x * tz;
    if (kz0 >= a.kx) return;  // uniform
    const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
    for (int k = threadIdx.x; k < a.ny; k += blockDim.x) tw[k] = twg[k];
    cplx<T>* b2 = spec + (size_t)(rhs * 3 + 2) * cgrid + (size_t)z * a.ny * a.kxs;
    const int c = threadIdx.x % tz, ystep = blockDim.x / tz;
    const int kx = kz0 + c;
    const bool act = (int)threadIdx.x < ystep * tz && kx < a.kx;
    cplx<T>* l = buf + c * a.lstride;
    if (act) {
      for (int y0 = threadIdx.x / tz; y0 < a.ny; y0 += 8 * ystep) {
        cplx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int y = y0 + u * ystep;
          v[u] = y < a.ny ? b2[(size_t)y * a.kxs + kx] : cmk<T>(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int y = y0 + u * ystep;
          if (y < a.ny) l[pidx(y)] = v[u];
        }
      }
    }
    __syncthreads();
    fft_any<T, E, N>(buf, a.lstride, min(tz, a.kx - kz0), a.plan, tw, inverse, scratch);
    if (act)
      for (int y = threadIdx.x / tz; y < a.ny; y += ystep) b2[(size_t)y * a.kxs + kx] = l[pidx(y)];
    return;
  }
  const int kx0 = blockIdx.x * a.tx;
  const cplx<T>* twg = reinterpret_cast<const cplx<T>*>(a.tw);
  for (int k = threadIdx.x; k < a.ny; k += blockDim.x) tw[k] = twg[k];
  cplx<T>* b0 = spec + (size_t)(rhs * 3) * cgrid + (size_t)z * a.ny * a.kxs;
  cplx<T>* b1 = b0 + cgrid;
  const int c = threadIdx.x % a.tx, ystep = blockDim.x / a.tx;
  const int kx = kx0 + c;
  const bool colok = kx < a.kx;
  const bool fullcol = kx == a.kx - 1;
  const int nyq = a.ny % 2 == 0 ? a.ny / 2 : -1;
  const T k0 = (T)(fft_freq(kx, a.nx) * sp.twopi_over_l[0]);
  const T ky1 = (T)sp.twopi_over_l[1];
  cplx<T>* lx = buf + c * a.lstride;
  cplx<T>* ly = buf + (a.tx + c) * a.lstride;
  const bool act = (int)threadIdx.x < ystep * a.tx;
  if (act) {
    // batches of 8 rows (16 loads) in flight before the LDS writes; the
    // inverse reads slot 1 only in the last column (slot 1 of the row
    // ky = ny/2 is patched after the loop by the lane that holds that row)
    const bool two = !inverse || fullcol;
    for (int y0 = threadIdx.x / a.tx; y0 < a.ny; y0 += 8 * ystep) {
      cplx<T> vx[8], vy[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int y = y0 + u * ystep;
        const bool ok = colok && y < a.ny;
        const size_t o = (size_t)y * a.kxs + kx;
        vx[u] = ok ? b0[o] : cmk<T>(0, 0);
        vy[u] = (ok && two) ? b1[o] : cmk<T>(0, 0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int y = y0 + u * ystep;
        if (y < a.ny) {
          if (inverse && !fullcol) {
            const T k1 = (T)fft_freq(y, a.ny) * ky1;
            lx[pidx(y)] = cmk<T>(k0 * vx[u].x, k0 * vx[u].y);
            ly[pidx(y)] = cmk<T>(k1 * vx[u].x, k1 * vx[u].y);
          } else {
            lx[pidx(y)] = vx[u];
            ly[pidx(y)] = vy[u];
          }
        }
      }
    }
    if (inverse && !fullcol && colok && nyq >= 0 && (int)threadIdx.x / a.tx == nyq % ystep) {
      const size_t o = (size_t)nyq * a.kxs + kx;
      lx[pidx(nyq)] = b0[o];
      ly[pidx(nyq)] = b1[o];
    }
  }
  __syncthreads();
  fft_any<T, E, N>(buf, a.lstride, 2 * a.tx, a.plan, tw, inverse, scratch);
  if (act && colok) {
    for (int y = threadIdx.x / a.tx; y < a.ny; y += ystep) {
      const size_t o = (size_t)y * a.kxs + kx;
      const cplx<T> X = lx[pidx(y)], Y = ly[pidx(y)];
      if (inverse || fullcol || y == nyq) {
        b0[o] = X;
        b1[o] = Y;
      } else {
        const T k1 = (T)fft_freq(y, a.ny) * ky1;
        b0[o] = cmk<T>(k0 * X.x + k1 * Y.x, k0 * X.y + k1 * Y.y);
      }
    }
  }
}

}  // namespace mpol
