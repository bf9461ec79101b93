// Shared device/host helpers for the spectral-Ewald mutual-polarization library.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

namespace mpol {

template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return cmk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return cmk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return cmk(a.x, -a.y); }
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return cmk(a.x * s, a.y * s); }
// multiply by -i
template <typename T>
__device__ __forceinline__ cplx<T> cmul_mi(cplx<T> a) { return cmk(a.y, -a.x); }
// multiply by +i
template <typename T>
__device__ __forceinline__ cplx<T> cmul_pi(cplx<T> a) { return cmk(-a.y, a.x); }

// Padded LDS index: one spare slot every 16 elements keeps the strided
// Stockham stores and the row-parallel spread/contract accesses bank-conflict
// free.
__device__ __forceinline__ int pidx(int a) { return a + (a >> 4); }
__host__ __device__ constexpr int padded_len(int n) { return (n + (n >> 4)) | 1; }

// Sum over groups of R consecutive lanes (R a power of two <= 16, groups
// inside one 16-lane DPP row); every lane of a group gets the group's sum.
// Uses row mirrors and quad permutations (no LDS traffic).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppSwap1 = 0xB1, kDppSwap2 = 0x4E;
// row_newbcast:n (0x150 + n): every lane of a 16-lane row reads lane n of
// its row
constexpr int kDppRowBcast = 0x150;

template <typename T>
__device__ __forceinline__ T group_sum(T v, int R) {
  if (R >= 16) v += dpp_mov<kDppRowMirror>(v);
  if (R >= 8) v += dpp_mov<kDppRowHalfMirror>(v);
  if (R >= 4) v += dpp_mov<kDppSwap2>(v);
  if (R >= 2) v += dpp_mov<kDppSwap1>(v);
  return v;
}

// Workgroups go to the 8 XCDs round-robin by linear id.  xcd_block maps a
// linear id to a logical block so that each XCD runs one contiguous range of
// logical blocks (a bijection on [0, total)): blocks over neighbouring
// particles then share that XCD's L2.
__device__ __forceinline__ int xcd_block(int lin, int total) {
  constexpr int kXcd = 8;
  const int q = total / kXcd, r = total % kXcd;
  const int x = lin % kXcd, k = lin / kXcd;
  return x * q + min(x, r) + k;
}

// Per-system state of the batched GMRES.
enum RhsState : int { kDone = 0, kIter = 1, kSolve = 2, kWait = 3 };

// Operator batch entry b (a right-hand side, or a right-hand side's grid
// component) is skipped when a gate is given and its system, b / gdiv, is
// not in an Arnoldi step: inside a cycle the systems that have finished it
// cost no operator work.
__device__ __forceinline__ bool rhs_idle(const int* gate, int gdiv, int b) {
  return gate != nullptr && gate[b / gdiv] != kIter;
}

// Frequency of natural-order FFT index i on an n-point grid: i for
// i <= floor((n-1)/2), else i - n (even n: index n/2 maps to -n/2).
__host__ __device__ inline int fft_freq(int i, int n) { return (i <= (n - 1) / 2) ? i : i - n; }

// Maximum number of radix stages in an FFT plan.
constexpr int kMaxStages = 16;

struct FftPlan {
  int n;                    // transform length
  int e;                    // elements per thread (0: generic O(n^2) DFT path)
  int nstages;
  int radix[kMaxStages];
  int ns[kMaxStages];       // product of the radices of earlier stages
};

}  // namespace mpol
