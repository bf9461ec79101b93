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
__host__ __device__ inline int padded_len(int n) { return (n + (n >> 4)) | 1; }

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
