// LDS-resident batched 1D FFTs for gfx950.
//
// Lines live in LDS with the padded index pidx(); a workgroup transforms
// `nlines` lines of length n.  The fast path is a mixed-radix Stockham
// autosort FFT (radices 2, 3, 4, 5, 8, 16): every thread keeps E elements of
// one line in registers, reads its butterfly inputs, waits for the whole
// workgroup, and writes the outputs in place, so one LDS buffer suffices.
// Lengths with a prime factor above 5 (or too short for the fast path) use a
// direct O(n^2) DFT with an LDS scratch buffer.
//
// Sign convention: forward = exp(-2 pi i jk/n), inverse = exp(+2 pi i jk/n),
// both unnormalised (the caller folds 1/G into the k-space scale), matching
// numpy.fft.fftn / ifftn up to that factor.
#pragma once

#include "mpol_common.h"

namespace mpol {

template <typename T>
struct FftConst {
  static constexpr double r2 = 0.70710678118654752440084436210485;   // sqrt(2)/2
  static constexpr double s3 = 0.86602540378443864676372317075294;   // sqrt(3)/2
  static constexpr double c51 = 0.30901699437494742410229341718282;  // cos(2pi/5)
  static constexpr double c52 = -0.80901699437494742410229341718282; // cos(4pi/5)
  static constexpr double s51 = 0.95105651629515357211643933337938;  // sin(2pi/5)
  static constexpr double s52 = 0.58778525229247312916870595463907;  // sin(4pi/5)
  static constexpr double c161 = 0.92387953251128675612818318939679; // cos(pi/8)
  static constexpr double s161 = 0.38268343236508977172845998403040; // sin(pi/8)
};

template <typename T>
__device__ __forceinline__ void dft2(cplx<T>* v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T>
__device__ __forceinline__ void dft3(cplx<T>* v) {
  const T s3 = (T)FftConst<T>::s3;
  cplx<T> t1 = cadd(v[1], v[2]);
  cplx<T> t2 = csub(v[0], cscale(t1, (T)0.5));
  cplx<T> t3 = cscale(cmul_mi(csub(v[1], v[2])), s3);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>* v) {
  cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  cplx<T> t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, cmul_mi(t3));
  v[3] = csub(t1, cmul_mi(t3));
}

template <typename T>
__device__ __forceinline__ void dft5(cplx<T>* v) {
  const T c1 = (T)FftConst<T>::c51, c2 = (T)FftConst<T>::c52;
  const T s1 = (T)FftConst<T>::s51, s2 = (T)FftConst<T>::s52;
  cplx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  cplx<T> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  cplx<T> a1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
  cplx<T> a2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
  cplx<T> b1 = cmul_mi(cadd(cscale(t3, s1), cscale(t4, s2)));
  cplx<T> b2 = cmul_mi(csub(cscale(t3, s2), cscale(t4, s1)));
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

// 8-point DFT: one decimation-in-frequency radix-2 step, then two 4-point DFTs.
template <typename T>
__device__ __forceinline__ void dft8(cplx<T>* v) {
  const T r2 = (T)FftConst<T>::r2;
  cplx<T> u[4], d[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    u[k] = cadd(v[k], v[k + 4]);
    d[k] = csub(v[k], v[k + 4]);
  }
  d[1] = cmk(r2 * (d[1].x + d[1].y), r2 * (d[1].y - d[1].x));   // * (1 - i)/sqrt2
  d[2] = cmul_mi(d[2]);                                         // * -i
  d[3] = cmk(r2 * (d[3].y - d[3].x), -r2 * (d[3].x + d[3].y));  // * (-1 - i)/sqrt2
  dft4(u);
  dft4(d);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = u[k];
    v[2 * k + 1] = d[k];
  }
}

// 16-point DFT: one radix-2 DIF step with w16^k twiddles, then two 8-point DFTs.
template <typename T>
__device__ __forceinline__ void dft16(cplx<T>* v) {
  const T c1 = (T)FftConst<T>::c161, s1 = (T)FftConst<T>::s161, r2 = (T)FftConst<T>::r2;
  cplx<T> u[8], d[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    u[k] = cadd(v[k], v[k + 8]);
    d[k] = csub(v[k], v[k + 8]);
  }
  // w16^k = cos(pi k/8) - i sin(pi k/8)
  d[1] = cmul(d[1], cmk(c1, -s1));
  d[2] = cmul(d[2], cmk(r2, -r2));
  d[3] = cmul(d[3], cmk(s1, -c1));
  d[4] = cmul_mi(d[4]);
  d[5] = cmul(d[5], cmk(-s1, -c1));
  d[6] = cmul(d[6], cmk(-r2, -r2));
  d[7] = cmul(d[7], cmk(-c1, -s1));
  dft8(u);
  dft8(d);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[2 * k] = u[k];
    v[2 * k + 1] = d[k];
  }
}

template <typename T, int R>
__device__ __forceinline__ void dft_small(cplx<T>* v) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3(v);
  else if constexpr (R == 4) dft4(v);
  else if constexpr (R == 5) dft5(v);
  else if constexpr (R == 8) dft8(v);
  else if constexpr (R == 16) dft16(v);
}

template <typename T, int E, int R>
__device__ __forceinline__ void stage_load(const cplx<T>* line, int n, int t, int tl, cplx<T>* v, bool conj_in) {
  if constexpr (E % R == 0) {
    const int nr = n / R;
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
      const int j = t + tl * b;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        cplx<T> a = line[pidx(j + k * nr)];
        v[b * R + k] = conj_in ? cconj(a) : a;
      }
    }
  }
}

template <typename T, int E, int R>
__device__ __forceinline__ void stage_store(cplx<T>* line, int n, int ns, int t, int tl, cplx<T>* v,
                                            const cplx<T>* tw, bool conj_out) {
  if constexpr (E % R == 0) {
    const int step = n / (ns * R);
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
      const int j = t + tl * b;
      const int jm = j % ns;
      cplx<T>* w = v + b * R;
      if (ns > 1) {
#pragma unroll
        for (int k = 1; k < R; ++k) w[k] = cmul(w[k], tw[jm * k * step]);
      }
      dft_small<T, R>(w);
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int k = 0; k < R; ++k) line[pidx(base + k * ns)] = conj_out ? cconj(w[k]) : w[k];
    }
  }
}

// Transform nlines lines (line l starts at buf + l*lstride) in place.
// All threads of the block must call this.
template <typename T, int E>
__device__ void fft_lines(cplx<T>* buf, int lstride, int nlines, const FftPlan& p, const cplx<T>* tw,
                          bool inverse, cplx<T>* scratch) {
  const int n = p.n;
  if constexpr (E == 0) {
    // direct DFT, out of place through scratch
    const int total = nlines * n;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int l = idx / n, k = idx - l * n;
      const cplx<T>* line = buf + l * lstride;
      T ar = 0, ai = 0;
      int m = 0;
      for (int j = 0; j < n; ++j) {
        cplx<T> w = tw[m];
        if (inverse) w.y = -w.y;
        cplx<T> a = line[pidx(j)];
        ar += a.x * w.x - a.y * w.y;
        ai += a.x * w.y + a.y * w.x;
        m += k;
        if (m >= n) m -= n;
      }
      scratch[l * lstride + pidx(k)] = cmk(ar, ai);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int l = idx / n, k = idx - l * n;
      buf[l * lstride + pidx(k)] = scratch[l * lstride + pidx(k)];
    }
    __syncthreads();
  } else {
    const int tl = n / E;
    const int lpp = blockDim.x / tl;
    cplx<T> v[E];
    for (int s = 0; s < p.nstages; ++s) {
      const int R = p.radix[s], ns = p.ns[s];
      const bool cin = inverse && s == 0;
      const bool cout = inverse && s == p.nstages - 1;
      for (int l0 = 0; l0 < nlines; l0 += lpp) {
        const int l = l0 + (int)threadIdx.x / tl;
        const int t = (int)threadIdx.x % tl;
        const bool act = ((int)threadIdx.x < lpp * tl) && (l < nlines);
        cplx<T>* line = buf + l * lstride;
        if (act) {
          switch (R) {
            case 2: stage_load<T, E, 2>(line, n, t, tl, v, cin); break;
            case 3: stage_load<T, E, 3>(line, n, t, tl, v, cin); break;
            case 4: stage_load<T, E, 4>(line, n, t, tl, v, cin); break;
            case 5: stage_load<T, E, 5>(line, n, t, tl, v, cin); break;
            case 8: stage_load<T, E, 8>(line, n, t, tl, v, cin); break;
            case 16: stage_load<T, E, 16>(line, n, t, tl, v, cin); break;
            default: break;
          }
        }
        __syncthreads();
        if (act) {
          switch (R) {
            case 2: stage_store<T, E, 2>(line, n, ns, t, tl, v, tw, cout); break;
            case 3: stage_store<T, E, 3>(line, n, ns, t, tl, v, tw, cout); break;
            case 4: stage_store<T, E, 4>(line, n, ns, t, tl, v, tw, cout); break;
            case 5: stage_store<T, E, 5>(line, n, ns, t, tl, v, tw, cout); break;
            case 8: stage_store<T, E, 8>(line, n, ns, t, tl, v, tw, cout); break;
            case 16: stage_store<T, E, 16>(line, n, ns, t, tl, v, tw, cout); break;
            default: break;
          }
        }
        __syncthreads();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Compile-time plans: the same Stockham stages with the radix sequence, the
// stride products and the elements per thread fixed at compile time, so every
// stage is straight-line code (about a third of the registers of the
// run-time plan).  Used for the grid sizes listed in is_ct_size().
// ---------------------------------------------------------------------------
struct CtPlan {
  int e, nst;
  int r[kMaxStages];
  int ns[kMaxStages];
};

constexpr CtPlan ct_plan(int n) {
  CtPlan p{};
  p.e = (n % 5 == 0) ? 20 : ((n % 3 == 0) ? 24 : 16);
  const int cands[6] = {16, 8, 4, 2, 3, 5};
  int rem = n, ns = 1, s = 0;
  while (rem > 1 && s < kMaxStages) {
    int r = 0;
    for (int c = 0; c < 6; ++c)
      if (rem % cands[c] == 0 && p.e % cands[c] == 0) {
        r = cands[c];
        break;
      }
    if (r == 0) break;
    p.r[s] = r;
    p.ns[s] = ns;
    ns *= r;
    rem /= r;
    ++s;
  }
  p.nst = s;
  return p;
}

__host__ __device__ constexpr bool is_ct_size(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 96 || n == 192 || n == 160 || n == 320;
}

template <typename T, int N, int S>
__device__ __forceinline__ void ct_stage(cplx<T>* buf, int lstride, int nlines, const cplx<T>* tw, bool cin,
                                         bool cout) {
  constexpr CtPlan p = ct_plan(N);
  constexpr int E = p.e, R = p.r[S], NS = p.ns[S], TL = N / E;
  const int LPP = (int)blockDim.x / TL;
  static_assert(E % R == 0, "radix must divide the elements per thread");
  cplx<T> v[E];
  for (int l0 = 0; l0 < nlines; l0 += LPP) {
    const int l = l0 + (int)threadIdx.x / TL;
    const int t = (int)threadIdx.x % TL;
    const bool act = ((int)threadIdx.x < LPP * TL) && (l < nlines);
    cplx<T>* line = buf + l * lstride;
    if (act) {
#pragma unroll
      for (int b = 0; b < E / R; ++b) {
        const int j = t + TL * b;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          const cplx<T> a = line[pidx(j + k * (N / R))];
          v[b * R + k] = cin ? cconj(a) : a;
        }
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int b = 0; b < E / R; ++b) {
        const int j = t + TL * b;
        const int jm = (NS == 1) ? 0 : j % NS;
        cplx<T>* w = v + b * R;
        if constexpr (NS > 1) {
#pragma unroll
          for (int k = 1; k < R; ++k) w[k] = cmul(w[k], tw[jm * k * (N / (NS * R))]);
        }
        dft_small<T, R>(w);
        const int base = (j - jm) * R + jm;
#pragma unroll
        for (int k = 0; k < R; ++k) line[pidx(base + k * NS)] = cout ? cconj(w[k]) : w[k];
      }
    }
    __syncthreads();
  }
}

template <typename T, int N, int S = 0>
__device__ __forceinline__ void fft_ct(cplx<T>* buf, int lstride, int nlines, const cplx<T>* tw, bool inverse) {
  constexpr CtPlan p = ct_plan(N);
  if constexpr (S < p.nst) {
    ct_stage<T, N, S>(buf, lstride, nlines, tw, inverse && S == 0, inverse && S == p.nst - 1);
    fft_ct<T, N, S + 1>(buf, lstride, nlines, tw, inverse);
  }
}

// Kernels are templated on <E, N>: N > 0 selects the compile-time plan for
// that length, N == 0 the run-time plan with E elements per thread.
template <typename T, int E, int N>
__device__ __forceinline__ void fft_any(cplx<T>* buf, int lstride, int nlines, const FftPlan& p, const cplx<T>* tw,
                                        bool inverse, cplx<T>* scratch) {
  if constexpr (N > 0) fft_ct<T, N>(buf, lstride, nlines, tw, inverse);
  else fft_lines<T, E>(buf, lstride, nlines, p, tw, inverse, scratch);
}

}  // namespace mpol
