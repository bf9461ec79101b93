// Internal host-side definitions of libmpol.so shared by its translation
// units: error handling, device buffers, FFT plans, launch-shape choices, the
// context structure and the operator launch sequences (templates, instantiated
// once each in mpol_ops_*.hip).
#pragma once

#include "mpol.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpol_forces.h"
#include "mpol_gmres.h"
#include "mpol_kernels.h"
#include "mpol_mfma.h"
#include "mpol_sizes.h"
#include "mpol_zreg.h"

namespace mpol {
void real_space_tables(double xi, const double* r, int64_t n, double* out, bool regularize = true);
void real_space_tables_quad(double xi, const double* r, int64_t n, double* out, bool regularize = true);
}

using namespace mpol;

namespace mpol {

// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(MPOL_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HIPALLOC(buf, cnt)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (buf).reserve(cnt);                                                     \
    if (e_ != hipSuccess) return fail(MPOL_ENOMEM, std::string("device allocation failed: ") + #buf); \
  } while (0)

#define CHECK_STATUS(expr)        \
  do {                            \
    int s_ = (expr);              \
    if (s_ != MPOL_OK) return s_; \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t reserve(size_t cnt) {
    if (cnt <= n && p) return hipSuccess;
    release();
    if (cnt == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), cnt * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      n = 0;
      return e;
    }
    n = cnt;
    return hipSuccess;
  }
  void release() {
    // a buffer is only replaced between calls; wait for kernels that may still read it
    if (p) (void)hipDeviceSynchronize();
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// Page-locked host buffer (grows only), so device-to-host copies of solver
// state are queued on the stream without a staging wait.
template <typename T>
struct HostBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t reserve(size_t cnt) {
    if (cnt <= n && p) return hipSuccess;
    release();
    if (cnt == 0) return hipSuccess;
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), cnt * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      n = 0;
      return e;
    }
    n = cnt;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipDeviceSynchronize();
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  ~HostBuf() { release(); }
};

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Ascending order inside every segment [start[k], start[k+1]) of vals: short
// segments one thread each, long ones one workgroup each.
inline void sort_segments(int nseg, const int* start, int* vals, hipStream_t st) {
  hipLaunchKernelGGL(k_sort_segments, dim3(cdiv(nseg, 256)), dim3(256), 0, st, nseg, start, vals);
  hipLaunchKernelGGL(k_sort_segments_big, dim3(std::min(nseg, 1024)), dim3(256), 0, st, nseg, start, vals);
}

inline FftPlan make_plan(int n) {
  FftPlan p{};
  p.n = n;
  p.e = 0;
  p.nstages = 0;
  int m = n;
  for (int q : {2, 3, 5})
    while (m % q == 0) m /= q;
  const bool smooth = (m == 1);
  if (!smooth || n < 12) return p;
  for (int e : {16, 24, 20, 12}) {
    if (n % e != 0 || n / e > 256) continue;
    bool ok = true;
    for (int q : {2, 3, 5})
      if (n % q == 0 && e % q != 0) ok = false;
    if (!ok) continue;
    FftPlan t{};
    t.n = n;
    t.e = e;
    int rem = n, ns = 1, s = 0;
    while (rem > 1 && s < kMaxStages) {
      int r = 0;
      for (int cand : {16, 8, 4, 2, 3, 5})
        if (rem % cand == 0 && e % cand == 0) {
          r = cand;
          break;
        }
      if (r == 0) break;
      t.radix[s] = r;
      t.ns[s] = ns;
      ns *= r;
      rem /= r;
      ++s;
    }
    if (rem != 1) continue;
    t.nstages = s;
    return t;
  }
  return p;
}

inline void make_twiddles(int n, std::vector<double>& tw) {
  tw.resize(2 * (size_t)n);
  for (int k = 0; k < n; ++k) {
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    tw[2 * k] = (double)std::cos(ang);
    tw[2 * k + 1] = -(double)std::sin(ang);
  }
}

struct TileCfg {
  int ty = 1, tz = 1, rows = 1, ncl = 2, lstride = 1, ch = 64;
  size_t lds = 0;
};

struct ColCfg {
  int tx = 1, lstride = 1;
  size_t lds = 0;
};

// Column workgroups of the y / register z passes before the tail set
// (ColArgs::tail): kx / tx when one column is left over (kx = nx/2 + 1 with
// nx/2 a multiple of tx), else 0 (every workgroup takes tx columns).
inline int tail_blocks(int kx, int tx) { return (tx > 1 && kx % tx == 1) ? kx / tx : 0; }

inline ColCfg col_cfg(int n, bool generic, int kx, int lines_per_col, size_t csize, size_t budget, int tx_pref) {
  ColCfg c;
  int tx = std::min(tx_pref, kx);
  for (;; tx = std::max(1, tx / 2)) {
    c.tx = tx;
    c.lstride = padded_len(n);
    c.lds = ((size_t)n + (size_t)lines_per_col * tx * c.lstride * (generic ? 2 : 1)) * csize;
    if (c.lds <= budget || tx == 1) break;
  }
  return c;
}

constexpr size_t kLdsTile = 48 * 1024;
constexpr size_t kLdsMfma = 80 * 1024;  // two matrix-core tile workgroups per CU
constexpr int kTuneKnobs = 11;
constexpr size_t kLastLevelCache = (size_t)256 << 20;  // MI355X Infinity Cache bytes
constexpr int kRhsGroup = 4;  // automatic right-hand sides per matrix-core tile workgroup

// Candidates of every tile, sorted by stencil start along x within chunks of
// ch, for one tile shape and precision (built on first use after
// mpol_set_geometry; see k_img_fill).
struct TileImage {
  bool valid = false;
  int ty = 0, tz = 0, ch = 0;
  int64_t total = 0;
  DevBuf<int> cnt, off;
  DevBuf<int2> gs;
  DevBuf<unsigned char> fw;
  DevBuf<int> pcnt, poff, pslot;  // slots of each sorted particle
  DevBuf<unsigned char> part;     // contraction partials
};
constexpr size_t kLdsMax = 150 * 1024;  // dynamic LDS per workgroup (gfx950 has 160 KB per CU)
constexpr size_t kLdsCol = 112 * 1024;

}  // namespace mpol

struct mpol_ctx {
  ~mpol_ctx() {
    for (hipEvent_t& e : step_ev)
      if (e) (void)hipEventDestroy(e);
  }
  int device = 0;
  hipStream_t stream = nullptr;  // launch stream; nullptr is the default stream
  // optional operator of the mixed scheme's correction solves and first
  // residual (mpol_set_inner_operator): another context on the same
  // particles, typically a coarser grid at a looser Ewald tolerance; its
  // correction solves stop at inner_floor at the lowest
  mpol_ctx* inner = nullptr;
  double inner_floor = 0.0;
  // tables
  int ntab = 0;
  double self_perp = 0.0;
  DevBuf<double> perp, para, rvals;
  // geometry
  bool have_geo = false;
  int n = 0;
  double box[3] = {0, 0, 0}, h[3] = {0, 0, 0}, eta[3] = {0, 0, 0};
  double xi = 0, rc = 0;
  int ng[3] = {0, 0, 0}, kx = 0, kxs = 0, P = 0, half = 0, nc[3] = {0, 0, 0};
  int64_t npairs = 0;
  DevBuf<double> x, base;
  DevBuf<int> bin_key, bin_count, bin_start, bin_cursor, cell_key, cell_count, cell_start, cell_cursor,
      cell_members, g0u, perm, rs_count, rs_off, rs_j, err;
  DevBuf<int4> rec;
  DevBuf<double4> base64;
  DevBuf<float4> base32;
  DevBuf<double> ms;
  double norm = 0.0;
  DevBuf<double> rs_A, rs_BA, rs_rh;
  DevBuf<float> rs_A32, rs_BA32, rs_rh32;
  FftPlan plan[3];
  DevBuf<double> ktab;      // e_x, e_y, e_z, then (cubic boxes) the Bessel factor table
  DevBuf<float> ktab32;     // cubic boxes: e_z, then the dipole table over |k|^2, in FP32
  bool cubic = false;
  DevBuf<double> tw64[3];
  DevBuf<float> tw32[3];
  // permittivity: diagonal of the dipole rows (and, when quadrupole tables are
  // set, of the quadrupole rows); imaginary parts when complex
  int nsets = 0;
  bool cplx_eps = false;
  // the pass that applies the anti-Hermitian (Nyquist) part of the wave-space
  // operator for complex permittivities (run_operator), and its output
  bool odd_pass = false;
  DevBuf<double> odd_tmp;
  DevBuf<double> eps_re, eps_im;
  DevBuf<double> diag, diag_im, diag_q, diag_qim;
  bool diag_q_valid = false;
  // dipole + quadrupole operator: 7 tables on the rvals grid, per-pair coefficients
  bool have_qtab = false;
  double self_gq2 = 0.0;
  DevBuf<double> qtab;
  bool quad_pairs_valid = false;
  DevBuf<double> rs_Q;
  DevBuf<float> rs_Q32;
  int64_t jq_off = -1;      // offset of the quadrupole Bessel table in ktab (cubic boxes)
  int64_t jk_off = -1;      // offset of the dipole table divided by |k|^2 in ktab (cubic boxes)
  // polydisperse operator: radii (caller's order), radius types and pair tables
  // (ntypes > 0) or closed-form couplings (ntypes == 0)
  bool have_sizes = false;
  int sz_ntypes = 0;
  DevBuf<double> sz_radius, sz_rad_sorted, sz_perp, sz_para, sz_A, sz_BA, sz_diag, sz_diag_im, sz_slab;
  DevBuf<float> sz_A32, sz_BA32;
  DevBuf<int> sz_type, sz_col;
  bool sz_pairs_valid = false, sz_diag_valid = false;
  DevBuf<unsigned char> sz_grid;
  // node-major spread: per-node segments of sorted particle indices and their
  // w r entries ([0] FP32, [1] FP64), particle-major dipoles of the batch
  bool sz_nodes_valid = false, sz_nw_valid[2] = {false, false};
  int sz_entries = 0;
  DevBuf<int> sz_ncnt, sz_noff, sz_nidx;
  DevBuf<float4> sz_nw32, sz_pw32;
  DevBuf<double4> sz_nw64, sz_pw64;  // node-major (nw) and particle-major (pw) entries
  DevBuf<unsigned char> sz_pm, sz_gridT;
  // vector-Jacobian product workspaces
  DevBuf<double> vjp_in, vjp_out, vjp_grid;
  // workspaces
  DevBuf<unsigned char> spec;
  DevBuf<double> zsum;
  // launch-shape overrides (0 = automatic): tile ty, tile tz, particle chunk,
  // y-pass columns, z-pass columns, z-pass threads, spread/contract engine
  // (1 matrix cores, 2 vector loops)
  int tune[2][kTuneKnobs] = {};  // [0] FP64 operator, [1] FP32 operator
  // per-tile candidate images of the matrix-core spread/contract ([0] FP32, [1] FP64)
  TileImage img[2];
  // gmres
  DevBuf<double> gV, gW, gR, gAx, gH, gcs, gsn, gg, gy, gh1, gh2, gpart, gpart2, gbn, gbeta, grel, grn2, gx, gd,
      gnorm;
  DevBuf<int> gstate, gkdim;
  // while set (inside an Arnoldi step), the per-system GMRES state: operator
  // work of systems not in kIter is skipped
  const int* gate = nullptr;
  HostBuf<double> hscr;  // page-locked staging of small reads
  // Three slots of per-system GMRES state and residual estimates that the
  // kernels write into page-locked host memory (two alternating Arnoldi-step
  // slots, read one step behind, and the cycle-start slot), their device
  // addresses and the events the host waits on before reading a slot.
  HostBuf<int> hstep;
  HostBuf<double> hsteprel;
  int* dstep = nullptr;
  double* dsteprel = nullptr;
  hipEvent_t step_ev[3] = {nullptr, nullptr, nullptr};
  // truncated fixed-dipole forces: cell list
  DevBuf<int> f_cell, f_count, f_start, f_cursor, f_members;
  DevBuf<int> drop_keep, drop_pos, drop_err;  // drop-in pair filter (mpol_dropin.hip)
  DevBuf<double> ftab1, ftab2;                // pair-force tables (mpol_ewald_forces)
  int* scratch = nullptr;  // this device's allocation-free scratch (mpol_compute_forces_noalloc)
};

namespace mpol {

// Raise a kernel's dynamic-LDS limit to at least `bytes` on the current
// device.  The limit only grows, and the attribute is set once per (device,
// kernel, larger size) under a process-wide lock: launches from several
// threads (one context each) never meet a concurrent attribute change of the
// kernel they launch after its first use, and launches make no attribute call.
void raise_lds_limit(const void* kernel, size_t bytes);

template <typename K>
void set_lds(K kernel, size_t bytes) {
  raise_lds_limit(reinterpret_cast<const void*>(kernel), bytes);
}

// Launch KERNEL(E, N, ...) with the compile-time plan for length n when there
// is one, else with the run-time plan of E elements per thread.
#define DISPATCH_FFT(n, e, KERNEL, ...)                          \
  switch (n) {                                                   \
    case 32: KERNEL(0, 32, __VA_ARGS__); break;                  \
    case 64: KERNEL(0, 64, __VA_ARGS__); break;                  \
    case 96: KERNEL(0, 96, __VA_ARGS__); break;                  \
    case 128: KERNEL(0, 128, __VA_ARGS__); break;                \
    case 160: KERNEL(0, 160, __VA_ARGS__); break;                \
    case 192: KERNEL(0, 192, __VA_ARGS__); break;                \
    case 256: KERNEL(0, 256, __VA_ARGS__); break;                \
    case 320: KERNEL(0, 320, __VA_ARGS__); break;                \
    case 512: KERNEL(0, 512, __VA_ARGS__); break;                \
    default:                                                     \
      switch (e) {                                               \
        case 0: KERNEL(0, 0, __VA_ARGS__); break;                \
        case 12: KERNEL(12, 0, __VA_ARGS__); break;              \
        case 16: KERNEL(16, 0, __VA_ARGS__); break;              \
        case 20: KERNEL(20, 0, __VA_ARGS__); break;              \
        case 24: KERNEL(24, 0, __VA_ARGS__); break;              \
        default: return fail(MPOL_EINVAL, "unsupported FFT plan"); \
      }                                                          \
  }

inline bool generic_fft(const FftPlan& p) { return !is_ct_size(p.n) && p.e == 0; }

// Launch KERNEL(N, ...) for a length with a compile-time FFT plan.
#define DISPATCH_CT(n, KERNEL, ...)                                  \
  switch (n) {                                                       \
    case 32: KERNEL(32, __VA_ARGS__); break;                         \
    case 64: KERNEL(64, __VA_ARGS__); break;                         \
    case 96: KERNEL(96, __VA_ARGS__); break;                         \
    case 128: KERNEL(128, __VA_ARGS__); break;                       \
    case 160: KERNEL(160, __VA_ARGS__); break;                       \
    case 192: KERNEL(192, __VA_ARGS__); break;                       \
    case 256: KERNEL(256, __VA_ARGS__); break;                       \
    case 320: KERNEL(320, __VA_ARGS__); break;                       \
    case 512: KERNEL(512, __VA_ARGS__); break;                       \
    default: return fail(MPOL_EINVAL, "no compile-time FFT plan");   \
  }

template <typename T>
const void* tw_ptr(mpol_ctx* c, int axis) {
  if constexpr (sizeof(T) == 8) return c->tw64[axis].p;
  else return c->tw32[axis].p;
}

template <typename T>
const void* base_ptr(mpol_ctx* c) {
  if constexpr (sizeof(T) == 8) return c->base64.p;
  else return c->base32.p;
}

// out = A m (or its parts when hr/hslab are given) in precision T, for NC = 3
// (dipole operator: hr <- real space, hslab <- slab term, out <- wave space)
// or NC = 8 (dipole + quadrupole operator: hr <- real space, out <- wave
// space).  Stage boundaries recorded when `ev` is given (kStages + 1 events).
constexpr int kStages = 8;
constexpr int kZsumBlocks = 64;
enum { kOpDipole = 3, kOpQuad = 8, kOpSizes = 103 };
inline int unknowns_per_particle(int kind) { return kind == kOpQuad ? 8 : 3; }

// Largest tile (rows of nx nodes, NC components) and particle chunk that fit
// the LDS budget.
template <int NC>
TileCfg tile_cfg_nc(int nx, bool generic, size_t csize, int P, size_t budget, bool sort_arrays = false) {
  const int opts[][2] = {{4, 4}, {4, 2}, {2, 2}, {2, 1}, {1, 1}};
  const size_t rsize = csize / 2;
  TileCfg c;
  for (const auto& o : opts) {
    c.ty = o[0];
    c.tz = o[1];
    c.rows = c.ty * c.tz;
    c.ncl = (NC * c.rows + 1) / 2;
    c.lstride = padded_len(nx);
    const size_t grid = ((size_t)nx + (size_t)c.ncl * c.lstride * (generic ? 2 : 1)) * csize;
    const size_t per = (NC + (size_t)P + c.rows) * rsize + 2 * sizeof(int);
    const size_t fixed = sort_arrays ? ((size_t)nx + 1) * sizeof(int) : 0;
    for (c.ch = 256; c.ch >= 64; c.ch /= 2) {
      c.lds = grid + fixed + c.ch * per;
      if (c.lds <= budget) return c;
    }
  }
  return c;
}

// LDS bytes of a tile kernel with the given shape.
template <int NC>
size_t tile_lds(int nx, bool generic, size_t csize, int P, const TileCfg& t, bool sort_arrays) {
  return ((size_t)nx + (size_t)t.ncl * t.lstride * (generic ? 2 : 1)) * csize +
         (sort_arrays ? ((size_t)nx + 1) * sizeof(int) : 0) +
         (size_t)t.ch * ((NC + (size_t)P + t.rows) * (csize / 2) + 2 * sizeof(int));
}

// Launch-shape knobs of the operator in precision T.
template <typename T>
const int* knobs(const mpol_ctx* c) {
  return c->tune[sizeof(T) == 4 ? 1 : 0];
}

template <int NC>
void apply_tile_tuning(const int* tn, TileCfg& t) {
  if (tn[0] > 0 && tn[1] > 0) {
    t.ty = tn[0];
    t.tz = tn[1];
    t.rows = t.ty * t.tz;
    t.ncl = (NC * t.rows + 1) / 2;
  }
  if (tn[2] > 0) t.ch = tn[2];
}

// Tile shape and chunk of the matrix-core spread/contract: the largest tile
// (rows first) whose chunk of 256 (FP32) or 128 (FP64) candidates, else 128,
// fits kLdsMfma, else the largest chunk that fits.  ok = false when the grid or stencil does not suit
// that path (compile-time x plan, nx % 16 == 0, P + 16 <= nx) or the engine
// knob asks for the vector loops.
template <typename T, int NC>
TileCfg mfma_tile_cfg(const mpol_ctx* c, bool& ok) {
  const int nx = c->ng[0];
  TileCfg t;
  const int* tn = knobs<T>(c);
  ok = tn[6] != 2 && is_ct_size(nx) && nx % 16 == 0 && c->P + 16 <= nx;
  if (!ok) return t;
  const size_t cs = sizeof(cplx<T>);
  // 4 or 8 rows (the kernels are instantiated for those)
  const int opts[][2] = {{4, 2}, {2, 2}};
  bool found = false;
  for (int pass = 0; pass < 3 && !found; ++pass) {
    const size_t budget = pass < 2 ? kLdsMfma : kLdsMax;
    for (const auto& o : opts) {
      t.ty = o[0];
      t.tz = o[1];
      t.rows = t.ty * t.tz;
      t.ncl = (NC * t.rows + 1) / 2;
      t.lstride = padded_len(nx);
      // FP32: 256 candidates per chunk (one chunk for most tiles at phi = 0.4,
      // 256^3, P = 10: 0.17 ms against 0.18 ms with 128 for the spread, three
      // workgroups per CU); FP64: 128 (two workgroups per CU; 256 leaves one)
      const int ch0 = sizeof(T) == 4 ? 256 : 128;
      for (t.ch = ch0; t.ch >= (pass == 0 ? 128 : 32); t.ch /= 2) {
        t.lds = mfma_tile_lds<NC>(nx, cs, c->P, t.ncl, t.lstride, t.ch, t.rows);
        if (t.lds <= budget) {
          found = true;
          break;
        }
      }
      if (found) break;
    }
  }
  const TileCfg chosen = t;
  apply_tile_tuning<NC>(tn, t);
  // 4 or 8 rows; 16 rows (4 x 4 tiles) for the dipole operator only: a tile
  // shape tuned on the dipole operator that this operator lacks keeps the
  // default shape (with the tuned chunk)
  const bool rows_ok = t.rows == 4 || t.rows == 8 || (NC == 3 && t.rows == 16 && t.ty == 4 && t.tz == 4);
  if (!rows_ok || t.ty > 4 || t.tz > 4) {
    const int ch = t.ch;
    t = chosen;
    t.ch = ch;
  }
  t.ncl = (NC * t.rows + 1) / 2;
  t.lds = mfma_tile_lds<NC>(nx, cs, c->P, t.ncl, t.lstride, t.ch, t.rows);
  if (t.lds > kLdsMax) ok = false;
  // candidate slots of the image are int32
  const double redund = (double)(c->P + t.ty) / t.ty * (double)(c->P + t.tz) / t.tz;
  if ((double)c->n * redund >= 2.0e9) ok = false;
  return t;
}

// Build the tile image of precision T for tile shape ta (no-op when the
// current one matches).
template <typename T>
int ensure_tile_image(mpol_ctx* c, const TileArgs& ta, TileImage& im) {
  if (im.valid && im.ty == ta.ty && im.tz == ta.tz && im.ch == ta.ch) return MPOL_OK;
  im.valid = false;
  hipStream_t st = c->stream;
  const int ntx = cdiv(c->ng[1], ta.ty), ntz = cdiv(c->ng[2], ta.tz), ntiles = ntx * ntz;
  HIPALLOC(im.cnt, (size_t)ntiles);
  HIPALLOC(im.off, (size_t)ntiles + 1);
  hipLaunchKernelGGL(k_img_count, dim3(cdiv(ntiles, 256)), dim3(256), 0, st, ta, ntx, ntiles, im.cnt.p);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, im.cnt.p, im.off.p, ntiles);
  int total = 0;
  HIPCHK(hipMemcpyAsync(&total, im.off.p + ntiles, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (total < 0) return fail(MPOL_EINVAL, "tile image too large");
  const size_t slots = total > 0 ? (size_t)total : 1;
  HIPALLOC(im.gs, slots);
  HIPALLOC(im.fw, slots * (size_t)(ta.P + ta.rows) * sizeof(T));
  hipLaunchKernelGGL((k_img_fill<T>), dim3(ntx, ntz), dim3(256), 0, st, ta, im.off.p, im.gs.p,
                     reinterpret_cast<T*>(im.fw.p));
  // per-particle slot lists (the contraction's reduction order)
  const int n = c->n;
  HIPALLOC(im.pcnt, (size_t)n);
  HIPALLOC(im.poff, (size_t)n + 1);
  HIPALLOC(im.pslot, slots);
  HIPCHK(hipMemsetAsync(im.pcnt.p, 0, (size_t)n * sizeof(int), st));
  hipLaunchKernelGGL(k_img_pcount, dim3(cdiv(total, 256)), dim3(256), 0, st, im.gs.p, (int64_t)total, im.pcnt.p);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, im.pcnt.p, im.poff.p, n);
  HIPCHK(hipMemsetAsync(im.pcnt.p, 0, (size_t)n * sizeof(int), st));
  hipLaunchKernelGGL(k_img_pscatter, dim3(cdiv(total, 256)), dim3(256), 0, st, im.gs.p, (int64_t)total, im.poff.p,
                     im.pcnt.p, im.pslot.p);
  hipLaunchKernelGGL(k_img_psort, dim3(cdiv(n, 256)), dim3(256), 0, st, n, im.poff.p, im.pslot.p);
  HIPCHK(hipGetLastError());
  im.valid = true;
  im.ty = ta.ty;
  im.tz = ta.tz;
  im.ch = ta.ch;
  im.total = total;
  return MPOL_OK;
}

// 16-row (4 x 4) tiles of the matrix-core spread / contraction: instantiated
// for the dipole operator only (NC = 3: three column blocks of 16).
template <typename T, int N, int NC>
void launch_mfma16(bool spread, dim3 grid, size_t lds, hipStream_t st, const TileArgs& ta, const ImgArgs& im,
                   cplx<T>* spec) {
  if constexpr (NC == 3) {
    if (spread) {
      set_lds(k_spread_mfma_xfft<T, N, NC, 16>, lds);
      hipLaunchKernelGGL((k_spread_mfma_xfft<T, N, NC, 16>), grid, dim3(256), lds, st, ta, im, spec);
    } else {
      set_lds(k_xifft_contract_mfma<T, N, NC, 16>, lds);
      hipLaunchKernelGGL((k_xifft_contract_mfma<T, N, NC, 16>), grid, dim3(256), lds, st, ta, im, spec);
    }
  }
}

template <typename T, int NC>
int run_operator(mpol_ctx* c, const double* m, double* out, double* hr, double* hslab, int nrhs, double slab,
                 hipEvent_t* ev = nullptr, int stop_after = kStages) {
  const size_t cs = sizeof(cplx<T>);
  const size_t spec_bytes = (size_t)NC * nrhs * c->ng[2] * c->ng[1] * c->kxs * cs;
  HIPALLOC(c->spec, spec_bytes);
  HIPALLOC(c->zsum, (size_t)nrhs * (kZsumBlocks + 1));
  cplx<T>* spec = reinterpret_cast<cplx<T>*>(c->spec.p);
  const int n = c->n;
  hipStream_t st = c->stream;
  const int* gate = c->gate;
  const int rps = c->cplx_eps ? 2 : 1;  // operator right-hand sides per system
  const bool odd = c->odd_pass;         // wave space only, anti-Hermitian part

  if (ev) HIPCHK(hipEventRecord(ev[0], st));
  const T* pA;
  const T* pBA;
  const T* prh;
  if constexpr (sizeof(T) == 8) {
    pA = c->rs_A.p; pBA = c->rs_BA.p; prh = c->rs_rh.p;
  } else {
    pA = c->rs_A32.p; pBA = c->rs_BA32.p; prh = c->rs_rh32.p;
  }
  if (odd) {
    // the contraction adds onto zeros
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    hipLaunchKernelGGL(k_fill_double, dim3(cdiv((size_t)nrhs * NC * n, 256)), dim3(256), 0, st, out,
                       (int64_t)nrhs * NC * n, 0.0);
  } else if constexpr (NC == kOpDipole) {
    if (slab != 0.0) {
      double* zpart = c->zsum.p + nrhs;
      hipLaunchKernelGGL(k_sum_z, dim3(kZsumBlocks, nrhs), dim3(256), 0, st, m, n, zpart);
      hipLaunchKernelGGL(k_sum_finish, dim3(cdiv(nrhs, 256)), dim3(256), 0, st, zpart, kZsumBlocks, nrhs, c->zsum.p);
    }
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    const double V = c->box[0] * c->box[1] * c->box[2];
    hipLaunchKernelGGL((k_real_space<T>), dim3(cdiv((size_t)n * 16, 256), nrhs), dim3(256), 0, st, n, nrhs, c->nsets, m, out,
                       c->diag.p, c->cplx_eps ? c->diag_im.p : nullptr, c->rs_off.p, c->rs_j.p, pA, pBA, prh,
                       c->zsum.p, slab / V, hr != nullptr ? 1 : 0, hr, hslab, gate, rps);
  } else {
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    const T* pq;
    if constexpr (sizeof(T) == 8) pq = c->rs_Q.p;
    else pq = c->rs_Q32.p;
    hipLaunchKernelGGL((k_real_space_quad<T>), dim3(cdiv((size_t)n * 16, 256), nrhs), dim3(256), 0, st, n, nrhs, c->nsets, m,
                       out, c->diag.p, c->diag_q.p, c->cplx_eps ? c->diag_im.p : nullptr,
                       c->cplx_eps ? c->diag_qim.p : nullptr, c->rs_off.p, c->rs_j.p, pA, pBA, prh, pq,
                       hr != nullptr ? 1 : 0, hr, gate, rps);
  }
  // the stage events bracket the real-space sum alone; the permutation of
  // the moments into sorted order counts with the spread
  if (ev) HIPCHK(hipEventRecord(ev[2], st));
  HIPALLOC(c->ms, (size_t)nrhs * NC * n);
  hipLaunchKernelGGL(k_permute_m<NC>, dim3(cdiv(n, 256), nrhs), dim3(256), 0, st, m, c->perm.p, n, c->ms.p, gate,
                     rps);

  // tile (x axis) kernels
  const bool gen_x = generic_fft(c->plan[0]);
  TileCfg tc = tile_cfg_nc<NC>(c->ng[0], gen_x, cs, c->P, kLdsTile);
  apply_tile_tuning<NC>(knobs<T>(c), tc);
  tc.lds = tile_lds<NC>(c->ng[0], gen_x, cs, c->P, tc, false);
  TileArgs ta{};
  ta.nx = c->ng[0]; ta.ny = c->ng[1]; ta.nz = c->ng[2]; ta.kx = c->kx; ta.kxs = c->kxs;
  ta.ty = tc.ty; ta.tz = tc.tz; ta.rows = tc.rows; ta.ncl = tc.ncl; ta.lstride = tc.lstride; ta.ch = tc.ch;
  ta.P = c->P; ta.half = c->half; ta.n = n; ta.nrhs = nrhs;
  ta.rec = c->rec.p; ta.base = base_ptr<T>(c); ta.ms = c->ms.p; ta.bin_start = c->bin_start.p;
  for (int d = 0; d < 3; ++d) {
    ta.h[d] = c->h[d];
    ta.cexp[d] = 2.0 * c->xi * c->xi / c->eta[d];
  }
  ta.norm = c->norm;
  ta.tw = tw_ptr<T>(c, 0); ta.plan = c->plan[0];
  ta.hvol = c->h[0] * c->h[1] * c->h[2];
  ta.gate = gate;
  ta.gdiv = rps;
  const dim3 tgrid(cdiv(c->ng[1], tc.ty), cdiv(c->ng[2], tc.tz), nrhs);
  if (tc.lds > kLdsMax) return fail(MPOL_EINVAL, "spread/contract tile does not fit in LDS (reduce tile or chunk)");

  // matrix-core spread/contract over the tile image when the grid suits it
  bool use_mfma = false;
  const TileCfg tm = mfma_tile_cfg<T, NC>(c, use_mfma);
  if (knobs<T>(c)[6] == 1 && !use_mfma) return fail(MPOL_EINVAL, "matrix-core spread/contract does not apply to this grid");
  TileArgs tma = ta;
  ImgArgs ima{};
  dim3 mgrid(1, 1, 1);
  if (use_mfma) {
    tma.ty = tm.ty; tma.tz = tm.tz; tma.rows = tm.rows; tma.ncl = tm.ncl; tma.lstride = tm.lstride; tma.ch = tm.ch;
    TileImage& im = c->img[sizeof(T) == 8 ? 1 : 0];
    CHECK_STATUS(ensure_tile_image<T>(c, tma, im));
    ima.off = im.off.p;
    ima.gs = im.gs.p;
    ima.fw = im.fw.p;
    ima.perm = c->perm.p;
    ima.total = im.total;
    HIPALLOC(im.part, (size_t)nrhs * (size_t)std::max<int64_t>(im.total, 1) * NC * sizeof(T));
    ima.part = im.part.p;
    // right-hand sides per workgroup: one staged tile image serves them in
    // turn (knob 7; automatic: all of them, up to kRhsGroup)
    const int rg_knob = knobs<T>(c)[7];
    tma.rg = std::max(1, std::min(nrhs, rg_knob > 0 ? rg_knob : kRhsGroup));
    mgrid = dim3(cdiv(c->ng[1], tm.ty), cdiv(c->ng[2], tm.tz), cdiv(nrhs, tma.rg));
  }

  // the spread's staged chunk shares LDS with the tile lines when a
  // workgroup serves one right-hand side and keeps its accumulators in
  // registers (with more, aliasing would restage the image per right-hand
  // side, which the groups exist to avoid).  With the group knob automatic,
  // the FP32 dipole spread takes one right-hand side per workgroup where it
  // can alias (more workgroups per CU: three-right-hand-side FP32 operator
  // 1.55 -> 1.51 ms at N = 1e5, 256^3; the FP64 one 3.21 -> 3.30 ms, so FP64
  // keeps the group), the contraction keeps the group.
  TileArgs tms = tma;
  dim3 msgrid = mgrid;
  size_t spread_lds = tm.lds;
  tms.alias = 0;
  if (use_mfma && mfma_reg_acc(c->ng[0], NC, tm.rows, sizeof(T))) {
    if (NC == kOpDipole && sizeof(T) == 4 && knobs<T>(c)[7] == 0) {
      tms.rg = 1;
      msgrid.z = nrhs;
    }
    if (tms.rg == 1) {
      tms.alias = 1;
      spread_lds = mfma_tile_lds_alias<NC>(c->ng[0], cs, c->P, tm.ncl, tm.lstride, tm.ch, tm.rows);
    }
  }
  // FP32: atomic-free gather spread; FP64: LDS FP64 atomics (faster there)
  const bool gather = sizeof(T) == 4 && is_ct_size(c->ng[0]) && c->P + 3 <= c->ng[0];
  if (use_mfma) {
#define LAUNCH_SPREAD_MFMA(N, ...)                                                              \
  do {                                                                                          \
    if (tm.rows == 16) {                                                                        \
      launch_mfma16<T, N, NC>(true, msgrid, spread_lds, st, tms, ima, spec);                    \
    } else if (tm.rows == 8) {                                                                  \
      set_lds(k_spread_mfma_xfft<T, N, NC, 8>, spread_lds);                                    \
      hipLaunchKernelGGL((k_spread_mfma_xfft<T, N, NC, 8>), msgrid, dim3(256), spread_lds, st, tms, ima, spec); \
    } else {                                                                                    \
      set_lds(k_spread_mfma_xfft<T, N, NC, 4>, spread_lds);                                    \
      hipLaunchKernelGGL((k_spread_mfma_xfft<T, N, NC, 4>), msgrid, dim3(256), spread_lds, st, tms, ima, spec); \
    }                                                                                           \
  } while (0)
    DISPATCH_CT(c->ng[0], LAUNCH_SPREAD_MFMA, 0);
#undef LAUNCH_SPREAD_MFMA
  } else if (gather) {
    TileCfg tg = tile_cfg_nc<NC>(c->ng[0], false, cs, c->P, kLdsTile, true);
    apply_tile_tuning<NC>(knobs<T>(c), tg);
    tg.lds = tile_lds<NC>(c->ng[0], false, cs, c->P, tg, true);
    if (tg.lds > kLdsMax) return fail(MPOL_EINVAL, "spread tile does not fit in LDS (reduce tile or chunk)");
    TileArgs tga = ta;
    tga.ty = tg.ty; tga.tz = tg.tz; tga.rows = tg.rows; tga.ncl = tg.ncl; tga.lstride = tg.lstride; tga.ch = tg.ch;
    const dim3 ggrid(cdiv(c->ng[1], tg.ty), cdiv(c->ng[2], tg.tz), nrhs);
#define LAUNCH_GATHER(N, ...)                                                                   \
  do {                                                                                          \
    set_lds(k_spread_gather_xfft<T, 0, N, 4, NC>, tg.lds);                                      \
    hipLaunchKernelGGL((k_spread_gather_xfft<T, 0, N, 4, NC>), ggrid, dim3(256), tg.lds, st, tga, m, spec); \
  } while (0)
    DISPATCH_CT(c->ng[0], LAUNCH_GATHER, 0);
#undef LAUNCH_GATHER
  } else {
    // the spread tile accumulates in FP64 in both precisions
    using TA = double;
    TileCfg tsc = tile_cfg_nc<NC>(c->ng[0], gen_x, sizeof(cplx<TA>), c->P, kLdsTile);
    apply_tile_tuning<NC>(knobs<T>(c), tsc);
    const size_t slds = tile_lds<NC>(c->ng[0], gen_x, sizeof(cplx<TA>), c->P, tsc, false);
    if (slds > kLdsMax) return fail(MPOL_EINVAL, "spread tile does not fit in LDS (reduce tile or chunk)");
    TileArgs tsa = ta;
    tsa.ty = tsc.ty; tsa.tz = tsc.tz; tsa.rows = tsc.rows; tsa.ncl = tsc.ncl; tsa.lstride = tsc.lstride;
    tsa.ch = tsc.ch;
    tsa.base = c->base64.p;
    tsa.tw = c->tw64[0].p;
    const dim3 sgrid(cdiv(c->ng[1], tsa.ty), cdiv(c->ng[2], tsa.tz), nrhs);
#define LAUNCH_SPREAD(E, N, ...)                                                                \
  do {                                                                                          \
    set_lds(k_spread_xfft<TA, E, N, T, NC>, slds);                                              \
    hipLaunchKernelGGL((k_spread_xfft<TA, E, N, T, NC>), sgrid, dim3(256), slds, st, tsa, m, spec); \
  } while (0)
    DISPATCH_FFT(c->ng[0], ta.plan.e, LAUNCH_SPREAD, 0);
#undef LAUNCH_SPREAD
  }
  if (ev) HIPCHK(hipEventRecord(ev[3], st));
  if (stop_after <= 2) return MPOL_OK;

  ScaleParams sp{};
  for (int d = 0; d < 3; ++d) {
    sp.twopi_over_l[d] = 2.0 * kPi / c->box[d];
    sp.one_minus_eta[d] = 1.0 - c->eta[d];
  }
  sp.inv_4xi2 = 1.0 / (4.0 * c->xi * c->xi);
  sp.inv_g = 1.0 / ((double)c->ng[0] * (double)c->ng[1] * (double)c->ng[2]);
  sp.ex = c->ktab.p;
  sp.ey = sp.ex + c->ng[0];
  sp.ez = sp.ey + c->ng[1];
  sp.jt = c->cubic ? sp.ez + c->ng[2] : nullptr;
  sp.jq = (c->cubic && c->jq_off >= 0) ? c->ktab.p + c->jq_off : nullptr;
  sp.jk = (c->cubic && c->jk_off >= 0) ? c->ktab.p + c->jk_off : nullptr;
  sp.ez32 = sp.jk ? c->ktab32.p : nullptr;
  sp.jk32 = sp.jk ? c->ktab32.p + c->ng[2] : nullptr;

  // the rank-one scheme (mpol_zreg.h) for nz = 256 dipole grids of cubic
  // boxes: the register z pass in columns of TXR with one left over, not
  // forced off (z_rank1 = 2) or to the LDS z pass (z_cols), a compile-time
  // y plan; the full operator only (the stage stops read the standard
  // three-component grids)
  constexpr int TXR = sizeof(T) == 4 ? 16 : 8;
  const bool zreg = NC == 3 && c->ng[2] == 256 && knobs<T>(c)[4] == 0 && !odd;
  sp.odd = odd ? 1 : 0;
  const bool rank1 = zreg && tail_blocks(c->kx, TXR) > 0 && knobs<T>(c)[8] != 2 && sp.jk != nullptr &&
                     is_ct_size(c->ng[1]) && stop_after >= kStages;

  // Several systems: the y, z and inverse y passes run system by system, so
  // that each pass finds much of the spectrum the one before it wrote still
  // in the last-level cache (256 MB; a system's half spectra are 200 MB in
  // FP32 at 256^3, three systems' 600 MB).  Knob 10: 1 system by system, 2
  // each pass over all systems; automatic: system by system when all the
  // systems' spectra overflow that cache and one system's fill a quarter of
  // it (smaller launches leave the chip part idle: the three-system FP32
  // operator at 128^3 is slower system by system), except in timed runs,
  // whose stage events need the pass order.  The full operator only (the
  // stage stops return between passes).
  const int nsys = nrhs / rps;
  const int fft_order = knobs<T>(c)[10];
  const bool big = spec_bytes > kLastLevelCache && spec_bytes / nsys >= kLastLevelCache / 4;
  const bool by_sys = nsys > 1 && stop_after >= kStages &&
                      (fft_order == 1 || (fft_order == 0 && ev == nullptr && big));
  const int nbatch = by_sys ? nsys : 1;  // launch batches of the three passes
  const int rb = by_sys ? rps : nrhs;    // right-hand sides per batch
  const size_t batch_stride = (size_t)rb * NC * c->ng[2] * c->ng[1] * c->kxs;
  cplx<T>* const spec_all = spec;

  // y axis forward
  const ColCfg cy = col_cfg(c->ng[1], generic_fft(c->plan[1]), c->kx, 1, cs, kLdsCol,
                            knobs<T>(c)[3] > 0 ? knobs<T>(c)[3] : (sizeof(T) == 8 ? 8 : 16));  // 128-byte rows
  ColArgs ay{};
  ay.nx = c->ng[0]; ay.ny = c->ng[1]; ay.nz = c->ng[2]; ay.kx = c->kx; ay.kxs = c->kxs; ay.tx = cy.tx; ay.lstride = cy.lstride;
  ay.plan = c->plan[1]; ay.tw = tw_ptr<T>(c, 1);
  ay.gate = gate;
  ay.gdiv = NC * rps;  // batch index = rhs * NC + component
  const dim3 ygrid(cdiv(c->kx, cy.tx), c->ng[2], NC * rb);
  if (cy.lds > kLdsMax) return fail(MPOL_EINVAL, "y-pass columns do not fit in LDS");
  // rank-one scheme: x and y components together, the z component in
  // workgroups of the other role (k_fft_y_xy)
  ColArgs axy = ay;
  axy.gdiv = rps;  // right-hand side = blockIdx.z / 2
  const dim3 xygrid(cdiv(c->kx, cy.tx), c->ng[2], 2 * rb);
  const size_t xylds = ((size_t)c->ng[1] + 2 * (size_t)cy.tx * cy.lstride * (generic_fft(c->plan[1]) ? 2 : 1)) * cs;
  // the 2 tx lines in one pass of the line FFTs (ny / e threads per line), 256 to 512 threads
  const int xythreads = std::min(512, std::max(256, 2 * cy.tx * (c->ng[1] / ct_plan(c->ng[1]).e)));
  if (rank1 && xylds > kLdsMax) return fail(MPOL_EINVAL, "y-pass columns do not fit in LDS");
#define LAUNCH_FFTY(E, N, INV)                                                                  \
  do {                                                                                          \
    set_lds(k_fft_y<T, E, N>, cy.lds);                                                          \
    if constexpr (NC == 3 && N > 0) {                                                           \
      if (rank1) {                                                                              \
        if (INV) {                                                                              \
          set_lds(k_fft_y_xy<T, E, N, true>, xylds);                                            \
          hipLaunchKernelGGL((k_fft_y_xy<T, E, N, true>), xygrid, dim3(xythreads), xylds, st, axy, sp, spec); \
        } else {                                                                                \
          set_lds(k_fft_y_xy<T, E, N, false>, xylds);                                           \
          hipLaunchKernelGGL((k_fft_y_xy<T, E, N, false>), xygrid, dim3(xythreads), xylds, st, axy, sp, spec); \
        }                                                                                       \
        break;                                                                                  \
      }                                                                                         \
    }                                                                                           \
    hipLaunchKernelGGL((k_fft_y<T, E, N>), ygrid, dim3(256), cy.lds, st, ay, spec, INV);       \
  } while (0)

  // z axis forward + scale + inverse
  const ColCfg cz = col_cfg(c->ng[2], generic_fft(c->plan[2]), c->kx, NC, cs, kLdsCol,
                            knobs<T>(c)[4] > 0 ? knobs<T>(c)[4] : (sizeof(T) == 8 ? 4 : 8));
  ColArgs az{};
  az.nx = c->ng[0]; az.ny = c->ng[1]; az.nz = c->ng[2]; az.kx = c->kx; az.kxs = c->kxs; az.tx = cz.tx; az.lstride = cz.lstride;
  az.plan = c->plan[2]; az.tw = tw_ptr<T>(c, 2);
  az.gate = gate;
  az.gdiv = rps;
  const dim3 zgrid(cdiv(c->kx, cz.tx), c->ng[1], rb);
  if (cz.lds > kLdsMax) return fail(MPOL_EINVAL, "z-pass columns do not fit in LDS");
  // one thread group per line of the z transform, all NC*tx lines at once
  int zthreads = 256;
  {
    const int e = is_ct_size(c->ng[2]) ? ct_plan(c->ng[2]).e : c->plan[2].e;
    if (e > 0) zthreads = std::min(1024, std::max(512, NC * cz.tx * (c->ng[2] / e)));
    zthreads = (zthreads + 63) / 64 * 64;
    if (knobs<T>(c)[5] > 0) zthreads = knobs<T>(c)[5];
  }
#define LAUNCH_FFTZ(E, N, ...)                                                                  \
  do {                                                                                          \
    set_lds(k_fft_z_scale<T, E, N, NC>, cz.lds);                                                \
    hipLaunchKernelGGL((k_fft_z_scale<T, E, N, NC>), zgrid, dim3(zthreads), cz.lds, st, az, sp, spec); \
  } while (0)
  for (int sb = 0; sb < nbatch; ++sb) {
    // this batch's spectra and system states (the kernels count right-hand
    // sides from the batch's first; the launch macros read spec)
    cplx<T>* const spec = spec_all + (size_t)sb * batch_stride;
    const int* const gate_b = gate != nullptr ? gate + (size_t)sb * (rb / rps) : nullptr;
    ay.gate = gate_b;
    axy.gate = gate_b;
    az.gate = gate_b;
    DISPATCH_FFT(c->ng[1], ay.plan.e, LAUNCH_FFTY, 0);
    if (ev && !by_sys) HIPCHK(hipEventRecord(ev[4], st));
    if (stop_after <= 3) return MPOL_OK;

    // nz = 256 dipole grids: transforms in registers (unless a z column width is forced)
    if (zreg) {
      ColArgs ar = az;
      ar.tx = TXR;
      ar.tail = tail_blocks(c->kx, TXR);
      if (rank1) {
        // rank-one form on the lines with non-Nyquist x and y wave numbers, the
        // full form on the last column and the row ky = ny/2 (eight lines per
        // workgroup) in the same launch
        const int nfull = (c->ng[1] % 2 == 0 ? c->kx - 1 : 0) + c->ng[1];
        const dim3 r1grid(ar.tail + 1, std::max(c->ng[1], cdiv(nfull, 8)), rb);
        if constexpr (sizeof(T) == 4)
          hipLaunchKernelGGL((k_fft_z_scale_reg256_rank1<T, TXR, false>), r1grid, dim3(256), 0, st, ar, sp, spec);
        else
          hipLaunchKernelGGL((k_fft_z_scale_reg256_rank1<T, TXR, true>), r1grid, dim3(256), 0, st, ar, sp, spec);
      } else {
        const dim3 rgrid(ar.tail > 0 ? ar.tail + 1 : cdiv(c->kx, TXR), c->ng[1], rb);
        if constexpr (sizeof(T) == 4)
          hipLaunchKernelGGL((k_fft_z_scale_reg256<T, TXR>), rgrid, dim3(TXR * 16), 0, st, ar, sp, spec);
        else
          hipLaunchKernelGGL((k_fft_z_scale_reg256_pair<T, TXR>), rgrid, dim3(TXR * 32), 0, st, ar, sp, spec);
      }
    } else {
      DISPATCH_FFT(c->ng[2], az.plan.e, LAUNCH_FFTZ, 0);
    }
    if (ev && !by_sys) HIPCHK(hipEventRecord(ev[5], st));
    if (stop_after <= 4) return MPOL_OK;

    DISPATCH_FFT(c->ng[1], ay.plan.e, LAUNCH_FFTY, 1);
  }
  if (ev && by_sys) {
    // system by system: the forward y stage carries the three passes
    HIPCHK(hipEventRecord(ev[4], st));
    HIPCHK(hipEventRecord(ev[5], st));
  }
  if (ev) HIPCHK(hipEventRecord(ev[6], st));
  if (stop_after <= 5) return MPOL_OK;

#define LAUNCH_CONTRACT(E, N, ...)                                                              \
  do {                                                                                          \
    set_lds(k_xifft_contract<T, E, N, NC>, tc.lds);                                             \
    hipLaunchKernelGGL((k_xifft_contract<T, E, N, NC>), tgrid, dim3(256), tc.lds, st, ta, spec, out); \
  } while (0)
#define LAUNCH_CONTRACT_MFMA(N, ...)                                                            \
  do {                                                                                          \
    if (tm.rows == 16) {                                                                        \
      launch_mfma16<T, N, NC>(false, mgrid, tm.lds, st, tma, ima, spec);                        \
    } else if (tm.rows == 8) {                                                                  \
      set_lds(k_xifft_contract_mfma<T, N, NC, 8>, tm.lds);                                      \
      hipLaunchKernelGGL((k_xifft_contract_mfma<T, N, NC, 8>), mgrid, dim3(256), tm.lds, st, tma, ima, spec); \
    } else {                                                                                    \
      set_lds(k_xifft_contract_mfma<T, N, NC, 4>, tm.lds);                                      \
      hipLaunchKernelGGL((k_xifft_contract_mfma<T, N, NC, 4>), mgrid, dim3(256), tm.lds, st, tma, ima, spec); \
    }                                                                                           \
  } while (0)
  if (use_mfma) {
    DISPATCH_CT(c->ng[0], LAUNCH_CONTRACT_MFMA, 0);
    if (ev) HIPCHK(hipEventRecord(ev[7], st));
    const TileImage& im = c->img[sizeof(T) == 8 ? 1 : 0];
    hipLaunchKernelGGL((k_contract_reduce<T, NC>), dim3(cdiv((size_t)n * 16, 256), nrhs), dim3(256), 0, st, n, ima.total,
                       im.poff.p, im.pslot.p, c->perm.p, reinterpret_cast<const T*>(ima.part), ta.hvol, out, gate,
                       rps);
  } else {
    DISPATCH_FFT(c->ng[0], ta.plan.e, LAUNCH_CONTRACT, 0);
    if (ev) HIPCHK(hipEventRecord(ev[7], st));
  }
#undef LAUNCH_CONTRACT_MFMA
  if (ev) HIPCHK(hipEventRecord(ev[8], st));
#undef LAUNCH_FFTY
#undef LAUNCH_FFTZ
#undef LAUNCH_CONTRACT
  HIPCHK(hipGetLastError());
  // complex permittivities: the reference's wave-space operator, applied to
  // the complex moments p = a + i b, adds i B p to the Hermitian part's
  // result (B real, nonzero on the Nyquist planes only), so the real output
  // gains -B b and the imaginary output +B a (knob 9 = 2 keeps the Hermitian
  // part only)
  if (c->cplx_eps && !odd && stop_after >= kStages && knobs<T>(c)[9] != 2) {
    HIPALLOC(c->odd_tmp, (size_t)nrhs * NC * n);
    c->odd_pass = true;
    const int s = run_operator<T, NC>(c, m, c->odd_tmp.p, nullptr, nullptr, nrhs, 0.0, nullptr, kStages);
    c->odd_pass = false;
    CHECK_STATUS(s);
    hipLaunchKernelGGL(k_add_odd_part, dim3(cdiv((size_t)NC * n, 256), nrhs / 2), dim3(256), 0, st,
                       (int64_t)NC * n, c->odd_tmp.p, out, gate);
    HIPCHK(hipGetLastError());
  }
  return MPOL_OK;
}

// Right-hand sides per system: 2 (real, imaginary part) with complex permittivities.
inline int rhs_per_system(const mpol_ctx* c) { return c->cplx_eps ? 2 : 1; }

// Node-major spread structure of the polydisperse operator (index once per
// geometry and radii, entries once per precision).
template <typename T>
int ensure_sizes_nodes(mpol_ctx* c, const SizesGridArgs& sa) {
  hipStream_t st = c->stream;
  const int nodes = sa.nx * sa.ny * sa.nz;
  if (!c->sz_nodes_valid) {
    HIPALLOC(c->sz_ncnt, (size_t)nodes);
    HIPALLOC(c->sz_noff, (size_t)nodes + 1);
    HIPCHK(hipMemsetAsync(c->sz_ncnt.p, 0, (size_t)nodes * sizeof(int), st));
    hipLaunchKernelGGL(k_sizes_node_count, dim3(sa.n), dim3(256), 0, st, sa, c->sz_ncnt.p);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, c->sz_ncnt.p, c->sz_noff.p, nodes);
    int total = 0;
    HIPCHK(hipMemcpyAsync(&total, c->sz_noff.p + nodes, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    HIPALLOC(c->sz_nidx, (size_t)std::max(1, total));
    HIPCHK(hipMemsetAsync(c->sz_ncnt.p, 0, (size_t)nodes * sizeof(int), st));
    hipLaunchKernelGGL(k_sizes_node_fill, dim3(sa.n), dim3(256), 0, st, sa, c->sz_noff.p, c->sz_ncnt.p,
                       c->sz_nidx.p);
    hipLaunchKernelGGL(k_sizes_node_sort, dim3(cdiv(nodes, 256)), dim3(256), 0, st, nodes, c->sz_noff.p,
                       c->sz_nidx.p);
    HIPCHK(hipGetLastError());
    c->sz_entries = total;
    c->sz_nodes_valid = true;
    c->sz_nw_valid[0] = c->sz_nw_valid[1] = false;
  }
  const int pi = sizeof(T) == 8 ? 1 : 0;
  if (!c->sz_nw_valid[pi]) {
    typename Vec4<T>::type* ent;
    typename Vec4<T>::type* pent;
    const size_t pcount = (size_t)sa.n * sa.P * sa.P * sa.P;
    if constexpr (sizeof(T) == 8) {
      HIPALLOC(c->sz_nw64, (size_t)std::max(1, c->sz_entries));
      HIPALLOC(c->sz_pw64, pcount);
      ent = c->sz_nw64.p;
      pent = c->sz_pw64.p;
    } else {
      HIPALLOC(c->sz_nw32, (size_t)std::max(1, c->sz_entries));
      HIPALLOC(c->sz_pw32, pcount);
      ent = c->sz_nw32.p;
      pent = c->sz_pw32.p;
    }
    hipLaunchKernelGGL((k_sizes_node_weights<T>), dim3(cdiv(nodes, 256)), dim3(256), 0, st, sa, c->sz_noff.p,
                       c->sz_nidx.p, ent);
    hipLaunchKernelGGL((k_sizes_particle_weights<T>), dim3(sa.n), dim3(256), 0, st, sa, pent);
    HIPCHK(hipGetLastError());
    c->sz_nw_valid[pi] = true;
  }
  return MPOL_OK;
}

// Polydisperse operator: real space (pair couplings of the two radii,
// dielectric 3/(4 pi a^3 (eps-1))) + one scalar grid per right-hand side.
// hr != nullptr: hr <- real space, out <- wave space.
constexpr size_t kSizesContractLds = 64 * 1024;  // stencil staging of the polydisperse contraction

template <typename T>
int run_operator_sizes(mpol_ctx* c, const double* m, double* out, double* hr, int nrhs) {
  const int n = c->n;
  hipStream_t st = c->stream;
  const size_t G = (size_t)c->ng[0] * c->ng[1] * c->ng[2];
  const size_t cs = sizeof(cplx<T>);
  HIPALLOC(c->spec, (size_t)nrhs * c->ng[2] * c->ng[1] * c->kxs * cs);
  HIPALLOC(c->sz_grid, (size_t)nrhs * G * sizeof(T));
  HIPALLOC(c->zsum, (size_t)nrhs * (kZsumBlocks + 1));
  HIPALLOC(c->ms, (size_t)nrhs * 3 * n);
  cplx<T>* spec = reinterpret_cast<cplx<T>*>(c->spec.p);
  T* grid = reinterpret_cast<T*>(c->sz_grid.p);
  const int* gate = c->gate;
  const int rps = c->cplx_eps ? 2 : 1;
  // real space (no slab term: the real-space kernel does not read zsum)
  const T* pA;
  const T* pBA;
  const T* prh;
  if constexpr (sizeof(T) == 8) {
    pA = c->sz_A.p; pBA = c->sz_BA.p; prh = c->rs_rh.p;
  } else {
    pA = c->sz_A32.p; pBA = c->sz_BA32.p; prh = c->rs_rh32.p;
  }
  if (hr) HIPALLOC(c->sz_slab, (size_t)nrhs * 3 * n);
  hipLaunchKernelGGL((k_real_space<T>), dim3(cdiv((size_t)n * 16, 256), nrhs), dim3(256), 0, st, n, nrhs, c->nsets, m, out,
                     c->sz_diag.p, c->cplx_eps ? c->sz_diag_im.p : nullptr, c->rs_off.p, c->rs_j.p, pA, pBA, prh,
                     c->zsum.p, 0.0, hr != nullptr ? 1 : 0, hr, hr ? c->sz_slab.p : nullptr, gate, rps);
  // spread into real grids
  SizesGridArgs sa{};
  sa.gate = gate;
  sa.gdiv = rps;
  sa.nx = c->ng[0]; sa.ny = c->ng[1]; sa.nz = c->ng[2];
  sa.P = c->P; sa.half = c->half; sa.n = n; sa.nrhs = nrhs;
  for (int d = 0; d < 3; ++d) sa.h[d] = c->h[d];
  sa.xi = c->xi;
  sa.eta = c->eta[0];
  sa.hvol = c->h[0] * c->h[1] * c->h[2];
  sa.rec = c->rec.p;
  sa.base = c->base64.p;
  sa.rad = c->sz_rad_sorted.p;
  const int P3 = c->P * c->P * c->P;
  // node-major spread and particle-major contraction from stored entries
  // (2 x 16 B per stencil point in FP32, 2 x 32 B in FP64)
  const bool by_entries = knobs<T>(c)[6] != 2 && (double)n * P3 <= 4.0e8 && (double)G < 2.0e9;
  if (by_entries) {
    // gather by node: every live right-hand side's grid is written whole
    CHECK_STATUS(ensure_sizes_nodes<T>(c, sa));
    using V = typename Vec4<T>::type;
    HIPALLOC(c->sz_pm, (size_t)nrhs * n * sizeof(V));
    V* pm = reinterpret_cast<V*>(c->sz_pm.p);
    const V* ent;
    if constexpr (sizeof(T) == 8) ent = c->sz_nw64.p;
    else ent = c->sz_nw32.p;
    hipLaunchKernelGGL((k_sizes_permute_pm<T>), dim3(cdiv((size_t)n * nrhs, 256)), dim3(256), 0, st, m, c->perm.p,
                       n, nrhs, pm, gate, rps);
    // one right-hand side per lane: more waves in flight hide the entry ->
    // dipole load chain better than several right-hand sides per lane
#define LAUNCH_SPREAD_NODES(W)                                                                         \
  hipLaunchKernelGGL((k_sizes_spread_nodes<T, W, 1>), dim3(cdiv((int64_t)G * W, 256), cdiv(nrhs, W)), dim3(256), 0, \
                     st, sa, c->sz_noff.p, ent, pm, grid)
    if (nrhs <= 4) LAUNCH_SPREAD_NODES(4);
    else if (nrhs <= 8) LAUNCH_SPREAD_NODES(8);
    else
      hipLaunchKernelGGL((k_sizes_spread_rows<T>), dim3(cdiv((int64_t)G * 16, 256), cdiv(nrhs, 16)), dim3(256), 0, st,
                         sa, c->sz_noff.p, ent, pm, grid);
#undef LAUNCH_SPREAD_NODES
  } else {
    // per-particle scatter with atomics (engine knob 2, or more stencil
    // points than the entries are kept for)
    hipLaunchKernelGGL(k_permute_m<3>, dim3(cdiv(n, 256), nrhs), dim3(256), 0, st, m, c->perm.p, n, c->ms.p, gate,
                       rps);
    HIPCHK(hipMemsetAsync(grid, 0, (size_t)nrhs * G * sizeof(T), st));
    hipLaunchKernelGGL((k_sizes_spread<T>), dim3(n), dim3(256), 0, st, sa, c->ms.p, grid);
  }
  // x transform of the real rows
  const bool gen_x = generic_fft(c->plan[0]);
  TileCfg tr = tile_cfg_nc<1>(c->ng[0], gen_x, cs, 0, kLdsTile);
  TileArgs ta{};
  ta.nx = c->ng[0]; ta.ny = c->ng[1]; ta.nz = c->ng[2]; ta.kx = c->kx; ta.kxs = c->kxs;
  ta.ty = tr.ty; ta.tz = tr.tz; ta.rows = tr.rows; ta.ncl = tr.ncl; ta.lstride = tr.lstride;
  ta.tw = tw_ptr<T>(c, 0); ta.plan = c->plan[0];
  ta.gate = gate;
  ta.gdiv = rps;
  const size_t rlds = ((size_t)c->ng[0] + (size_t)tr.ncl * tr.lstride * (gen_x ? 2 : 1)) * cs;
  if (rlds > kLdsMax) return fail(MPOL_EINVAL, "row transform tile does not fit in LDS");
  const dim3 tgrid(cdiv(c->ng[1], tr.ty), cdiv(c->ng[2], tr.tz), nrhs);
#define LAUNCH_R2C(E, N, ...)                                                                   \
  do {                                                                                          \
    set_lds(k_rows_r2c<T, E, N>, rlds);                                                         \
    hipLaunchKernelGGL((k_rows_r2c<T, E, N>), tgrid, dim3(256), rlds, st, ta, grid, spec);      \
  } while (0)
  DISPATCH_FFT(c->ng[0], ta.plan.e, LAUNCH_R2C, 0);
#undef LAUNCH_R2C
  // y forward, z forward + scale + inverse, y inverse
  const ColCfg cy = col_cfg(c->ng[1], generic_fft(c->plan[1]), c->kx, 1, cs, kLdsCol,
                            knobs<T>(c)[3] > 0 ? knobs<T>(c)[3] : (sizeof(T) == 8 ? 8 : 16));
  ColArgs ay{};
  ay.nx = c->ng[0]; ay.ny = c->ng[1]; ay.nz = c->ng[2]; ay.kx = c->kx; ay.kxs = c->kxs; ay.tx = cy.tx; ay.lstride = cy.lstride;
  ay.plan = c->plan[1]; ay.tw = tw_ptr<T>(c, 1);
  ay.gate = gate;
  ay.gdiv = rps;
  const dim3 ygrid(cdiv(c->kx, cy.tx), c->ng[2], nrhs);
  if (cy.lds > kLdsMax) return fail(MPOL_EINVAL, "y-pass columns do not fit in LDS");
#define LAUNCH_FFTY(E, N, INV)                                                                  \
  do {                                                                                          \
    set_lds(k_fft_y<T, E, N>, cy.lds);                                                          \
    hipLaunchKernelGGL((k_fft_y<T, E, N>), ygrid, dim3(256), cy.lds, st, ay, spec, INV);       \
  } while (0)
  const ColCfg cz = col_cfg(c->ng[2], generic_fft(c->plan[2]), c->kx, 1, cs, kLdsCol,
                            knobs<T>(c)[4] > 0 ? knobs<T>(c)[4] : (sizeof(T) == 8 ? 8 : 16));
  ColArgs az{};
  az.nx = c->ng[0]; az.ny = c->ng[1]; az.nz = c->ng[2]; az.kx = c->kx; az.kxs = c->kxs; az.tx = cz.tx; az.lstride = cz.lstride;
  az.plan = c->plan[2]; az.tw = tw_ptr<T>(c, 2);
  az.gate = gate;
  az.gdiv = rps;
  ScaleParams sp{};
  for (int d = 0; d < 3; ++d) {
    sp.twopi_over_l[d] = 2.0 * kPi / c->box[d];
    sp.one_minus_eta[d] = 1.0 - c->eta[d];
  }
  sp.inv_4xi2 = 1.0 / (4.0 * c->xi * c->xi);
  sp.inv_g = 1.0 / (double)G;
  sp.ex = c->ktab.p;
  sp.ey = sp.ex + c->ng[0];
  sp.ez = sp.ey + c->ng[1];
  const dim3 zgrid(cdiv(c->kx, cz.tx), c->ng[1], nrhs);
  if (cz.lds > kLdsMax) return fail(MPOL_EINVAL, "z-pass columns do not fit in LDS");
  int zthreads = 256;
  {
    const int e = is_ct_size(c->ng[2]) ? ct_plan(c->ng[2]).e : c->plan[2].e;
    if (e > 0) zthreads = std::min(1024, std::max(256, cz.tx * (c->ng[2] / e)));
    zthreads = (zthreads + 63) / 64 * 64;
  }
#define LAUNCH_FFTZ(E, N, ...)                                                                  \
  do {                                                                                          \
    set_lds(k_fft_z_scale<T, E, N, 1>, cz.lds);                                                 \
    hipLaunchKernelGGL((k_fft_z_scale<T, E, N, 1>), zgrid, dim3(zthreads), cz.lds, st, az, sp, spec); \
  } while (0)
  // ny == nz == 32: forward y, forward z + scale + inverse z and inverse y
  // in one kernel per group of y-z planes (engine knob 2: the separate passes)
  constexpr int kPlaneTx = 2;
  if (c->ng[1] == 32 && c->ng[2] == 32 && knobs<T>(c)[6] != 2) {
    const size_t plds = (32 + 2 * (size_t)kPlaneTx * 32 * padded_len(32)) * cs;
    // one thread per line pair position: kPlaneTx * 32 lines of 32 / ct_plan(32).e threads
    constexpr int threads = kPlaneTx * 32 * (32 / ct_plan(32).e);
    set_lds(k_sizes_yz_plane<T, 32, kPlaneTx>, plds);
    hipLaunchKernelGGL((k_sizes_yz_plane<T, 32, kPlaneTx>), dim3(cdiv(c->kx, kPlaneTx), 1, nrhs), dim3(threads), plds,
                       st, ay, sp, spec);
  } else {
    DISPATCH_FFT(c->ng[1], ay.plan.e, LAUNCH_FFTY, 0);
    DISPATCH_FFT(c->ng[2], az.plan.e, LAUNCH_FFTZ, 0);
    DISPATCH_FFT(c->ng[1], ay.plan.e, LAUNCH_FFTY, 1);
  }
#undef LAUNCH_FFTZ
#undef LAUNCH_FFTY
#define LAUNCH_C2R(E, N, ...)                                                                   \
  do {                                                                                          \
    set_lds(k_rows_c2r<T, E, N>, rlds);                                                         \
    hipLaunchKernelGGL((k_rows_c2r<T, E, N>), tgrid, dim3(256), rlds, st, ta, spec, grid);      \
  } while (0)
  DISPATCH_FFT(c->ng[0], ta.plan.e, LAUNCH_C2R, 0);
#undef LAUNCH_C2R
  const size_t clds = (size_t)c->P * c->P * c->P * (3 * sizeof(T) + sizeof(int));
  if (by_entries) {
    const int ldg = (nrhs + 3) / 4 * 4;  // aligned 4-vector reads of a node's right-hand sides
    HIPALLOC(c->sz_gridT, (size_t)ldg * G * sizeof(T));
    T* gT = reinterpret_cast<T*>(c->sz_gridT.p);
    hipLaunchKernelGGL((k_sizes_grid_transpose<T>), dim3(cdiv(G, 64), cdiv(nrhs, 16)), dim3(256), 0, st, grid, (int)G,
                       nrhs, ldg, gT);
    using V = typename Vec4<T>::type;
    const V* pent;
    if constexpr (sizeof(T) == 8) pent = c->sz_pw64.p;
    else pent = c->sz_pw32.p;
    // a wave: 64 / (K W) particles x K stencil splits x W lanes of 4 right-hand sides
#define LAUNCH_CONTRACT_NODES(W, K)                                                                        \
  hipLaunchKernelGGL((k_sizes_contract_nodes<T, W, K>), dim3(cdiv(cdiv(n, 64 / ((K) * (W))) * (int64_t)64, 256),  \
                     cdiv(nrhs, 4 * (W))), dim3(256), 0, st, sa, pent, gT, ldg, out)
    if (nrhs <= 4) LAUNCH_CONTRACT_NODES(1, 32);
    else if (nrhs <= 8) LAUNCH_CONTRACT_NODES(2, 16);
    else LAUNCH_CONTRACT_NODES(4, 8);
#undef LAUNCH_CONTRACT_NODES
  } else if (clds <= kSizesContractLds) {
    set_lds(k_sizes_contract_lds<T>, clds);
    hipLaunchKernelGGL((k_sizes_contract_lds<T>), dim3(n), dim3(256), clds, st, sa, grid, out);
  } else {
    hipLaunchKernelGGL((k_sizes_contract<T>), dim3(n), dim3(256), 0, st, sa, grid, out);
  }
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

extern template int run_operator<float, kOpDipole>(mpol_ctx*, const double*, double*, double*, double*, int, double,
                                                   hipEvent_t*, int);
extern template int run_operator<double, kOpDipole>(mpol_ctx*, const double*, double*, double*, double*, int, double,
                                                    hipEvent_t*, int);
extern template int run_operator<float, kOpQuad>(mpol_ctx*, const double*, double*, double*, double*, int, double,
                                                 hipEvent_t*, int);
extern template int run_operator<double, kOpQuad>(mpol_ctx*, const double*, double*, double*, double*, int, double,
                                                  hipEvent_t*, int);
extern template int run_operator_sizes<float>(mpol_ctx*, const double*, double*, double*, int);
extern template int run_operator_sizes<double>(mpol_ctx*, const double*, double*, double*, int);

}  // namespace mpol

