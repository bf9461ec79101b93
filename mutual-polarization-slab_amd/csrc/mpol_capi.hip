// C ABI of the MI355X spectral-Ewald mutual-polarization library.
//
// Host side of libmpol.so: context and workspace management, geometry
// precompute, the device-resident restarted GMRES driver and every extern "C"
// entry point.  See include/mpol.h for the contract of every entry point.  The
// operator launch sequences are instantiated in mpol_ops_*.hip.
#include "mpol_internal.h"

#include <map>
#include <mutex>
#include <utility>

namespace mpol {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

void raise_lds_limit(const void* kernel, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> limit;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  size_t& cur = limit[{dev, kernel}];
  if (bytes <= cur) return;
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess) cur = bytes;
}
}  // namespace mpol

namespace {

// Scratch of mpol_compute_forces_noalloc, one copy per GPU (held until the
// process ends): cell index and cell members of at most kNoAllocMaxN
// particles, then the per-cell counts, placement cursors and starts of at
// most kNoAllocMaxCells cells -- 2 x 1 MiB + 3 x 8 MiB of int32.  Every call
// on a GPU shares its copy, so calls that overlap on one GPU race.
constexpr int64_t kNoAllocMaxN = 1 << 18;
constexpr int64_t kNoAllocMaxCells = 8 * kNoAllocMaxN;  // the cell-list rule allows at most max(1024, 8 N)
constexpr int64_t kScratchCell = 0;
constexpr int64_t kScratchMembers = kScratchCell + kNoAllocMaxN;
constexpr int64_t kScratchCount = kScratchMembers + kNoAllocMaxN;
constexpr int64_t kScratchCursor = kScratchCount + kNoAllocMaxCells;
constexpr int64_t kScratchStart = kScratchCursor + kNoAllocMaxCells;
constexpr int64_t kScratchInts = kScratchStart + kNoAllocMaxCells + 1;
static __device__ __align__(256) int scratch[kScratchInts];

// Device address of `scratch` on each GPU (the host-side name is only a
// handle for the runtime), looked up once per device with it current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_dev[kMaxDevices] = {};

int scratch_address(int device, int** out) {
  if (device < 0 || device >= kMaxDevices) return fail(MPOL_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = static_cast<int*>(ptr);
  }
  *out = g_scratch_dev[device];
  return MPOL_OK;
}

// Operator `kind` (kOpDipole / kOpQuad) on nsys systems.
int apply(mpol_ctx* c, int kind, const double* m, double* out, int nsys, int precision, double slab) {
  const int nrhs = nsys * rhs_per_system(c);
  if (kind == kOpSizes) {
    if (precision == MPOL_FP32) return run_operator_sizes<float>(c, m, out, nullptr, nrhs);
    return run_operator_sizes<double>(c, m, out, nullptr, nrhs);
  }
  if (kind == kOpQuad) {
    if (precision == MPOL_FP32) return run_operator<float, kOpQuad>(c, m, out, nullptr, nullptr, nrhs, 0.0);
    return run_operator<double, kOpQuad>(c, m, out, nullptr, nullptr, nrhs, 0.0);
  }
  if (precision == MPOL_FP32) return run_operator<float, kOpDipole>(c, m, out, nullptr, nullptr, nrhs, slab);
  return run_operator<double, kOpDipole>(c, m, out, nullptr, nullptr, nrhs, slab);
}

int check_ready(mpol_ctx* c) {
  if (!c) return fail(MPOL_EINVAL, "null context");
  if (!c->have_geo) return fail(MPOL_ESTATE, "geometry not set (call mpol_set_geometry first)");
  if (c->nsets == 0) return fail(MPOL_ESTATE, "permittivity not set (call mpol_set_permittivity first)");
  return MPOL_OK;
}

// Partial-sum blocks per right-hand side of the reductions over vectors of
// n elements: at least 256 elements per block (one per thread) and about
// 1024 blocks in all (at most kDotBlocks per right-hand side), so that short
// batched vectors still keep enough loads in flight.
int dot_blocks(int64_t n, int nrhs) {
  const int64_t by_total = (1024 + nrhs - 1) / nrhs, by_len = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>({by_total, by_len, (int64_t)kDotBlocks}));
}

// squared norms / dot products of `cnt` vectors A_i (stride stride_i) with w, per rhs -> out[rhs*cnt+i]
int dots(mpol_ctx* c, const double* A, size_t stride_i, int cnt, const double* w, int64_t n, int nrhs, double* out) {
  const int nb = dot_blocks(n, nrhs);
  hipLaunchKernelGGL(k_dots, dim3(nb, nrhs), dim3(256), 0, c->stream, A, stride_i, cnt, w, n, nrhs, c->gpart.p);
  hipLaunchKernelGGL(k_dots_finish, dim3(cdiv((int64_t)nrhs * cnt, 4)), dim3(256), 0, c->stream, c->gpart.p, nb,
                     nrhs * cnt, out);
  return MPOL_OK;
}

// Copy cnt doubles from the device to host memory (through a page-locked
// buffer) and wait for them.
int read_doubles(mpol_ctx* c, const double* dev, int cnt, double* host) {
  if (c->hscr.reserve((size_t)cnt) != hipSuccess) return fail(MPOL_ENOMEM, "page-locked host allocation failed");
  HIPCHK(hipMemcpyAsync(c->hscr.p, dev, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::copy(c->hscr.p, c->hscr.p + cnt, host);
  return MPOL_OK;
}

// Sets the operator gate of a context for one scope.
struct GateScope {
  mpol_ctx* c;
  GateScope(mpol_ctx* ctx, const int* state) : c(ctx) { c->gate = state; }
  ~GateScope() { c->gate = nullptr; }
};

// Slot layout of the page-locked state the GMRES kernels post: slots 0 and 1
// alternate between Arnoldi steps (the host reads step j - 1 while step j
// runs), slot 2 holds the cycle start.
constexpr int kCycleSlot = 2;

int ensure_step_reads(mpol_ctx* c, int nrhs) {
  const int* had = c->hstep.p;
  const double* hadrel = c->hsteprel.p;
  if (c->hstep.reserve((size_t)3 * nrhs) != hipSuccess || c->hsteprel.reserve((size_t)3 * nrhs) != hipSuccess)
    return fail(MPOL_ENOMEM, "page-locked host allocation failed");
  if (c->hstep.p != had || !c->dstep)
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->dstep), c->hstep.p, 0));
  if (c->hsteprel.p != hadrel || !c->dsteprel)
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->dsteprel), c->hsteprel.p, 0));
  for (hipEvent_t& e : c->step_ev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return MPOL_OK;
}

// Restarted GMRES on x (in: initial guess, out: solution) for A x = b.
// nrhs here counts systems; a system's vector holds all its right-hand sides
// (kind components per particle, times 2 with complex permittivities).
// Krylov dimension of the FP32 correction solves of the mixed scheme (at
// least the caller's restart): one cycle usually reaches the inner tolerance.
constexpr int kInnerRestart = 32;
// Smallest relative tolerance of a correction solve, and its restart cycles
// (a correction that stalls returns to the FP64 residual, which decides).
constexpr double kInnerFloor = 5e-7;
constexpr int kInnerCycles = 4;

// Restarted GMRES.  inner = true is the correction solve of the mixed scheme:
// x starts at zero (the first residual is b, no operator application) and
// the solve ends on the Givens estimate of the residual without a final
// true-residual application (the caller's FP64 residual certifies it).
int gmres_core(mpol_ctx* c, int kind, const double* b, double* x, int nrhs, double tol, int restart, int maxit,
               int precision, double slab, int64_t* nmv, double* relres_host, int* converged_host,
               bool inner = false) {
  const int64_t n3 = (int64_t)unknowns_per_particle(kind) * c->n * rhs_per_system(c);
  const size_t tot = (size_t)nrhs * n3;
  const size_t sv = tot;  // stride between basis vectors
  HIPALLOC(c->gV, (size_t)(restart + 1) * tot);
  HIPALLOC(c->gW, tot);
  HIPALLOC(c->gR, tot);
  HIPALLOC(c->gAx, tot);
  HIPALLOC(c->gH, (size_t)nrhs * (restart + 1) * restart);
  HIPALLOC(c->gcs, (size_t)nrhs * restart);
  HIPALLOC(c->gsn, (size_t)nrhs * restart);
  HIPALLOC(c->gg, (size_t)nrhs * (restart + 1));
  HIPALLOC(c->gy, (size_t)nrhs * restart);
  HIPALLOC(c->gh1, (size_t)nrhs * (restart + 1));
  HIPALLOC(c->gh2, (size_t)nrhs * (restart + 1));
  HIPALLOC(c->gpart, (size_t)nrhs * (restart + 1) * kDotBlocks);
  HIPALLOC(c->gpart2, (size_t)nrhs * (restart + 1) * kDotBlocks);
  HIPALLOC(c->gbn, (size_t)nrhs);
  HIPALLOC(c->gbeta, (size_t)nrhs);
  HIPALLOC(c->grel, (size_t)nrhs);
  HIPALLOC(c->grn2, (size_t)nrhs);
  HIPALLOC(c->gstate, (size_t)nrhs);
  HIPALLOC(c->gkdim, (size_t)nrhs);
  hipStream_t st = c->stream;
  const dim3 vgrid(cdiv(n3, 256), nrhs);
  const int sgrid = cdiv(nrhs, 256);

  CHECK_STATUS(ensure_step_reads(c, nrhs));
  // ||b|| and the first states stay on the device
  dots(c, b, 0, 1, b, n3, nrhs, c->grn2.p);
  hipLaunchKernelGGL(k_norms_init, dim3(sgrid), dim3(256), 0, st, nrhs, c->grn2.p, c->gbn.p, c->gstate.p);
  std::vector<int> state(nrhs, kWait);
  std::vector<double> rel(nrhs, 0.0);
  // wait for a slot the kernels posted and take its states and residuals
  auto take_slot = [&](int slot) -> int {
    HIPCHK(hipEventSynchronize(c->step_ev[slot]));
    std::copy(c->hstep.p + (size_t)slot * nrhs, c->hstep.p + (size_t)(slot + 1) * nrhs, state.begin());
    std::copy(c->hsteprel.p + (size_t)slot * nrhs, c->hsteprel.p + (size_t)(slot + 1) * nrhs, rel.begin());
    return MPOL_OK;
  };
  // a cycle that ended on its residual estimate usually ends the solve at
  // the next cycle start, so that start is read before anything else is
  // enqueued; any other cycle start is read while the first Arnoldi step of
  // the new cycle runs
  bool likely_done = false;
  bool all_done = false;
  for (int outer = 0; outer <= maxit; ++outer) {
    // true residual at the start of each cycle (and once after the last)
    if (inner && outer == 0) {
      hipLaunchKernelGGL(k_copy, dim3(cdiv((int64_t)tot, 256)), dim3(256), 0, st, b, c->gR.p, (int64_t)tot);
    } else {
      CHECK_STATUS(apply(c, kind, x, c->gAx.p, nrhs, precision, slab));
      ++*nmv;
      hipLaunchKernelGGL(k_residual, dim3(cdiv((int64_t)tot, 256)), dim3(256), 0, st, b, c->gAx.p, c->gR.p,
                         (int64_t)tot);
    }
    dots(c, c->gR.p, 0, 1, c->gR.p, n3, nrhs, c->grn2.p);
    hipLaunchKernelGGL(k_cycle_start, dim3(sgrid), dim3(256), 0, st, nrhs, restart, c->grn2.p, c->gbn.p, tol,
                       c->gstate.p, c->gg.p, c->gbeta.p, c->grel.p, c->dstep + (size_t)kCycleSlot * nrhs,
                       c->dsteprel + (size_t)kCycleSlot * nrhs);
    HIPCHK(hipEventRecord(c->step_ev[kCycleSlot], st));
    bool cycle_read = false;
    if (likely_done || outer == maxit) {
      CHECK_STATUS(take_slot(kCycleSlot));
      cycle_read = true;
      all_done = std::all_of(state.begin(), state.end(), [](int s) { return s == kDone; });
      if (all_done || outer == maxit) break;
    }
    hipLaunchKernelGGL(k_scale_into, vgrid, dim3(256), 0, st, c->gR.p, c->gbeta.p, c->gV.p, n3, c->gstate.p);
    // The host stays one Arnoldi step ahead of the state it reads: step j is
    // enqueued before the state after step j - 1 (or the cycle start, for
    // j = 0) is waited for.  Every kernel of a step is gated per system
    // (kIter), so a step enqueued for a system whose cycle had just ended
    // does nothing for it; the x update acts on kSolve systems only and
    // moves them to kWait, so enqueuing it again is harmless.
    bool stop = false;
    bool updated = false;  // the x update of this cycle's last step is enqueued
    auto take_state = [&](int jj) -> int {
      CHECK_STATUS(take_slot(jj & 1));
      if (!updated && std::any_of(state.begin(), state.end(), [](int s) { return s == kSolve; })) {
        // a cycle ended for some right-hand sides: x += V y, then they wait for the next cycle
        hipLaunchKernelGGL(k_update_x, vgrid, dim3(256), 0, st, c->gV.p, sv, c->gy.p, c->gkdim.p, c->gstate.p,
                           restart, x, n3);
        hipLaunchKernelGGL(k_solve_to_wait, dim3(sgrid), dim3(256), 0, st, nrhs, c->gstate.p);
      }
      stop = std::none_of(state.begin(), state.end(), [](int s) { return s == kIter; });
      return MPOL_OK;
    };
    int pending = -1;  // step whose state is in flight
    for (int j = 0; j < restart; ++j) {
      double* vj = c->gV.p + (size_t)j * sv;
      {
        // systems whose cycle has ended (or that converged) skip the operator
        GateScope gs(c, c->gstate.p);
        CHECK_STATUS(apply(c, kind, vj, c->gW.p, nrhs, precision, slab));
      }
      ++*nmv;
      // classical Gram-Schmidt, twice; each subtraction pass finishes its
      // coefficients from the previous launch's partial sums and forms the
      // partial sums of the next reduction (second-pass coefficients into
      // gpart2, ||w||^2 into gpart); the last reduction's finish runs with
      // the Arnoldi step
      const int cj = j + 1;
      const int nb = dot_blocks(n3, nrhs);
      const dim3 dgrid(nb, nrhs);
      hipLaunchKernelGGL(k_dots, dgrid, dim3(256), 0, st, c->gV.p, sv, cj, c->gW.p, n3, nrhs, c->gpart.p);
      hipLaunchKernelGGL(k_sub_dots, dgrid, dim3(256), 0, st, c->gV.p, sv, cj, c->gpart.p, c->gh1.p, c->gW.p, n3,
                         c->gstate.p, 0, c->gpart2.p);
      hipLaunchKernelGGL(k_sub_dots, dgrid, dim3(256), 0, st, c->gV.p, sv, cj, c->gpart2.p, c->gh2.p, c->gW.p, n3,
                         c->gstate.p, 1, c->gpart.p);
      hipLaunchKernelGGL(k_finish_arnoldi, dim3(nrhs), dim3(64), 0, st, c->gpart.p, nb, restart, j, c->gh1.p,
                         c->gh2.p, c->grn2.p, c->gH.p, c->gcs.p, c->gsn.p, c->gg.p, c->gbn.p, tol, c->gstate.p,
                         c->gkdim.p, c->grel.p, c->gy.p, c->dstep + (size_t)(j & 1) * nrhs,
                         c->dsteprel + (size_t)(j & 1) * nrhs);
      hipLaunchKernelGGL(k_normalize_next, vgrid, dim3(256), 0, st, c->gW.p, c->gV.p + (size_t)(j + 1) * sv,
                         c->grn2.p, n3, c->gstate.p);
      HIPCHK(hipEventRecord(c->step_ev[j & 1], st));
      if (!cycle_read) {
        // the cycle start, read while step 0 runs
        CHECK_STATUS(take_slot(kCycleSlot));
        cycle_read = true;
        all_done = std::all_of(state.begin(), state.end(), [](int s) { return s == kDone; });
        if (all_done) {
          --*nmv;  // step 0 had no iterating system
          break;
        }
      } else if (pending >= 0) {
        CHECK_STATUS(take_state(pending));
        if (stop) {
          --*nmv;  // step j had no iterating system left
          pending = -1;
          break;
        }
      }
      pending = j;
    }
    if (all_done) break;
    if (pending == restart - 1) {
      // the last step ends the cycle of every system still iterating: the
      // x update goes ahead of the read
      hipLaunchKernelGGL(k_update_x, vgrid, dim3(256), 0, st, c->gV.p, sv, c->gy.p, c->gkdim.p, c->gstate.p,
                         restart, x, n3);
      hipLaunchKernelGGL(k_solve_to_wait, dim3(sgrid), dim3(256), 0, st, nrhs, c->gstate.p);
      updated = true;
    }
    if (pending >= 0) CHECK_STATUS(take_state(pending));
    // the cycle ended on its estimate for every system still iterating
    likely_done = true;
    for (int r = 0; r < nrhs; ++r)
      if (state[r] != kDone && !(rel[r] <= tol)) likely_done = false;
    if (inner && likely_done) break;  // the caller's FP64 residual decides
  }
  HIPCHK(hipGetLastError());
  int conv = 0;
  for (int r = 0; r < nrhs; ++r) {
    if (relres_host) relres_host[r] = rel[r];
    if (rel[r] <= tol) ++conv;
  }
  if (converged_host) *converged_host = conv;
  return MPOL_OK;
}

// Argument checks of compute_forces (the reference's messages) and the
// launch parameters: cells no smaller than r_cut, the all-pairs loop when
// the cells would outnumber max(1024, 8 N).
int force_setup(mpol_ctx* c, const double* pos, const double* dip, int64_t n, const double box[3], double r_cut,
                double prefactor, int32_t periodic_z, int32_t use_cell_list, double* forces, double* pe,
                ForceParams& p, int64_t& ncells) {
  if (!c || !box) return fail(MPOL_EINVAL, "null argument");
  if (n < 0 || n > (1LL << 30)) return fail(MPOL_EINVAL, "Invalid shapes: pos(N,3), dip(N,3), box(3) required.");
  if (n > 0 && (!pos || !dip || !forces || !pe)) return fail(MPOL_EINVAL, "null argument");
  if (!(box[0] > 0.0 && box[1] > 0.0 && box[2] > 0.0)) return fail(MPOL_EINVAL, "Box lengths must be positive.");
  if (!(r_cut > 0.0)) return fail(MPOL_EINVAL, "r_cut must be > 0.");
  if (r_cut >= std::min(0.5 * box[0], 0.5 * box[1]))
    return fail(MPOL_EINVAL, "r_cut too large for minimum-image periodic x,y. Require r_cut < 0.5*min(Lx,Ly).");
  if (periodic_z && r_cut >= 0.5 * box[2])
    return fail(MPOL_EINVAL,
                "r_cut too large for minimum-image periodic z. Require r_cut < 0.5*Lz when periodic_z=True.");
  p.n = (int)n;
  double cells = 1.0;
  for (int d = 0; d < 3; ++d) {
    p.box[d] = box[d];
    // cells no smaller than r_cut, so every neighbour within r_cut is in an adjacent cell
    const double nc = std::max(1.0, std::min(std::floor(box[d] / r_cut), 1e9));
    p.nc[d] = (int)nc;
    cells *= nc;
  }
  p.rc2 = r_cut * r_cut;
  p.pref = prefactor;
  p.periodic_z = periodic_z ? 1 : 0;
  // very dilute systems with tiny r_cut: the all-pairs loop rather than more
  // cells than a few per particle
  p.use_cells =
      (use_cell_list && cells <= (double)std::max<int64_t>(1024, 8 * n) && cells < (double)(1LL << 30)) ? 1 : 0;
  ncells = p.use_cells ? (int64_t)cells : 0;
  return MPOL_OK;
}

// Cell list (counts, scan, placement, then the caller's order inside every
// cell, so that the pair sums run in the same order on every call) and the
// pair sums.  count and cursor must hold zeros.
int force_launch(const ForceParams& p, int64_t ncells, const double* pos, const double* dip, int* cell,
                 int* members, int* count, int* cursor, int* start, double* forces, double* pe, hipStream_t st) {
  const int blocks = cdiv(p.n, 256);
  if (p.use_cells) {
    hipLaunchKernelGGL(k_force_cells, dim3(blocks), dim3(256), 0, st, p, pos, cell, count);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, count, start, (int)ncells);
    hipLaunchKernelGGL(k_force_scatter, dim3(blocks), dim3(256), 0, st, p.n, cell, start, cursor, members);
    sort_segments((int)ncells, start, members, st);
  }
  hipLaunchKernelGGL(k_forces, dim3(blocks), dim3(256), 0, st, p, pos, dip, cell, start, members, forces, pe);
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

}  // namespace

extern "C" {

int mpol_version(void) { return 100; }

const char* mpol_last_error(void) { return g_err.c_str(); }

int mpol_device_count(int* count) {
  if (!count) return fail(MPOL_EINVAL, "null count");
  HIPCHK(hipGetDeviceCount(count));
  return MPOL_OK;
}

int mpol_create(mpol_ctx** out, int device, void* stream) {
  if (!out) return fail(MPOL_EINVAL, "null output pointer");
  int nd = 0;
  HIPCHK(hipGetDeviceCount(&nd));
  if (device < 0 || device >= nd) return fail(MPOL_EINVAL, "device index out of range");
  HIPCHK(hipSetDevice(device));
  int* scr = nullptr;
  CHECK_STATUS(scratch_address(device, &scr));
  mpol_ctx* c = new mpol_ctx();
  c->device = device;
  c->scratch = scr;
  c->stream = reinterpret_cast<hipStream_t>(stream);  // NULL: the device's default stream
  *out = c;
  return MPOL_OK;
}

int mpol_destroy(mpol_ctx* c) {
  if (!c) return MPOL_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete c;
  return MPOL_OK;
}

int mpol_set_inner_operator(mpol_ctx* c, mpol_ctx* inner, double floor) {
  if (!c) return fail(MPOL_EINVAL, "null context");
  if (inner == c) return fail(MPOL_EINVAL, "a context cannot be its own inner operator");
  if (inner != nullptr) {
    if (inner->inner != nullptr) return fail(MPOL_EINVAL, "the inner operator has an inner operator of its own");
    if (!(floor >= 0.0 && floor < 1.0)) return fail(MPOL_EINVAL, "inner floor must be in [0, 1)");
    if (inner->device != c->device) return fail(MPOL_EINVAL, "inner operator on another device");
    if (!inner->have_geo || !c->have_geo) return fail(MPOL_ESTATE, "set the geometry of both contexts first");
    if (inner->n != c->n) return fail(MPOL_EINVAL, "inner operator on another number of particles");
  }
  c->inner = inner;
  c->inner_floor = inner != nullptr ? floor : 0.0;
  return MPOL_OK;
}

int mpol_set_stream(mpol_ctx* c, void* stream) {
  if (!c) return fail(MPOL_EINVAL, "null context");
  c->stream = reinterpret_cast<hipStream_t>(stream);
  return MPOL_OK;
}

int mpol_synchronize(mpol_ctx* c) {
  if (!c) return fail(MPOL_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return MPOL_OK;
}

int mpol_set_tables(mpol_ctx* c, const double* perp, const double* para, const double* rvals, int64_t ntab) {
  if (!c || !perp || !para || !rvals) return fail(MPOL_EINVAL, "null table pointer");
  if (ntab < 2) return fail(MPOL_EINVAL, "tables need at least two entries");
  for (int64_t i = 1; i < ntab; ++i)
    if (!(rvals[i] > rvals[i - 1])) return fail(MPOL_EINVAL, "rvals must be strictly increasing");
  HIPCHK(hipSetDevice(c->device));
  HIPALLOC(c->perp, (size_t)ntab);
  HIPALLOC(c->para, (size_t)ntab);
  HIPALLOC(c->rvals, (size_t)ntab);
  HIPCHK(hipMemcpyAsync(c->perp.p, perp, ntab * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->para.p, para, ntab * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->rvals.p, rvals, ntab * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ntab = (int)ntab;
  c->self_perp = perp[0];
  c->have_geo = false;   // pair data depends on the tables
  c->have_qtab = false;  // quadrupole tables live on the same rvals
  c->have_sizes = false; // and so do the polydisperse tables
  return MPOL_OK;
}

int mpol_set_geometry(mpol_ctx* c, const double* x, int64_t n, const double box[3], const int32_t ngrid[3],
                      int32_t P, double xi, const double eta[3], double rc) {
  if (!c || !x || !box || !ngrid || !eta) return fail(MPOL_EINVAL, "null argument");
  if (c->ntab == 0) return fail(MPOL_ESTATE, "tables not set (call mpol_set_tables first)");
  if (n < 1 || n > (1LL << 30)) return fail(MPOL_EINVAL, "particle count out of range");
  for (int d = 0; d < 3; ++d) {
    if (!(box[d] > 0.0)) return fail(MPOL_EINVAL, "box lengths must be positive");
    if (ngrid[d] < 2 || ngrid[d] > 4096) return fail(MPOL_EINVAL, "grid size must be in [2, 4096]");
    if (!(eta[d] > 0.0)) return fail(MPOL_EINVAL, "eta must be positive");
  }
  if (P < 1 || P > 64) return fail(MPOL_EINVAL, "P must be in [1, 64]");
  if (!(xi > 0.0) || !(rc > 0.0)) return fail(MPOL_EINVAL, "xi and rc must be positive");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  c->have_geo = false;
  c->img[0].valid = false;
  c->img[1].valid = false;
  c->n = (int)n;
  for (int d = 0; d < 3; ++d) {
    c->box[d] = box[d];
    c->ng[d] = ngrid[d];
    c->h[d] = box[d] / (double)ngrid[d];
    c->eta[d] = eta[d];
    int ncell = (int)std::floor(box[d] / rc);
    c->nc[d] = ncell < 3 ? 3 : ncell;
  }
  c->kx = c->ng[0] / 2 + 1;
  c->kxs = (c->kx + 15) / 16 * 16;
  c->P = P;
  c->half = (P - 1) / 2;
  c->xi = xi;
  c->rc = rc;
  const int nbins = c->ng[1] * c->ng[2];
  const int64_t ncells = (int64_t)c->nc[0] * c->nc[1] * c->nc[2];
  if (ncells > (1LL << 28)) return fail(MPOL_EINVAL, "too many real-space cells");
  HIPALLOC(c->x, 3 * (size_t)n);
  HIPALLOC(c->base, 3 * (size_t)n);
  HIPALLOC(c->bin_key, (size_t)n);
  HIPALLOC(c->bin_count, (size_t)nbins);
  HIPALLOC(c->bin_start, (size_t)nbins + 1);
  HIPALLOC(c->bin_cursor, (size_t)nbins);
  HIPALLOC(c->cell_key, (size_t)n);
  HIPALLOC(c->cell_count, (size_t)ncells);
  HIPALLOC(c->cell_start, (size_t)ncells + 1);
  HIPALLOC(c->cell_cursor, (size_t)ncells);
  HIPALLOC(c->cell_members, (size_t)n);
  HIPALLOC(c->g0u, 3 * (size_t)n);
  HIPALLOC(c->perm, (size_t)n);
  HIPALLOC(c->rec, (size_t)n);
  HIPALLOC(c->base64, (size_t)n);
  HIPALLOC(c->base32, (size_t)n);
  HIPALLOC(c->rs_count, (size_t)n);
  HIPALLOC(c->rs_off, (size_t)n + 1);
  HIPALLOC(c->err, 1);
  HIPCHK(hipMemcpyAsync(c->x.p, x, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemsetAsync(c->bin_count.p, 0, nbins * sizeof(int), st));
  HIPCHK(hipMemsetAsync(c->bin_cursor.p, 0, nbins * sizeof(int), st));
  HIPCHK(hipMemsetAsync(c->cell_count.p, 0, ncells * sizeof(int), st));
  HIPCHK(hipMemsetAsync(c->cell_cursor.p, 0, ncells * sizeof(int), st));
  HIPCHK(hipMemsetAsync(c->err.p, 0, sizeof(int), st));

  GeoParams g{};
  g.n = (int)n;
  g.nx = c->ng[0]; g.ny = c->ng[1]; g.nz = c->ng[2];
  g.P = P; g.half = c->half;
  double peta = 1.0;
  for (int d = 0; d < 3; ++d) {
    g.box[d] = box[d];
    g.h[d] = c->h[d];
    g.eta[d] = eta[d];
    g.nc[d] = c->nc[d];
  }
  peta = (eta[0] * eta[1]) * eta[2];
  g.xi = xi;
  g.norm = std::pow(2.0 * xi * xi / 3.14159265358979323846, 1.5) * std::sqrt(1.0 / peta);
  c->norm = g.norm;
  const int pg = cdiv(n, 256);
  hipLaunchKernelGGL(k_geo_keys, dim3(pg), dim3(256), 0, st, g, c->x.p, c->bin_key.p, c->bin_count.p,
                     c->cell_key.p, c->cell_count.p, c->base.p, c->g0u.p);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, c->bin_count.p, c->bin_start.p, nbins);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, c->cell_count.p, c->cell_start.p, (int)ncells);
  hipLaunchKernelGGL(k_geo_scatter, dim3(pg), dim3(256), 0, st, g, c->bin_key.p, c->bin_start.p, c->bin_cursor.p,
                     c->cell_key.p, c->cell_start.p, c->cell_cursor.p, c->perm.p, c->cell_members.p);
  // caller's order inside every bin and cell: the sorted order, and with it
  // every summation order downstream, is the same on every call
  sort_segments(nbins, c->bin_start.p, c->perm.p, st);
  sort_segments((int)ncells, c->cell_start.p, c->cell_members.p, st);
  hipLaunchKernelGGL(k_geo_records, dim3(pg), dim3(256), 0, st, (int)n, c->perm.p, c->base.p, c->g0u.p, c->rec.p,
                     c->base64.p, c->base32.p);
  PairParams pp{};
  pp.n = (int)n;
  for (int d = 0; d < 3; ++d) {
    pp.box[d] = box[d];
    pp.nc[d] = c->nc[d];
  }
  pp.rc = rc;
  pp.ntab = c->ntab;
  hipLaunchKernelGGL((k_pairs<false>), dim3(pg), dim3(256), 0, st, pp, c->x.p, c->cell_key.p, c->cell_start.p,
                     c->cell_members.p, c->rvals.p, c->perp.p, c->para.p, c->rs_count.p, nullptr, nullptr, nullptr,
                     nullptr, nullptr, nullptr, nullptr, nullptr, c->err.p);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, c->rs_count.p, c->rs_off.p, (int)n);
  int total = 0, errflag = 0;
  HIPCHK(hipMemcpyAsync(&total, c->rs_off.p + n, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&errflag, c->err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  if (errflag) {
    return fail(MPOL_ERANGE,
                "RealSpace interpolation: pair separations fall outside rvals table. Increase r_table max "
                "(in RunCapacitance) or adjust xi/errortol.");
  }
  c->npairs = total;
  const size_t np = total > 0 ? (size_t)total : 1;
  HIPALLOC(c->rs_j, np);
  HIPALLOC(c->rs_A, np);
  HIPALLOC(c->rs_BA, np);
  HIPALLOC(c->rs_rh, 3 * np);
  HIPALLOC(c->rs_A32, np);
  HIPALLOC(c->rs_BA32, np);
  HIPALLOC(c->rs_rh32, 3 * np);
  hipLaunchKernelGGL((k_pairs<true>), dim3(pg), dim3(256), 0, st, pp, c->x.p, c->cell_key.p, c->cell_start.p,
                     c->cell_members.p, c->rvals.p, c->perp.p, c->para.p, nullptr, c->rs_off.p, c->rs_j.p,
                     c->rs_A.p, c->rs_BA.p, c->rs_rh.p, c->rs_A32.p, c->rs_BA32.p, c->rs_rh32.p, c->err.p);
  // FFT plans and twiddles
  for (int d = 0; d < 3; ++d) {
    c->plan[d] = make_plan(c->ng[d]);
    std::vector<double> tw;
    make_twiddles(c->ng[d], tw);
    std::vector<float> tw32(tw.begin(), tw.end());
    HIPALLOC(c->tw64[d], tw.size());
    HIPALLOC(c->tw32[d], tw32.size());
    HIPCHK(hipMemcpyAsync(c->tw64[d].p, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->tw32[d].p, tw32.data(), tw32.size() * sizeof(float), hipMemcpyHostToDevice, st));
  }
  // k-space coefficient tables
  {
    c->cubic = (box[0] == box[1] && box[1] == box[2]);
    std::vector<double> kt;
    for (int d = 0; d < 3; ++d) {
      for (int i = 0; i < c->ng[d]; ++i) {
        const double k = fft_freq(i, c->ng[d]) * (2.0 * kPi / box[d]);
        kt.push_back(std::exp(-(1.0 - eta[d]) * k * k / (4.0 * xi * xi)));
      }
    }
    if (c->cubic) {
      int64_t nmax = 0;
      for (int d = 0; d < 3; ++d) nmax += (int64_t)(c->ng[d] / 2) * (c->ng[d] / 2);
      const double dk = 2.0 * kPi / box[0];
      for (int64_t q = 0; q <= nmax; ++q) {
        if (q == 0) {
          kt.push_back(0.0);
          continue;
        }
        const double kk = dk * dk * (double)q;
        const double kn = std::sqrt(kk);
        const double j = std::sqrt(2.0 / (kPi * kn)) * (std::sin(kn) / kn - std::cos(kn));
        kt.push_back(9.0 * kPi / (2.0 * kn) * (j * j) / kk);
      }
      // (j1^2, j1 j2, j2^2)/k^2 of the dipole + quadrupole blocks
      c->jq_off = (int64_t)kt.size();
      for (int64_t q = 0; q <= nmax; ++q) {
        if (q == 0) {
          kt.push_back(0.0); kt.push_back(0.0); kt.push_back(0.0);
          continue;
        }
        const double kk = dk * dk * (double)q;
        double j1, j2;
        sph_j12(std::sqrt(kk), j1, j2);
        kt.push_back(j1 * j1 / kk);
        kt.push_back(j1 * j2 / kk);
        kt.push_back(j2 * j2 / kk);
      }
      // the dipole table (after the three separable ones) over |k|^2, for
      // the rank-one register z pass
      c->jk_off = (int64_t)kt.size();
      const size_t jt_off = (size_t)c->ng[0] + c->ng[1] + c->ng[2];
      for (int64_t q = 0; q <= nmax; ++q) kt.push_back(q == 0 ? 0.0 : kt[jt_off + (size_t)q] / (dk * dk * (double)q));
      // FP32 copies of e_z and that table
      std::vector<float> kt32;
      const size_t ez_off = (size_t)c->ng[0] + c->ng[1];
      for (int i = 0; i < c->ng[2]; ++i) kt32.push_back((float)kt[ez_off + i]);
      for (int64_t q = 0; q <= nmax; ++q) kt32.push_back((float)kt[(size_t)c->jk_off + (size_t)q]);
      HIPALLOC(c->ktab32, kt32.size());
      HIPCHK(hipMemcpy(c->ktab32.p, kt32.data(), kt32.size() * sizeof(float), hipMemcpyHostToDevice));
    } else {
      c->jq_off = -1;
      c->jk_off = -1;
    }
    HIPALLOC(c->ktab, kt.size());
    HIPCHK(hipMemcpyAsync(c->ktab.p, kt.data(), kt.size() * sizeof(double), hipMemcpyHostToDevice, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  c->have_geo = true;
  c->quad_pairs_valid = false;
  c->sz_pairs_valid = false;
  c->sz_nodes_valid = false;
  return MPOL_OK;
}

}  // extern "C"

namespace {
__global__ void k_diag(const double* __restrict__ eps, int64_t total, double self_perp, double* __restrict__ diag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < total) diag[k] = -3.0 / (4.0 * kPi * (1.0 - eps[k])) + self_perp;
}

// Complex permittivity eps = er + i ei: c/(1 - eps) = c (a + i b)/(a^2 + b^2)
// with a = 1 - er, b = ei; diag = -coef/(1 - eps) + self (self term real).
__global__ void k_diag_cplx(const double* __restrict__ er, const double* __restrict__ ei, int64_t total,
                            double coef, double self, double* __restrict__ dre, double* __restrict__ dim) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const double a = 1.0 - er[k], b = ei[k];
  const double inv = 1.0 / (a * a + b * b);
  dre[k] = -coef * a * inv + self;
  dim[k] = -coef * b * inv;
}

// Quadrupole rows: 15/(8 pi (eps - 1)) + grad_quad_2(0)/2.
__global__ void k_diag_q(const double* __restrict__ eps, int64_t total, double self, double* __restrict__ diag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < total) diag[k] = -15.0 / (8.0 * kPi * (1.0 - eps[k])) + self;
}

int check_sets(mpol_ctx* c, int nsys) {
  if (c->nsets != 1 && c->nsets != nsys)
    return fail(MPOL_EINVAL, "permittivity sets must be 1 or one per right-hand side");
  return MPOL_OK;
}

// Quadrupole pair coefficients and diagonal, (re)built when the geometry,
// the quadrupole tables or the permittivities changed.
int prepare_quad(mpol_ctx* c) {
  if (!c->have_qtab) return fail(MPOL_ESTATE, "quadrupole tables not set (call mpol_set_quad_tables first)");
  hipStream_t st = c->stream;
  if (!c->quad_pairs_valid) {
    const int64_t np = c->npairs;
    HIPALLOC(c->rs_Q, (size_t)std::max<int64_t>(1, np) * kQuadTabs);
    HIPALLOC(c->rs_Q32, (size_t)std::max<int64_t>(1, np) * kQuadTabs);
    PairParams pp{};
    pp.n = c->n;
    for (int d = 0; d < 3; ++d) {
      pp.box[d] = c->box[d];
      pp.nc[d] = c->nc[d];
    }
    pp.rc = c->rc;
    pp.ntab = c->ntab;
    hipLaunchKernelGGL(k_pairs_quad, dim3(cdiv(c->n, 256)), dim3(256), 0, st, pp, c->x.p, c->rs_off.p, c->rs_j.p,
                       c->rvals.p, c->qtab.p, c->rs_Q.p, c->rs_Q32.p);
    HIPCHK(hipGetLastError());
    c->quad_pairs_valid = true;
  }
  if (!c->diag_q_valid) {
    const int64_t total = (int64_t)c->nsets * c->n;
    HIPALLOC(c->diag_q, (size_t)total);
    if (c->cplx_eps) {
      HIPALLOC(c->diag_qim, (size_t)total);
      hipLaunchKernelGGL(k_diag_cplx, dim3(cdiv(total, 256)), dim3(256), 0, st, c->eps_re.p, c->eps_im.p, total,
                         15.0 / (8.0 * kPi), 0.5 * c->self_gq2, c->diag_q.p, c->diag_qim.p);
    } else {
      hipLaunchKernelGGL(k_diag_q, dim3(cdiv(total, 256)), dim3(256), 0, st, c->eps_re.p, total, 0.5 * c->self_gq2,
                         c->diag_q.p);
    }
    HIPCHK(hipGetLastError());
    c->diag_q_valid = true;
  }
  return MPOL_OK;
}

// Polydisperse pair coefficients, sorted radii and diagonal, (re)built when the
// geometry, the radii/tables or the permittivities changed.
int prepare_sizes(mpol_ctx* c) {
  if (!c->have_sizes) return fail(MPOL_ESTATE, "radii not set (call mpol_set_sizes first)");
  hipStream_t st = c->stream;
  for (int d = 0; d < 3; ++d)
    if (c->P > c->ng[d]) return fail(MPOL_EINVAL, "polydisperse operator needs P <= grid size on every axis");
  if (!c->sz_pairs_valid) {
    const size_t np = (size_t)std::max<int64_t>(1, c->npairs);
    HIPALLOC(c->sz_A, np);
    HIPALLOC(c->sz_BA, np);
    HIPALLOC(c->sz_A32, np);
    HIPALLOC(c->sz_BA32, np);
    HIPALLOC(c->sz_rad_sorted, (size_t)c->n);
    hipLaunchKernelGGL(k_permute_scalar, dim3(cdiv(c->n, 256)), dim3(256), 0, st, c->sz_radius.p, c->perm.p, c->n,
                       c->sz_rad_sorted.p);
    PairParams pp{};
    pp.n = c->n;
    for (int d = 0; d < 3; ++d) {
      pp.box[d] = c->box[d];
      pp.nc[d] = c->nc[d];
    }
    pp.rc = c->rc;
    pp.ntab = c->ntab;
    hipLaunchKernelGGL(k_pairs_sizes, dim3(cdiv(c->n, 256)), dim3(256), 0, st, pp, c->x.p, c->rs_off.p, c->rs_j.p,
                       c->rvals.p, c->sz_radius.p, c->sz_type.p, c->sz_ntypes, c->sz_col.p, c->sz_perp.p,
                       c->sz_para.p, c->xi, c->sz_A.p, c->sz_BA.p, c->sz_A32.p, c->sz_BA32.p);
    HIPCHK(hipGetLastError());
    c->sz_pairs_valid = true;
  }
  if (!c->sz_diag_valid) {
    const int64_t total = (int64_t)c->nsets * c->n;
    HIPALLOC(c->sz_diag, (size_t)total);
    if (c->cplx_eps) HIPALLOC(c->sz_diag_im, (size_t)total);
    hipLaunchKernelGGL(k_diag_sizes, dim3(cdiv(total, 256)), dim3(256), 0, st, c->eps_re.p,
                       c->cplx_eps ? c->eps_im.p : nullptr, c->n, total, c->sz_radius.p, c->sz_type.p, c->sz_ntypes,
                       c->sz_col.p, c->sz_perp.p, c->ntab, c->xi, c->sz_diag.p,
                       c->cplx_eps ? c->sz_diag_im.p : nullptr);
    HIPCHK(hipGetLastError());
    c->sz_diag_valid = true;
  }
  return MPOL_OK;
}

// Restarted GMRES, plain FP64 or FP64 iterative refinement around an
// FP32-operator GMRES, for operator `kind` on nsys systems.
int solve_impl(mpol_ctx* c, int kind, const double* b, const double* x0, double* x, int nsys,
               const mpol_solve_opts* o, mpol_solve_info* info) {
  hipStream_t st = c->stream;
  const int64_t n3 = (int64_t)unknowns_per_particle(kind) * c->n * rhs_per_system(c);
  const int64_t tot = n3 * nsys;
  const int nrhs = nsys;
  const double slab = kind == kOpDipole ? o->slab_factor : 0.0;
  mpol_solve_info inf{};
  std::vector<double> rel(nrhs, 0.0);
  int conv = 0;
  if (x != x0) HIPCHK(hipMemcpyAsync(x, x0, tot * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (!o->mixed) {
    int64_t nmv = 0;
    CHECK_STATUS(gmres_core(c, kind, b, x, nrhs, o->tol, o->restart, o->maxit, MPOL_FP64, slab, &nmv, rel.data(),
                            &conv));
    inf.matvecs_fp64 = nmv;
  } else {
    // the correction operator: the inner context when one is set (dipole and
    // dipole + quadrupole operators), launching on this context's stream for
    // the solve
    mpol_ctx* ci = ((kind == kOpDipole || kind == kOpQuad) && c->inner != nullptr) ? c->inner : c;
    if (ci != c) {
      CHECK_STATUS(check_ready(ci));
      if (ci->n != c->n || ci->nsets != c->nsets || ci->cplx_eps != c->cplx_eps || ci->device != c->device)
        return fail(MPOL_EINVAL, "inner operator: particles, permittivity sets and device must match the context's");
    }
    struct StreamSwap {
      mpol_ctx* ctx;
      hipStream_t saved;
      ~StreamSwap() { ctx->stream = saved; }
    } swap{ci, ci->stream};
    ci->stream = st;
    if (ci != c && kind == kOpQuad) CHECK_STATUS(prepare_quad(ci));
    const double floor_in = ci != c ? std::max(c->inner_floor, kInnerFloor) : kInnerFloor;
    HIPALLOC(c->gd, (size_t)tot);
    HIPALLOC(c->gx, (size_t)tot);
    HIPALLOC(c->gR, (size_t)tot);
    HIPALLOC(c->gAx, (size_t)tot);
    HIPALLOC(c->gnorm, (size_t)2 * nrhs);
    HIPALLOC(c->gpart, (size_t)nrhs * (o->restart + 1) * kDotBlocks);
    // ||b||^2 (gnorm[0, nrhs)), read together with the first residual's
    std::vector<double> bn2(nrhs), rn2(nrhs), both(2 * (size_t)nrhs);
    dots(c, b, 0, 1, b, n3, nrhs, c->gnorm.p);
    bool have_bn2 = false;
    int64_t nmv64 = 0, nmv32 = 0;
    int ref = 0;
    // The first residual (from the caller's guess) only seeds the first
    // correction solve, so it is taken with the FP32 operator of the
    // correction solves.  With one operator every later residual is FP64.
    // With an inner operator the later ones are FP32 on this operator while
    // they only seed the next correction (its rounding, ~4e-8 relative, is
    // far below what a sweep removes) and FP64 once the residual may meet
    // tol: predicted, per sweep, as the largest residual times the larger of
    // the correction's tolerance and the reduction the last sweep achieved.
    // Any FP32 residual that meets tol is certified by an FP64 one, so the
    // loop ends on FP64 residuals only.
    enum { kResFp64, kResInner, kResFp32 };
    int rkind = o->max_refine > 0 ? kResInner : kResFp64;
    int64_t nmv_inner = 0;
    double last_max = 0.0, reduction = floor_in;
    for (;; ++ref) {
      mpol_ctx* const rc = rkind == kResInner ? ci : c;
      CHECK_STATUS(apply(rc, kind, x, c->gAx.p, nrhs, rkind == kResFp64 ? MPOL_FP64 : MPOL_FP32, slab));
      if (rkind == kResFp64) ++nmv64;
      else ++nmv32;
      if (rkind == kResInner && ci != c) ++nmv_inner;
      hipLaunchKernelGGL(k_residual, dim3(cdiv(tot, 256)), dim3(256), 0, st, b, c->gAx.p, c->gd.p, tot);
      dots(c, c->gd.p, 0, 1, c->gd.p, n3, nrhs, c->gnorm.p + nrhs);
      if (!have_bn2) {
        CHECK_STATUS(read_doubles(c, c->gnorm.p, 2 * nrhs, both.data()));
        std::copy(both.begin(), both.begin() + nrhs, bn2.begin());
        std::copy(both.begin() + nrhs, both.end(), rn2.begin());
        have_bn2 = true;
      } else {
        CHECK_STATUS(read_doubles(c, c->gnorm.p + nrhs, nrhs, rn2.data()));
      }
      double ratio = 1e300, max_rel = 0.0;
      conv = 0;
      for (int r = 0; r < nrhs; ++r) {
        rel[r] = bn2[r] > 0.0 ? std::sqrt(rn2[r] / bn2[r]) : 0.0;
        max_rel = std::max(max_rel, rel[r]);
        if (rel[r] <= o->tol) ++conv;
        else if (rn2[r] > 0.0) ratio = std::min(ratio, std::sqrt(bn2[r] / rn2[r]));
      }
      if (rkind != kResFp64) {
        if (conv == nrhs || ref >= o->max_refine) {  // certify with an FP64 residual
          rkind = kResFp64;
          --ref;
          continue;
        }
      } else if (conv == nrhs || ref >= o->max_refine) {
        break;
      }
      if (last_max > 0.0 && max_rel > 0.0) reduction = std::min(1.0, max_rel / last_max);
      last_max = max_rel;
      // inner tolerance relative to the current residual, not below
      // kInnerFloor: the FP32 operator's rounding (~1e-7 relative) limits how
      // far its own residual estimate can usefully fall (an inner operator's
      // distance from this one, by the floor set with it)
      double tin = 0.3 * o->tol * ratio;
      tin = std::min(0.1, std::max(tin, floor_in));
      // the residual becomes the correction solve's rhs (gx); the correction starts at zero in gd
      hipLaunchKernelGGL(k_take, dim3(cdiv(tot, 256)), dim3(256), 0, st, c->gd.p, c->gx.p, tot);
      int64_t nmv = 0;
      CHECK_STATUS(gmres_core(ci, kind, c->gx.p, c->gd.p, nrhs, tin, std::max(o->restart, kInnerRestart),
                              std::min(o->maxit, kInnerCycles), MPOL_FP32, slab, &nmv, nullptr, nullptr, true));
      nmv32 += nmv;
      if (ci != c) nmv_inner += nmv;
      hipLaunchKernelGGL(k_axpy_all, dim3(cdiv(tot, 256)), dim3(256), 0, st, 1.0, c->gd.p, x, tot);
      rkind = (ci == c || max_rel * std::max(tin, reduction) <= o->tol) ? kResFp64 : kResFp32;
    }
    inf.matvecs_inner = nmv_inner;
    inf.matvecs_fp64 = nmv64;
    inf.matvecs_fp32 = nmv32;
    inf.refinements = ref;
  }
  HIPCHK(hipStreamSynchronize(st));
  inf.converged = conv;
  inf.max_relres = 0.0;
  for (double v : rel) inf.max_relres = std::max(inf.max_relres, v);
  if (info) *info = inf;
  return MPOL_OK;
}

int check_solve_args(mpol_ctx* c, const double* b, const double* x0, double* x, int32_t nrhs,
                     const mpol_solve_opts* o) {
  if (!b || !x0 || !x || !o || nrhs < 1) return fail(MPOL_EINVAL, "bad solve arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  if (o->restart < 1 || o->restart > 200 || o->maxit < 0) return fail(MPOL_EINVAL, "bad restart/maxit");
  if (!(o->tol > 0.0)) return fail(MPOL_EINVAL, "tolerance must be positive");
  return MPOL_OK;
}
}  // namespace

extern "C" {

int mpol_set_permittivity(mpol_ctx* c, const double* eps, int32_t nsets) {
  if (!c || !eps) return fail(MPOL_EINVAL, "null argument");
  if (!c->have_geo) return fail(MPOL_ESTATE, "geometry not set (call mpol_set_geometry first)");
  if (nsets < 1) return fail(MPOL_EINVAL, "nsets must be >= 1");
  HIPCHK(hipSetDevice(c->device));
  const int64_t total = (int64_t)nsets * c->n;
  HIPALLOC(c->diag, (size_t)total);
  HIPALLOC(c->eps_re, (size_t)total);
  HIPCHK(hipMemcpyAsync(c->eps_re.p, eps, total * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_diag, dim3(cdiv(total, 256)), dim3(256), 0, c->stream, eps, total, c->self_perp, c->diag.p);
  HIPCHK(hipGetLastError());
  c->nsets = nsets;
  c->cplx_eps = false;
  c->diag_q_valid = false;
  c->sz_diag_valid = false;
  return MPOL_OK;
}

int mpol_set_permittivity_complex(mpol_ctx* c, const double* eps_re, const double* eps_im, int32_t nsets) {
  if (!c || !eps_re || !eps_im) return fail(MPOL_EINVAL, "null argument");
  if (!c->have_geo) return fail(MPOL_ESTATE, "geometry not set (call mpol_set_geometry first)");
  if (nsets < 1) return fail(MPOL_EINVAL, "nsets must be >= 1");
  HIPCHK(hipSetDevice(c->device));
  const int64_t total = (int64_t)nsets * c->n;
  HIPALLOC(c->diag, (size_t)total);
  HIPALLOC(c->diag_im, (size_t)total);
  HIPALLOC(c->eps_re, (size_t)total);
  HIPALLOC(c->eps_im, (size_t)total);
  HIPCHK(hipMemcpyAsync(c->eps_re.p, eps_re, total * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->eps_im.p, eps_im, total * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_diag_cplx, dim3(cdiv(total, 256)), dim3(256), 0, c->stream, eps_re, eps_im, total,
                     3.0 / (4.0 * kPi), c->self_perp, c->diag.p, c->diag_im.p);
  HIPCHK(hipGetLastError());
  c->nsets = nsets;
  c->cplx_eps = true;
  c->diag_q_valid = false;
  c->sz_diag_valid = false;
  return MPOL_OK;
}

int mpol_matvec(mpol_ctx* c, const double* m, double* out, int32_t nrhs, int32_t precision, double slab) {
  CHECK_STATUS(check_ready(c));
  if (!m || !out || nrhs < 1) return fail(MPOL_EINVAL, "bad matvec arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  if (precision != MPOL_FP64 && precision != MPOL_FP32) return fail(MPOL_EINVAL, "unknown precision");
  HIPCHK(hipSetDevice(c->device));
  return apply(c, kOpDipole, m, out, nrhs, precision, slab);
}

int mpol_matvec_timed(mpol_ctx* c, const double* m, double* out, int32_t nrhs, int32_t precision, double slab,
                      int32_t repeats, double* stage_ms) {
  CHECK_STATUS(check_ready(c));
  if (!m || !out || nrhs < 1 || repeats < 1 || !stage_ms) return fail(MPOL_EINVAL, "bad matvec arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  HIPCHK(hipSetDevice(c->device));
  const int nr = nrhs * rhs_per_system(c);
  hipEvent_t ev[kStages + 1];
  for (int k = 0; k <= kStages; ++k) HIPCHK(hipEventCreate(&ev[k]));
  for (int k = 0; k < kStages; ++k) stage_ms[k] = 0.0;
  int status = MPOL_OK;
  for (int r = 0; r < repeats && status == MPOL_OK; ++r) {
    status = precision == MPOL_FP32 ? run_operator<float, kOpDipole>(c, m, out, nullptr, nullptr, nr, slab, ev)
                                    : run_operator<double, kOpDipole>(c, m, out, nullptr, nullptr, nr, slab, ev);
    if (status != MPOL_OK) break;
    if (hipEventSynchronize(ev[kStages]) != hipSuccess) {
      status = fail(MPOL_EHIP, "hipEventSynchronize failed");
      break;
    }
    for (int k = 0; k < kStages; ++k) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
      stage_ms[k] += ms / repeats;
    }
  }
  for (int k = 0; k <= kStages; ++k) (void)hipEventDestroy(ev[k]);
  return status;
}

int mpol_debug_spectrum(mpol_ctx* c, const double* m, int32_t nrhs, int32_t precision, int32_t stage,
                        void* spec_out) {
  CHECK_STATUS(check_ready(c));
  if (!m || !spec_out || nrhs < 1 || stage < 2 || stage > 5) return fail(MPOL_EINVAL, "bad debug arguments");
  if (c->cplx_eps) return fail(MPOL_ESTATE, "spectrum dump needs real permittivities");
  HIPCHK(hipSetDevice(c->device));
  DevBuf<double> tmp;
  HIPALLOC(tmp, (size_t)nrhs * 3 * c->n);
  const int st = precision == MPOL_FP32
                     ? run_operator<float, kOpDipole>(c, m, tmp.p, nullptr, nullptr, nrhs, 0.0, nullptr, stage)
                     : run_operator<double, kOpDipole>(c, m, tmp.p, nullptr, nullptr, nrhs, 0.0, nullptr, stage);
  CHECK_STATUS(st);
  const size_t cs = precision == MPOL_FP32 ? sizeof(cplx<float>) : sizeof(cplx<double>);
  const size_t rows = (size_t)3 * nrhs * c->ng[2] * c->ng[1];
  // drop the row padding of the internal layout
  HIPCHK(hipMemcpy2DAsync(spec_out, (size_t)c->kx * cs, c->spec.p, (size_t)c->kxs * cs, (size_t)c->kx * cs, rows,
                          hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MPOL_OK;
}

int mpol_matvec_components(mpol_ctx* c, const double* m, double* hk, double* hr, double* hslab, int32_t nrhs,
                           int32_t precision, double slab) {
  CHECK_STATUS(check_ready(c));
  if (!m || !hk || !hr || !hslab || nrhs < 1) return fail(MPOL_EINVAL, "bad matvec arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  HIPCHK(hipSetDevice(c->device));
  const int nr = nrhs * rhs_per_system(c);
  if (precision == MPOL_FP32) return run_operator<float, kOpDipole>(c, m, hk, hr, hslab, nr, slab);
  return run_operator<double, kOpDipole>(c, m, hk, hr, hslab, nr, slab);
}

int mpol_solve(mpol_ctx* c, const double* b, const double* x0, double* x, int32_t nrhs, const mpol_solve_opts* o,
               mpol_solve_info* info) {
  CHECK_STATUS(check_ready(c));
  CHECK_STATUS(check_solve_args(c, b, x0, x, nrhs, o));
  HIPCHK(hipSetDevice(c->device));
  return solve_impl(c, kOpDipole, b, x0, x, nrhs, o, info);
}

int mpol_set_sizes(mpol_ctx* c, const double* radius, const int32_t* type, int32_t ntypes, const double* perp,
                   const double* para) {
  if (!c || !radius) return fail(MPOL_EINVAL, "null argument");
  if (!c->have_geo) return fail(MPOL_ESTATE, "geometry not set (call mpol_set_geometry first)");
  if (ntypes < 0 || ntypes > 4096) return fail(MPOL_EINVAL, "ntypes must be in [0, 4096]");
  if (ntypes > 0 && (!type || !perp || !para)) return fail(MPOL_EINVAL, "tables and types needed when ntypes > 0");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int n = c->n;
  HIPALLOC(c->sz_radius, (size_t)n);
  HIPCHK(hipMemcpyAsync(c->sz_radius.p, radius, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (ntypes > 0) {
    const int64_t m = (int64_t)ntypes * (ntypes + 1) / 2;
    std::vector<int> col((size_t)ntypes * ntypes);
    int lin = 0;
    for (int i = 0; i < ntypes; ++i)
      for (int j = i; j < ntypes; ++j) {
        col[(size_t)i * ntypes + j] = lin;
        col[(size_t)j * ntypes + i] = lin;
        ++lin;
      }
    HIPALLOC(c->sz_type, (size_t)n);
    HIPALLOC(c->sz_col, col.size());
    HIPALLOC(c->sz_perp, (size_t)m * c->ntab);
    HIPALLOC(c->sz_para, (size_t)m * c->ntab);
    HIPCHK(hipMemcpyAsync(c->sz_type.p, type, n * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c->sz_col.p, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->sz_perp.p, perp, (size_t)m * c->ntab * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->sz_para.p, para, (size_t)m * c->ntab * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  c->sz_ntypes = ntypes;
  c->have_sizes = true;
  c->sz_pairs_valid = false;
  c->sz_nodes_valid = false;
  c->sz_diag_valid = false;
  return MPOL_OK;
}

int mpol_sizes_matvec(mpol_ctx* c, const double* m, double* out, int32_t nrhs, int32_t precision) {
  CHECK_STATUS(check_ready(c));
  if (!m || !out || nrhs < 1) return fail(MPOL_EINVAL, "bad matvec arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  if (precision != MPOL_FP64 && precision != MPOL_FP32) return fail(MPOL_EINVAL, "unknown precision");
  HIPCHK(hipSetDevice(c->device));
  CHECK_STATUS(prepare_sizes(c));
  return apply(c, kOpSizes, m, out, nrhs, precision, 0.0);
}

int mpol_sizes_matvec_components(mpol_ctx* c, const double* m, double* hk, double* hr, int32_t nrhs,
                                 int32_t precision) {
  CHECK_STATUS(check_ready(c));
  if (!m || !hk || !hr || nrhs < 1) return fail(MPOL_EINVAL, "bad matvec arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  HIPCHK(hipSetDevice(c->device));
  CHECK_STATUS(prepare_sizes(c));
  const int nr = nrhs * rhs_per_system(c);
  if (precision == MPOL_FP32) return run_operator_sizes<float>(c, m, hk, hr, nr);
  return run_operator_sizes<double>(c, m, hk, hr, nr);
}

int mpol_sizes_solve(mpol_ctx* c, const double* b, const double* x0, double* x, int32_t nrhs,
                     const mpol_solve_opts* o, mpol_solve_info* info) {
  CHECK_STATUS(check_ready(c));
  CHECK_STATUS(check_solve_args(c, b, x0, x, nrhs, o));
  HIPCHK(hipSetDevice(c->device));
  CHECK_STATUS(prepare_sizes(c));
  return solve_impl(c, kOpSizes, b, x0, x, nrhs, o, info);
}

int mpol_set_quad_tables(mpol_ctx* c, const double* qtab, int64_t ntab) {
  if (!c || !qtab) return fail(MPOL_EINVAL, "null argument");
  if (c->ntab == 0) return fail(MPOL_ESTATE, "tables not set (call mpol_set_tables first)");
  if (ntab != c->ntab) return fail(MPOL_EINVAL, "quadrupole tables must have the length of rvals");
  HIPCHK(hipSetDevice(c->device));
  HIPALLOC(c->qtab, (size_t)kQuadTabs * ntab);
  HIPCHK(hipMemcpyAsync(c->qtab.p, qtab, (size_t)kQuadTabs * ntab * sizeof(double), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->self_gq2 = qtab[4 * ntab];  // grad_quad_2 at r = 0: its self term
  c->have_qtab = true;
  c->quad_pairs_valid = false;
  c->diag_q_valid = false;
  return MPOL_OK;
}

int mpol_quad_matvec(mpol_ctx* c, const double* m, double* out, int32_t nrhs, int32_t precision) {
  CHECK_STATUS(check_ready(c));
  if (!m || !out || nrhs < 1) return fail(MPOL_EINVAL, "bad matvec arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  if (precision != MPOL_FP64 && precision != MPOL_FP32) return fail(MPOL_EINVAL, "unknown precision");
  HIPCHK(hipSetDevice(c->device));
  CHECK_STATUS(prepare_quad(c));
  return apply(c, kOpQuad, m, out, nrhs, precision, 0.0);
}

int mpol_quad_matvec_components(mpol_ctx* c, const double* m, double* hk, double* hr, int32_t nrhs,
                                int32_t precision) {
  CHECK_STATUS(check_ready(c));
  if (!m || !hk || !hr || nrhs < 1) return fail(MPOL_EINVAL, "bad matvec arguments");
  CHECK_STATUS(check_sets(c, nrhs));
  HIPCHK(hipSetDevice(c->device));
  CHECK_STATUS(prepare_quad(c));
  const int nr = nrhs * rhs_per_system(c);
  if (precision == MPOL_FP32) return run_operator<float, kOpQuad>(c, m, hk, hr, nullptr, nr, 0.0);
  return run_operator<double, kOpQuad>(c, m, hk, hr, nullptr, nr, 0.0);
}

int mpol_quad_solve(mpol_ctx* c, const double* b, const double* x0, double* x, int32_t nrhs,
                    const mpol_solve_opts* o, mpol_solve_info* info) {
  CHECK_STATUS(check_ready(c));
  CHECK_STATUS(check_solve_args(c, b, x0, x, nrhs, o));
  HIPCHK(hipSetDevice(c->device));
  CHECK_STATUS(prepare_quad(c));
  return solve_impl(c, kOpQuad, b, x0, x, nrhs, o, info);
}

int mpol_set_tuning_for(mpol_ctx* c, int32_t precision, const int32_t* knobs, int32_t nknobs) {
  if (!c || (nknobs > 0 && !knobs) || nknobs < 0 || nknobs > kTuneKnobs) return fail(MPOL_EINVAL, "bad tuning arguments");
  if (precision != MPOL_FP32 && precision != MPOL_FP64 && precision != -1)
    return fail(MPOL_EINVAL, "precision must be MPOL_FP32, MPOL_FP64 or -1 (both)");
  int t[kTuneKnobs] = {};
  for (int i = 0; i < nknobs; ++i) t[i] = knobs[i];
  const bool pow2_rows = (t[0] == 0 && t[1] == 0) ||
                         (t[0] >= 1 && t[0] <= 4 && t[1] >= 1 && t[1] <= 4 && ((t[0] * t[1]) & (t[0] * t[1] - 1)) == 0);
  if (!pow2_rows) return fail(MPOL_EINVAL, "tile rows ty*tz must be a power of two with ty, tz in [1, 4]");
  if (t[2] != 0 && (t[2] < 32 || t[2] > 256)) return fail(MPOL_EINVAL, "particle chunk must be in [32, 256]");
  if (t[3] < 0 || t[3] > 64 || t[4] < 0 || t[4] > 64) return fail(MPOL_EINVAL, "column widths must be in [1, 64]");
  if (t[5] != 0 && (t[5] < 64 || t[5] > 1024 || t[5] % 64 != 0))
    return fail(MPOL_EINVAL, "z-pass threads must be a multiple of 64 in [64, 1024]");
  if (t[6] < 0 || t[6] > 2) return fail(MPOL_EINVAL, "engine must be 0 (automatic), 1 (matrix cores) or 2 (vector loops)");
  if (t[7] < 0 || t[7] > 64) return fail(MPOL_EINVAL, "right-hand sides per workgroup must be in [1, 64]");
  if (t[8] < 0 || t[8] > 2) return fail(MPOL_EINVAL, "z register pass must be 0 (automatic), 1 (rank-one) or 2 (full)");
  if (t[9] < 0 || t[9] > 2)
    return fail(MPOL_EINVAL, "complex Nyquist part must be 0 (automatic), 1 (reference operator) or 2 (Hermitian part)");
  if (t[10] < 0 || t[10] > 2)
    return fail(MPOL_EINVAL, "y/z pass order must be 0 (automatic), 1 (system by system) or 2 (pass by pass)");
  for (int p = 0; p < 2; ++p) {
    if (precision != -1 && p != (precision == MPOL_FP32 ? 1 : 0)) continue;
    for (int i = 0; i < kTuneKnobs; ++i) c->tune[p][i] = t[i];
  }
  return MPOL_OK;
}

int mpol_set_tuning(mpol_ctx* c, const int32_t* knobs, int32_t nknobs) {
  return mpol_set_tuning_for(c, -1, knobs, nknobs);
}

int mpol_geometry_info(mpol_ctx* c, mpol_geometry_summary* info) {
  if (!c || !info) return fail(MPOL_EINVAL, "null argument");
  if (!c->have_geo) return fail(MPOL_ESTATE, "geometry not set");
  info->n = c->n;
  info->npairs = c->npairs;
  for (int d = 0; d < 3; ++d) {
    info->ngrid[d] = c->ng[d];
    info->ncell[d] = c->nc[d];
    info->fft_e[d] = c->plan[d].e;
  }
  info->P = c->P;
  info->tile_rows = tile_cfg_nc<kOpDipole>(c->ng[0], generic_fft(c->plan[0]), sizeof(cplx<double>), c->P, kLdsTile).rows;
  return MPOL_OK;
}

int mpol_real_space_table(double xi, const double* r, int64_t n, double* out) {
  if (!r || !out || n < 0) return fail(MPOL_EINVAL, "bad table arguments");
  if (!(xi > 0.0)) return fail(MPOL_EINVAL, "xi must be positive");
  for (int64_t i = 0; i < n; ++i)
    if (!(r[i] > 0.0)) return fail(MPOL_EINVAL, "table separations must be positive");
  real_space_tables(xi, r, n, out);
  return MPOL_OK;
}

int mpol_real_space_table_field(double xi, const double* r, int64_t n, double* out) {
  if (!r || !out || n < 0) return fail(MPOL_EINVAL, "bad table arguments");
  if (!(xi > 0.0)) return fail(MPOL_EINVAL, "xi must be positive");
  for (int64_t i = 0; i < n; ++i)
    if (!(r[i] > 0.0)) return fail(MPOL_EINVAL, "table separations must be positive");
  real_space_tables(xi, r, n, out, false);
  return MPOL_OK;
}

int mpol_real_space_table_sizes(double xi, const double* r, int64_t n, const double* radii, int32_t M,
                                double* perp, double* para) {
  if (!r || !radii || !perp || !para || n < 1 || M < 1) return fail(MPOL_EINVAL, "bad table arguments");
  if (!(xi > 0.0)) return fail(MPOL_EINVAL, "xi must be positive");
  for (int32_t k = 0; k < M; ++k)
    if (!(radii[k] > 0.0)) return fail(MPOL_EINVAL, "radii must be positive");
  for (int64_t i = 1; i < n; ++i)
    if (!(r[i] > 0.0)) return fail(MPOL_EINVAL, "table separations after the first must be positive");
  int64_t lin = 0;
  for (int32_t i = 0; i < M; ++i)
    for (int32_t j = i; j < M; ++j, ++lin) {
      double* tp = perp + lin * n;
      double* tq = para + lin * n;
      tp[0] = tq[0] = sizes_self(radii[i], radii[j], xi);  // row 0 is the r -> 0 value
      for (int64_t k = 1; k < n; ++k) sizes_coupling(r[k], radii[i], radii[j], xi, tp[k], tq[k]);
    }
  return MPOL_OK;
}

int mpol_real_space_table_sizes_field(double xi, const double* r, int64_t n, const double* radii, int32_t M,
                                      double* perp, double* para) {
  if (!r || !radii || !perp || !para || n < 1 || M < 1) return fail(MPOL_EINVAL, "bad table arguments");
  if (!(xi > 0.0)) return fail(MPOL_EINVAL, "xi must be positive");
  for (int32_t k = 0; k < M; ++k)
    if (!(radii[k] > 0.0)) return fail(MPOL_EINVAL, "radii must be positive");
  for (int64_t i = 0; i < n; ++i)
    if (!(r[i] > 0.0)) return fail(MPOL_EINVAL, "table separations must be positive");
  int64_t lin = 0;
  for (int32_t i = 0; i < M; ++i)
    for (int32_t j = i; j < M; ++j, ++lin)
      for (int64_t k = 0; k < n; ++k)
        sizes_coupling(r[k], radii[i], radii[j], xi, perp[lin * n + k], para[lin * n + k], false);
  return MPOL_OK;
}

int mpol_real_space_table_quad(double xi, const double* r, int64_t n, double* out) {
  if (!r || !out || n < 0) return fail(MPOL_EINVAL, "bad table arguments");
  if (!(xi > 0.0)) return fail(MPOL_EINVAL, "xi must be positive");
  for (int64_t i = 0; i < n; ++i)
    if (!(r[i] > 0.0)) return fail(MPOL_EINVAL, "table separations must be positive");
  real_space_tables_quad(xi, r, n, out);
  return MPOL_OK;
}

int mpol_real_space_table_quad_field(double xi, const double* r, int64_t n, double* out) {
  if (!r || !out || n < 0) return fail(MPOL_EINVAL, "bad table arguments");
  if (!(xi > 0.0)) return fail(MPOL_EINVAL, "xi must be positive");
  for (int64_t i = 0; i < n; ++i)
    if (!(r[i] > 0.0)) return fail(MPOL_EINVAL, "table separations must be positive");
  real_space_tables_quad(xi, r, n, out, false);
  return MPOL_OK;
}

int mpol_compute_forces(mpol_ctx* c, const double* pos, const double* dip, int64_t n, const double box[3],
                        double r_cut, double prefactor, int32_t periodic_z, int32_t use_cell_list, double* forces,
                        double* pe) {
  ForceParams p{};
  int64_t ncells = 0;
  CHECK_STATUS(force_setup(c, pos, dip, n, box, r_cut, prefactor, periodic_z, use_cell_list, forces, pe, p, ncells));
  if (n == 0) return MPOL_OK;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  if (p.use_cells) {
    HIPALLOC(c->f_cell, (size_t)n);
    HIPALLOC(c->f_members, (size_t)n);
    HIPALLOC(c->f_count, (size_t)ncells);
    HIPALLOC(c->f_cursor, (size_t)ncells);
    HIPALLOC(c->f_start, (size_t)ncells + 1);
    HIPCHK(hipMemsetAsync(c->f_count.p, 0, ncells * sizeof(int), st));
    HIPCHK(hipMemsetAsync(c->f_cursor.p, 0, ncells * sizeof(int), st));
  }
  return force_launch(p, ncells, pos, dip, c->f_cell.p, c->f_members.p, c->f_count.p, c->f_cursor.p, c->f_start.p,
                      forces, pe, st);
}

int mpol_compute_forces_noalloc(mpol_ctx* c, const double* pos, const double* dip, int64_t n, const double box[3],
                                double r_cut, double prefactor, int32_t periodic_z, int32_t use_cell_list,
                                double* forces, double* pe) {
  if (n > kNoAllocMaxN)
    return fail(MPOL_EINVAL, "mpol_compute_forces_noalloc: N = " + std::to_string(n) + " exceeds its maximum of " +
                                 std::to_string(kNoAllocMaxN) + " particles (use mpol_compute_forces)");
  ForceParams p{};
  int64_t ncells = 0;
  CHECK_STATUS(force_setup(c, pos, dip, n, box, r_cut, prefactor, periodic_z, use_cell_list, forces, pe, p, ncells));
  if (n == 0) return MPOL_OK;
  // force_setup keeps the cells at most max(1024, 8 N) <= kNoAllocMaxCells
  if (p.use_cells && ncells > kNoAllocMaxCells)
    return fail(MPOL_EINVAL, "mpol_compute_forces_noalloc: too many cells for its scratch");
  HIPCHK(hipSetDevice(c->device));
  if (!c->scratch) return fail(MPOL_ESTATE, "scratch address of this device not set up");
  hipStream_t st = c->stream;
  int* cell = c->scratch + kScratchCell;
  int* members = c->scratch + kScratchMembers;
  int* count = c->scratch + kScratchCount;
  int* cursor = c->scratch + kScratchCursor;
  int* start = c->scratch + kScratchStart;
  if (p.use_cells) {
    // counts and cursors start at zero (a fill kernel: no memset node in a captured graph)
    hipLaunchKernelGGL(k_fill_int, dim3(cdiv(ncells, 256)), dim3(256), 0, st, count, ncells, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(cdiv(ncells, 256)), dim3(256), 0, st, cursor, ncells, 0);
  }
  return force_launch(p, ncells, pos, dip, cell, members, count, cursor, start, forces, pe, st);
}

}  // extern "C"
