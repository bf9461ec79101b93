// Entry points of the drop-in precompute and fast-path functions of
// "CPU - Ewald test/ewald.py" (kernels in mpol_dropin.h).  Stateless apart
// from the context's device, stream and a few small work buffers: every
// array is the caller's (device memory).  See include/mpol.h.
#include "mpol_dropin.h"
#include "mpol_internal.h"

namespace {

int ready(mpol_ctx* c) {
  if (!c) return fail(MPOL_EINVAL, "null context");
  HIPCHK(hipSetDevice(c->device));
  return MPOL_OK;
}

int grid_dims(const int32_t ngrid[3], int64_t* G) {
  if (!ngrid) return fail(MPOL_EINVAL, "null grid size");
  for (int d = 0; d < 3; ++d)
    if (ngrid[d] < 1) return fail(MPOL_EINVAL, "grid sizes must be positive");
  *G = (int64_t)ngrid[0] * ngrid[1] * ngrid[2];
  return MPOL_OK;
}

int blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

// Zero the int flag, run `launch`, read the flag back.
template <typename F>
int with_flag(mpol_ctx* c, F launch, int* flag) {
  HIPALLOC(c->drop_err, 1);
  HIPCHK(hipMemsetAsync(c->drop_err.p, 0, sizeof(int), c->stream));
  launch(c->drop_err.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(flag, c->drop_err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MPOL_OK;
}

DropKParams kparams(const int32_t ngrid[3], const double box[3], double xi, const double eta[3]) {
  DropKParams p{};
  p.nx = ngrid[0];
  p.ny = ngrid[1];
  p.nz = ngrid[2];
  p.k0x = ngrid[0] / 2;  // ceil((N - 1)/2): the k = 0 entry in fftshift order
  p.k0y = ngrid[1] / 2;
  p.k0z = ngrid[2] / 2;
  for (int d = 0; d < 3; ++d) {
    p.one_m_eta[d] = 1.0 - eta[d];
    p.kunit[d] = box ? (2.0 * kDropPi) / box[d] : 0.0;
  }
  p.four_xi2 = 4.0 * (xi * xi);
  return p;
}

}  // namespace

extern "C" {

int mpol_grid_kernel(mpol_ctx* c, const double* x, int64_t n, const int32_t ngrid[3], const double h[3], double xi,
                     const double eta[3], int32_t P, const int64_t* offset, const double* offsetxyz, int64_t* linnode,
                     double* w_spread, double* w_contract) {
  CHECK_STATUS(ready(c));
  int64_t G = 0;
  CHECK_STATUS(grid_dims(ngrid, &G));
  if (!h || !eta) return fail(MPOL_EINVAL, "null argument");
  if (n < 0 || P < 1 || P > 64) return fail(MPOL_EINVAL, "bad particle count or stencil size");
  if (n == 0) return MPOL_OK;
  if (!x || !offset || !offsetxyz || !linnode || !w_spread || !w_contract) return fail(MPOL_EINVAL, "null argument");
  DropGridParams p{};
  p.n = n;
  p.P3 = (int64_t)P * P * P;
  p.nx = ngrid[0];
  p.ny = ngrid[1];
  p.nz = ngrid[2];
  double peta = 1.0, ph = 1.0;
  for (int d = 0; d < 3; ++d) {
    p.h[d] = h[d];
    p.inv_eta[d] = 1.0 / eta[d];
    peta *= eta[d];
    ph *= h[d];
  }
  p.m2xi2 = -2.0 * (xi * xi);
  p.cnst = std::pow(2.0 * (xi * xi) / kDropPi, 1.5) * std::sqrt(1.0 / peta);
  p.prodh = ph;
  hipLaunchKernelGGL(k_drop_grid_kernel, dim3(blocks(n * p.P3)), dim3(256), 0, c->stream, p, x, offset, offsetxyz,
                     linnode, w_spread, w_contract);
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

int mpol_spread_precomp(mpol_ctx* c, const double* m, int64_t n, int32_t P3, const int64_t* linnode,
                        const double* w_spread, const int32_t ngrid[3], double* H) {
  CHECK_STATUS(ready(c));
  int64_t G = 0;
  CHECK_STATUS(grid_dims(ngrid, &G));
  if (n < 0 || P3 < 0 || !H) return fail(MPOL_EINVAL, "bad spread arguments");
  HIPCHK(hipMemsetAsync(H, 0, (size_t)G * 3 * 2 * sizeof(double), c->stream));
  if (n == 0 || P3 == 0) return MPOL_OK;
  if (!m || !linnode || !w_spread) return fail(MPOL_EINVAL, "null argument");
  int bad = 0;
  CHECK_STATUS(with_flag(
      c,
      [&](int* err) {
        hipLaunchKernelGGL(k_drop_spread, dim3(blocks(n * P3)), dim3(256), 0, c->stream, n, (int64_t)P3, ngrid[0],
                           ngrid[1], ngrid[2], linnode, w_spread, reinterpret_cast<const double2*>(m),
                           reinterpret_cast<double2*>(H), err);
      },
      &bad));
  if (bad) return fail(MPOL_EINVAL, "linnode holds node indices outside the grid");
  return MPOL_OK;
}

int mpol_contract_precomp(mpol_ctx* c, const double* H, const int32_t ngrid[3], const int64_t* linnode,
                          const double* w_contract, int64_t n, int32_t P3, double* Hk) {
  CHECK_STATUS(ready(c));
  int64_t G = 0;
  CHECK_STATUS(grid_dims(ngrid, &G));
  if (n < 0 || P3 < 0) return fail(MPOL_EINVAL, "bad contract arguments");
  if (n == 0) return MPOL_OK;
  if (!H || !linnode || !w_contract || !Hk) return fail(MPOL_EINVAL, "null argument");
  int bad = 0;
  CHECK_STATUS(with_flag(
      c,
      [&](int* err) {
        hipLaunchKernelGGL(k_drop_contract, dim3(cdiv(n, 4)), dim3(256), 0, c->stream, n, (int64_t)P3, ngrid[0],
                           ngrid[1], ngrid[2], linnode, w_contract, reinterpret_cast<const double2*>(H),
                           reinterpret_cast<double2*>(Hk), err);
      },
      &bad));
  if (bad) return fail(MPOL_EINVAL, "linnode holds node indices outside the grid");
  return MPOL_OK;
}

int mpol_scale_kgrid(mpol_ctx* c, const double* fH, const double* k, const int32_t ngrid[3], double xi,
                     const double eta[3], double* out) {
  CHECK_STATUS(ready(c));
  int64_t G = 0;
  CHECK_STATUS(grid_dims(ngrid, &G));
  if (!fH || !k || !eta || !out) return fail(MPOL_EINVAL, "null argument");
  const DropKParams p = kparams(ngrid, nullptr, xi, eta);
  hipLaunchKernelGGL(k_drop_scale_k, dim3(blocks(G)), dim3(256), 0, c->stream, p, k,
                     reinterpret_cast<const double2*>(fH), reinterpret_cast<double2*>(out));
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

int mpol_kspace_table(mpol_ctx* c, const int32_t ngrid[3], const double box[3], double xi, const double eta[3],
                      double* khat, double* coeff) {
  CHECK_STATUS(ready(c));
  int64_t G = 0;
  CHECK_STATUS(grid_dims(ngrid, &G));
  if (!box || !eta || !khat || !coeff) return fail(MPOL_EINVAL, "null argument");
  const DropKParams p = kparams(ngrid, box, xi, eta);
  hipLaunchKernelGGL(k_drop_kspace, dim3(blocks(G)), dim3(256), 0, c->stream, p, reinterpret_cast<double2*>(khat),
                     coeff);
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

int mpol_scale_precomp(mpol_ctx* c, const double* fH, const double* khat, const double* coeff, int64_t G,
                       double* out) {
  CHECK_STATUS(ready(c));
  if (G < 0) return fail(MPOL_EINVAL, "bad grid size");
  if (G == 0) return MPOL_OK;
  if (!fH || !khat || !coeff || !out) return fail(MPOL_EINVAL, "null argument");
  hipLaunchKernelGGL(k_drop_scale_fast, dim3(blocks(G)), dim3(256), 0, c->stream, G,
                     reinterpret_cast<const double2*>(fH), reinterpret_cast<const double2*>(khat), coeff,
                     reinterpret_cast<double2*>(out));
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

int mpol_real_space_pairs(mpol_ctx* c, const double* x, const double box[3], const int64_t* p1, const int64_t* p2,
                          int64_t m, double rc, const double* perp, const double* para, const double* rvals,
                          int64_t ntab, int64_t* p1f, int64_t* p2f, double* rhat, double* A, double* B,
                          int64_t* kept_host) {
  CHECK_STATUS(ready(c));
  if (!box || !kept_host) return fail(MPOL_EINVAL, "null argument");
  if (m < 0 || m >= (1LL << 31) - 1 || ntab < 0 || ntab > (1LL << 30)) return fail(MPOL_EINVAL, "bad pair count");
  *kept_host = 0;
  if (m == 0) return MPOL_OK;
  if (!x || !p1 || !p2 || !perp || !para || !rvals || !p1f || !p2f || !rhat || !A || !B)
    return fail(MPOL_EINVAL, "null argument");
  DropPairParams p{};
  p.m = m;
  p.ntab = (int)ntab;
  p.rc = rc;
  for (int d = 0; d < 3; ++d) p.box[d] = box[d];
  HIPALLOC(c->drop_keep, (size_t)m);
  HIPALLOC(c->drop_pos, (size_t)m + 1);
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_drop_pairs_keep, dim3(blocks(m)), dim3(256), 0, st, p, x, p1, p2, c->drop_keep.p);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, c->drop_keep.p, c->drop_pos.p, (int)m);
  int bad = 0;
  CHECK_STATUS(with_flag(
      c,
      [&](int* err) {
        hipLaunchKernelGGL(k_drop_pairs_fill, dim3(blocks(m)), dim3(256), 0, st, p, x, p1, p2, c->drop_keep.p,
                           c->drop_pos.p, perp, para, rvals, p1f, p2f, rhat, A, B, err);
      },
      &bad));
  int kept = 0;
  HIPCHK(hipMemcpyAsync(&kept, c->drop_pos.p + m, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *kept_host = kept;
  if (bad)
    return fail(MPOL_ERANGE,
                "RealSpace interpolation: pair separations fall outside rvals table. Increase r_table max "
                "(in RunCapacitance) or adjust xi/errortol.");
  return MPOL_OK;
}

int mpol_real_space_precomp(mpol_ctx* c, const double* m, const double* lambdap, int64_t n, const int64_t* p1,
                            const int64_t* p2, const double* rhat, const double* A, const double* B, int64_t npairs,
                            double self_perp, double* Hr) {
  CHECK_STATUS(ready(c));
  if (n < 0 || npairs < 0) return fail(MPOL_EINVAL, "bad sizes");
  if (n == 0) return MPOL_OK;
  if (!m || !lambdap || !Hr) return fail(MPOL_EINVAL, "null argument");
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_drop_rs_diag, dim3(blocks(n)), dim3(256), 0, st, n, reinterpret_cast<const double2*>(m),
                     lambdap, self_perp, reinterpret_cast<double2*>(Hr));
  if (npairs > 0) {
    if (!p1 || !p2 || !rhat || !A || !B) return fail(MPOL_EINVAL, "null argument");
    hipLaunchKernelGGL(k_drop_rs_fast, dim3(blocks(npairs)), dim3(256), 0, st, npairs,
                       reinterpret_cast<const double2*>(m), p1, p2, rhat, A, B, reinterpret_cast<double2*>(Hr));
  }
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

int mpol_real_space_direct(mpol_ctx* c, const double* x, const double* m, const double* lambdap, int64_t n,
                           const double box[3], const int64_t* p1, const int64_t* p2, int64_t ncand, double rc,
                           const double* perp, const double* para, const double* rvals, int64_t ntab, double* Hr) {
  CHECK_STATUS(ready(c));
  if (!box) return fail(MPOL_EINVAL, "null argument");
  if (n < 0 || ncand < 0 || ntab < 1 || ntab > (1LL << 30)) return fail(MPOL_EINVAL, "bad sizes");
  if (n == 0) return MPOL_OK;
  if (!x || !m || !lambdap || !perp || !para || !rvals || !Hr) return fail(MPOL_EINVAL, "null argument");
  hipStream_t st = c->stream;
  double self_perp = 0.0;
  HIPCHK(hipMemcpyAsync(&self_perp, perp, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  hipLaunchKernelGGL(k_drop_rs_diag, dim3(blocks(n)), dim3(256), 0, st, n, reinterpret_cast<const double2*>(m),
                     lambdap, self_perp, reinterpret_cast<double2*>(Hr));
  if (ncand == 0) {
    HIPCHK(hipGetLastError());
    return MPOL_OK;
  }
  if (!p1 || !p2) return fail(MPOL_EINVAL, "null argument");
  DropPairParams p{};
  p.m = ncand;
  p.ntab = (int)ntab;
  p.rc = rc;
  for (int d = 0; d < 3; ++d) p.box[d] = box[d];
  int bad = 0;
  CHECK_STATUS(with_flag(
      c,
      [&](int* err) {
        hipLaunchKernelGGL(k_drop_rs_direct, dim3(blocks(ncand)), dim3(256), 0, st, p, x,
                           reinterpret_cast<const double2*>(m), p1, p2, perp, para, rvals,
                           reinterpret_cast<double2*>(Hr), err);
      },
      &bad));
  if (bad)
    return fail(MPOL_ERANGE,
                "RealSpace interpolation: pair separation d exceeds the maximum tabulated rvals. Increase r_table "
                "max (in RunCapacitance) or reduce rc/tolerance.");
  return MPOL_OK;
}

}  // extern "C"
