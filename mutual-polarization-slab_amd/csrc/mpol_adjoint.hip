// Vector-Jacobian product of the dipole operator (position and permittivity
// gradients of lam^T A p), the backward pass of a polarization solve.
#include "mpol_adjoint.h"
#include "mpol_internal.h"
#include "mpol_sizes.h"

namespace {

// Adds scale * d/dx of the wave-space part of sum_rhs lam^T A(x) p (the
// Gaussian weights' position derivative at both the spread and the
// contraction of every particle) to grad (device, N x 3, caller's order).
int wave_vjp_add(mpol_ctx* c, const double* m, const double* lam, int nrhs, double* grad, double scale) {
  hipStream_t st = c->stream;
  const int n = c->n;
  const size_t vec = (size_t)nrhs * 3 * n;
  const size_t G = (size_t)c->ng[0] * c->ng[1] * c->ng[2];
  HIPALLOC(c->vjp_in, 2 * vec);
  HIPALLOC(c->vjp_out, 2 * vec);
  HIPALLOC(c->vjp_grid, (size_t)2 * nrhs * 3 * G);
  HIPCHK(hipMemcpyAsync(c->vjp_in.p, m, vec * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(c->vjp_in.p + vec, lam, vec * sizeof(double), hipMemcpyDeviceToDevice, st));
  // field grids of [p, lam] up to the inverse y transform (half spectra along x)
  const int fwd = run_operator<double, kOpDipole>(c, c->vjp_in.p, c->vjp_out.p, nullptr, nullptr, 2 * nrhs, 0.0,
                                                  nullptr, 5);
  if (fwd != MPOL_OK) return fwd;
  // inverse x transform of every (input, component) plane into real grids
  const size_t cs = sizeof(cplx<double>);
  const bool gen_x = generic_fft(c->plan[0]);
  TileCfg tr = tile_cfg_nc<1>(c->ng[0], gen_x, cs, 0, kLdsTile);
  TileArgs ta{};
  ta.nx = c->ng[0]; ta.ny = c->ng[1]; ta.nz = c->ng[2]; ta.kx = c->kx; ta.kxs = c->kxs;
  ta.ty = tr.ty; ta.tz = tr.tz; ta.rows = tr.rows; ta.ncl = tr.ncl; ta.lstride = tr.lstride;
  ta.tw = c->tw64[0].p; ta.plan = c->plan[0];
  const size_t rlds = ((size_t)c->ng[0] + (size_t)tr.ncl * tr.lstride * (gen_x ? 2 : 1)) * cs;
  if (rlds > kLdsMax) return fail(MPOL_EINVAL, "row transform tile does not fit in LDS");
  const dim3 tgrid(cdiv(c->ng[1], tr.ty), cdiv(c->ng[2], tr.tz), 2 * nrhs * 3);
  const cplx<double>* spec = reinterpret_cast<const cplx<double>*>(c->spec.p);
  double* grid = c->vjp_grid.p;
#define LAUNCH_C2R(E, N, ...)                                                                   \
  do {                                                                                          \
    set_lds(k_rows_c2r<double, E, N>, rlds);                                                    \
    hipLaunchKernelGGL((k_rows_c2r<double, E, N>), tgrid, dim3(256), rlds, st, ta, spec, grid); \
  } while (0)
  DISPATCH_FFT(c->ng[0], ta.plan.e, LAUNCH_C2R, 0);
#undef LAUNCH_C2R
  WaveVjpArgs wa{};
  wa.nx = c->ng[0]; wa.ny = c->ng[1]; wa.nz = c->ng[2];
  wa.P = c->P; wa.half = c->half; wa.n = n; wa.nrhs = nrhs;
  for (int d = 0; d < 3; ++d) {
    wa.h[d] = c->h[d];
    wa.cexp[d] = 2.0 * c->xi * c->xi / c->eta[d];
  }
  wa.norm = c->norm;
  wa.hvol = c->h[0] * c->h[1] * c->h[2];
  wa.rec = c->rec.p;
  wa.base = c->base64.p;
  hipLaunchKernelGGL(k_wave_vjp, dim3(n), dim3(256), 0, st, wa, grid, m, lam, grad, scale);
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

PairParams pair_params(const mpol_ctx* c) {
  PairParams pp{};
  pp.n = c->n;
  for (int d = 0; d < 3; ++d) {
    pp.box[d] = c->box[d];
    pp.nc[d] = c->nc[d];
  }
  pp.rc = c->rc;
  pp.ntab = c->ntab;
  return pp;
}

}  // namespace

extern "C" int mpol_matvec_vjp(mpol_ctx* c, const double* m, const double* lam, int32_t nrhs, double* grad_x,
                               double* grad_eps) {
  if (!c) return fail(MPOL_EINVAL, "null context");
  if (!c->have_geo) return fail(MPOL_ESTATE, "geometry not set (call mpol_set_geometry first)");
  if (c->nsets == 0) return fail(MPOL_ESTATE, "permittivity not set (call mpol_set_permittivity first)");
  if (c->cplx_eps) return fail(MPOL_ESTATE, "gradients need real permittivities");
  if (!m || !lam || !grad_x || nrhs < 1) return fail(MPOL_EINVAL, "bad vjp arguments");
  if (c->nsets != 1 && c->nsets != nrhs) return fail(MPOL_EINVAL, "permittivity sets must be 1 or nrhs");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int n = c->n;
  // real space (writes grad_x), then the wave-space weights (adds)
  hipLaunchKernelGGL(k_real_space_vjp, dim3(cdiv(n, 256)), dim3(256), 0, st, pair_params(c), c->x.p, c->rs_off.p,
                     c->rs_j.p, c->rvals.p, c->perp.p, c->para.p, m, lam, nrhs, grad_x);
  CHECK_STATUS(wave_vjp_add(c, m, lam, nrhs, grad_x, 1.0));
  if (grad_eps) {
    const int64_t total = (int64_t)c->nsets * n;
    hipLaunchKernelGGL(k_eps_vjp, dim3(cdiv(total, 256)), dim3(256), 0, st, c->eps_re.p, n, c->nsets, nrhs, m, lam,
                       grad_eps);
  }
  HIPCHK(hipGetLastError());
  return MPOL_OK;
}

extern "C" int mpol_ewald_forces(mpol_ctx* c, const double* p, int32_t nrhs, const double* f1_host,
                                 const double* f2_host, double* forces) {
  if (!c) return fail(MPOL_EINVAL, "null context");
  if (!c->have_geo) return fail(MPOL_ESTATE, "geometry not set (call mpol_set_geometry first)");
  if (c->nsets == 0) return fail(MPOL_ESTATE, "permittivity not set (call mpol_set_permittivity first)");
  if (c->cplx_eps) return fail(MPOL_ESTATE, "forces need real permittivities");
  if (!p || !f1_host || !f2_host || !forces || nrhs < 1) return fail(MPOL_EINVAL, "bad force arguments");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int n = c->n;
  HIPALLOC(c->ftab1, (size_t)c->ntab);
  HIPALLOC(c->ftab2, (size_t)c->ntab);
  HIPCHK(hipMemcpyAsync(c->ftab1.p, f1_host, c->ntab * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ftab2.p, f2_host, c->ntab * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_ewald_forces_real, dim3(cdiv(n, 256)), dim3(256), 0, st, pair_params(c), c->x.p,
                     c->rs_off.p, c->rs_j.p, c->rvals.p, c->ftab1.p, c->ftab2.p, p, nrhs, forces);
  CHECK_STATUS(wave_vjp_add(c, p, p, nrhs, forces, -0.5));
  // the host tables are read by the copies above before this call returns
  HIPCHK(hipStreamSynchronize(st));
  return MPOL_OK;
}
