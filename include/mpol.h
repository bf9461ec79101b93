/* C ABI of the MI355X spectral-Ewald mutual-polarization library (libmpol.so).
 *
 * Plain C types only: device pointers are `const double*` / `double*` to
 * memory on the context's GPU (for example the data_ptr() of a torch tensor),
 * host pointers are marked "host".  Every call returns an mpol_status; on
 * failure mpol_last_error() returns a message for the calling thread.
 *
 * Layouts (all C order):
 *   positions    x[N][3]           FP64, device
 *   permittivity eps[nsets][N]     FP64, device (nsets = 1 or nrhs)
 *   dipoles/fields m[nrhs][N][3]   FP64, device -- per right-hand side the
 *                                  3N interleaved vector [p1x p1y p1z p2x ...]
 *                                  of the reference solver.
 *   dipole + quadrupole moments m[nrhs][N][8] = (S_x, S_y, S_z, Q_xx, Q_xy,
 *                                  Q_xz, Q_yy, Q_yz) per particle; the
 *                                  operator returns (E_x, E_y, E_z, G_xx, G_xy,
 *                                  G_xz, G_yy, G_yz) in the same layout.
 *   complex permittivities         (after mpol_set_permittivity_complex) every
 *                                  right-hand side is a complex vector stored
 *                                  as its real part followed by its imaginary
 *                                  part: m[nrhs][2][N][C], C = 3 or 8.
 *
 * A context launches on the caller's HIP stream and owns the geometry
 * precompute and all workspaces.  A context is not re-entrant: calls on one
 * context must not overlap.
 */
#ifndef MPOL_H_
#define MPOL_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpol_ctx mpol_ctx;

typedef enum {
  MPOL_OK = 0,
  MPOL_EINVAL = 1,   /* bad argument or shape        -> Python ValueError   */
  MPOL_ERANGE = 2,   /* pair separation outside table -> Python ValueError  */
  MPOL_EHIP = 3,     /* HIP runtime failure          -> Python RuntimeError */
  MPOL_ENOMEM = 4,   /* device allocation failed     -> Python MemoryError  */
  MPOL_ESTATE = 5    /* call out of order (no geometry/tables yet)          */
} mpol_status;

typedef enum { MPOL_FP64 = 0, MPOL_FP32 = 1 } mpol_precision;

typedef struct {
  double tol;          /* relative residual tolerance ||b - A x|| <= tol ||b||   */
  int restart;         /* Krylov dimension per cycle                            */
  int maxit;           /* maximum number of restart cycles                      */
  int mixed;           /* 0: FP64 operator; 1: FP32 operator inside FP64
                          iterative refinement (FP64 Arnoldi/Givens always):
                          each sweep takes an FP64 residual, then an FP32
                          correction solve from zero with Krylov dimension
                          max(restart, 32) that stops on its GMRES residual
                          estimate                                             */
  int max_refine;      /* refinement sweeps in mixed mode                       */
  double slab_factor;  /* coefficient of the slab correction                    */
} mpol_solve_opts;

typedef struct {
  int64_t matvecs_fp64;   /* batched operator applications in FP64          */
  int64_t matvecs_fp32;   /* batched operator applications in FP32          */
  int32_t converged;      /* number of right-hand sides that met tol        */
  int32_t refinements;    /* refinement sweeps performed (mixed mode)       */
  double max_relres;      /* max over rhs of the final true ||b-Ax||/||b||  */
  int64_t matvecs_inner;  /* of matvecs_fp32, applications of the inner
                             operator (mpol_set_inner_operator)             */
} mpol_solve_info;

typedef struct {
  int64_t n;              /* particles                                       */
  int64_t npairs;         /* directed real-space pairs with d < rc           */
  int32_t ngrid[3];
  int32_t P;
  int32_t ncell[3];
  int32_t tile_rows;      /* (y,z) grid rows per spread/contract tile         */
  int32_t fft_e[3];       /* elements/thread of each axis FFT (0 = direct DFT) */
} mpol_geometry_summary;

/* Library version (major*10000 + minor*100 + patch). */
int mpol_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* mpol_last_error(void);

/* Number of visible HIP devices. */
int mpol_device_count(int* count);

/* Create a context on `device`.  `stream` is the hipStream_t to launch on
 * (NULL: the device's default stream). */
int mpol_create(mpol_ctx** ctx, int device, void* stream);
int mpol_destroy(mpol_ctx* ctx);
/* Change the stream later calls launch on (NULL: the default stream). */
int mpol_set_stream(mpol_ctx* ctx, void* stream);
/* Wait for all work the context has queued. */
int mpol_synchronize(mpol_ctx* ctx);

/* Real-space coupling tables (host, FP64, ntab entries each).  perp is the
 * I - rr coupling, para the rr coupling, rvals the separations (ascending,
 * rvals[0] = 0 carrying the self term perp[0]).
 * Replaces the table arguments Hperp/Hpara/rvals of
 * magnetic_field ("CPU - Ewald test/ewald.py":1146-1168). */
int mpol_set_tables(mpol_ctx* ctx, const double* perp_host, const double* para_host,
                    const double* rvals_host, int64_t ntab);

/* Geometry precompute for positions x (device, FP64 N x 3): cell list and
 * directed pair list with d < rc (counting sorts on the GPU), stencil start
 * nodes and separable Gaussian factors of every particle.
 * Replaces cell_list / neighbor_list / prepare_real_space_pairs /
 * prepare_grid_kernel / precompute_kspace ("CPU - Ewald test/ewald.py":84-199,
 * 939-1016, 233-296, 523-582) as called from capacitance()
 * ("CPU - Ewald test/capacitance.py":370-385).
 * Returns MPOL_ERANGE when a pair separation falls outside the table. */
int mpol_set_geometry(mpol_ctx* ctx, const double* x_dev, int64_t n, const double box[3],
                      const int32_t ngrid[3], int32_t P, double xi, const double eta[3], double rc);

/* Particle permittivities, nsets x N (device FP64); set s applies to rhs s
 * (nsets == nrhs) or to every rhs (nsets == 1). */
int mpol_set_permittivity(mpol_ctx* ctx, const double* eps_dev, int32_t nsets);

/* Complex particle permittivities eps_re + i eps_im, nsets x N each (device
 * FP64).  From then on every right-hand side of the operator and the solves
 * is complex (layout above), as in the MATLAB spectrum drivers, which take
 * the dielectric function at each frequency
 * (Capacitance_Hetero/CapacitanceTensorSpectrum.m:16-62,
 * Capacitance_Quadrupole/CapacitanceTensorSpectrum.m:1-64).  The operator off
 * the diagonal is real, so it acts on both parts; the diagonal couples them.
 * mpol_set_permittivity returns to real right-hand sides. */
int mpol_set_permittivity_complex(mpol_ctx* ctx, const double* eps_re_dev, const double* eps_im_dev,
                                  int32_t nsets);

/* out = A m for nrhs right-hand sides: wave-space + real-space + slab terms.
 * Replaces magnetic_field ("CPU - Ewald test/ewald.py":1146-1236). */
int mpol_matvec(mpol_ctx* ctx, const double* m_dev, double* out_dev, int32_t nrhs, int32_t precision,
                double slab_factor);

/* Diagnostic timing of mpol_matvec (no counterpart in the reference): runs it
 * `repeats` times and returns in stage_ms (host, 8 entries) the mean
 * milliseconds of each stage measured with HIP events on the context stream:
 * [0] slab sum, [1] real-space sum, [2] spread + forward x-FFT (with the
 * permutation of the moments into the sorted order),
 * [3] forward y-FFT, [4] z-FFT + k-space scale + inverse z-FFT,
 * [5] inverse y-FFT, [6] inverse x-FFT + contraction, [7] per-particle sum of
 * the contraction's tile partials (matrix-core engine; 0 otherwise). */
int mpol_matvec_timed(mpol_ctx* ctx, const double* m_dev, double* out_dev, int32_t nrhs, int32_t precision,
                      double slab_factor, int32_t repeats, double* stage_ms_host);

/* Diagnostic (no counterpart in the reference): run the operator on m up to
 * and including `stage` (2 = spread + forward x-FFT, 3 = forward y-FFT,
 * 4 = z-FFT + scale + inverse z-FFT, 5 = inverse y-FFT) and copy the grid
 * spectrum spec[rhs*3+c][z][y][kx] (kx = nx/2+1, complex FP64 or FP32 by
 * precision) into spec_out (device). */
int mpol_debug_spectrum(mpol_ctx* ctx, const double* m_dev, int32_t nrhs, int32_t precision, int32_t stage,
                        void* spec_out_dev);

/* The three parts separately: hk (wave space), hr (real space incl. the
 * dielectric and self terms) and hslab (= -slab_factor sum(m_z)/V on z, so
 * A m = hk + hr - hslab).
 * Replaces magnetic_field_components ("CPU - Ewald test/ewald.py":1239-1327). */
int mpol_matvec_components(mpol_ctx* ctx, const double* m_dev, double* hk_dev, double* hr_dev,
                           double* hslab_dev, int32_t nrhs, int32_t precision, double slab_factor);

/* Restarted GMRES on the device for nrhs right-hand sides b (device), initial
 * guesses x0 (device), result x (device).  info (host) may be NULL.
 * Replaces magnetic_dipole / gmres_restarted
 * ("CPU - Ewald test/capacitance.py":205-321, 60-199). */
int mpol_solve(mpol_ctx* ctx, const double* b_dev, const double* x0_dev, double* x_dev, int32_t nrhs,
               const mpol_solve_opts* opts, mpol_solve_info* info);

/* Operator of the mixed scheme's correction solves (no counterpart in the
 * reference, whose solves have one operator): `inner` is another context on
 * the same device and particles, with the same permittivity sets, typically
 * the same system on a coarser grid at a looser Ewald tolerance.  With it,
 * mpol_solve in mixed mode takes the first residual and runs every
 * correction solve on the inner context's FP32 operator (on this context's
 * stream), while every later residual stays on this context's operator: in
 * FP32 while it only seeds the next correction, in FP64 once it may meet tol
 * (and the solve ends on FP64 residuals only), so the result meets tol for
 * this context's operator.  A correction solve stops at `floor` (relative)
 * at the lowest, about the relative distance between the two operators
 * (refining past it gains nothing).  matvecs_inner counts the inner
 * operator's applications (also in matvecs_fp32).  The dipole and the
 * dipole + quadrupole solves (mpol_solve, mpol_quad_solve; for the latter the
 * inner context needs its quadrupole tables); the inner context must outlive
 * its use and not be used concurrently.  inner = NULL removes it. */
int mpol_set_inner_operator(mpol_ctx* ctx, mpol_ctx* inner, double floor);

/* Dipole + quadrupole operator (8 unknowns per particle).  The seven
 * quadrupole tables (host FP64, 7 x ntab, on the rvals of mpol_set_tables, in
 * the order field_quad_1, field_quad_2, field_quad_3, grad_quad_1,
 * grad_quad_2, grad_quad_3, grad_quad_4, entry 0 the self term) complete the
 * field_dip_1 / field_dip_2 pair given to mpol_set_tables.
 * Replaces the table arguments of MatrixVectorMultiply
 * (Capacitance_Quadrupole/MatrixVectorMultiply.m:1-3). */
int mpol_set_quad_tables(mpol_ctx* ctx, const double* qtab_host, int64_t ntab);

/* out = A m, the fields E and gradients G at the particles from their dipoles
 * S and quadrupoles Q: wave space (four k-space blocks) + real space (nine
 * tables, dielectric terms 3/(4 pi (eps-1)) and 15/(8 pi (eps-1))).
 * Replaces MatrixVectorMultiply (Capacitance_Quadrupole/MatrixVectorMultiply.m:31-77). */
int mpol_quad_matvec(mpol_ctx* ctx, const double* m_dev, double* out_dev, int32_t nrhs, int32_t precision);

/* Wave-space part hk and real-space part hr separately (A m = hk + hr); no
 * counterpart in the reference (used to check the Ewald splitting). */
int mpol_quad_matvec_components(mpol_ctx* ctx, const double* m_dev, double* hk_dev, double* hr_dev,
                                int32_t nrhs, int32_t precision);

/* Restarted GMRES for the dipole + quadrupole system (b = [E0; 0] per
 * particle for the reference's problem).  opts.slab_factor is ignored.
 * Replaces MatrixVectorSolve (Capacitance_Quadrupole/MatrixVectorSolve.m:35-43). */
int mpol_quad_solve(mpol_ctx* ctx, const double* b_dev, const double* x0_dev, double* x_dev, int32_t nrhs,
                    const mpol_solve_opts* opts, mpol_solve_info* info);

/* Polydisperse operator (spheres of different radii, Capacitance_Sizes):
 * radius (device FP64, N, caller's order) and either
 *   ntypes > 0: radius types type (device int32, N, values in [0, ntypes))
 *               and the pair tables perp / para (host FP64, one column of
 *               ntab entries per type pair i <= j in the order
 *               (0,0), (0,1), ..., (0,M-1), (1,1), ...; on the rvals of
 *               mpol_set_tables, entry 0 the r -> 0 value), interpolated as
 *               the reference does (Capacitance_Sizes/RealSpace.m:19-76), or
 *   ntypes == 0: closed-form pair couplings of the particles' own radii
 *               (any radius distribution; type, perp, para unused).
 * The geometry must use the method's isotropic eta and P
 * (Capacitance_Sizes/ComputeField.m:23-29) with P <= the grid size.
 * Replaces the radius and table arguments of ComputeField
 * (Capacitance_Sizes/ComputeField.m:1-2). */
int mpol_set_sizes(mpol_ctx* ctx, const double* radius_dev, const int32_t* type_dev, int32_t ntypes,
                   const double* perp_host, const double* para_host);

/* out = A p for dipoles p (nrhs x N x 3): dielectric 3 p/(4 pi a^3 (eps-1)),
 * real-space pair sum, and one scalar grid per right-hand side (radius-
 * dependent spread, exp(-(1-eta) q^2/(4 xi^2))/q^2 scale, vector contraction).
 * Replaces ComputeField (Capacitance_Sizes/ComputeField.m:22-57). */
int mpol_sizes_matvec(mpol_ctx* ctx, const double* m_dev, double* out_dev, int32_t nrhs, int32_t precision);

/* Wave-space part hk and real-space part hr separately (A p = hk + hr); no
 * counterpart in the reference (used to check the Ewald splitting). */
int mpol_sizes_matvec_components(mpol_ctx* ctx, const double* m_dev, double* hk_dev, double* hr_dev,
                                 int32_t nrhs, int32_t precision);

/* Restarted GMRES for the polydisperse dipoles.  opts.slab_factor is ignored.
 * Replaces ComputeDipole (Capacitance_Sizes/ComputeDipole.m:25-38). */
int mpol_sizes_solve(mpol_ctx* ctx, const double* b_dev, const double* x0_dev, double* x_dev, int32_t nrhs,
                     const mpol_solve_opts* opts, mpol_solve_info* info);

/* Vector-Jacobian product of the dipole operator for nrhs pairs (p, lam)
 * (device, nrhs x N x 3 each): grad_x (device, N x 3) = d/dx of
 * sum_rhs lam^T A(x) p at the current positions -- the real-space pair sum
 * through the derivative of the interpolated tables and the wave-space part
 * through the derivative of the Gaussian weights (slab and diagonal terms do
 * not depend on x) -- and, when grad_eps (device, nsets x N) is not NULL,
 * d/d eps of the same sum.  With the adjoint lam of a solve A p = b for an
 * objective L(p) (A lam = dL/dp, A is symmetric), dL/dx = -grad_x.  Real
 * permittivities only.  No counterpart in the reference (its HOOMD forces
 * come from truncated point dipoles, dipole_forces_ext/hoomd_force.py:192-226);
 * it gives the forces of the Ewald-summed polarization energy. */
int mpol_matvec_vjp(mpol_ctx* ctx, const double* p_dev, const double* lam_dev, int32_t nrhs, double* grad_x_dev,
                    double* grad_eps_dev);

/* Forces F = -dU/dx (device, N x 3) of the Ewald-summed polarisation energy
 * U = 1/2 p^T A p - b^T p at solved dipoles p (device, nrhs x N x 3; summed
 * over the right-hand sides), computed forward: the real-space part from the
 * pair-force couplings f1 = field_dip_force_1, f2 = field_dip_force_2 (host,
 * ntab entries on the rvals of mpol_set_tables, as mpol_real_space_table
 * returns them: f1 = dA/dr = (B - A)/r, f2 = d(B - A)/dr of the couplings A, B
 * given to mpol_set_tables), interpolated per pair; the wave-space part from
 * the position derivative of the Gaussian weights.  The reference tabulates
 * these couplings ("CPU - Ewald test/ewald.py":798-916, self terms :928-929)
 * but computes no forces from them; its HOOMD forces come from truncated point
 * dipoles (dipole_forces_ext/hoomd_force.py:192-226).  Real permittivities
 * only. */
int mpol_ewald_forces(mpol_ctx* ctx, const double* p_dev, int32_t nrhs, const double* f1_host,
                      const double* f2_host, double* forces_dev);

/* Launch-shape tuning (no counterpart in the reference; used by the
 * autotuner): knobs[0..10] = spread/contract tile rows along y and z (ty*tz a
 * power of two, each <= 4), particles staged per chunk (32..256), y-pass and
 * z-pass columns per workgroup (<= 64), z-pass threads (multiple of 64),
 * spread/contract engine (1 matrix cores over per-tile candidate images,
 * 2 vector loops; for the polydisperse operator 2 selects the per-particle
 * kernels (atomic scatter, LDS-staged contraction) and the separate y/z
 * passes instead of the stored-entry gather/contraction and the one-kernel
 * y/z passes), right-hand sides per matrix-core spread/contract workgroup
 * (1..64: one staged tile image serves them in turn; automatic: up to four,
 * and one for the dipole spread where its staged chunk can share LDS with
 * the tile lines), the dipole register
 * z pass for nz = 256 (1 rank-one form on the lines with non-Nyquist x and y
 * wave numbers, 2 the full three-component form everywhere), the wave-space
 * operator for complex permittivities (1 the reference's, whose Nyquist
 * planes carry k = -N/2 on both halves of the spectrum: the Hermitian part
 * plus an extra wave-space pass for the anti-Hermitian part; 2 the
 * Hermitian part only, one pass), the order of the y, z and inverse y
 * passes over several systems (1 system by system, so that each pass reads
 * what the one before wrote while it is still in the last-level cache; 2 each
 * pass over all systems).  0 keeps the automatic choice (1 for the z pass
 * and the complex operator; for the pass order, system by system when the
 * systems' spectra overflow the 256 MB last-level cache and one system's
 * fill a quarter of it, except in timed stage runs, mpol_matvec_timed);
 * nknobs may be shorter than 11. */
int mpol_set_tuning(mpol_ctx* ctx, const int32_t* knobs, int32_t nknobs);

/* The same knobs for the operator in one precision only (MPOL_FP32 or
 * MPOL_FP64; -1 = both, as mpol_set_tuning).  No counterpart in the
 * reference; the autotuner (mpol_hip/autotune.py) sets them per precision. */
int mpol_set_tuning_for(mpol_ctx* ctx, int32_t precision, const int32_t* knobs, int32_t nknobs);

/* Geometry summary (host). */
int mpol_geometry_info(mpol_ctx* ctx, mpol_geometry_summary* info);

/* The six real-space coupling tables at separations r (host, n values > 0),
 * each with its r = 0 self term first: out (host) is 6 x (n+1) in the order
 * pot_charge, pot_dip, field_dip_1, field_dip_2, field_dip_force_1,
 * field_dip_force_2.  Replaces real_space_table
 * ("CPU - Ewald test/ewald.py":604-931). */
int mpol_real_space_table(double xi, const double* r_host, int64_t n, double* out_host);

/* The same six couplings without the overlap terms of r < 2: the field at a
 * point (not a sphere) from a sphere's dipole, as the field-map tables are
 * built.  Replaces RealSpaceTable (Compute_Field/RealSpaceTable.m:1-163,
 * whose (r < 2) regularisation lines are commented out). */
int mpol_real_space_table_field(double xi, const double* r_host, int64_t n, double* out_host);

/* The eight couplings the quadrupole method adds, at separations r (host, n
 * values > 0), each with its r = 0 self term first: out (host) is 8 x (n+1)
 * in the order pot_quad, field_quad_1, field_quad_2, field_quad_3,
 * grad_quad_1, grad_quad_2, grad_quad_3, grad_quad_4.  Replaces those outputs
 * of RealSpaceTable (Capacitance_Quadrupole/RealSpaceTable.m:1-380). */
int mpol_real_space_table_quad(double xi, const double* r_host, int64_t n, double* out_host);

/* The same eight quadrupole couplings without the overlap terms of r < 2: the
 * field and gradient at a point from a sphere's moments, as the field-map
 * tables are built.  Replaces RealSpaceTable
 * (Compute_Field_Quadrupole/RealSpaceTable.m:1-380, whose regularisation
 * lines are commented out). */
int mpol_real_space_table_quad_field(double xi, const double* r_host, int64_t n, double* out_host);

/* Polydisperse pair tables: for the M radii (host) and separations r (host,
 * n values, r[i] > 0 for i > 0), perp and para (host, M(M+1)/2 x n, pair
 * order as in mpol_set_sizes) get the I - rr and rr couplings; entry 0 of
 * every column is the r -> 0 value whatever r[0] is.
 * Replaces RealSpaceTable (Capacitance_Sizes/RealSpaceTable.m:1-150). */
int mpol_real_space_table_sizes(double xi, const double* r_host, int64_t n, const double* radii_host, int32_t M,
                                double* perp_host, double* para_host);

/* The same pair couplings without the overlap terms and without an r -> 0
 * entry (every r[i] > 0): the field at a point treated as a sphere of radius
 * radii[i] from a sphere of radius radii[j], as the field-map tables are built.
 * Replaces RealSpaceTable (Compute_Field_Sizes/RealSpaceTable.m:1-110). */
int mpol_real_space_table_sizes_field(double xi, const double* r_host, int64_t n, const double* radii_host,
                                      int32_t M, double* perp_host, double* para_host);

/* Truncated point-dipole forces and per-particle energies of fixed dipoles:
 * pos, dip (n x 3) and the outputs forces (n x 3), pe (n) are FP64 device
 * arrays; box (3 lengths) is host memory.  x and y are periodic by minimum
 * image (rint), z when periodic_z != 0; pairs with r^2 > r_cut^2 or
 * r^2 < 1e-24 are skipped; each particle receives half of every pair energy.
 * use_cell_list selects the cell-list neighbour search (cells of size
 * L / floor(L / r_cut)) over the all-pairs loop; both give the same pairs.
 * Rejects the arguments the reference rejects, with its messages.  Only the
 * context's device, stream and cell-list buffers are used (no tables or
 * geometry needed).  Replaces compute_forces
 * ("CPU - HOOMD simulation test/dipole_forces_ext/_dipole_forces_ext.cpp":119-497,
 * binding at :454-497). */
int mpol_compute_forces(mpol_ctx* ctx, const double* pos_dev, const double* dip_dev, int64_t n,
                        const double box_host[3], double r_cut, double prefactor, int32_t periodic_z,
                        int32_t use_cell_list, double* forces_dev, double* pe_dev);

/* ---- Step-by-step functions of the reference's fast path --------------
 * Stateless (the context gives only the device and stream); every array is
 * device memory owned by the caller, in the reference's layouts: complex
 * arrays as interleaved (re, im) FP64 pairs; grids (Nx, Ny, Nz, 3) in C order
 * (the reference's H[..., c] reshaped in Fortran order from the linear node
 * index ix + Nx iy + Nx Ny iz); int64 indices.  The operator (mpol_matvec)
 * never forms these arrays; these entry points serve callers of the
 * reference's precompute API. */

/* linnode, w_spread, w_contract (n x P^3 each) from positions x (n x 3) and the
 * stencil offsets offset (int64, P^3 x 3) / offsetxyz (P^3 x 3).
 * Replaces prepare_grid_kernel ("CPU - Ewald test/ewald.py":233-296). */
int mpol_grid_kernel(mpol_ctx* ctx, const double* x_dev, int64_t n, const int32_t ngrid[3], const double h[3],
                     double xi, const double eta[3], int32_t P, const int64_t* offset_dev, const double* offsetxyz_dev,
                     int64_t* linnode_dev, double* w_spread_dev, double* w_contract_dev);

/* H (complex grid) = sum over particles and stencil points of w m_c at node
 * linnode; MPOL_EINVAL if a node index lies outside the grid.
 * Replaces spread_fast ("CPU - Ewald test/ewald.py":299-331) and, with
 * mpol_grid_kernel, spread (:370-418). */
int mpol_spread_precomp(mpol_ctx* ctx, const double* m_dev, int64_t n, int32_t P3, const int64_t* linnode_dev,
                        const double* w_spread_dev, const int32_t ngrid[3], double* H_dev);

/* Hk (complex n x 3) = sum_k w_contract[n, k] H[linnode[n, k], c].
 * Replaces contract_fast ("CPU - Ewald test/ewald.py":334-368) and, with
 * mpol_grid_kernel, contract (:421-467). */
int mpol_contract_precomp(mpol_ctx* ctx, const double* H_dev, const int32_t ngrid[3], const int64_t* linnode_dev,
                          const double* w_contract_dev, int64_t n, int32_t P3, double* Hk_dev);

/* fHtilde = c(k) (khat . fH) khat for a given wave-vector grid k (real,
 * Nx x Ny x Nz x 3, fftshift order); zero at the k = 0 entry, non-finite
 * values mapped as numpy.nan_to_num.
 * Replaces scale ("CPU - Ewald test/ewald.py":473-516). */
int mpol_scale_kgrid(mpol_ctx* ctx, const double* fH_dev, const double* k_dev, const int32_t ngrid[3], double xi,
                     const double eta[3], double* out_dev);

/* khat (complex Nx x Ny x Nz x 3) and Htildecoeff (real Nx x Ny x Nz) on the
 * fftshift-ordered grid of box.
 * Replaces precompute_kspace ("CPU - Ewald test/ewald.py":523-582). */
int mpol_kspace_table(mpol_ctx* ctx, const int32_t ngrid[3], const double box[3], double xi, const double eta[3],
                      double* khat_dev, double* coeff_dev);

/* fHtilde = (coeff (khat . fH)) khat over G grid points (complex arithmetic).
 * Replaces scale_fast ("CPU - Ewald test/ewald.py":585-599). */
int mpol_scale_precomp(mpol_ctx* ctx, const double* fH_dev, const double* khat_dev, const double* coeff_dev,
                       int64_t G, double* out_dev);

/* Of the m candidate pairs (p1, p2), those with minimum-image d < rc, in their
 * input order: p1f, p2f, rhat (kept x 3), A, B (arrays of m entries; the
 * first *kept_host are written); MPOL_ERANGE if a kept separation falls
 * outside rvals.  Replaces prepare_real_space_pairs
 * ("CPU - Ewald test/ewald.py":939-1016). */
int mpol_real_space_pairs(mpol_ctx* ctx, const double* x_dev, const double box[3], const int64_t* p1_dev,
                          const int64_t* p2_dev, int64_t m, double rc, const double* perp_dev, const double* para_dev,
                          const double* rvals_dev, int64_t ntab, int64_t* p1f_dev, int64_t* p2f_dev, double* rhat_dev,
                          double* A_dev, double* B_dev, int64_t* kept_host);

/* Hr (complex n x 3) from the filtered pairs: the dielectric and self terms
 * plus A m_j + (B - A)(m_j . rhat) rhat at p1.
 * Replaces real_space_fast ("CPU - Ewald test/ewald.py":1019-1072). */
int mpol_real_space_precomp(mpol_ctx* ctx, const double* m_dev, const double* lambdap_dev, int64_t n,
                            const int64_t* p1_dev, const int64_t* p2_dev, const double* rhat_dev, const double* A_dev,
                            const double* B_dev, int64_t npairs, double self_perp, double* Hr_dev);

/* Hr from the candidate pairs directly (d = 0 skipped, d < rc kept,
 * interpolation as the reference's loop); MPOL_ERANGE past the table.
 * Replaces real_space ("CPU - Ewald test/ewald.py":1076-1139). */
int mpol_real_space_direct(mpol_ctx* ctx, const double* x_dev, const double* m_dev, const double* lambdap_dev,
                           int64_t n, const double box[3], const int64_t* p1_dev, const int64_t* p2_dev, int64_t ncand,
                           double rc, const double* perp_dev, const double* para_dev, const double* rvals_dev,
                           int64_t ntab, double* Hr_dev);

/* The same truncated forces without any device allocation: no counterpart
 * in the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops such as a HOOMD step, harnesses that forbid
 * hipMalloc).  Arguments, checks, results (bit for bit) and launches are those
 * of mpol_compute_forces, except that its cell-list buffers live in a fixed
 * per-GPU scratch array of the library (27 MiB, held until the process ends)
 * instead of the context's buffers, and it clears them with a fill kernel, not
 * a memset.  n at most 262144 (2^18); larger n returns MPOL_EINVAL before any
 * launch.  Every call on one GPU shares that scratch: calls that overlap on one
 * GPU (several streams, threads or contexts) race and must be serialised by
 * the caller. */
int mpol_compute_forces_noalloc(mpol_ctx* ctx, const double* pos_dev, const double* dip_dev, int64_t n,
                                const double box_host[3], double r_cut, double prefactor, int32_t periodic_z,
                                int32_t use_cell_list, double* forces_dev, double* pe_dev);

#ifdef __cplusplus
}
#endif

#endif /* MPOL_H_ */
