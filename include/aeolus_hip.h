/*
 * aeolus_hip: MI355X (gfx950) spectral time step behind the spherical-backend
 * interface of Aeolus.
 *
 * C ABI.  Every function returns an int status (AEO_OK = 0, negative on
 * failure) and leaves a thread-local message for aeo_last_error().  All
 * array arguments of execute calls are DEVICE pointers (for example
 * torch.Tensor.data_ptr() of a CUDA/HIP tensor); arguments documented as
 * "host" are host pointers read during the call.  Complex arrays are
 * interleaved FP64 pairs (the complex128 layout).  `stream` is a hipStream_t
 * (NULL = the default stream).  Plans and models allocate their device
 * memory when created; execute calls allocate nothing and only enqueue work
 * on `stream`.  A plan or model is not safe for concurrent use from several
 * host threads or streams; distinct objects are.
 *
 * Spectral layout (the reference's shared coefficient layout): one field is
 * a dense (L+1) x (L+1) complex128 array a[l][m], C order, m >= 0 only;
 * entries with m > l are padding (synthesis ignores them, analysis writes
 * 0).  A real field is implied by a_{l,-m} = (-1)^m conj(a_lm).  Several
 * fields are stacked contiguously (field stride (L+1)^2).
 *
 * Grid layout: Gauss grids are (nlat, nlon) FP64, C order, latitudes north
 * to south, longitudes 2 pi k / nlon.  Point sets are (npoints,) FP64.
 *
 * Reference paths cited below are relative to the Aeolus repository root;
 * num/ = src/planetary_sandbox/numerics/, phy/ = src/planetary_sandbox/physics/,
 * run/ = src/planetary_sandbox/run/.
 */
#ifndef AEOLUS_HIP_H
#define AEOLUS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define AEO_ABI_VERSION 1

enum aeo_status {
    AEO_OK = 0,
    AEO_ERR_INVALID = -1,      /* bad argument (size, null pointer, layout) */
    AEO_ERR_UNSUPPORTED = -2,  /* valid request outside what the library implements */
    AEO_ERR_HIP = -3,          /* HIP runtime or launch failure */
    AEO_ERR_STATE = -4         /* object used in a state that forbids the call */
};

/* Spectral operator codes for aeo_spec_op. */
enum aeo_spec_op_code {
    AEO_OP_LAP = 0,         /* -l(l+1)/R^2 a                      num/spectral_operators.py:213-218 */
    AEO_OP_INVLAP = 1,      /* a / lap, lap_0 = 1/R^2            num/spectral_operators.py:275-292 */
    AEO_OP_DLAMBDA = 2,     /* (i m / R) a, m = 0 column zero     num/spectral_operators.py:220-226 */
    AEO_OP_SINTH_DTH = 3,   /* S a (sin(theta) d/dtheta, l+-1)    num/spectral_operators.py:228-241 */
    AEO_OP_SINTH_DTH_T = 4, /* S^T a                              num/spectral_operators.py:243-270 */
    AEO_OP_TRUNC23 = 5,     /* zero rows and columns > 2L/3       num/spectral_operators.py:329-334 */
    AEO_OP_INVLAP_ZERO = 6  /* a / lap with the l = 0 row zeroed  phy/barotropic.py:68-93 */
};

typedef struct aeo_plan_s* aeo_plan_t;
typedef struct aeo_model_s* aeo_model_t;
typedef struct aeo_ensemble_s* aeo_ensemble_t;
typedef struct aeo_pe_s* aeo_pe_t;

/* ABI version (AEO_ABI_VERSION) of the loaded library. */
int aeo_version(void);

/* Message of the last failed call on this host thread ("" if none). */
const char* aeo_last_error(void);

/* Gauss-Legendre x uniform-longitude transform plan.
 * Replaces GaussLatLonSphericalHarmonics.__init__ + PointSetSphericalHarmonics
 * basis construction (num/latlon_grid.py:160-184, num/fast_geodesic_sh.py:15-66,
 * cu/sh_matrix.cu:7-76).  The FP64 normalised associated-Legendre table is
 * built on the device by a stable recurrence.
 *   x_nodes   host, nlat: cos(colatitude) of each row, north to south, symmetric
 *             (x[j] = -x[nlat-1-j], as numpy.polynomial.legendre.leggauss gives)
 *   w_rows    host, nlat: solid-angle weight of every point of row j
 *             (Gauss weight * 2 pi / nlon), symmetric
 *   coslat_rows host, nlat: cos(latitude) of each row (the kernels clamp it at
 *             1e-8 for products and mask |cos| < 1e-6 for velocities)
 *   max_fields: fields per launch sequence (>= 1); the workspaces are sized
 *             for it, and execute calls with more fields run in chunks of
 *             max_fields. */
int aeo_plan_create_gauss(int nlat, int nlon, int l_max, const double* x_nodes,
                          const double* w_rows, const double* coslat_rows,
                          int max_fields, aeo_plan_t* out);

/* Arbitrary point-set transform plan (dense basis, FP64).
 * Replaces PointSetSphericalHarmonics.__init__ (num/fast_geodesic_sh.py:15-66)
 * and, with Voronoi weights, GeodesicSphericalHarmonics (num/optimized_geodesic_sh.py:42-133).
 *   lat, lon, weights, coslat: host, npoints each (weights NULL = uniform 4 pi / N,
 *   num/fast_geodesic_sh.py:35-37; coslat = cos(lat), clamped in the kernels);
 *   max_fields as for aeo_plan_create_gauss. */
int aeo_plan_create_points(int npoints, int l_max, const double* lat, const double* lon,
                           const double* weights, const double* coslat, int max_fields,
                           aeo_plan_t* out);

int aeo_plan_destroy(aeo_plan_t plan);

/* Build the split-precision tables of a plan: the FP64 Legendre tables (Gauss)
 * or the dense Y basis (point sets) cut into `nslices` INT8 radix-254 digits
 * per entry with one power-of-two exponent per table row (2 <= nslices <= 8).
 * Models on this plan can then run their Legendre / Y-basis contractions as
 * INT8 MFMA products with exact INT32 accumulation and FP64 recombination
 * (aeo_model_set_precision).  No counterpart in the reference (its transforms
 * are FP64 only). */
int aeo_plan_enable_split(aeo_plan_t plan, int nslices);

/* Query: l_max, number of grid points, number of basis functions. */
int aeo_plan_info(aeo_plan_t plan, int* l_max, int* npoints, int* nbasis);

/* Copy the plan's dense basis row-major (npoints, B) complex128 into `Y_out`
 * (device).  Y[i, l(l+1)/2 + m] = Y_lm(point i).  Replaces the Y_matrix
 * attribute (num/fast_geodesic_sh.py:48-56). */
int aeo_plan_basis(aeo_plan_t plan, double* Y_out, void* stream);

/* Analysis of `nfields` grid fields -> dense coefficients.
 * Replaces PointSetSphericalHarmonics.transform (num/fast_geodesic_sh.py:68-87). */
int aeo_analysis(aeo_plan_t plan, const double* f, int nfields, double* a_out, void* stream);

/* Synthesis of `nfields` coefficient fields -> grid fields (uses Re(a_l0) only).
 * Replaces PointSetSphericalHarmonics.inv_transform (num/fast_geodesic_sh.py:89-107). */
int aeo_synthesis(aeo_plan_t plan, const double* a, int nfields, double* f_out, void* stream);

/* Spectral algebra on `nfields` dense coefficient fields (in may equal out
 * except for AEO_OP_SINTH_DTH and AEO_OP_SINTH_DTH_T).
 * Replaces the SpectralOperators methods named at each op code. */
int aeo_spec_op(int op, int l_max, double radius, const double* in, double* out,
                int nfields, void* stream);

/* Pseudospectral Jacobian J(a, b) on the product plan, analysed back to
 * coefficients and (dealias != 0) 2/3-truncated once.
 * Replaces SpectralOperators.jacobian_pseudospectral(a, b, dealias,
 * return_spectral=True) (num/spectral_operators.py:590-658). */
int aeo_jacobian(aeo_plan_t prod, double radius, const double* a, const double* b,
                 double* out, int dealias, void* stream);

/* Velocity (u east, v north) on the plan's points from streamfunction
 * coefficients, with the |cos(lat)| < 1e-6 -> 0 mask.
 * Replaces SpectralOperators.velocity_from_streamfunction (num/spectral_operators.py:517-538). */
int aeo_velocity(aeo_plan_t plan, double radius, const double* psi, double* u, double* v,
                 void* stream);

/* Gradient (d/dx, d/dy) of a scalar on the plan's points (unclamped cos(lat)).
 * Replaces SpectralOperators.grad_from_scalar (num/spectral_operators.py:542-558). */
int aeo_gradient(aeo_plan_t plan, double radius, const double* q, double* dqdx, double* dqdy,
                 void* stream);

/* Barotropic vorticity model on a product plan.
 * Replaces BarotropicVorticity.__init__ (phy/barotropic.py:21-66). */
int aeo_bve_create(aeo_plan_t prod, double radius, double omega, double viscosity,
                   aeo_model_t* out);

/* A batch of `nmembers` independent BVE states on one Gauss product plan
 * (same L, grid, radius, rotation, viscosity).  States are stacked
 * (nmembers, L+1, L+1) complex128.  Replaces nmembers BarotropicVorticity
 * instances stepped in lockstep (phy/barotropic.py:21-167). */
int aeo_bve_create_batch(aeo_plan_t prod, double radius, double omega, double viscosity,
                         int nmembers, aeo_model_t* out);

/* Shallow-water model on a product plan.  phi_s_lm: device (L+1)^2 complex
 * surface geopotential g*h_s, or NULL for a flat bottom.
 * Replaces ShallowWaterModel.__init__ (phy/shallow_water.py:172-283). */
int aeo_swe_create(aeo_plan_t prod, double radius, double omega, double gravity,
                   double mean_depth, double hyperdiffusion_nu4, const double* phi_s_lm,
                   aeo_model_t* out);

int aeo_model_destroy(aeo_model_t model);

/* 0 = FP64 Legendre / Y-basis contractions; k > 0 = split INT8 digits with k slices
 * (the plan must have aeo_plan_enable_split(plan, k)).  Everything else of the
 * step (recurrence tables, FFTs, products, spectral algebra, RK4 update)
 * stays FP64.  No counterpart in the reference. */
int aeo_model_set_precision(aeo_model_t model, int nslices);

/* Members per workgroup of the Legendre stage kernels (1, 2, 4 or 8): the
 * tile shape the autotuner chooses.  0 selects the member-batched GEMM form
 * of the split-precision BVE stage (analysis, epilogue and synthesis kernels
 * whose MFMA tiles span the member batch; bitwise the fused kernel's
 * results): BVE models on Gauss plans only, it allocates its workspaces
 * here, and steps fail with a state error while the precision is FP64.
 * No counterpart in the reference. */
int aeo_model_set_tiles(aeo_model_t model, int members_per_block);

/* Tile shape of the GEMM form (members_per_block 0 above): synthesis
 * columns (member x field x re/im) per workgroup, 64 or 128; members per
 * epilogue workgroup, 4 or 8; 16-column MFMA tiles per synthesis wave, 2 or 4
 * (defaults 64, 8, 4).  Results do not depend on it.  No counterpart in the
 * reference. */
int aeo_model_set_gemm_tiles(aeo_model_t model, int syn_cols, int epi_members, int syn_ntiles);

/* Tendency of one state (BVE: one field; SWE: the (3, L+1, L+1) stack zeta,
 * delta, phi).  forcing: device, same shape, or NULL (BVE only).
 * Replaces BarotropicVorticity.tendency (phy/barotropic.py:120-167) and
 * ShallowWaterModel.tendency (phy/shallow_water.py:347-418). */
int aeo_model_tendency(aeo_model_t model, const double* y, const double* forcing, double* out,
                       void* stream);

/* `nsteps` classical RK4 steps of size dt: y_out = RK4^nsteps(y).  y_out may
 * equal y.  forcing as for aeo_model_tendency.
 * Replaces rk4_step (run/bve/runner.py:211-216) and rk4_step_array without a
 * stage validator (run/engine.py:406-431), called nsteps times. */
int aeo_model_rk4(aeo_model_t model, const double* y, double dt, int nsteps,
                  const double* forcing, double* y_out, void* stream);

/* Adjoint of the BVE RK4 step (Gauss plans; models from aeo_bve_create or
 * aeo_bve_create_batch).  No counterpart in the reference: its only
 * adjoint-flavoured primitive is adjoint_sin_theta_d_theta_coeffs
 * (num/spectral_operators.py:243-270), which these kernels apply in place.
 *
 * Cotangents use the real Euclidean inner product of the dense (L+1, L+1)
 * complex128 arrays, <x, y> = sum Re(conj(x) y) (the convention of
 * torch.autograd for complex tensors).
 *
 * aeo_model_enable_adjoint: set-up call that allocates the model's adjoint
 * workspaces (all of them or, on failure, none); aeo_model_rk4_adjoint and
 * aeo_model_tendency_adjoint allocate nothing and return AEO_ERR_STATE
 * until it has succeeded.  Calling it again is a no-op.
 * aeo_model_checkpoint_elems: complex128 entries a checkpoint buffer for
 * nsteps steps needs (nsteps * 4 * members * (L+1)^2).
 * aeo_model_rk4_record: aeo_model_rk4 (bit-identical y_out) that also writes
 * the input state of every RK4 stage into `checkpoints` (device, caller
 * owned, opaque internal layout).
 * aeo_model_rk4_adjoint: given ybar_out = dJ/dy_out, ybar = dJ/dy for the
 * recorded run (same model, dt, nsteps; the constant forcing drops out).
 * ybar may equal ybar_out.
 * aeo_model_tendency_adjoint: ybar = T'(y)^T kbar for one tendency call
 * (forcing drops out). */
int aeo_model_enable_adjoint(aeo_model_t model);
int aeo_model_checkpoint_elems(aeo_model_t model, int nsteps, long long* out);
int aeo_model_rk4_record(aeo_model_t model, const double* y, double dt, int nsteps,
                         const double* forcing, double* y_out, double* checkpoints,
                         void* stream);
int aeo_model_rk4_adjoint(aeo_model_t model, const double* checkpoints, double dt, int nsteps,
                          const double* ybar_out, double* ybar, void* stream);
int aeo_model_tendency_adjoint(aeo_model_t model, const double* y, const double* kbar,
                               double* ybar, void* stream);

/* Bracket every launch of one kernel class of the model's stages with a
 * pair of HIP events on the launch stream (0 = Legendre synthesis, 1 = FFT
 * grid stage with the pointwise products, 2 = Legendre analysis with the
 * tendency/RK4 epilogue; -1 disables).  Events are created here, not in the
 * execute calls.  No counterpart in the reference (measurement support). */
int aeo_model_profile(aeo_model_t model, int kernel_class, int max_launches);

/* Sum of the bracketed launch durations (ms) and their count since the last
 * read; resets the count. */
int aeo_model_profile_read(aeo_model_t model, double* total_ms, int* launches);

/* An ensemble of member batches with different L (one Gauss model per L
 * class, e.g. from aeo_bve_create_batch); the ensemble does not own them.
 * No counterpart in the reference (it runs one state at a time). */
int aeo_ensemble_create(int ngroups, const aeo_model_t* models, aeo_ensemble_t* out);
int aeo_ensemble_destroy(aeo_ensemble_t ens);

/* `nsteps` RK4 steps of every group: y[g] -> y_out[g] (device, boundary
 * layout; y_out[g] may equal y[g]).  With use_graph != 0 two steps of all
 * groups are captured once (per dt) into one hipGraph with one concurrent
 * branch per group, and replayed nsteps/2 times on `stream`; an odd last
 * step runs eagerly.  Each group computes exactly what aeo_model_rk4 does.
 * Replaces running rk4_step (run/bve/runner.py:211-216) per member. */
int aeo_ensemble_rk4(aeo_ensemble_t ens, const double* const* y, double dt, int nsteps,
                     double* const* y_out, int use_graph, void* stream);

/* Spectral invariants of `nfields` vorticity fields with compensated sums:
 * out[4 f + 0..3] = circulation, energy, relative enstrophy, absolute
 * enstrophy (device FP64).
 * Replaces spectral_diagnostics (run/bve/diagnostics.py:52-97). */
int aeo_spectral_diag(int l_max, double radius, double omega, const double* zeta, int nfields,
                      double* out, void* stream);

enum aeo_grid_stats_mode {
    AEO_GRID_STATS_BVE = 0,    /* a = zeta, b = u, c = v */
    AEO_GRID_STATS_SWE = 1,    /* a = u, b = v, c = phi anomaly, d = phi_s or NULL */
    AEO_GRID_STATS_MINMAX = 2  /* a = field */
};

/* Per-step grid statistics of one sampling of `npts` points with quadrature
 * weights w (device FP64; one workgroup, fixed reduction order, compensated
 * sums), written to the device array out:
 *   BVE:    {max|zeta|, sum w zeta^2, sum w, max |u|}
 *           (DiagnosticsRecorder.record, run/bve/diagnostics.py:162-211)
 *   SWE:    k = {Phi0, mean phi_s, gravity}, Phi = Phi0 + phi:
 *           {max|u|, min Phi, max Phi, sum w [Phi|u|^2/2 + Phi^2/2 + Phi phi_s],
 *            min eta, max eta}, eta = (Phi - Phi0 + phi_s - mean phi_s)/gravity
 *           (SWEDiagnosticsRecorder.record, run/swe/diagnostics.py:113-173,
 *            characteristic_fields, phy/shallow_water.py:456-484)
 *   MINMAX: {min a, max a} (total_geopotential_extrema, phy/shallow_water.py:404-426)
 * One device-to-host read of `out` then gives the whole row. */
int aeo_grid_stats(int mode, int npts, const double* w, const double* a, const double* b,
                   const double* c, const double* d, const double* k, double* out, void* stream);

/* Weak-form vector analysis of `nfields` tangent fields (east, north
 * components, device (nfields, npoints) each, on the points of `ext`):
 *   a = analysis_{L+1}(F_east / cos lat), b = analysis_{L+1}(F_north / cos lat),
 *   div_lm  = (i m / R) a_lm + (S^T_ext b)_lm / R,
 *   curl_lm = (i m / R) b_lm - (S^T_ext a)_lm / R,
 * (S^T_ext c)_lm = C+_lm c_{l+1,m} + C-_lm c_{l-1,m} with the unclipped C+ at
 * l = L; outputs (nfields, L+1, L+1) complex128, 2/3-truncated when
 * `truncate` != 0; cos lat clamped at 1e-8.  `ext` is a plan of degree L+1 on
 * the product sampling (same points and weights).  `work`: device, at least
 * aeo_vector_curl_div_work(ext, nfields) doubles.  The analysis uses the
 * plan's own workspace: calls on one plan must not overlap on different
 * streams.
 * Replaces SpectralOperators.vector_curl_div_spectral and
 * _vector_analysis_pieces (num/spectral_operators.py:399-502). */
int aeo_vector_curl_div_work(aeo_plan_t ext, int nfields, long long* n_doubles);
int aeo_vector_curl_div(aeo_plan_t ext, double radius, const double* f_east,
                        const double* f_north, int nfields, int truncate, double* curl_out,
                        double* div_out, double* work, void* stream);

/* aeo_vector_curl_div without the `work` argument, for callers that cannot
 * allocate device memory (code under stream/graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a fixed 16 MiB device array inside the library, one copy per
 * GPU, held from the library's load until the process ends; the call
 * allocates nothing and returns AEO_ERR_UNSUPPORTED when
 * aeo_vector_curl_div_work(ext, nfields) exceeds 2^21 doubles (for example
 * more than 21 fields on a 128 x 256 product grid at L = 85).  Every call on
 * one GPU shares that array (and the plan's own workspace): calls that
 * overlap on one GPU, on any streams, race.  Results equal
 * aeo_vector_curl_div's bitwise. */
int aeo_vector_curl_div_noalloc(aeo_plan_t ext, double radius, const double* f_east,
                                const double* f_north, int nfields, int truncate,
                                double* curl_out, double* div_out, void* stream);

/* Dry hydrostatic primitive-equation model on a product plan `prod` (degree
 * L) and its degree-(L+1) companion `ext` on the same points.
 *   interfaces: host, nlev + 1 sigma interfaces, exactly 0 first and 1 last,
 *               strictly increasing (phy/sigma_coordinate.py:43-81)
 *   phi_surface_lm: device (L+1)^2 complex128 surface geopotential, or NULL
 *               for zero (its band limit is the caller's check,
 *               phy/primitive_equations.py:209-238)
 * Replaces PrimitiveEquationsModel.__init__ (phy/primitive_equations.py:180-264). */
int aeo_pe_create(aeo_plan_t prod, aeo_plan_t ext, int nlev, const double* interfaces,
                  double radius, double omega, double r_dry, double cp_dry,
                  const double* phi_surface_lm, aeo_pe_t* out);
int aeo_pe_destroy(aeo_pe_t pe);

/* d/dt of the (3 nlev + 1, L+1, L+1) complex128 stack
 * [zeta_1..K, delta_1..K, T_1..K, ln p_s] (device; out must not alias y).
 * The model's scratch and its two plans' workspaces are shared by its calls:
 * calls on one model must not overlap on different streams.
 * Replaces PrimitiveEquationsModel.tendency (phy/primitive_equations.py:482-753). */
int aeo_pe_tendency(aeo_pe_t pe, const double* y, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AEOLUS_HIP_H */
