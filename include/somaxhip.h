/*
 * somaxhip -- C ABI of the MI355X-native somax QG / shallow-water time step.
 *
 * Plain C: pointers, sizes and scalars only.  Every entry point returns an
 * int status (SOMAXHIP_OK == 0); the text of the last error on the calling
 * thread is available from somaxhip_last_error().  No C++ exception crosses
 * this boundary.  All device buffers are caller-owned and must stay valid
 * until the work queued on `stream` has finished; nothing is retained past
 * return except inside a plan.  Work is queued on the caller's hipStream_t
 * (pass NULL for the default stream) and nothing here synchronises the host.
 *
 * Layouts are those of the reference state pytrees: a field of a model with
 * nx x ny interior cells and nl layers is float32[nl][ny+2][nx+2], row-major,
 * x fastest, ghost ring included (somax/_src/models/qg/baroclinic.py:34,
 * swm/multilayer.py:38-40).
 *
 * Each entry point names the reference interface it stands in for.  The
 * reference is Python/JAX (jejjohnson/somax v0.0.6); its arithmetic lives in
 * finitevolX v0.0.39 and diffrax, called from the lines cited.
 */
#ifndef SOMAXHIP_H
#define SOMAXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t without pulling in the HIP headers. */
typedef struct ihipStream_t* somaxhip_stream_t;

enum {
  SOMAXHIP_OK = 0,
  SOMAXHIP_ERR_ARG = 1,     /* invalid argument (shape, pointer, flag) */
  SOMAXHIP_ERR_HIP = 2,     /* a HIP runtime call failed */
  SOMAXHIP_ERR_ALLOC = 3,   /* device allocation failed */
  SOMAXHIP_ERR_DEVICE = 4   /* no gfx950 device / wrong architecture */
};

/* Advection of total PV. */
enum {
  SOMAXHIP_ADV_ARAKAWA = 0, /* reference: finitevolx.arakawa_jacobian, baroclinic.py:175 */
  SOMAXHIP_ADV_WENO = 1     /* extension: flux-form WENO-Z (5th order, 3rd/1st at walls) */
};

/* Arithmetic of the DST-I Helmholtz inversion. */
enum {
  SOMAXHIP_PREC_FP32 = 0,   /* plain: FP32 MFMA sine-transform GEMMs */
  SOMAXHIP_PREC_MIXED = 1,  /* BF16 hi/lo 3-product MFMA, FP32 accumulate, one
                               iterative-refinement pass on the Helmholtz residual */
  SOMAXHIP_PREC_BF16X6 = 2, /* extension: BF16 hi/mid/lo three-piece operands, six
                               products per element with FP64 accumulation flushes,
                               one direct solve (no refinement pass) */
  SOMAXHIP_PREC_FP64 = 3    /* the whole step in FP64, as the reference runs on the CPU
                               with x64 enabled: FP64 stencils, projections and DST-I
                               GEMMs; single models only, through the *_f64 calls */
};

/* ------------------------------------------------------------------------ */
/* Library                                                                  */
/* ------------------------------------------------------------------------ */

/* Copies the last error message of the calling thread (NUL-terminated,
 * truncated to len).  Returns the full message length. */
size_t somaxhip_last_error(char* buf, size_t len);

/* Checks that the current device is gfx950 and warms the runtime.
 * No reference counterpart (JAX initialises XLA lazily). */
int somaxhip_init(void);

/* Library version as major*10000 + minor*100 + patch. */
int somaxhip_version(void);

/* ------------------------------------------------------------------------ */
/* Multilayer / barotropic QG                                               */
/* ------------------------------------------------------------------------ */

/* Immutable per-model tables (sine matrices, eigenvalue tables, modal
 * matrices, forcing profiles, homogeneous solutions).  Built once per model
 * by the host from create() (baroclinic.py:231-332, barotropic.py:186-248). */
typedef struct somaxhip_qg_plan_s* somaxhip_qg_plan;

typedef struct {
  int nx, ny;                 /* interior cells */
  int nl;                     /* layers (1 for BarotropicQG) */
  double dx, dy;              /* grid spacing (ArakawaCGrid2D.from_interior) */
  const double* cl2m;         /* nl*nl, row-major: modal[m] = sum_l cl2m[m*nl+l]*layer[l] */
  const double* cm2l;         /* nl*nl, row-major: layer[l] = sum_m cm2l[l*nl+m]*modal[m] */
  const double* lambdas;      /* nl Helmholtz parameters, f0^2 * eigenvalue (baroclinic.py:296) */
  const double* beta_y_row;   /* ny+2 values beta*(y_j - Ly/2), y_j = j*dy (baroclinic.py:307-311) */
  const double* wind_row;     /* ny+2 values of the wind-curl pattern (baroclinic.py:313-318) */
  const double* wind_scale;   /* nl multipliers of tau0*pattern: 1/H0 on layer 0 (baroclinic.py:181)
                                 or 1 (barotropic.py:143); 0 elsewhere */
  int zero_psi_ghost;         /* 1: zero_boundaries(psi) after to_layer (baroclinic.py:158);
                                 0: keep the DST ghost values (barotropic.py:119-121) */
  int advection;              /* SOMAXHIP_ADV_* */
  int precision;              /* SOMAXHIP_PREC_* */
  int mass_correction;        /* extension: add alpha_m * homsol_m so each mode has zero mean */
  const float* homsol;        /* nl*(ny+2)*(nx+2) host array, required iff mass_correction */
  const double* homsol_mean;  /* nl means over the full array, required iff mass_correction */
} somaxhip_qg_desc;

/* Differentiable parameters (BaroclinicQGParams, baroclinic.py:37-48), passed
 * per call so they may change between calls without a new plan. */
typedef struct {
  double nu;     /* lateral_viscosity: + nu * laplacian(q) (baroclinic.py:188) */
  double nu4;    /* extension: - nu4 * laplacian(laplacian(q)) */
  double kappa;  /* bottom_drag: - kappa * laplacian(psi[-1]) (baroclinic.py:184-185) */
  double tau0;   /* wind_amplitude (baroclinic.py:181) */
} somaxhip_qg_params;

/* Builds the plan on the current device (host tables are copied).  A plane of
 * (nx + 2) * (ny + 2) cells must stay below 2^31 (SOMAXHIP_ERR_ARG otherwise). */
int somaxhip_qg_plan_create(const somaxhip_qg_desc* desc, somaxhip_qg_plan* plan);
int somaxhip_qg_plan_destroy(somaxhip_qg_plan plan);

/* Bytes of caller-provided device workspace that the calls below need. */
size_t somaxhip_qg_workspace_bytes(somaxhip_qg_plan plan);

/* psi = _invert_pv(q), or _invert_pv(apply_boundary_conditions(q)) when
 * apply_bc != 0 (the form every RHS evaluation uses).
 * Replaces BaroclinicQG._invert_pv (baroclinic.py:135-159) /
 * BarotropicQG._invert_pv (barotropic.py:113-121). */
int somaxhip_qg_invert(somaxhip_qg_plan plan, const float* q, float* psi, int apply_bc,
                       void* workspace, somaxhip_stream_t stream);

/* dq = vector_field(t, state).q with state.q = q, or, when apply_bc != 0,
 * vector_field(t, apply_boundary_conditions(state)).q as composed by
 * SomaxModel.build_terms (core/model.py:47-49).  psi_out (nullable) receives
 * the streamfunction.  Replaces BaroclinicQG.vector_field
 * (baroclinic.py:161-190) / BarotropicQG.vector_field (barotropic.py:123-152). */
int somaxhip_qg_rhs(somaxhip_qg_plan plan, const float* q, float* dq, float* psi_out,
                    const somaxhip_qg_params* params, int apply_bc, void* workspace,
                    somaxhip_stream_t stream);

/* diagnose(state) (baroclinic.py:197-228 / barotropic.py:159-183): psi =
 * _invert_pv(q); u = -d(psi)/dy on V faces, v = d(psi)/dx on U faces, zeta =
 * laplacian(psi) (u, v, zeta nullable); ke_enstrophy (device, 2*nl doubles)
 * receives per layer the kinetic energy 0.5 sum (uT^2 + vT^2) dx dy and the
 * enstrophy 0.5 sum q^2 dx dy over the interior, summed in FP64 in a fixed
 * order. */
int somaxhip_qg_diagnose(somaxhip_qg_plan plan, const float* q, float* psi, float* u, float* v, float* zeta,
                         double* ke_enstrophy, void* workspace, somaxhip_stream_t stream);

/* Tsit5 time stepping: the solver SomaxModel.integrate uses by default
 * (core/model.py:75, diffrax.Tsit5 with ConstantStepSize).  Advances q in
 * place by nsteps steps of dt with the boundary condition applied before
 * every RHS; six RHS evaluations per step (first same as last), each an
 * inversion plus the fused RHS kernel.  tsit5_workspace: a second caller
 * buffer of somaxhip_qg_tsit5_workspace_bytes (stage state and the seven
 * stage derivatives). */
size_t somaxhip_qg_tsit5_workspace_bytes(somaxhip_qg_plan plan);
int somaxhip_qg_step_tsit5(somaxhip_qg_plan plan, float* q, int nsteps, double dt, const somaxhip_qg_params* params,
                           void* workspace, void* tsit5_workspace, somaxhip_stream_t stream);

/* Ghost ring of nl fields of (ny+2) x (nx+2) set to 0 in place.
 * Replaces finitevolx.zero_boundaries as used by
 * BaroclinicQG.apply_boundary_conditions (baroclinic.py:192-195). */
int somaxhip_zero_ghost(float* field, int nl, int ny, int nx, somaxhip_stream_t stream);

/* FP64 forms of somaxhip_qg_invert, _rhs, _step (SSP-RK3, below), _step_tsit5,
 * _diagnose and somaxhip_zero_ghost, for a plan created with precision
 * SOMAXHIP_PREC_FP64 (the FP32 / BF16 calls, the masked solve and the
 * reparameterized step refuse such a plan, and these refuse any other): the
 * same arguments with double fields; the workspace sizes
 * come from somaxhip_qg_workspace_bytes / somaxhip_qg_tsit5_workspace_bytes
 * of the plan.  They replace the same reference functions evaluated under
 * jax_enable_x64: BarotropicQG / BaroclinicQG._invert_pv (barotropic.py:113-121,
 * baroclinic.py:135-159), vector_field (barotropic.py:123-152,
 * baroclinic.py:161-190), diagnose (barotropic.py:159-183,
 * baroclinic.py:197-228), SomaxModel.integrate with Tsit5 (core/model.py:53-88),
 * and zero_boundaries (baroclinic.py:192-195). */
int somaxhip_qg_invert_f64(somaxhip_qg_plan plan, const double* q, double* psi, int apply_bc,
                           void* workspace, somaxhip_stream_t stream);
int somaxhip_qg_rhs_f64(somaxhip_qg_plan plan, const double* q, double* dq, double* psi_out,
                        const somaxhip_qg_params* params, int apply_bc, void* workspace,
                        somaxhip_stream_t stream);
int somaxhip_qg_step_f64(somaxhip_qg_plan plan, double* q, double* psi_out, int nsteps, double dt,
                         const somaxhip_qg_params* params, void* workspace, somaxhip_stream_t stream);
int somaxhip_qg_step_tsit5_f64(somaxhip_qg_plan plan, double* q, int nsteps, double dt,
                               const somaxhip_qg_params* params, void* workspace, void* tsit5_workspace,
                               somaxhip_stream_t stream);
int somaxhip_qg_diagnose_f64(somaxhip_qg_plan plan, const double* q, double* psi, double* u, double* v,
                             double* zeta, double* ke_enstrophy, void* workspace, somaxhip_stream_t stream);
int somaxhip_zero_ghost_f64(double* field, int nl, int ny, int nx, somaxhip_stream_t stream);

/* Advances q in place by nsteps SSP-RK3 steps of size dt (Shu-Osher form of
 * the ssrk3_update_stage_{1,2,3} scheme, notebooks/dev/QG_Multilayer/
 * mqg_demo_retangular.ipynb cell 31), applying the boundary condition before
 * every RHS like SomaxModel.build_terms.  Stands in for the hot loop of
 * SomaxModel.integrate (core/model.py:53-88).  psi_out (nullable) receives
 * the streamfunction of the last stage's input state. */
int somaxhip_qg_step(somaxhip_qg_plan plan, float* q, float* psi_out, int nsteps, double dt,
                     const somaxhip_qg_params* params, void* workspace,
                     somaxhip_stream_t stream);

/* Size of the library-owned device array that somaxhip_qg_step_noalloc uses
 * as its workspace (64 MiB: plans up to 3 layers at 512 x 512). */
#define SOMAXHIP_NOALLOC_SCRATCH_BYTES (64u << 20)

/* somaxhip_qg_step without a caller workspace.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The step's workspace is a static device array of
 * SOMAXHIP_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds from
 * module load until the process ends; the call itself allocates nothing.
 * Results equal somaxhip_qg_step's bit for bit.  A plan whose
 * somaxhip_qg_workspace_bytes exceeds the array is refused with
 * SOMAXHIP_ERR_ARG.  Every call on one GPU shares that GPU's array, which
 * keeps the last call's contents (the step needs no zeroed workspace): calls
 * that overlap in time on one GPU -- other streams or other host threads --
 * race and must be serialised by the caller. */
int somaxhip_qg_step_noalloc(somaxhip_qg_plan plan, float* q, float* psi_out, int nsteps, double dt,
                             const somaxhip_qg_params* params, somaxhip_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Multilayer shallow water                                                 */
/* ------------------------------------------------------------------------ */

/* Boundary conditions of MultilayerShallowWater2D (multilayer.py:203-223). */
enum {
  SOMAXHIP_BC_PERIODIC = 0, /* enforce_periodic on h, u, v (multilayer.py:213-216) */
  SOMAXHIP_BC_WALL = 1,     /* free-slip walls (multilayer.py:217-222, 383-410) */
  SOMAXHIP_BC_CHANNEL = 2   /* extension: periodic in x, the wall rules in y */
};

/* PV flux of the momentum equations. */
enum {
  SOMAXHIP_VFLUX_AVERAGE = 0,          /* reference: q averaged to the face times the four-point average of
                                          the other mass flux (multilayer.py:172-187) */
  SOMAXHIP_VFLUX_ENERGY_ENSTROPHY = 1  /* extension: Arakawa & Lamb (1981) triads (the "EEN" form); does no
                                          work, and keeps the potential enstrophy of a non-divergent mass
                                          flux */
};

typedef struct somaxhip_swm_plan_s* somaxhip_swm_plan;

typedef struct {
  int nx, ny;               /* interior cells */
  int nl;                   /* layers */
  double dx, dy;
  const double* g_prime;    /* nl reduced gravities; p_k = sum_{j<=k} g'_j h_j (multilayer.py:180-182) */
  const double* H;          /* nl layer thicknesses; the wind is divided by H[0] (multilayer.py:190) */
  const double* f_row;      /* ny+2 values f0 + beta*(y_j - Ly/2), y_j = j*dy (multilayer.py:345-349) */
  const double* wind_row;   /* ny+2 values of the x wind-stress pattern (multilayer.py:352-356) */
  int bc;                   /* SOMAXHIP_BC_* */
  int vorticity_flux;       /* SOMAXHIP_VFLUX_* */
} somaxhip_swm_desc;

/* MultilayerSW2DParams (multilayer.py:43-55), passed per call. */
typedef struct {
  double nu;     /* lateral_viscosity: + nu * laplacian of u and v (multilayer.py:194-195) */
  double kappa;  /* bottom_drag: - kappa * (u, v) of the bottom layer (multilayer.py:198-199) */
  double tau0;   /* wind_amplitude (multilayer.py:190-191) */
} somaxhip_swm_params;

/* Plan of one MultilayerShallowWater2D.create (multilayer.py:259-377).  A
 * plane of (nx + 2) * (ny + 2) cells must stay below 2^31. */
int somaxhip_swm_plan_create(const somaxhip_swm_desc* desc, somaxhip_swm_plan* plan);
int somaxhip_swm_plan_destroy(somaxhip_swm_plan plan);
size_t somaxhip_swm_workspace_bytes(somaxhip_swm_plan plan);

/* (dh, du, dv) = vector_field(t, state) (multilayer.py:150-201) with state
 * (h, u, v), each float32[nl][ny+2][nx+2]; apply_bc != 0 applies
 * apply_boundary_conditions first, as SomaxModel.build_terms does. */
int somaxhip_swm_rhs(somaxhip_swm_plan plan, const float* h, const float* u, const float* v, float* dh,
                     float* du, float* dv, const somaxhip_swm_params* params, int apply_bc,
                     somaxhip_stream_t stream);

/* (ho, uo, vo) = apply_boundary_conditions(h, u, v) (multilayer.py:203-223);
 * outputs must not alias inputs. */
int somaxhip_swm_apply_bc(somaxhip_swm_plan plan, const float* h, const float* u, const float* v, float* ho,
                          float* uo, float* vo, somaxhip_stream_t stream);

/* diagnose(state) (multilayer.py:225-255): PV q = (zeta + f)/h_X and relative
 * vorticity zeta at X points, kinetic energy at T points (each nullable);
 * energy_enstrophy (device, 2*nl doubles) receives per layer the energy
 * sum (KE + 0.5 g' h^2) dx dy and the potential enstrophy sum 0.5 q^2 h_X dx dy
 * over the interior, summed in FP64 in a fixed order.  Uses the workspace. */
int somaxhip_swm_diagnose(somaxhip_swm_plan plan, const float* h, const float* u, const float* v, float* q,
                          float* zeta, float* ke, double* energy_enstrophy, void* workspace,
                          somaxhip_stream_t stream);

/* Advances (h, u, v) in place by nsteps SSP-RK3 steps of dt with the
 * boundary condition applied before every RHS (the hot loop of
 * SomaxModel.integrate, core/model.py:53-88). */
int somaxhip_swm_step(somaxhip_swm_plan plan, float* h, float* u, float* v, int nsteps, double dt,
                      const somaxhip_swm_params* params, void* workspace, somaxhip_stream_t stream);

/* Tsit5 time stepping of (h, u, v) in place (the default solver of
 * SomaxModel.integrate, core/model.py:75), as somaxhip_qg_step_tsit5;
 * tsit5_workspace of somaxhip_swm_tsit5_workspace_bytes. */
size_t somaxhip_swm_tsit5_workspace_bytes(somaxhip_swm_plan plan);
int somaxhip_swm_step_tsit5(somaxhip_swm_plan plan, float* h, float* u, float* v, int nsteps, double dt,
                            const somaxhip_swm_params* params, void* tsit5_workspace, somaxhip_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Reparameterized QG (QG = shallow water + geostrophic projection)          */
/* ------------------------------------------------------------------------ */

/* ReparameterizedQG.project (qg/reparameterized.py:142-178): P = G (QG)^-1 Q
 * of a shallow-water state (h, u, v), each float32[nl][ny+2][nx+2]:
 *   q = zeta(u, v) - f0 (h - H) / H   (zeta at X points, forward differences)
 *   psi = the Helmholtz inversion of `plan`, a QG plan of the same grid and
 *         layers with zero_psi_ghost = 1 and lambdas = f0^2 * eigenvalues
 *   (u, v) = grad_perp(psi) = (-dpsi/dy, dpsi/dx) (backward differences)
 *   h = H (1 + f0 M psi),  M = Cm2l diag(eigenvalues) Cl2m (nl x nl, row-major)
 * h_deviation != 0: h and ho hold h - H instead.  psi_out nullable.  Uses the
 * plan's workspace. */
int somaxhip_qg_project(somaxhip_qg_plan plan, double f0, const double* H, const double* M, const float* h,
                        const float* u, const float* v, int h_deviation, float* ho, float* uo, float* vo,
                        float* psi_out, void* workspace, somaxhip_stream_t stream);

enum {
  SOMAXHIP_SOLVER_SSPRK3 = 0,
  SOMAXHIP_SOLVER_TSIT5 = 1
};

/* Advances a ReparameterizedQG state (h, u, v) in place by nsteps steps of
 * dt of y' = RHS(P(BC(y))): the shallow-water RHS of `swm` after its boundary
 * condition and the projection of `qg` (ReparameterizedQG.build_terms with
 * apply_boundary_conditions = project o swm BC, reparameterized.py:186-189).
 * qg_workspace: the QG plan's; tsit5_workspace of somaxhip_rqg_workspace_bytes. */
size_t somaxhip_rqg_workspace_bytes(somaxhip_swm_plan swm);
int somaxhip_rqg_step(somaxhip_swm_plan swm, somaxhip_qg_plan qg, int solver, double f0, const double* H,
                      const double* M, float* h, float* u, float* v, int nsteps, double dt,
                      const somaxhip_swm_params* params, void* qg_workspace, void* workspace,
                      somaxhip_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Helmholtz solve on a masked domain (capacitance-matrix method)           */
/* ------------------------------------------------------------------------ */

/* Replaces inverse_elliptic_dst_cmm / compute_capacitance_matrices of the
 * former somax._src.operators.dst, used in
 * notebooks/dev/QG_Multilayer/helmholtz_decomp.ipynb (cells 19-33) and
 * mqg_demo_octagonal.ipynb (cell 12); the module is no longer in the tree.
 * Solves (lap - lam_p) psi_p = rhs_p on the points where mask = 1, with
 * psi = 0 elsewhere, for every plane p of float32[P][Ny][Nx] fields.
 * plan: a QG plan whose inversion is the rectangular solve G: identity
 * modal matrices (cl2m = cm2l = I), lambdas = lam_p, zero_psi_ghost = 0,
 * no mass correction; it must outlive the solver.
 * mask: host float[Ny*Nx] of 0 / 1.  bidx: host int[nb], the flat indices
 * (row * Nx + column) of the points outside the mask that touch it in a 3x3
 * window.  cap: host double[P][nb][nb], per plane the inverse of the matrix
 * G[i][j] = (G e_{b_j})[b_i].  A solve is
 *   r0 = rhs * mask;  psi1 = G r0;  r0[b] = -cap psi1[b];  psi = (G r0) * mask. */
typedef struct somaxhip_cmm_s* somaxhip_cmm;
int somaxhip_cmm_create(somaxhip_qg_plan plan, const float* mask, int nb, const int* bidx, const double* cap,
                        somaxhip_cmm* cmm);
int somaxhip_cmm_destroy(somaxhip_cmm cmm);
/* Workspace of a solve: the plan's workspace followed by two fields. */
size_t somaxhip_cmm_workspace_bytes(somaxhip_cmm cmm);
int somaxhip_cmm_solve(somaxhip_cmm cmm, const float* rhs, float* psi, void* workspace, somaxhip_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Run diagnostics (NaN flags and field ranges between integration chunks)  */
/* ------------------------------------------------------------------------ */

#define SOMAXHIP_STATS_MAX_FIELDS 64

/* The per-field reduction of the runner's chunk protocol: _state_diagnostics
 * (cli/_run.py:408-439, min / mean / max / NaN count of each State field),
 * _has_non_finite (:499-501) and the non-finite counts of _assert_finite_state
 * (:703-753), computed on the device instead of after a full-field host copy.
 * fields[f]: device float32 pointer (host array of nfields pointers, any
 * alignment), counts[f] elements.  stats: device double[5 * nfields], per
 * field { min, max, sum, n_nan, n_nonfinite }, where min, max and sum skip
 * NaN only (numpy's nanmin / nanmax / nansum: +-Inf take part) and
 * n_nonfinite counts NaN and +-Inf (numpy's ~isfinite).  An empty field gives
 * { +Inf, -Inf, 0, 0, 0 }.  Deterministic; asynchronous on stream.
 * workspace: somaxhip_stats_workspace_bytes(nfields) bytes (0 if nfields is
 * out of range). */
size_t somaxhip_stats_workspace_bytes(int nfields);
int somaxhip_field_stats(const float* const* fields, const size_t* counts, int nfields, double* stats,
                         void* workspace, somaxhip_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Ensembles                                                                */
/* ------------------------------------------------------------------------ */

/* Extension with no reference counterpart: the reference advances one model
 * per SomaxModel.integrate call (core/model.py:53-88); an ensemble there is a
 * Python loop (or jax.vmap over identical grids).  Here members of any grid
 * size and layer count advance together.  Members with the same grid, advection
 * and precision form a class; each class runs every kernel of the step in one
 * launch over all its planes, and classes are queued largest grid first. */
typedef struct somaxhip_ens_s* somaxhip_ens;

enum {
  SOMAXHIP_ENS_CONCURRENT = 1, /* each class (or launch group) on its own stream, forked from and joined to the caller's */
  SOMAXHIP_ENS_GRAPH = 2,      /* capture one step as a hipGraph and replay it */
  SOMAXHIP_ENS_MERGED = 4      /* classes that share precision, advection, biharmonic flag and stage tile
                                  height form a launch group (up to 8 classes): every kernel phase of a
                                  Runge-Kutta stage is one launch over the whole group (a work table of
                                  the classes' blocks); the arithmetic per member is unchanged */
};

/* members[i] / params[i], i < n: one QG model each (as for plan_create);
 * flags: SOMAXHIP_ENS_* bits. */
int somaxhip_ens_create(const somaxhip_qg_desc* members, const somaxhip_qg_params* params, int n, int flags,
                        somaxhip_ens* ens);
int somaxhip_ens_destroy(somaxhip_ens ens);

/* The packed ensemble state: one float32 buffer of state_floats elements;
 * member i's field float32[nl][ny+2][nx+2] starts at member_offset(i). */
size_t somaxhip_ens_state_floats(somaxhip_ens ens);
long somaxhip_ens_member_offset(somaxhip_ens ens, int member); /* -1 on error */
size_t somaxhip_ens_workspace_bytes(somaxhip_ens ens);
int somaxhip_ens_num_classes(somaxhip_ens ens);

/* Class cls (0 .. num_classes-1, largest grid first): interior size, planes,
 * members, whether its transforms are folded, its stage tile height. */
int somaxhip_ens_class_info(somaxhip_ens ens, int cls, int* nx, int* ny, int* planes, int* members, int* fold,
                            int* stage_rows);

/* Tile height (8 or 16 rows of 64 columns) of class cls's Arakawa stage
 * kernel: a knob for the ensemble autotuner; results do not depend on it.
 * Only before the ensemble's first step: afterwards SOMAXHIP_ERR_ARG. */
int somaxhip_ens_set_stage_rows(somaxhip_ens ens, int cls, int rows);

/* Replaces every member's parameters (n entries, in member order). */
int somaxhip_ens_set_params(somaxhip_ens ens, const somaxhip_qg_params* params, somaxhip_stream_t stream);

/* Advances every member by nsteps SSP-RK3 steps of size dt, in place, with
 * the arithmetic of somaxhip_qg_step.  Under SOMAXHIP_ENS_GRAPH the captured
 * step is kept and replayed while q, workspace and dt stay the same. */
int somaxhip_ens_step(somaxhip_ens ens, float* q, int nsteps, double dt, void* workspace,
                      somaxhip_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Measurement                                                              */
/* ------------------------------------------------------------------------ */

/* Per-kernel-class totals of one profiled run: launches, summed duration
 * (HIP events around every launch on the caller's stream) and the
 * algorithmic flops and bytes of those launches. */
typedef struct {
  char name[32];
  int launches;
  double total_ms;
  double flops;
  double bytes;
} somaxhip_kernel_stat;

/* somaxhip_qg_step with every launch bracketed by HIP events; synchronises
 * the stream before returning.  No reference counterpart (the reference only
 * times whole chunks with time.perf_counter, cli/_run.py:289-307). */
int somaxhip_qg_step_profile(somaxhip_qg_plan plan, float* q, int nsteps, double dt,
                             const somaxhip_qg_params* params, void* workspace,
                             somaxhip_stream_t stream, somaxhip_kernel_stat* stats,
                             int max_stats, int* nstats);

/* somaxhip_swm_step with every stage launch bracketed by HIP events, read
 * after the step (one synchronisation at the end, none between launches);
 * stats[0] = "swm_stage". */
int somaxhip_swm_step_profile(somaxhip_swm_plan plan, float* h, float* u, float* v, int nsteps, double dt,
                              const somaxhip_swm_params* params, void* workspace, somaxhip_stream_t stream,
                              somaxhip_kernel_stat* stats, int max_stats, int* nstats);

/* somaxhip_ens_step run eagerly on the caller's stream, classes one after
 * another, every launch timed; synchronises the stream before returning. */
int somaxhip_ens_step_profile(somaxhip_ens ens, float* q, int nsteps, double dt, void* workspace,
                              somaxhip_stream_t stream, somaxhip_kernel_stat* stats, int max_stats,
                              int* nstats);

#ifdef __cplusplus
}
#endif

#endif /* SOMAXHIP_H */
