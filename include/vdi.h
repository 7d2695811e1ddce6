/* vdi.h -- C ABI of the MI355X (gfx950) per-timestep fluid solve.
 *
 * The reference (chr1shr/vdropimpact) has no FFI layer: its boundary is the
 * C++ class API of fluid_2d, mgs_mac, mgs_fem, visco_impl and gas_layer.
 * Every entry point below replaces one of those calls; the comment on each
 * cites the reference interface it stands in for.  Plain C types only: no
 * HIP or torch types cross this boundary.
 *
 * Conventions
 *   - Every function returns an int status (VDI_OK on success).  The
 *     reference is fail-stop (fatal_error -> exit(1), common.cc:44-47);
 *     here the caller decides, and vdi_status_string() gives the reference's
 *     message for each status.
 *   - A context owns all of its device memory and one HIP stream.  Host
 *     buffers passed in are copied, never retained.
 *   - A context is used by one host thread at a time (like the reference's
 *     objects, which share the src array between solvers).
 *   - Padded cell arrays have the reference layout: (m+4) x (n+4) doubles,
 *     cell (i,j), i in [-2,m+1], j in [-2,n+1], at (i+2) + (m+4)(j+2)
 *     (fluid_2d.cc:18, fbase/fm).
 */
#ifndef VDI_ABI_H
#define VDI_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define VDI_OK 0
#define VDI_ERR_MG_NOCONV 1     /* "Multigrid routine for ... failed" (mgs_mac.hh:68-70, mgs_fem.hh:75-77, visco_impl.cc:34-37) */
#define VDI_ERR_NEWTON_NOCONV 2 /* "Too many Newton steps" (gas_layer_solve.cc:204) */
#define VDI_ERR_TRIDIAG 3       /* "LAPACK routine failed" (tri_solve.cc:29-32) */
#define VDI_ERR_HIP 4           /* HIP runtime error (no reference counterpart) */
#define VDI_ERR_ARG 5           /* invalid argument (no reference counterpart) */

/* ---- linear-system kinds (the four solvers of fluid_2d, fluid_2d.hh:135-144) ---- */
#define VDI_MAC 0
#define VDI_FEM 1
#define VDI_VX 2
#define VDI_VY 3

/* ---- arithmetic forms ---- */
#define VDI_FORM_FP64 0   /* plain FP64 V-cycles */
#define VDI_FORM_MIXED 1  /* FP32 smoothing and transfers inside FP64 defect correction */

/* ---- array identifiers for vdi_get_array / vdi_set_array ---- */
enum {
    VDI_U = 0, VDI_V, VDI_P,                        /* padded (m+4)(n+4): field::u,v,p */
    VDI_UL, VDI_VL, VDI_UD, VDI_VD, VDI_C0, VDI_C1, /* padded: edge velocities, c0/c1 */
    VDI_SRC,                                        /* (m+1)(n+1): shared RHS, s_bc = src+mn */
    VDI_GP, VDI_GTAU, VDI_GP_OLD, VDI_GTAU_OLD,     /* gas_layer p[me], tau[ml], p_old, tau_old */
    VDI_H, VDI_HX, VDI_HT, VDI_FBD, VDI_PG, VDI_PG_OLD, /* gas_layer_solve h[m],hx[me],ht[m],fbd[m],pg[me],pg_old[me] */
    VDI_Z_MAC, VDI_Z_FEM, VDI_Z_VX, VDI_Z_VY,       /* solver solutions mgs_common::z */
    VDI_NARRAYS
};

/* Constants the step needs, derived from a reference .cfg exactly as
 * fileinfo.cc:141-152 and fluid_2d.cc:15-17 derive them.  Field order is
 * part of the ABI. */
typedef struct vdi_params {
    int m, n;              /* grid cells (fileinfo::m, n) */
    int x_sym;             /* fileinfo::x_sym */
    int implicit_visc;     /* fileinfo::implicit_visc */
    double ax, bx, ay, by; /* domain (fileinfo.cc:148-152) */
    double rho, mu;        /* liquid density and dynamic viscosity mu = nul*rho */
    double dt_reg;         /* fluid_2d::dt_reg (fluid_2d.cc:17) */
    double alpha, mug, Pamb, h0, R, V, sigma; /* gl_params (gas_layer_solve.hh:20-41) */
} vdi_params;

/* Per-solver and per-step counters, mirroring the reference's machine-
 * readable timing line (fluid_2d.cc:66-77). */
typedef struct vdi_stats {
    int last_cycles[4];    /* V-cycles of the most recent solve, by VDI_MAC.. */
    long total_cycles[4];  /* since creation or the last vdi_reset_stats */
    long solves[4];
    int last_newton;       /* Newton steps of the most recent gas-layer update */
    long steps;
    double time;           /* fluid_2d::time */
    int last_height_passes; /* Jacobi passes of the most recent height update
                               (0: the serial sweep ran); no counterpart in the
                               reference, whose sweep is serial */
} vdi_stats;

/* Kernel classes for vdi_time_kernels (measurement only). */
#define VDI_KCLASS_MG_DOWN0 0   /* level-0 smooth+restrict of every solve */
#define VDI_KCLASS_MG_UP0 1     /* level-0 prolong+smooth+residual of every solve */
#define VDI_KCLASS_MG_MID 2     /* down/up kernels of levels 1..lc-1 */
#define VDI_KCLASS_MG_COARSE 3  /* single-workgroup exact solve of the coarsest level */
#define VDI_KCLASS_GAS 4        /* gas-layer update */
#define VDI_KCLASS_SIMPSON 5    /* far-field boundary integrals */
/* The same launches grouped by kernel function (every level), as a profiler
 * names them: */
#define VDI_KCLASS_CC_DOWN 6    /* k_down<double, false>: cell-centred smooth + restrict */
#define VDI_KCLASS_CC_UP 7      /* k_up<double, false>: cell-centred prolong + smooth */
#define VDI_KCLASS_FEM_DOWN 8   /* k_down<double, true>: nodal smooth + restrict */
#define VDI_KCLASS_FEM_UP 9     /* k_up<double, true>: nodal prolong + smooth */
#define VDI_KCLASS_XSOLVE 10    /* k_xsolve<double>: the three phases of a larger coarsest level's exact solve */
#define VDI_KCLASS_SCC_DOWN 11  /* k_sdown<double, nu>: streaming cell-centred smooth + restrict */
#define VDI_KCLASS_SCC_UP 12    /* k_sup<double, nu>: streaming cell-centred prolong + smooth */
#define VDI_KCLASS_SFEM_DOWN 13 /* k_sndown<double, nu>: streaming nodal smooth + restrict */
#define VDI_KCLASS_SFEM_UP 14   /* k_snup<double, nu>: streaming nodal prolong + smooth */

typedef struct vdi_ctx vdi_ctx;
typedef struct vdi_mg vdi_mg;
typedef struct vdi_batch vdi_batch;
typedef struct vdi_nn vdi_nn;

const char *vdi_status_string(int status);
/* Text of the last HIP error seen by this library (or ""). */
const char *vdi_last_error(void);

/* fluid_2d::fluid_2d(const char*) (fluid_2d.cc:14-24): allocate the state,
 * the four solvers and the gas layer on the current HIP device. */
int vdi_create(const vdi_params *p, vdi_ctx **out);
/* fluid_2d::~fluid_2d (fluid_2d.cc:27-36) */
int vdi_destroy(vdi_ctx *ctx);
/* Solver schedule and stopping rule for all four solvers: nu1/nu2 smoothing
 * sweeps (nu1 = nu2 = 0 keeps each solver's own), relative residual
 * tolerance ||b-Az||/||b||, V-cycle cap (replaces the TGMG constants and
 * tgmg_accuracy, mgs_mac.cc:11, mgs_fem.cc:11, visco_impl.cc:13). */
int vdi_set_solver(vdi_ctx *ctx, int nu1, int nu2, double rel_tol, int max_cycles);
/* nu1/nu2 of one solver (VDI_MAC..VDI_VY) only.  No single reference
 * counterpart: each reference solver has its own TGMG instance and constants
 * (mgs_mac.cc:13-14, mgs_fem.cc:13-14, visco_impl.cc:13). */
int vdi_set_schedule(vdi_ctx *ctx, int which_solver, int nu1, int nu2);
/* The schedule in use on the finest level (defaults: V(3,2) for MAC and
 * FEM, V(4,2) for the viscous solves). */
int vdi_get_schedule(vdi_ctx *ctx, int which_solver, int *nu1, int *nu2);
/* V(nu1, nu2) on the finest level and V(nu1s, nu2s) on the restricted
 * levels of one solver (vdi_set_schedule sets both to the same pair).  No
 * reference counterpart (the TGMG schedule is internal to the absent
 * library); defaults: MAC V(3,2) everywhere, FEM V(3,2) then V(2,2),
 * viscous V(4,2) then V(0,1). */
int vdi_set_schedule_levels(vdi_ctx *ctx, int which_solver, int nu1, int nu2, int nu1s, int nu2s);
/* out = {nu1, nu2, nu1s, nu2s} of one solver. */
int vdi_get_schedule_levels(vdi_ctx *ctx, int which_solver, int out[4]);
/* Arithmetic form of the V-cycles (VDI_FORM_*). */
int vdi_set_form(vdi_ctx *ctx, int form);
/* Capture steps in HIP graphs (default on): after the first two steps each
 * step is replayed from a graph with predicted V-cycle budgets, checked once
 * at its end and redone on the synchronising path if a solve fell short.
 * No reference counterpart (launch mechanics only; results are the same). */
int vdi_set_graphs(vdi_ctx *ctx, int enable);
/* Run the u- and v-velocity Helmholtz solves of a captured step as two
 * parallel graph branches (default on; the solves are independent).  No
 * reference counterpart (scheduling only; results are the same). */
int vdi_set_concurrency(vdi_ctx *ctx, int enable);
/* Steps run as captured graphs, and of those the steps redone on the direct
 * path because a solve needed more cycles than its budget.  No counterpart
 * in the reference. */
int vdi_get_counters(vdi_ctx *ctx, long *graph_steps, long *redo_steps);
/* The V-cycle budgets [MAC, VX, VY, FEM] the next captured step enqueues
 * (the largest count among each system's recent solves).  No counterpart in
 * the reference. */
int vdi_get_budgets(vdi_ctx *ctx, int *out);
/* fluid_2d::init_fields (fluid_2d_sio.cc:14-35) */
int vdi_init_fields(vdi_ctx *ctx);
/* fluid_2d::step_forward(double dt) (fluid_2d.cc:93-405) */
int vdi_step_forward(vdi_ctx *ctx, double dt);
/* gas_layer::update(double t, double dt) (gas_layer.cc:36-46) */
int vdi_gas_layer_update(vdi_ctx *ctx, double t, double dt);
/* mgs_mac::solve_v_cycle(double t) (mgs_mac.hh:54-59) and
 * mgs_fem::solve_v_cycle(double t) (mgs_fem.hh:61-66): time-extrapolated
 * initial guess, then V-cycles on the shared src.  which = VDI_MAC/VDI_FEM. */
int vdi_mg_solve_v_cycle(vdi_ctx *ctx, int which, double t, int *vcycles);
/* visco_impl::setup_and_solve(double vfx, double vfy, double t)
 * (visco_impl.cc:21-39).  which = VDI_VX or VDI_VY. */
int vdi_visco_setup_and_solve(vdi_ctx *ctx, int which, double vfx, double vfy, double t, int *vcycles);
/* fluid_2d::tm and init_tracers (fluid_2d.cc:495-505): tracer positions
 * (x0, y0, x1, y1, ...); ntrace = 0 removes them.  Each vdi_step_forward then
 * moves them first, by the bilinear velocity at the step's start
 * (fluid_2d::update_tracers, fluid_2d.cc:507-525).  vdi_get_tracers reads
 * them back (ntrace must match). */
int vdi_set_tracers(vdi_ctx *ctx, const double *pos, int ntrace);
int vdi_get_tracers(vdi_ctx *ctx, double *pos, int ntrace);
/* The advection passes alone, on the context's current arrays, reps times
 * (the average launch time in *ms, HIP events).  VDI_PASS_GODUNOV: limited
 * slopes, half-step face predictor, upwind choices and the MAC right-hand
 * side (fluid_2d.cc:105-232, gas pressure from VDI_GP/VDI_GP_OLD);
 * reads u, v, p, writes ul, vl, ud, vd, c0, c1 and src.
 * VDI_PASS_CORRECT_ADVECT: MAC correction of the faces by VDI_Z_MAC and the
 * advective update (fluid_2d.cc:236-261); writes c0, c1 (the corrected faces
 * are not stored).  No reference counterpart as separate calls. */
#define VDI_PASS_GODUNOV 0
#define VDI_PASS_CORRECT_ADVECT 1
int vdi_run_pass(vdi_ctx *ctx, int which, double dt, int reps, double *ms);
/* Host copies of state arrays (ids above); len must equal vdi_array_len. */
long vdi_array_len(vdi_ctx *ctx, int id);
int vdi_get_array(vdi_ctx *ctx, int id, double *host, long len);
int vdi_set_array(vdi_ctx *ctx, int id, const double *host, long len);
int vdi_get_time(vdi_ctx *ctx, double *t);
int vdi_set_time(vdi_ctx *ctx, double t);
int vdi_get_stats(vdi_ctx *ctx, vdi_stats *st);
int vdi_reset_stats(vdi_ctx *ctx);
/* Blocks until all work queued on the context's stream has finished. */
int vdi_synchronize(vdi_ctx *ctx);
/* Measurement (no reference counterpart): time every launch of one kernel
 * class (VDI_KCLASS_*, -1 to stop) with HIP events on the context's stream,
 * up to max_launches launches.  vdi_kernel_times returns the number of timed
 * launches, their summed duration in ms and their summed algorithmic bytes. */
int vdi_time_kernels(vdi_ctx *ctx, int kclass, int max_launches);
int vdi_kernel_times(vdi_ctx *ctx, long *launches, double *total_ms, double *alg_bytes);
/* Plain device copy of n doubles, repeated reps times, for calibrating
 * hardware byte counters against a kernel of known traffic. */
int vdi_copy_probe(long n, int reps, double *ms);
/* Device facts used to size launches: info[0..8] = LDS bytes per workgroup,
 * compute units, threads per workgroup, wavefront size, clock kHz, memory
 * clock kHz, memory bus bits, LDS bytes per compute unit, L2 bytes. */
int vdi_device_info(int *info, int n);
/* Restart-file body (fluid_2d_sio.cc:97-119 + gas_layer_solve.cc:326-332):
 * time, (n+4) rows x (m+4) x {u,v,p}, Vframe, p[me], tau[2..m+1], h[m],
 * ht[m], pg[me].  Length: vdi_state_len. */
long vdi_state_len(vdi_ctx *ctx);
int vdi_download_state(vdi_ctx *ctx, double *host, long len);
int vdi_upload_state(vdi_ctx *ctx, const double *host, long len);

/* ---- standalone linear systems (mg_test.cc:19-45, operator-level use) ----
 * kind VDI_MAC: (cx,cy) = (1/dx^2, 1/dy^2), m x n cells (mgs_mac.cc:23-27).
 * kind VDI_FEM: (cx,cy) = (dx,dy), m x n nodes (mgs_fem.cc:23-27).
 * kind VDI_VX/VDI_VY: (cx,cy) = (vfx,vfy), m x n cells, x_sym as in fluid_2d. */
int vdi_mg_create(int kind, int m, int n, double cx, double cy, int x_sym, vdi_mg **out);
int vdi_mg_destroy(vdi_mg *g);
int vdi_mg_set_schedule(vdi_mg *g, int nu1, int nu2);
int vdi_mg_set_form(vdi_mg *g, int form);
int vdi_mg_levels(vdi_mg *g);
/* Host <-> device copies of the right-hand side b and the solution z (m*n). */
int vdi_mg_set_b(vdi_mg *g, const double *host, long len);
int vdi_mg_set_z(vdi_mg *g, const double *host, long len);
int vdi_mg_get_z(vdi_mg *g, double *host, long len);
/* Device addresses of the system's level-0 solution z and right-hand side b
 * (m*n doubles each, owned by the system), for entry points that take device
 * arrays.  No reference counterpart. */
int vdi_mg_device_arrays(vdi_mg *g, double **z_dev, double **b_dev);
/* out = A z (host arrays, for operator checks). */
int vdi_mg_apply(vdi_mg *g, const double *z, double *out, long len);
/* tgmg::solve_v_cycle: V-cycles until ||b-Az|| <= rel_tol ||b||. */
int vdi_mg_solve(vdi_mg *g, double rel_tol, int max_cycles, int *vcycles);
/* tgmg::v_cycle (mg_test.cc:38): exactly k V-cycles, no convergence test. */
int vdi_mg_vcycles(vdi_mg *g, int k);
/* A solve with a fixed budget of V-cycles and no host read between them (the
 * form a captured step enqueues), then its outcome: *done = 1 converged, 3
 * zero right-hand side (z = 0), 0 budget spent, 2 NaN; *vcycles = V-cycles
 * run.  No reference counterpart (TGMG's solve_v_cycle always synchronises). */
int vdi_mg_solve_fixed(vdi_mg *g, double rel_tol, int budget, int *done, int *vcycles);
/* ||b - A z||_2 and ||b||_2 of the device arrays. */
int vdi_mg_residual(vdi_mg *g, double *rnorm, double *bnorm);
int vdi_mg_synchronize(vdi_mg *g);
/* Algorithmic bytes of one V-cycle in the current form (measurement). */
int vdi_mg_cycle_bytes(vdi_mg *g, double *bytes);

/* Size of the device scratch array of vdi_mg_solve_noalloc, per GPU. */
#define VDI_NOALLOC_SCRATCH_MB 24
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same solve
 * as vdi_mg_create(kind, m, n, cx, cy, x_sym) + vdi_mg_solve(rel_tol,
 * max_cycles) with the default schedule and tiles, on caller-owned device
 * arrays b_dev (read) and z_dev (initial guess in, solution out), on the null
 * stream, returning when done.  Every other buffer lives in one static device
 * array of VDI_NOALLOC_SCRATCH_MB MiB per GPU (enough for 1025 x 513
 * unknowns); larger systems return VDI_ERR_ARG.  FP64 form only.  Calls on one GPU share
 * that array: calls that overlap in time race. */
int vdi_mg_solve_noalloc(int kind, int m, int n, double cx, double cy, int x_sym, const double *b_dev,
                         double *z_dev, double rel_tol, int max_cycles, int *vcycles);

/* ---- tile shapes of the multigrid level kernels ----
 * No counterpart in the reference (TGMG's sweeps are not tiled).  Shapes are
 * indexes into a fixed candidate list (vdi_tile_shape gives tx, ty; the last
 * eight are streaming launches, one wavefront per strip of tx = 128 columns
 * less the halo and ty rows).
 * prec 0 = FP64 hierarchy, 1 = FP32 hierarchy of the mixed form; which 0 =
 * down (smooth + restrict), 1 = up (prolong + smooth), 2 = mixed level-0 up. */
int vdi_tile_shape(int shape, int *tx, int *ty);
/* 1 when `shape` is a streaming launch, 0 when it is a tile, negative when
 * out of range (no reference counterpart). */
int vdi_tile_stream(int shape);
int vdi_mg_set_tile(vdi_mg *g, int prec, int which, int level, int shape);
/* Times every candidate for every level kernel on the system's own arrays and
 * keeps the fastest (z and b are restored). */
int vdi_mg_autotune(vdi_mg *g, int reps);
/* Measurement (no reference counterpart): vdi_time_kernels and
 * vdi_kernel_times for a standalone system -- every launch of one kernel
 * class (VDI_KCLASS_*, -1 to stop) is stamped with the device clock, up to
 * max_launches launches; the totals cover the launches since the call that
 * started the timing. */
int vdi_mg_time_kernels(vdi_mg *g, int kclass, int max_launches);
int vdi_mg_kernel_times(vdi_mg *g, long *launches, double *total_ms, double *alg_bytes);
/* Same for the four solvers of a context. */
int vdi_autotune(vdi_ctx *ctx, int reps);
/* Writes the chosen shapes [prec][which][level] of solver `which_solver`
 * (VDI_MAC..VDI_VY) into out; returns the number written or a negative value. */
int vdi_get_tiles(vdi_ctx *ctx, int which_solver, int *out, int n);

/* ---- batches of independent simulations ----
 * No counterpart in the reference, which runs one simulation per process.
 * Each member is a full context (its own grid size and parameters; x_sym and
 * implicit_visc must agree).  A batched step replays one captured graph of
 * persistent kernels over all members. */
int vdi_batch_create(const vdi_params *params, int count, int form, vdi_batch **out);
int vdi_batch_destroy(vdi_batch *b);
int vdi_batch_size(vdi_batch *b);
/* Borrowed handle of member k (for vdi_get_array etc.; do not destroy). */
int vdi_batch_sim(vdi_batch *b, int k, vdi_ctx **out);
int vdi_batch_init_fields(vdi_batch *b);
/* Advances every member by its own dt_reg.  A member whose step fails (the
 * gas layer's "Too many Newton steps", where the reference's run of that
 * simulation would stop) does not advance; the others do, and the call
 * returns the first such status.  Until an ended member is restarted, later
 * calls return VDI_ERR_ARG. */
int vdi_batch_step(vdi_batch *b);
/* Status of each member's last step (n >= size). */
int vdi_batch_member_status(vdi_batch *b, int *status, int n);
/* Starts member k over as a new simulation with the same parameters: initial
 * fields and gas layer, time 0, empty time-extrapolation histories and zero
 * iterates, as a newly created context (a batch used as slots of a sweep). */
int vdi_batch_restart_member(vdi_batch *b, int k);
int vdi_batch_autotune(vdi_batch *b, int reps);
/* Last batched step: graph time (ms) and each member's workgroup busy time
 * (ms, summed over its work items; n >= size). */
int vdi_batch_step_times(vdi_batch *b, double *graph_ms, double *busy_ms, int n);
/* Per-member step time of the last batched step (ms, n >= size): from the
 * start of the member's first work item to the end of its last one, read from
 * the device clock (wall_clock64) inside the captured step. */
int vdi_batch_member_spans(vdi_batch *b, double *span_ms, int n);
int vdi_batch_counters(vdi_batch *b, long *batched_steps, long *redone_sims);
int vdi_batch_tiles(vdi_batch *b, int *out, int n);

/* ---- learned initial guess of the MAC solve ----
 * No counterpart in the reference, whose guess is the time extrapolation of
 * mgs_common.cc:17-50.  A 4-level BF16 U-Net (MFMA, FP32 accumulation) maps
 * the extrapolated guess's residual, the guess itself and the previous
 * step's correction to a correction; the
 * solve still iterates to the same tolerance, and keeps the extrapolated
 * guess whenever the network's guess has the larger residual.
 * m and n must be multiples of 64. */
int vdi_guess_attach(vdi_ctx *ctx, int capacity, unsigned long long seed);
/* mode 0 off, 1 record training pairs (direct steps), 2 use the network */
int vdi_guess_mode(vdi_ctx *ctx, int mode);
int vdi_guess_train(vdi_ctx *ctx, int epochs, int batch, double lr, unsigned long long seed, double *final_loss);
int vdi_guess_stats(vdi_ctx *ctx, long *pairs, long *accepted, long *rejected, double *kappa);
/* Copies the trained network (same grid) from src to dst. */
int vdi_guess_copy(vdi_ctx *dst, vdi_ctx *src);

/* The network alone (gradient checks): H x W pixels (multiples of 64),
 * inputs planar [3][H*W]. */
int vdi_nn_create(int H, int W, unsigned long long seed, vdi_nn **out);
int vdi_nn_destroy(vdi_nn *nn);
long vdi_nn_nparams(vdi_nn *nn);
int vdi_nn_get_params(vdi_nn *nn, float *out, long n);
int vdi_nn_set_params(vdi_nn *nn, const float *in, long n);
int vdi_nn_forward(vdi_nn *nn, const float *x2, float *y);
int vdi_nn_gradient(vdi_nn *nn, const float *x2, const float *t, float *grad, double *loss);

#ifdef __cplusplus
}
#endif
#endif
