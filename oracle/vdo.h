/* vdo.h -- CPU FP64 oracle for the vdropimpact per-timestep fluid solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (vdropimpact_amd/) never links or calls it.
 *
 * It restates, in plain C with OpenMP, the reference algorithm of
 * fluid_2d::step_forward (fluid_2d.cc:93-405) and everything it calls:
 * the Godunov passes, the MAC / FEM / implicit-viscosity linear systems
 * (mgs_mac.hh:25-65, mgs_fem.hh:42-72 + mgs_fem.cc:34-50, visco_impl.hh:28-53),
 * the boundary conditions (fluid_2d_sio.cc:256-405) and the gas layer
 * (gas_layer.cc, gas_layer_solve.cc:32-249, tri_solve.cc).  The reference's
 * multigrid library (TGMG) is not available, so the linear systems are solved
 * by this file's own geometric multigrid to a relative residual tolerance;
 * parity with the reference is therefore defined on converged solutions.
 *
 * Pinning status of each part (what checks it in tests/):
 *   operators (MAC, FEM, visco)  closed-form separable eigenvectors + dense solves
 *   multigrid solve              residual of the pinned operator; closed-form solutions
 *   exact coarsest-level solve   numpy dense solve of the assembled level operator,
 *                                both transform directions (test_oracle_pins.py)
 *   tridiagonal solve            tri_test closed form x(x-1)/2; scipy dgtsv
 *   mono_diff limiter            exact for linear data, 0 at extrema, odd symmetry,
 *                                fourth-order accuracy on smooth data
 *   Godunov selects              upwind/stagnation outcomes by case
 *   Simpson boundary integral    closed-form integral for constant pressure, O(dx^4)
 *   ENO2 height update           exact for quadratic h; upwind side by sign of u
 *   Newton gas layer             PDE residual written from the lubrication equation,
 *                                Jacobian checked by finite differences
 *   step passes K1-K15, BCs,     exactness on linear / bilinear / uniform fields,
 *   liquid fields, extrapolation discrete identities, Taylor-Green convergence,
 *                                mirror symmetry of the whole step
 *                                (test_oracle_step_pins.py; planted mistakes in
 *                                test_oracle_mutations.py)
 *                                here (TGMG and LAPACK are absent), so only the
 *                                components above are pinned.
 */
#ifndef VDO_ORACLE_H
#define VDO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as vdi_params in include/vdi.h. */
typedef struct vdo_params {
    int m, n;              /* grid cells */
    int x_sym;             /* symmetric about x=0 */
    int implicit_visc;     /* Crank-Nicolson viscosity */
    double ax, bx, ay, by; /* domain */
    double rho, mu;        /* liquid density and dynamic viscosity */
    double dt_reg;         /* regular timestep */
    double alpha, mug, Pamb, h0, R, V, sigma; /* gas layer */
} vdo_params;

typedef struct vdo_sim vdo_sim;
typedef struct vdo_mg vdo_mg;

/* Array identifiers shared with the GPU library (include/vdi.h). */
enum {
    VDO_U = 0, VDO_V, VDO_P,                       /* padded (m+4)x(n+4) */
    VDO_UL, VDO_VL, VDO_UD, VDO_VD, VDO_C0, VDO_C1, /* padded */
    VDO_SRC,                                       /* me*ne */
    VDO_GP, VDO_GTAU, VDO_GP_OLD, VDO_GTAU_OLD,     /* me, ml, me, ml */
    VDO_H, VDO_HX, VDO_HT, VDO_FBD, VDO_PG, VDO_PG_OLD, /* m, me, m, m, me, me */
    VDO_Z_MAC, VDO_Z_FEM, VDO_Z_VX, VDO_Z_VY,       /* mn, me*ne, mn, mn */
    VDO_NARRAYS,
    /* oracle only: the step's scratch fields (padded), for pinning the passes */
    VDO_UX = 64, VDO_UY, VDO_VXS, VDO_VYS, VDO_UR, VDO_VR, VDO_UU, VDO_VU, VDO_C2, VDO_C3
};

/* The passes of one step (fluid_2d.cc:93-405) in their order, for pinning
 * each against facts of its own; vdo_pass runs one on the current arrays. */
enum {
    VDO_PASS_K1 = 1,         /* limited slopes + normal face extrapolation, fluid_2d.cc:105-132 */
    VDO_PASS_K2,             /* Godunov choice (tangential form), :136-150 */
    VDO_PASS_K3,             /* transverse terms c0..c3, :153-160 */
    VDO_PASS_K4,             /* force predictor, :164-201 */
    VDO_PASS_K5,             /* Godunov choice, :205-215 */
    VDO_PASS_K6,             /* MAC right-hand side and s_bc, :218-232 */
    VDO_PASS_K7,             /* MAC correction, :236-248 */
    VDO_PASS_K8,             /* advective update, :253-261 */
    VDO_PASS_INVISCID_BC,    /* far-field ghosts, fluid_2d_sio.cc:256-290 */
    VDO_PASS_VISCOUS_BC,     /* left/bottom ghosts, fluid_2d_sio.cc:294-343 */
    VDO_PASS_VISCO_RHS_U,    /* fluid_2d.cc:272-299 */
    VDO_PASS_VISCO_RHS_V,    /* fluid_2d.cc:305-324 */
    VDO_PASS_VISCO_COPY,     /* fluid_2d.cc:330-334 */
    VDO_PASS_K14,            /* FEM right-hand side, fluid_2d.cc:338-377 */
    VDO_PASS_K15,            /* velocity and pressure update, fluid_2d.cc:383-396 */
    VDO_PASS_GAS_LIQUID_FIELDS, /* gas_layer_solve.cc:229-249 */
    VDO_PASS_GAS_TAU_BC,     /* gas_layer.cc:25-30 */
    VDO_PASS_GAS_HX,         /* gas_layer_solve.cc:58-62 */
    VDO_PASS_GAS_INIT,       /* gas_layer_solve.cc:32-53 */
    VDO_PASS_TRACERS         /* fluid_2d.cc:507-525 */
};

/* Multigrid operator kinds. */
enum { VDO_KIND_MAC = 0, VDO_KIND_FEM = 1, VDO_KIND_VX = 2, VDO_KIND_VY = 3 };

/* ---- whole simulation ---- */
vdo_sim *vdo_create(const vdo_params *p);
void vdo_destroy(vdo_sim *s);
void vdo_set_mg(vdo_sim *s, int nu1, int nu2, double rel_tol, int max_cycles);
/* nu1/nu2 of one solver (0 MAC, 1 FEM, 2 VX, 3 VY); nonzero on a bad argument. */
int vdo_set_schedule(vdo_sim *s, int which, int nu1, int nu2);
/* V(nu1, nu2) on the finest level, V(nu1s, nu2s) on the restricted levels. */
int vdo_set_schedule_levels(vdo_sim *s, int which, int nu1, int nu2, int nu1s, int nu2s);
void vdo_init_fields(vdo_sim *s);
int vdo_step_forward(vdo_sim *s, double dt);
int vdo_gas_update(vdo_sim *s, double t, double dt);
/* The two halves of the gas update, for pinning: ENO2 height update
 * (gas_layer_solve.cc:133-162) and the Newton pressure solve (:184-212). */
int vdo_gas_height(vdo_sim *s, double dt);
int vdo_gas_newton(vdo_sim *s, double dt);
double vdo_time(const vdo_sim *s);
double vdo_vframe(const vdo_sim *s);
long vdo_array_len(const vdo_sim *s, int id);
int vdo_get_array(const vdo_sim *s, int id, double *out, long len);
int vdo_set_array(vdo_sim *s, int id, const double *in, long len);
int vdo_set_time(vdo_sim *s, double t);
/* OpenMP threads of the oracle's loops (0 keeps the current count);
 * returns the count in use. */
int vdo_set_threads(int nthreads);
/* Tracer positions (x0, y0, x1, y1, ...), moved at the start of each step
 * (fluid_2d.cc:98, :507-525). */
int vdo_set_tracers(vdo_sim *s, const double *pos, int ntrace);
int vdo_get_tracers(const vdo_sim *s, double *pos, int ntrace);
/* Test switches for two reference quirks that break mirror symmetry: Q1
 * (the gas shear of cell i written at tau[i] and read at tau[i+2]) and Q13
 * (the height update sweeping in place).  1 = as in the reference (default);
 * 0 = shear written at tau[i+2], heights updated from the old values. */
int vdo_set_quirks(vdo_sim *s, int q1_tau_shift, int q13_inplace);
/* One pass of the step on the current arrays (VDO_PASS_*); nonzero when the
 * pass does not apply (viscous passes without implicit viscosity). */
int vdo_pass(vdo_sim *s, int pass, double dt);
/* mgs_common::initial_guess(t) (mgs_common.cc:17-50) on solver `which`'s z. */
int vdo_initial_guess(vdo_sim *s, int which, double t);
/* mgs_mac/mgs_fem::solve_v_cycle(t) (mgs_mac.hh:54-59, mgs_fem.hh:61-66) on the
 * shared src: time-extrapolated guess, then V-cycles.  Returns the V-cycles
 * or -1.  which = VDO_KIND_MAC or VDO_KIND_FEM. */
int vdo_solve_v_cycle(vdo_sim *s, int which, double t);
/* visco_impl::setup_and_solve(vfx, vfy, t) (visco_impl.cc:21-39). */
int vdo_visco_setup_and_solve(vdo_sim *s, int which, double vfx, double vfy, double t);
/* V-cycles spent by the last solve of each kind, and Newton steps of the
 * last gas update. */
int vdo_last_cycles(const vdo_sim *s, int kind);
int vdo_last_newton(const vdo_sim *s);

/* ---- standalone linear systems ----
 * kind MAC: (cx,cy) = (1/dx^2, 1/dy^2).  kind VX/VY: (cx,cy) = (vfx,vfy),
 * x_sym as in the simulation.  kind FEM: (cx,cy) = (dx,dy), grid m x n nodes.
 * m,n are the unknown counts (cells, or nodes for FEM). */
vdo_mg *vdo_mg_create(int kind, int m, int n, double cx, double cy, int x_sym);
void vdo_mg_destroy(vdo_mg *g);
void vdo_mg_set(vdo_mg *g, int nu1, int nu2, int ncs);
int vdo_mg_levels(const vdo_mg *g);
void vdo_mg_apply(vdo_mg *g, const double *z, double *out);           /* out = A z */
double vdo_mg_resnorm(vdo_mg *g, const double *b, const double *z);    /* ||b-Az||_2 */
/* Runs V-cycles from z until ||b-Az|| <= tol ||b|| (max_cycles cap).
 * Returns cycles used, or -1 on non-convergence. */
int vdo_mg_solve(vdo_mg *g, const double *b, double *z, double tol, int max_cycles);
/* Exactly k V-cycles (no convergence test), for cycle-by-cycle comparisons. */
void vdo_mg_vcycles(vdo_mg *g, const double *b, double *z, int k);

/* ---- small pieces exposed for pinning tests ---- */
double vdo_mono_diff(double f0, double f1, double f2, double f3, double f4);
void vdo_godunov_set(double *u0, double *u1, double *v0, double *v1);
void vdo_godunov_set_tang(double *u0, double *u1, double *v0, double *v1);
/* Composite-Simpson far-field velocity increment at (x,y) from pressure p[me]
 * on nodes ax + j dx; returns the increments that set_inviscid_boundaries adds. */
void vdo_simpson_velocity(int me, double ax, double dx, int x_sym, const double *p,
                          double x, double y, double prefac, double *du, double *dv);
/* Tridiagonal solve with partial pivoting (dgtsv semantics: dl[n-1], d[n],
 * du[n-1], b[n] overwritten; solution in b).  Returns 0 or the 1-based index
 * of a zero pivot. */
int vdo_tri_solve_pivot(int n, double *dl, double *d, double *du, double *b);
/* Thomas algorithm (tri_solve::solve semantics: a[n],b[n],c[n],d[n]). */
void vdo_tri_solve_thomas(int n, double *a, double *b, double *c, double *d);

#ifdef __cplusplus
}
#endif
#endif
