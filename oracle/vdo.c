/* vdo.c -- CPU FP64 oracle for the vdropimpact per-timestep fluid solve.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See vdo.h for the pinning status
 * of every function.  Each block below cites the reference file:line whose
 * algorithm it restates.  Reference quirks that change results are kept on
 * purpose and marked "quirk".
 */
#include "vdo.h"

#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define VDO_MAXLEV 24
/* Coarsening stops at the first level with at most VDO_EXACT_POINTS
 * unknowns (or earlier, on odd sizes).  That level is solved exactly by
 * separation of variables when its shorter side (without the nodal
 * Dirichlet row) has at most VDO_SEP_MAXT unknowns, else by ncs sweeps. */
#define VDO_EXACT_POINTS 3072
#define VDO_SEP_MAXT 64
/* A restricted level (not level 0) of at most VDO_EXACT_BIG unknowns, at most
 * 256 a side and at most VDO_SEP_MAXT unknown rows stops the coarsening too
 * (the GPU solves it with several workgroups). */
#define VDO_EXACT_BIG 9000

/* ------------------------------------------------------------------------ */
/* Multigrid                                                                 */
/* ------------------------------------------------------------------------ */

/* Boundary types of the cell-centred operators, per side.
 * N  : Neumann (ghost = interior)            -> contributes 0 to the diagonal
 * DF : Dirichlet on the face (ghost = -int.) -> contributes 2c
 * DC : Dirichlet at the ghost centre (= 0)   -> contributes c            */
enum { BC_N = 0, BC_DF = 1, BC_DC = 2 };

typedef struct {
    int m, n;        /* unknowns in x and y on this level */
    double cx, cy;   /* cell-centred couplings on this level */
    double *z, *b, *r;
} vdo_level;

struct vdo_mg {
    int kind;        /* VDO_KIND_* */
    int fem;         /* 1 for the nodal FEM operator */
    int nl;
    vdo_level L[VDO_MAXLEV];
    double sigma, s; /* cell-centred: A = sigma I + s Lap */
    int bc[4];       /* left, right, bottom, top */
    double fm, fex, fey, fc, hex, hey; /* FEM stencil (mgs_fem.cc:8-11) */
    int nu1, nu2, ncs;
    int nu1s, nu2s;  /* sweeps on the restricted levels 1 .. nl - 2 */
    /* solve history for the time-extrapolated initial guess (mgs_common.hh) */
    double *z_prev;
    int scount;
    double t0, t_prev;
    double vfx, vfy; /* visco: coefficients currently set up (visco_impl.cc) */
    int last_cycles;
    /* exact solve of the coarsest level: z = (V (x) I) (S + lambda_k Q)^-1 (V^T (x) I) b,
     * with unknown (p, t) at column p, row t + yoff (tdir 0) or column t, row p + yoff (tdir 1) */
    int sep, tdir, sT, sP, yoff;
    double *V;       /* sT x sT, V[t * sT + k] */
    double *lam;     /* sT generalized eigenvalues */
    double *Sd, *Qd; /* diagonals of S and Q along p (sP each) */
    double Se, Qe;   /* their constant off-diagonals */
};

static double bc_diag_factor(int bc) { return bc == BC_DF ? 2.0 : (bc == BC_DC ? 1.0 : 0.0); }
static double bc_ghost_factor(int bc) { return bc == BC_DF ? -1.0 : (bc == BC_DC ? 0.0 : 1.0); }

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) { fputs("vdo: out of memory\n", stderr); abort(); }
    return p;
}

/* Level hierarchy: halve while both cell counts are even and the coarse grid
 * keeps at least 2 cells per direction. */
static void mg_build_levels(vdo_mg *g, int m, int n, double cx, double cy) {
    int l = 0;
    int mc = g->fem ? m - 1 : m, nc = g->fem ? n - 1 : n;
    for (;;) {
        vdo_level *L = &g->L[l];
        L->m = g->fem ? mc + 1 : mc;
        L->n = g->fem ? nc + 1 : nc;
        L->cx = cx; L->cy = cy;
        size_t N = (size_t)L->m * L->n;
        L->z = l == 0 ? NULL : (double *)xcalloc(N, sizeof(double));
        L->b = l == 0 ? NULL : (double *)xcalloc(N, sizeof(double));
        L->r = (double *)xcalloc(N, sizeof(double));
        l++;
        if (l >= VDO_MAXLEV) break;
        {
            /* a restricted level of at most VDO_EXACT_BIG unknowns (at most
             * 256 a side, at most VDO_SEP_MAXT unknown rows) is the coarsest */
            const int rows = g->fem ? L->n - 1 : L->n;
            if (l >= 2 && N <= VDO_EXACT_BIG && (L->m > L->n ? L->m : L->n) <= 256 && rows >= 1 && rows <= VDO_SEP_MAXT)
                break;
        }
        if (N <= VDO_EXACT_POINTS) break;
        if (mc % 2 || nc % 2 || mc / 2 < 2 || nc / 2 < 2) break;
        mc /= 2; nc /= 2; cx *= 0.25; cy *= 0.25;
    }
    g->nl = l;
}

static void level_apply(const vdo_mg *g, const vdo_level *L, const double *z, double *out);

/* ---- exact solve of the coarsest level by separation of variables ----
 * Each level operator is a sum of two Kronecker products of 1-D symmetric
 * tridiagonal matrices, A = R (x) S + Pm (x) Q, R positive definite, the
 * first factor acting along the transformed direction t and the second along
 * the kept one p.  The generalized eigenvectors Pm v = lambda R v with
 * V^T R V = I split A z = b into (S + lambda_k Q) zh_k = (V^T b)_k, one
 * tridiagonal system per mode, and z = V zh.  Cell-centred
 * A = sigma I + s (cx Dx (x) I + cy I (x) Dy) (second differences with the
 * boundary types' diagonal factors); nodal Q1 on the rows above the
 * Dirichlet row: A = alpha Kx (x) My + beta Mx (x) Ky with 1-D stiffness
 * K = (-1, 2, -1) and mass M = (1, 4, 1), halved at the natural ends. */

static inline double fem_acc(const vdo_mg *g, const vdo_level *L, int i, int j);

/* Cyclic Jacobi rotations on the symmetric n x n matrix a (row-major,
 * overwritten); eigenvalues w, eigenvectors in the columns of q. */
static void sep_jacobi(int n, double *a, double *w, double *q) {
    for (int i = 0; i < n * n; i++) q[i] = 0;
    for (int i = 0; i < n; i++) q[i * n + i] = 1;
    double tot = 0;
    for (int i = 0; i < n * n; i++) tot += a[i] * a[i];
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) off += a[i * n + j] * a[i * n + j];
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n; p++)
            for (int r = p + 1; r < n; r++) {
                double apr = a[p * n + r];
                if (apr == 0) continue;
                double th = (a[r * n + r] - a[p * n + p]) / (2 * apr);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < n; k++) {
                    double x = a[k * n + p], y = a[k * n + r];
                    a[k * n + p] = c * x - sn * y;
                    a[k * n + r] = sn * x + c * y;
                }
                for (int k = 0; k < n; k++) {
                    double x = a[p * n + k], y = a[r * n + k];
                    a[p * n + k] = c * x - sn * y;
                    a[r * n + k] = sn * x + c * y;
                }
                a[p * n + r] = a[r * n + p] = 0;
                for (int k = 0; k < n; k++) {
                    double x = q[k * n + p], y = q[k * n + r];
                    q[k * n + p] = c * x - sn * y;
                    q[k * n + r] = sn * x + c * y;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = a[i * n + i];
}

/* 1-D line: diagonal d[0..n), off-diagonal e. */
static void cc_line(int n, double c, double sigma, int lo, int hi, double *d, double *e) {
    for (int i = 0; i < n; i++)
        d[i] = sigma - c * ((i > 0 ? 1.0 : bc_diag_factor(lo)) + (i < n - 1 ? 1.0 : bc_diag_factor(hi)));
    *e = c;
}

static void mg_sep_free(vdo_mg *g) {
    free(g->V); free(g->lam); free(g->Sd); free(g->Qd);
    g->V = g->lam = g->Sd = g->Qd = NULL;
    g->sep = 0;
}

static void mg_sep_setup(vdo_mg *g) {
    mg_sep_free(g);
    const vdo_level *L = &g->L[g->nl - 1];
    const int M = L->m, Nn = L->n, rows = g->fem ? Nn - 1 : Nn;
    const int tdir = rows <= M ? 0 : 1, T = tdir == 0 ? rows : M, P = tdir == 0 ? M : rows;
    if (T > VDO_SEP_MAXT || rows < 1) return;
    double *Rd = xcalloc(T, sizeof(double)), *Pd = xcalloc(T, sizeof(double)), Re = 0, Pe = 0;
    g->Sd = xcalloc(P, sizeof(double));
    g->Qd = xcalloc(P, sizeof(double));
    if (!g->fem) {
        const int lo_t = tdir == 0 ? g->bc[2] : g->bc[0], hi_t = tdir == 0 ? g->bc[3] : g->bc[1];
        const int lo_p = tdir == 0 ? g->bc[0] : g->bc[2], hi_p = tdir == 0 ? g->bc[1] : g->bc[3];
        const double ct = tdir == 0 ? L->cy : L->cx, cp = tdir == 0 ? L->cx : L->cy;
        cc_line(T, g->s * ct, g->sigma, lo_t, hi_t, Pd, &Pe);
        cc_line(P, g->s * cp, 0.0, lo_p, hi_p, g->Sd, &g->Se);
        for (int i = 0; i < T; i++) Rd[i] = 1;
        for (int i = 0; i < P; i++) g->Qd[i] = 1;
        Re = 0; g->Qe = 0;
    } else {
        /* scales from the edge couplings: fex = -4 alpha + 2 beta, fey = 2 alpha - 4 beta */
        const double beta = -(g->fex + 2 * g->fey) / 6, alpha = (2 * beta - g->fex) / 4;
        /* x line (all M nodes, natural ends) and y line (rows 1..Nn-1, natural top) */
        #define KX(i) ((i) == 0 || (i) == M - 1 ? 1.0 : 2.0)
        #define MX(i) ((i) == 0 || (i) == M - 1 ? 2.0 : 4.0)
        #define KY(q) ((q) == Nn - 2 ? 1.0 : 2.0)
        #define MY(q) ((q) == Nn - 2 ? 2.0 : 4.0)
        if (tdir == 0) { /* A = My (x) alpha Kx + Ky (x) beta Mx */
            for (int q = 0; q < T; q++) { Rd[q] = MY(q); Pd[q] = KY(q); }
            Re = 1; Pe = -1;
            for (int i = 0; i < P; i++) { g->Sd[i] = alpha * KX(i); g->Qd[i] = beta * MX(i); }
            g->Se = -alpha; g->Qe = beta;
        } else {         /* A = Mx (x) beta Ky + Kx (x) alpha My */
            for (int i = 0; i < T; i++) { Rd[i] = MX(i); Pd[i] = KX(i); }
            Re = 1; Pe = -1;
            for (int q = 0; q < P; q++) { g->Sd[q] = beta * KY(q); g->Qd[q] = alpha * MY(q); }
            g->Se = -beta; g->Qe = alpha;
        }
        #undef KX
        #undef MX
        #undef KY
        #undef MY
    }
    /* R = L L^T (bidiagonal L: diagonal l, subdiagonal gs), C = L^-1 Pm L^-T */
    double *l = xcalloc(T, sizeof(double)), *gs = xcalloc(T, sizeof(double));
    for (int i = 0; i < T; i++) {
        if (i > 0) gs[i] = Re / l[i - 1];
        l[i] = sqrt(Rd[i] - gs[i] * gs[i]);
    }
    double *C = xcalloc((size_t)T * T, sizeof(double)), *X = xcalloc((size_t)T * T, sizeof(double));
    for (int j = 0; j < T; j++) /* X = L^-1 Pm, column j */
        for (int i = 0; i < T; i++) {
            double v = i == j ? Pd[i] : (abs(i - j) == 1 ? Pe : 0.0);
            if (i > 0) v -= gs[i] * X[(i - 1) * T + j];
            X[i * T + j] = v / l[i];
        }
    for (int j = 0; j < T; j++) /* C^T = L^-1 X^T, so row j of C solves L c = (row j of X) */
        for (int i = 0; i < T; i++) {
            double v = X[j * T + i];
            if (i > 0) v -= gs[i] * C[j * T + i - 1];
            C[j * T + i] = v / l[i];
        }
    for (int i = 0; i < T; i++)
        for (int j = i + 1; j < T; j++) C[i * T + j] = C[j * T + i] = 0.5 * (C[i * T + j] + C[j * T + i]);
    double *W = xcalloc((size_t)T * T, sizeof(double));
    g->lam = xcalloc(T, sizeof(double));
    sep_jacobi(T, C, g->lam, W);
    g->V = xcalloc((size_t)T * T, sizeof(double));
    for (int k = 0; k < T; k++) /* V = L^-T W */
        for (int i = T - 1; i >= 0; i--) {
            double v = W[i * T + k];
            if (i + 1 < T) v -= gs[i + 1] * g->V[(i + 1) * T + k];
            g->V[i * T + k] = v / l[i];
        }
    free(Rd); free(Pd); free(l); free(gs); free(C); free(X); free(W);
    g->sep = 1; g->tdir = tdir; g->sT = T; g->sP = P; g->yoff = g->fem ? 1 : 0;
}

/* z = A^-1 b on the coarsest level (sep set). */
static void mg_sep_solve(vdo_mg *g, const vdo_level *L, const double *b, double *z) {
    const int M = L->m, N = M * L->n, T = g->sT, P = g->sP;
    double *rhs = xcalloc(N, sizeof(double));
    memcpy(rhs, b, sizeof(double) * N);
    memset(z, 0, sizeof(double) * N);
    if (g->fem) { /* Dirichlet row, then its coupling moved to the right-hand side */
        double *az = xcalloc(N, sizeof(double));
        for (int i = 0; i < M; i++) z[i] = b[i] / fem_acc(g, L, i, 0);
        level_apply(g, L, z, az);
        for (int k = M; k < N; k++) rhs[k] -= az[k];
        free(az);
    }
    double *h = xcalloc((size_t)T * P, sizeof(double));
    #define IDX(p, t) (g->tdir == 0 ? (p) + (size_t)M * ((t) + g->yoff) : (t) + (size_t)M * ((p) + g->yoff))
    for (int k = 0; k < T; k++)
        for (int p = 0; p < P; p++) {
            double v = 0;
            for (int t = 0; t < T; t++) v += g->V[t * T + k] * rhs[IDX(p, t)];
            h[(size_t)k * P + p] = v;
        }
    double *cpv = xcalloc(P, sizeof(double));
    for (int k = 0; k < T; k++) { /* Thomas on S + lambda_k Q */
        double *hk = h + (size_t)k * P, e = g->Se + g->lam[k] * g->Qe, y = 0;
        for (int p = 0; p < P; p++) {
            double den = g->Sd[p] + g->lam[k] * g->Qd[p] - (p > 0 ? e * cpv[p - 1] : 0.0);
            y = (hk[p] - (p > 0 ? e * y : 0.0)) / den;
            cpv[p] = e / den;
            hk[p] = y;
        }
        for (int p = P - 2; p >= 0; p--) hk[p] -= cpv[p] * hk[p + 1];
    }
    for (int t = 0; t < T; t++)
        for (int p = 0; p < P; p++) {
            double v = 0;
            for (int k = 0; k < T; k++) v += g->V[t * T + k] * h[(size_t)k * P + p];
            z[IDX(p, t)] = v;
        }
    #undef IDX
    free(h); free(cpv); free(rhs);
}

vdo_mg *vdo_mg_create(int kind, int m, int n, double cx, double cy, int x_sym) {
    vdo_mg *g = (vdo_mg *)xcalloc(1, sizeof(vdo_mg));
    g->kind = kind;
    g->fem = kind == VDO_KIND_FEM;
    g->nu1 = 2; g->nu2 = 2; g->ncs = 8;
    g->nu1s = 2; g->nu2s = 2;
    if (kind == VDO_KIND_MAC) {
        /* mgs_mac.hh:37-50: Neumann left/right/top, Dirichlet face at bottom */
        g->sigma = 0; g->s = 1;
        g->bc[0] = BC_N; g->bc[1] = BC_N; g->bc[2] = BC_DF; g->bc[3] = BC_N;
    } else if (kind == VDO_KIND_VX || kind == VDO_KIND_VY) {
        /* visco_impl.hh:32-53: A = I - vfx dxx - vfy dyy.  Left: x_sym ->
         * u odd (3 vfx), v even (vfx); else Dirichlet ghost (2 vfx).  Right
         * and top: Dirichlet ghosts moved to the RHS (2 vf).  Bottom: Neumann. */
        g->sigma = 1; g->s = -1;
        g->bc[0] = x_sym ? (kind == VDO_KIND_VX ? BC_DF : BC_N) : BC_DC;
        g->bc[1] = BC_DC; g->bc[2] = BC_N; g->bc[3] = BC_DC;
        g->vfx = cx; g->vfy = cy;
    } else {
        /* mgs_fem.cc:8-11 with r = dx/dy */
        double dxdy = cx / cy, dydx = cy / cx;
        g->fm = 4. / 3. * (dxdy + dydx);
        g->fey = 1. / 3. * (-2 * dxdy + dydx); g->hey = 0.5 * g->fey;
        g->fex = 1. / 3. * (-2 * dydx + dxdy); g->hex = 0.5 * g->fex;
        g->fc = -1. / 6. * (dxdy + dydx);
    }
    mg_build_levels(g, m, n, cx, cy);
    mg_sep_setup(g);
    g->z_prev = (double *)xcalloc((size_t)m * n, sizeof(double));
    return g;
}

void vdo_mg_destroy(vdo_mg *g) {
    if (!g) return;
    for (int l = 0; l < g->nl; l++) {
        if (l) { free(g->L[l].z); free(g->L[l].b); }
        free(g->L[l].r);
    }
    free(g->z_prev);
    mg_sep_free(g);
    free(g);
}

void vdo_mg_set(vdo_mg *g, int nu1, int nu2, int ncs) {
    g->nu1 = g->nu1s = nu1;
    g->nu2 = g->nu2s = nu2;
    g->ncs = ncs;
}
int vdo_mg_levels(const vdo_mg *g) { return g->nl; }

/* Re-scale the cell-centred couplings (visco re-setup, visco_impl.cc:26-30). */
static void mg_set_coeffs(vdo_mg *g, double cx, double cy) {
    for (int l = 0; l < g->nl; l++) { g->L[l].cx = cx; g->L[l].cy = cy; cx *= 0.25; cy *= 0.25; }
    mg_sep_setup(g);
}

/* ---- cell-centred operator ---- */
static inline double cc_diag(const vdo_mg *g, const vdo_level *L, int i, int j) {
    double d = 0;
    d += (i > 0 ? 1.0 : bc_diag_factor(g->bc[0])) * L->cx;
    d += (i < L->m - 1 ? 1.0 : bc_diag_factor(g->bc[1])) * L->cx;
    d += (j > 0 ? 1.0 : bc_diag_factor(g->bc[2])) * L->cy;
    d += (j < L->n - 1 ? 1.0 : bc_diag_factor(g->bc[3])) * L->cy;
    return g->sigma - g->s * d;
}

static inline double cc_off(const vdo_level *L, const double *z, int i, int j) {
    const double *w = z + i + (size_t)L->m * j;
    double sx = 0, sy = 0;
    if (i > 0) sx += w[-1];
    if (i < L->m - 1) sx += w[1];
    if (j > 0) sy += w[-L->m];
    if (j < L->n - 1) sy += w[L->m];
    return L->cx * sx + L->cy * sy;
}

/* ---- FEM operator (mgs_fem.hh:42-56, mgs_fem.cc:34-50) ---- */
static inline double fem_acc(const vdo_mg *g, const vdo_level *L, int i, int j) {
    return (i > 0 && i < L->m - 1 ? 1.0 : 0.5) * (j > 0 && j < L->n - 1 ? 1.0 : 0.5) * g->fm;
}

static inline double fem_off(const vdo_mg *g, const vdo_level *L, const double *z, int i, int j) {
    const int M = L->m;
    const double *w = z + i + (size_t)M * j, *wm = w - M;
    if (j == 0) return 0;
    if (j == L->n - 1) {
        if (i == 0) return g->fc * wm[1] + g->hey * wm[0] + g->hex * w[1];
        if (i == M - 1) return g->fc * wm[-1] + g->hey * wm[0] + g->hex * w[-1];
        return g->fc * (wm[-1] + wm[1]) + g->fey * wm[0] + g->hex * (w[-1] + w[1]);
    }
    const double *wp = w + M;
    if (i == 0) return g->fc * (wm[1] + wp[1]) + g->hey * (wm[0] + wp[0]) + g->fex * w[1];
    if (i == M - 1) return g->fc * (wm[-1] + wp[-1]) + g->hey * (wm[0] + wp[0]) + g->fex * w[-1];
    return g->fc * (wm[-1] + wm[1] + wp[-1] + wp[1]) + g->fey * (wm[0] + wp[0]) + g->fex * (w[-1] + w[1]);
}

static void level_apply(const vdo_mg *g, const vdo_level *L, const double *z, double *out) {
    const int M = L->m, N = L->n;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) {
            size_t ij = i + (size_t)M * j;
            if (g->fem) out[ij] = fem_acc(g, L, i, j) * z[ij] + fem_off(g, L, z, i, j);
            else out[ij] = g->s * cc_off(L, z, i, j) + cc_diag(g, L, i, j) * z[ij];
        }
}

static void level_residual(const vdo_mg *g, const vdo_level *L, const double *b, const double *z, double *r) {
    const int M = L->m, N = L->n;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) {
            size_t ij = i + (size_t)M * j;
            double az = g->fem ? fem_acc(g, L, i, j) * z[ij] + fem_off(g, L, z, i, j)
                               : g->s * cc_off(L, z, i, j) + cc_diag(g, L, i, j) * z[ij];
            r[ij] = b[ij] - az;
        }
}

/* One smoothing sweep.  Cell-centred: red-black Gauss-Seidel (red = i+j even
 * first).  FEM: four-colour Gauss-Seidel over (i&1)+2(j&1) = 0,1,2,3, which
 * is a true colouring of the 9-point stencil.  FEM row j=0 is Dirichlet and
 * is never smoothed. */
static void level_sweep(const vdo_mg *g, const vdo_level *L, const double *b, double *z) {
    const int M = L->m, N = L->n;
    if (!g->fem) {
        for (int c = 0; c < 2; c++) {
#pragma omp parallel for schedule(static)
            for (int j = 0; j < N; j++)
                for (int i = (j + c) & 1; i < M; i += 2) {
                    size_t ij = i + (size_t)M * j;
                    z[ij] = (b[ij] - g->s * cc_off(L, z, i, j)) / cc_diag(g, L, i, j);
                }
        }
    } else {
        for (int c = 0; c < 4; c++) {
            int ci = c & 1, cj = c >> 1;
#pragma omp parallel for schedule(static)
            for (int j = (cj ? 1 : 2); j < N; j += 2)
                for (int i = ci; i < M; i += 2) {
                    size_t ij = i + (size_t)M * j;
                    z[ij] = (b[ij] - fem_off(g, L, z, i, j)) / fem_acc(g, L, i, j);
                }
        }
    }
}

/* Restriction of r on level l to b on level l+1.  Cell-centred: average of
 * the four children.  FEM: transpose of bilinear interpolation (full
 * weighting, unnormalised), which makes the coarse Q1 stiffness equal to the
 * fine one (Galerkin); coarse Dirichlet row gets 0. */
static void level_restrict(const vdo_mg *g, const vdo_level *F, const double *r, vdo_level *C, double *bc) {
    const int Mc = C->m, Nc = C->n, Mf = F->m, Nf = F->n;
#pragma omp parallel for schedule(static)
    for (int J = 0; J < Nc; J++)
        for (int I = 0; I < Mc; I++) {
            double v = 0;
            if (!g->fem) {
                const double *q = r + 2 * I + (size_t)Mf * 2 * J;
                v = 0.25 * (q[0] + q[1] + q[Mf] + q[Mf + 1]);
            } else if (J > 0) {
                for (int b2 = -1; b2 <= 1; b2++) {
                    int jf = 2 * J + b2;
                    if (jf < 0 || jf >= Nf) continue;
                    double wy = b2 ? 0.5 : 1.0, sx = 0;
                    for (int a2 = -1; a2 <= 1; a2++) {
                        int i_f = 2 * I + a2;
                        if (i_f < 0 || i_f >= Mf) continue;
                        sx += (a2 ? 0.5 : 1.0) * r[i_f + (size_t)Mf * jf];
                    }
                    v += wy * sx;
                }
            }
            bc[I + (size_t)Mc * J] = v;
        }
}

/* z_fine += P e_coarse.  Cell-centred: bilinear (9/16, 3/16, 3/16, 1/16) with
 * ghost coarse values from the boundary type (N: copy, DF: negate, DC: 0).
 * FEM: nodal bilinear interpolation; Dirichlet row untouched. */
static void level_prolong(const vdo_mg *g, vdo_level *F, double *zf, const vdo_level *C, const double *ec) {
    const int Mf = F->m, Nf = F->n, Mc = C->m, Nc = C->n;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Nf; j++)
        for (int i = 0; i < Mf; i++) {
            double e;
            if (!g->fem) {
                int I = i >> 1, J = j >> 1;
                int I2 = (i & 1) ? I + 1 : I - 1, J2 = (j & 1) ? J + 1 : J - 1;
                double fx = 1, fy = 1;
                if (I2 < 0) { fx = bc_ghost_factor(g->bc[0]); I2 = I; }
                else if (I2 >= Mc) { fx = bc_ghost_factor(g->bc[1]); I2 = I; }
                if (J2 < 0) { fy = bc_ghost_factor(g->bc[2]); J2 = J; }
                else if (J2 >= Nc) { fy = bc_ghost_factor(g->bc[3]); J2 = J; }
                double e00 = ec[I + (size_t)Mc * J], e10 = fx * ec[I2 + (size_t)Mc * J];
                double e01 = fy * ec[I + (size_t)Mc * J2], e11 = fx * fy * ec[I2 + (size_t)Mc * J2];
                e = 0.5625 * e00 + 0.1875 * (e10 + e01) + 0.0625 * e11;
            } else {
                if (j == 0) continue;
                int I = i >> 1, J = j >> 1;
                const double *q = ec + I + (size_t)Mc * J;
                if (!(i & 1) && !(j & 1)) e = q[0];
                else if (!(j & 1)) e = 0.5 * (q[0] + q[1]);
                else if (!(i & 1)) e = 0.5 * (q[0] + q[Mc]);
                else e = 0.25 * (q[0] + q[1] + q[Mc] + q[Mc + 1]);
            }
            zf[i + (size_t)Mf * j] += e;
        }
}

static void mg_vcycle_level(vdo_mg *g, int l, const double *b, double *z) {
    vdo_level *L = &g->L[l];
    if (l == g->nl - 1) {
        if (g->sep) {
            mg_sep_solve(g, L, b, z);
        } else {
            for (int k = 0; k < g->ncs; k++) level_sweep(g, L, b, z);
        }
        return;
    }
    /* V(nu1, nu2) on the finest level, V(nu1s, nu2s) on the restricted ones */
    const int n1 = l == 0 ? g->nu1 : g->nu1s, n2 = l == 0 ? g->nu2 : g->nu2s;
    for (int k = 0; k < n1; k++) level_sweep(g, L, b, z);
    level_residual(g, L, b, z, L->r);
    vdo_level *C = &g->L[l + 1];
    level_restrict(g, L, L->r, C, C->b);
    memset(C->z, 0, sizeof(double) * (size_t)C->m * C->n);
    mg_vcycle_level(g, l + 1, C->b, C->z);
    level_prolong(g, L, z, C, C->z);
    for (int k = 0; k < n2; k++) level_sweep(g, L, b, z);
}

/* FEM Dirichlet row (mgs_fem.cc:38-41 + fluid_2d.cc:377): z = b / a_cc. */
static void mg_fem_dirichlet(vdo_mg *g, const double *b, double *z) {
    if (!g->fem) return;
    vdo_level *L = &g->L[0];
    for (int i = 0; i < L->m; i++) z[i] = b[i] / fem_acc(g, L, i, 0);
}

static double dot_self(const double *a, size_t n) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < n; i++) s += a[i] * a[i];
    return s;
}

void vdo_mg_apply(vdo_mg *g, const double *z, double *out) { level_apply(g, &g->L[0], z, out); }

double vdo_mg_resnorm(vdo_mg *g, const double *b, const double *z) {
    vdo_level *L = &g->L[0];
    level_residual(g, L, b, z, L->r);
    return sqrt(dot_self(L->r, (size_t)L->m * L->n));
}

void vdo_mg_vcycles(vdo_mg *g, const double *b, double *z, int k) {
    mg_fem_dirichlet(g, b, z);
    for (int c = 0; c < k; c++) mg_vcycle_level(g, 0, b, z);
}

int vdo_mg_solve(vdo_mg *g, const double *b, double *z, double tol, int max_cycles) {
    vdo_level *L = &g->L[0];
    size_t N = (size_t)L->m * L->n;
    mg_fem_dirichlet(g, b, z);
    double bn = sqrt(dot_self(b, N));
    if (bn == 0) { memset(z, 0, N * sizeof(double)); g->last_cycles = 0; return 0; }
    for (int c = 0;; c++) {
        double rn = vdo_mg_resnorm(g, b, z);
        if (!(rn == rn)) { g->last_cycles = c; return -1; }
        if (rn <= tol * bn) { g->last_cycles = c; return c; }
        if (c >= max_cycles) { g->last_cycles = c; return -1; }
        mg_vcycle_level(g, 0, b, z);
    }
}

/* mgs_common.cc:17-50: linear extrapolation in time from the last two
 * solutions, after two solves have been stored. */
static void mg_initial_guess(vdo_mg *g, double *z, double t) {
    size_t N = (size_t)g->L[0].m * g->L[0].n;
    if (g->scount < 2) {
        if (g->scount == 1) { memcpy(g->z_prev, z, N * sizeof(double)); g->t0 = t; }
        else g->t_prev = t;
        g->scount++;
    } else {
        if (fabs(g->t0 - g->t_prev) < DBL_EPSILON * 10) {
            fputs("Multigrid routine called for duplicate times\n", stderr);
            return;
        }
        double a = 1. / (g->t0 - g->t_prev), bb = (t - g->t_prev) * a, c = (t - g->t0) * a;
        double *zp = g->z_prev;
#pragma omp parallel for schedule(static)
        for (size_t ij = 0; ij < N; ij++) {
            double zz = z[ij];
            z[ij] = bb * zz - c * zp[ij];
            zp[ij] = zz;
        }
        g->t_prev = g->t0;
        g->t0 = t;
    }
}

/* ------------------------------------------------------------------------ */
/* Small pieces                                                              */
/* ------------------------------------------------------------------------ */

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* fluid_2d.cc:445-449 */
static inline double delta_lim(double f0, double f1, double f2) {
    double drm = f1 - f0, drp = f2 - f1;
    return drm > 0 ? (drp > 0 ? 2 * dmin(drm, drp) : 0) : (drp > 0 ? 0 : -2 * dmax(drm, drp));
}
/* fluid_2d.cc:434-437 */
static inline double delta_f(double f0, double f1, double f2) {
    double drc = 0.5 * (f2 - f0);
    return drc > 0 ? dmin(drc, delta_lim(f0, f1, f2)) : dmax(drc, -delta_lim(f0, f1, f2));
}
/* fluid_2d.cc:422-426: fourth-order monotonicity-limited slope */
double vdo_mono_diff(double f0, double f1, double f2, double f3, double f4) {
    double tdrc = f3 - f1, s = (1 / 6.) * (4 * tdrc - delta_f(f0, f1, f2) - delta_f(f2, f3, f4)),
           t = delta_lim(f1, f2, f3);
    return tdrc > 0 ? dmin(s, t) : dmax(s, -t);
}
/* fluid_2d.cc:455-461 */
void vdo_godunov_set(double *u0, double *u1, double *v0, double *v1) {
    if (*u0 > -*u1) {
        if (*u0 > 0) { *u1 = *u0; *v1 = *v0; return; }
    } else if (*u1 < 0) return;
    *u1 = 0;
    *v1 = 0.5 * (*v0 + *v1);
}
/* fluid_2d.cc:469-476 */
void vdo_godunov_set_tang(double *u0, double *u1, double *v0, double *v1) {
    if (*u0 > -*u1) {
        if (*u0 > 0) { *u1 = *u0; *v1 = *v0; return; }
    } else if (*u1 < 0) { *u0 = *u1; return; }
    *v1 = 0.5 * (*v0 + *v1);
    *u0 = 0.5 * (*u0 + *u1);
    *u1 = 0;
}

/* fluid_2d_sio.cc:395-405 */
static inline void simpsons_terms(int x_sym, double xms, double xps, double ysq, double p, double *fsu, double *fsv) {
    double rm = xms * xms + ysq;
    rm = 1 / (rm * rm);
    if (x_sym) {
        double rp = xps * xps + ysq;
        rp = 1 / (rp * rp);
        *fsu = p * (xms * rm + xps * rp);
        *fsv = -p * ((xms * xms - ysq) * rm + (xps * xps - ysq) * rp);
    } else {
        *fsu = p * xms * rm;
        *fsv = -p * (xms * xms - ysq) * rm;
    }
}
/* fluid_2d_sio.cc:366-385 (returns the increments instead of adding) */
void vdo_simpson_velocity(int me, double ax, double dx, int x_sym, const double *p,
                          double x, double y, double prefac, double *du, double *dv) {
    double su, sv, fsu, fsv, xms = x - ax, xps = x + ax, ysq = y * y;
    simpsons_terms(x_sym, xms, xps, ysq, p[0], &su, &sv);
    xms -= dx; xps += dx;
    for (int j = 1; j < me - 1; j++, xms -= dx, xps += dx) {
        simpsons_terms(x_sym, xms, xps, ysq, p[j], &fsu, &fsv);
        double fac = (j & 1) ? 4. : 2.;
        su += fac * fsu; sv += fac * fsv;
    }
    simpsons_terms(x_sym, xms, xps, ysq, p[me - 1], &fsu, &fsv);
    su += fsu; sv += fsv;
    *du = 2 * y * prefac * su;
    *dv = prefac * sv;
}

/* Gaussian elimination with partial pivoting for a tridiagonal system: the
 * algorithm LAPACK documents for dgtsv (called at tri_solve.cc:28). */
int vdo_tri_solve_pivot(int n, double *dl, double *d, double *du, double *b) {
    if (n <= 0) return 0;
    double *du2 = (double *)xcalloc((size_t)n, sizeof(double));
    for (int i = 0; i < n - 1; i++) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] == 0) { free(du2); return i + 1; }
            double fact = dl[i] / d[i];
            d[i + 1] -= fact * du[i];
            b[i + 1] -= fact * b[i];
            du2[i] = 0;
        } else {
            double fact = d[i] / dl[i];
            d[i] = dl[i];
            double tmp = d[i + 1];
            d[i + 1] = du[i] - fact * tmp;
            if (i < n - 2) { du2[i] = du[i + 1]; du[i + 1] = -fact * du[i + 1]; }
            du[i] = tmp;
            tmp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = tmp - fact * b[i + 1];
        }
    }
    if (d[n - 1] == 0) { free(du2); return n; }
    b[n - 1] /= d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / d[i];
    free(du2);
    return 0;
}

/* tri_solve.cc:37-49 (Thomas algorithm, no pivoting) */
void vdo_tri_solve_thomas(int m, double *a, double *b, double *c, double *d) {
    c[0] /= b[0];
    for (int i = 1; i < m - 1; i++) c[i] /= b[i] - a[i] * c[i - 1];
    d[0] /= b[0];
    for (int i = 1; i < m; i++) d[i] = (d[i] - a[i] * d[i - 1]) / (b[i] - a[i] * c[i - 1]);
    for (int i = m - 2; i >= 0; i--) d[i] -= c[i] * d[i + 1];
}

/* ------------------------------------------------------------------------ */
/* Simulation state                                                          */
/* ------------------------------------------------------------------------ */

/* The reference's 19-double cell record (fields.hh:7-19) as separate arrays. */
enum { F_U, F_V, F_P, F_UX, F_UY, F_VX, F_VY, F_UD, F_VD, F_UL, F_VL, F_UR, F_VR, F_UU, F_VU,
       F_C0, F_C1, F_C2, F_C3, NF };

struct vdo_sim {
    vdo_params P;
    int m, n, me, ne, ml, mn;
    double dx, dy, xsp, ysp, xxsp, yysp, rhoinv, muinv;
    double time;
    double *F[NF];           /* padded (m+4)x(n+4), cell (i,j) at (i+2)+ml(j+2) */
    double *src, *s_bc;      /* fluid_2d.cc:18-19 */
    /* gas layer (gas_layer.hh:26-34, gas_layer_solve.hh:163-173) */
    double *gp, *gtau, *gp_old, *gtau_old, *h, *hx, *ht, *fbd, *pg, *pg_old, Vframe;
    double gxsp, gxxsp;
    double *ta, *tb, *tc, *td;
    vdo_mg *mac, *fem, *vx, *vy;
    double *zmac, *zfem, *zvx, *zvy; /* solver solutions (mgs_common.hh:38) */
    double rel_tol;
    int max_cycles, last_newton;
    /* the reference's quirks Q1 (tau written 2 cells off) and Q13 (in-place
     * height sweep); 1 = as in the reference (the default).  Tests switch
     * them off to check the mirror symmetry they break. */
    int q1_tau_shift, q13_inplace;
    double *tm;  /* tracer positions (fluid_2d::tm), 2 ntrace */
    int ntrace;
};

#define AT(a, i, j) ((a)[((size_t)(j) + 2) * ml + (size_t)((i) + 2)])

vdo_sim *vdo_create(const vdo_params *p) {
    vdo_sim *s = (vdo_sim *)xcalloc(1, sizeof(vdo_sim));
    s->P = *p;
    int m = p->m, n = p->n;
    s->m = m; s->n = n; s->me = m + 1; s->ne = n + 1; s->ml = m + 4; s->mn = m * n;
    /* fluid_2d.cc:14-17 */
    s->dx = (p->bx - p->ax) / m; s->dy = (p->by - p->ay) / n;
    s->xsp = 1. / s->dx; s->ysp = 1. / s->dy; s->xxsp = s->xsp * s->xsp; s->yysp = s->ysp * s->ysp;
    s->rhoinv = 1. / p->rho; s->muinv = 1. / p->mu;
    size_t NP = (size_t)s->ml * (n + 4);
    for (int f = 0; f < NF; f++) s->F[f] = (double *)xcalloc(NP, sizeof(double));
    s->src = (double *)xcalloc((size_t)s->me * s->ne, sizeof(double));
    s->s_bc = s->src + s->mn; /* quirk Q9: s_bc aliases the tail of src */
    /* gas layer arrays, zero-initialised (quirks Q1/Q6) */
    s->gp = (double *)xcalloc(s->me, sizeof(double));
    s->gtau = (double *)xcalloc(s->ml, sizeof(double));
    s->gp_old = (double *)xcalloc(s->me, sizeof(double));
    s->gtau_old = (double *)xcalloc(s->ml, sizeof(double));
    s->h = (double *)xcalloc(m, sizeof(double));
    s->hx = (double *)xcalloc(s->me, sizeof(double));
    s->ht = (double *)xcalloc(m, sizeof(double));
    s->fbd = (double *)xcalloc(m, sizeof(double));
    s->pg = (double *)xcalloc(s->me, sizeof(double));
    s->pg_old = (double *)xcalloc(s->me, sizeof(double));
    s->ta = (double *)xcalloc(s->me, sizeof(double));
    s->tb = (double *)xcalloc(s->me, sizeof(double));
    s->tc = (double *)xcalloc(s->me, sizeof(double));
    s->td = (double *)xcalloc(s->me, sizeof(double));
    s->gxsp = 1. / ((p->bx - p->ax) / m); /* gas_layer.cc:12, gas_layer_solve.cc:15-17 */
    s->gxxsp = s->gxsp * s->gxsp;
    s->Vframe = p->V;
    s->mac = vdo_mg_create(VDO_KIND_MAC, m, n, s->xxsp, s->yysp, p->x_sym);
    s->fem = vdo_mg_create(VDO_KIND_FEM, s->me, s->ne, s->dx, s->dy, p->x_sym);
    s->zmac = (double *)xcalloc(s->mn, sizeof(double));
    s->zfem = (double *)xcalloc((size_t)s->me * s->ne, sizeof(double));
    if (p->implicit_visc) {
        /* visco_impl.cc:11-13: vfx starts at 0, set up on first solve.
         * quirk Q3: z is never cleared there; the oracle starts from zero. */
        s->vx = vdo_mg_create(VDO_KIND_VX, m, n, 0, 0, p->x_sym);
        s->vy = vdo_mg_create(VDO_KIND_VY, m, n, 0, 0, p->x_sym);
        s->zvx = (double *)xcalloc(s->mn, sizeof(double));
        s->zvy = (double *)xcalloc(s->mn, sizeof(double));
    }
    s->rel_tol = 1e-10; s->max_cycles = 200;
    s->q1_tau_shift = s->q13_inplace = 1;
    return s;
}

int vdo_set_tracers(vdo_sim *s, const double *pos, int ntrace) {
    free(s->tm);
    s->tm = NULL;
    s->ntrace = ntrace > 0 ? ntrace : 0;
    if (s->ntrace) {
        s->tm = (double *)xcalloc(2 * (size_t)ntrace, sizeof(double));
        memcpy(s->tm, pos, 2 * (size_t)ntrace * sizeof(double));
    }
    return 0;
}

int vdo_get_tracers(const vdo_sim *s, double *pos, int ntrace) {
    if (ntrace != s->ntrace) return -1;
    if (ntrace) memcpy(pos, s->tm, 2 * (size_t)ntrace * sizeof(double));
    return 0;
}

/* fluid_2d.cc:507-525: bilinear velocity at each tracer, from the cell whose
 * centre is at or below-left of it (C truncation toward zero, then the index
 * clamped to [-1, m-1] x [-1, n-1]) */
static void update_tracers(vdo_sim *s, double dt) {
    const int m = s->m, n = s->n, ml = s->ml;
    const double *u = s->F[F_U], *v = s->F[F_V];
    for (int k = 0; k < s->ntrace; k++) {
        double *tp = s->tm + 2 * k;
        double x = (tp[0] - s->P.ax) * s->xsp - 0.5, y = (tp[1] - s->P.ay) * s->ysp - 0.5;
        int i = (int)x, j = (int)y;
        if (i < -1) i = -1;
        if (i > m - 1) i = m - 1;
        if (j < -1) j = -1;
        if (j > n - 1) j = n - 1;
        x -= i;
        y -= j;
        tp[0] += dt * ((1 - y) * (AT(u, i, j) * (1 - x) + AT(u, i + 1, j) * x) +
                       y * (AT(u, i, j + 1) * (1 - x) + AT(u, i + 1, j + 1) * x));
        tp[1] += dt * ((1 - y) * (AT(v, i, j) * (1 - x) + AT(v, i + 1, j) * x) +
                       y * (AT(v, i, j + 1) * (1 - x) + AT(v, i + 1, j + 1) * x));
    }
}

int vdo_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

int vdo_set_quirks(vdo_sim *s, int q1_tau_shift, int q13_inplace) {
    s->q1_tau_shift = q1_tau_shift != 0;
    s->q13_inplace = q13_inplace != 0;
    return 0;
}

void vdo_destroy(vdo_sim *s) {
    if (!s) return;
    for (int f = 0; f < NF; f++) free(s->F[f]);
    free(s->src);
    free(s->gp); free(s->gtau); free(s->gp_old); free(s->gtau_old); free(s->h); free(s->hx);
    free(s->ht); free(s->fbd); free(s->pg); free(s->pg_old);
    free(s->ta); free(s->tb); free(s->tc); free(s->td);
    vdo_mg_destroy(s->mac); vdo_mg_destroy(s->fem); vdo_mg_destroy(s->vx); vdo_mg_destroy(s->vy);
    free(s->zmac); free(s->zfem); free(s->zvx); free(s->zvy);
    free(s->tm);
    free(s);
}

void vdo_set_mg(vdo_sim *s, int nu1, int nu2, double rel_tol, int max_cycles) {
    vdo_mg *gs[4] = {s->mac, s->fem, s->vx, s->vy};
    for (int k = 0; k < 4; k++) if (gs[k]) { gs[k]->nu1 = gs[k]->nu1s = nu1; gs[k]->nu2 = gs[k]->nu2s = nu2; }
    s->rel_tol = rel_tol; s->max_cycles = max_cycles;
}

int vdo_set_schedule(vdo_sim *s, int which, int nu1, int nu2) {
    return vdo_set_schedule_levels(s, which, nu1, nu2, nu1, nu2);
}

int vdo_set_schedule_levels(vdo_sim *s, int which, int nu1, int nu2, int nu1s, int nu2s) {
    vdo_mg *gs[4] = {s->mac, s->fem, s->vx, s->vy};
    if (which < 0 || which > 3 || !gs[which] || nu1 < 0 || nu2 < 0 || nu1 + nu2 == 0) return 1;
    if (nu1s < 0 || nu2s < 0 || nu1s + nu2s == 0) return 1;
    gs[which]->nu1 = nu1;
    gs[which]->nu2 = nu2;
    gs[which]->nu1s = nu1s;
    gs[which]->nu2s = nu2s;
    return 0;
}

double vdo_time(const vdo_sim *s) { return s->time; }
double vdo_vframe(const vdo_sim *s) { return s->Vframe; }
int vdo_set_time(vdo_sim *s, double t) { s->time = t; return 0; }
int vdo_last_newton(const vdo_sim *s) { return s->last_newton; }
int vdo_last_cycles(const vdo_sim *s, int kind) {
    vdo_mg *g = kind == VDO_KIND_MAC ? s->mac : kind == VDO_KIND_FEM ? s->fem : kind == VDO_KIND_VX ? s->vx : s->vy;
    return g ? g->last_cycles : -1;
}

static double *array_ptr(const vdo_sim *s, int id, long *len) {
    long NP = (long)s->ml * (s->n + 4);
    *len = 0;
    switch (id) {
    case VDO_U: *len = NP; return s->F[F_U];
    case VDO_V: *len = NP; return s->F[F_V];
    case VDO_P: *len = NP; return s->F[F_P];
    case VDO_UL: *len = NP; return s->F[F_UL];
    case VDO_VL: *len = NP; return s->F[F_VL];
    case VDO_UD: *len = NP; return s->F[F_UD];
    case VDO_VD: *len = NP; return s->F[F_VD];
    case VDO_C0: *len = NP; return s->F[F_C0];
    case VDO_C1: *len = NP; return s->F[F_C1];
    case VDO_SRC: *len = (long)s->me * s->ne; return s->src;
    case VDO_GP: *len = s->me; return s->gp;
    case VDO_GTAU: *len = s->ml; return s->gtau;
    case VDO_GP_OLD: *len = s->me; return s->gp_old;
    case VDO_GTAU_OLD: *len = s->ml; return s->gtau_old;
    case VDO_H: *len = s->m; return s->h;
    case VDO_HX: *len = s->me; return s->hx;
    case VDO_HT: *len = s->m; return s->ht;
    case VDO_FBD: *len = s->m; return s->fbd;
    case VDO_PG: *len = s->me; return s->pg;
    case VDO_PG_OLD: *len = s->me; return s->pg_old;
    case VDO_Z_MAC: *len = s->mn; return s->zmac;
    case VDO_Z_FEM: *len = (long)s->me * s->ne; return s->zfem;
    case VDO_Z_VX: *len = s->zvx ? s->mn : 0; return s->zvx;
    case VDO_Z_VY: *len = s->zvy ? s->mn : 0; return s->zvy;
    case VDO_UX: *len = NP; return s->F[F_UX];
    case VDO_UY: *len = NP; return s->F[F_UY];
    case VDO_VXS: *len = NP; return s->F[F_VX];
    case VDO_VYS: *len = NP; return s->F[F_VY];
    case VDO_UR: *len = NP; return s->F[F_UR];
    case VDO_VR: *len = NP; return s->F[F_VR];
    case VDO_UU: *len = NP; return s->F[F_UU];
    case VDO_VU: *len = NP; return s->F[F_VU];
    case VDO_C2: *len = NP; return s->F[F_C2];
    case VDO_C3: *len = NP; return s->F[F_C3];
    default: return NULL;
    }
}

long vdo_array_len(const vdo_sim *s, int id) { long len; array_ptr(s, id, &len); return len; }

int vdo_get_array(const vdo_sim *s, int id, double *out, long len) {
    long n; double *p = array_ptr(s, id, &n);
    if (!p || len != n) return -1;
    memcpy(out, p, (size_t)n * sizeof(double));
    return 0;
}

int vdo_set_array(vdo_sim *s, int id, const double *in, long len) {
    long n; double *p = array_ptr(s, id, &n);
    if (!p || len != n) return -1;
    memcpy(p, in, (size_t)n * sizeof(double));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Gas layer                                                                 */
/* ------------------------------------------------------------------------ */

/* gas_layer_solve.cc:58-62 */
static void gas_calculate_hx(vdo_sim *s) {
    int m = s->m;
    for (int i = 1; i < m; i++) s->hx[i] = s->gxsp * (s->h[i] - s->h[i - 1]);
    s->hx[0] = s->P.x_sym ? 0 : s->hx[1];
    s->hx[m] = s->hx[m - 1];
}

/* gas_layer_solve.cc:229-249 (tau written at [i], quirk Q1) */
static void gas_liquid_fields(vdo_sim *s) {
    const int m = s->m;
    const double fac = 0.5 * s->P.sigma * s->gxxsp, fac2 = s->P.sigma / s->P.R;
    double *p = s->gp, *pg = s->pg, *h = s->h;
    if (s->P.x_sym) {
        p[0] = pg[0] + 2 * fac * (h[1] - h[0]) - fac2;
        p[1] = pg[1] + fac * (h[2] - h[1]) - fac2;
    } else { p[0] = pg[0]; p[1] = pg[1]; }
    for (int i = 2; i < m - 1; i++) p[i] = pg[i] + fac * (h[i + 1] - h[i] - h[i - 1] + h[i - 2]) - fac2;
    p[m - 1] = pg[m - 1]; p[m] = pg[m];
    const int off = s->q1_tau_shift ? 0 : 2;
    for (int i = 0; i < m; i++) s->gtau[i + off] = -0.5 * s->gxsp * s->P.mug * (p[i + 1] - p[i]) * h[i];
}

/* gas_layer.cc:25-30 */
static void gas_apply_tau_bc(vdo_sim *s) {
    double *tau = s->gtau; int ml = s->ml;
    tau[0] = -tau[3]; tau[1] = -tau[2];
    tau[ml - 2] = -tau[ml - 3]; tau[ml - 1] = -tau[ml - 4];
}

/* gas_layer_solve.cc:32-53 */
static void gas_init_fields(vdo_sim *s) {
    const int m = s->m;
    const double dx = (s->P.bx - s->P.ax) / m;
    for (int i = 0; i < m; i++) s->fbd[i] = 0;
    for (int i = 0; i < m; i++) {
        double x = s->P.ax + (i + 0.5) * dx;
        s->h[i] = s->P.h0 + 0.5 / s->P.R * x * x;
    }
    gas_calculate_hx(s);
    for (int i = 0; i < m; i++) s->ht[i] = -s->P.V;
    for (int i = 0; i < s->me; i++) s->pg[i] = 0;
    gas_liquid_fields(s);
}

/* gas_layer_solve.cc:176-180 */
static inline double gas_h_ref(const vdo_sim *s, int i) {
    const double *h = s->h; int m = s->m;
    return i >= 0 ? (i < m ? h[i] : i == m ? 2 * h[m - 1] - h[m - 2] : 3 * h[m - 1] - 2 * h[m - 2])
                  : (s->P.x_sym ? (i == -1 ? h[0] : h[1]) : (i == -1 ? 2 * h[0] - h[1] : 3 * h[0] - 2 * h[1]));
}

/* gas_layer_solve.cc:168-170 */
static inline double eno2(double p0, double p1, double p2, double p3) {
    return fabs(p0 - 2 * p1 + p2) > fabs(p1 - 2 * p2 + p3) ? 3 * p1 - 4 * p2 + p3 : p0 - p2;
}

/* gas_layer_solve.cc:133-162.  Quirk: h is updated in place while i sweeps
 * upward, so the ENO2 stencil at i sees the new h[i-1], h[i-2]. */
static void gas_update_height(vdo_sim *s, double dt) {
    const int m = s->m, ml = s->ml;
    const double *u = s->F[F_U], *v = s->F[F_V];
    if (!s->q13_inplace) {
        /* every cell from the old heights, then all updated */
        for (int i = 0; i < m; i++) {
            double uc = 0.25 * s->gxsp * (AT(u, i, 0) + AT(u, i, -1));
            double vc = 0.5 * (AT(v, i, 0) + AT(v, i, -1));
            double dhdx = uc > 0 ? eno2(gas_h_ref(s, i + 1), s->h[i], gas_h_ref(s, i - 1), gas_h_ref(s, i - 2))
                                 : -eno2(gas_h_ref(s, i - 1), s->h[i], gas_h_ref(s, i + 1), gas_h_ref(s, i + 2));
            s->ht[i] = -uc * dhdx + vc - s->Vframe;
            s->fbd[i] += dt * vc;
        }
        for (int i = 0; i < m; i++) s->h[i] += dt * s->ht[i];
        gas_calculate_hx(s);
        return;
    }
    for (int i = 0; i < m; i++) {
        double uc = 0.25 * s->gxsp * (AT(u, i, 0) + AT(u, i, -1));
        double vc = 0.5 * (AT(v, i, 0) + AT(v, i, -1));
        double dhdx = uc > 0 ? eno2(gas_h_ref(s, i + 1), s->h[i], gas_h_ref(s, i - 1), gas_h_ref(s, i - 2))
                             : -eno2(gas_h_ref(s, i - 1), s->h[i], gas_h_ref(s, i + 1), gas_h_ref(s, i + 2));
        s->ht[i] = -uc * dhdx + vc - s->Vframe;
        s->h[i] += dt * s->ht[i];
        s->fbd[i] += dt * vc;
    }
    gas_calculate_hx(s);
}

/* gas_layer_solve.cc:67-129: assemble the Newton system for the implicit
 * lubrication equation, solve it, update pg; returns mean square residual
 * (quirk Q7: interior rows summed, divided by me).  *info is the tridiagonal
 * solver status. */
static double gas_newton_step(vdo_sim *s, double dtinv, int *info) {
    const int me = s->me, x_sym = s->P.x_sym;
    const double mu = s->P.mug, alpha = s->P.alpha, Pamb = s->P.Pamb, xsp = s->gxsp, xxsp = s->gxxsp;
    double *a = s->ta, *b = s->tb, *c = s->tc, *d = s->td, *pg = s->pg, *pg_old = s->pg_old, *h = s->h,
           *ht = s->ht, *hx = s->hx;
    double ressq = 0;
    /* quirk Q2: with x_sym the explicit row-0 formulas are overwritten by
     * the loop below; only the !x_sym Dirichlet row survives. */
    b[0] = 1; c[0] = 0; d[0] = -pg[0];
    a[me - 1] = 0; b[me - 1] = 1; d[me - 1] = -pg[me - 1];
    for (int i = x_sym ? 0 : 1; i < me - 1; i++) {
        double h_ns = i > 0 ? 0.5 * (h[i - 1] + h[i]) : h[i];
        double ht_ns = i > 0 ? 0.5 * (ht[i - 1] + ht[i]) : ht[i];
        double px = i > 0 ? 0.5 * xsp * (pg[i + 1] - pg[i - 1]) : 0;
        double pxx = i > 0 ? xxsp * (pg[i + 1] - 2 * pg[i] + pg[i - 1]) : xxsp * (pg[i + 1] - 2 * pg[i] + pg[i + 1]);
        double hsq = h_ns * h_ns, hcu = hsq * h_ns;
        double br = 3 * hsq * hx[i] * px + hcu * pxx - 12 * mu * alpha * h_ns * dtinv - 12 * mu * ht_ns;
        a[i] = -xsp * alpha * hcu * px + pg[i] * (-1.5 * xsp * hsq * hx[i] + hcu * xxsp) +
               Pamb * (-1.5 * xsp * hsq * hx[i] + hcu * xxsp);
        b[i] = br - 2 * pg[i] * hcu * xxsp - 2 * Pamb * hcu * xxsp;
        c[i] = xsp * alpha * hcu * px + pg[i] * (1.5 * xsp * hsq * hx[i] + hcu * xxsp) +
               Pamb * (1.5 * xsp * hsq * hx[i] + hcu * xxsp);
        if (i == 0) c[i] += a[i];
        d[i] = -(12 * mu * alpha * h_ns * dtinv * pg_old[i] + alpha * hcu * px * px + pg[i] * br +
                 Pamb * (-12 * mu * ht_ns + 3 * hsq * hx[i] * px + hcu * pxx));
        ressq += d[i] * d[i];
    }
    /* dgtsv_(&me,&nrhs,a+1,b,c,d,...) at tri_solve.cc:28 */
    *info = vdo_tri_solve_pivot(me, a + 1, b, c, d);
    for (int i = 0; i < me; i++) pg[i] += d[i];
    return ressq / me;
}

/* gas_layer_solve.cc:184-212: pg_old = pg, then Newton steps until two
 * successive mean square residuals are below (1e8 eps)^2. */
int vdo_gas_newton(vdo_sim *s, double dt) {
    memcpy(s->pg_old, s->pg, s->me * sizeof(double));
    const double pressure_r_error = DBL_EPSILON * 1e8;          /* gas_layer_solve.hh:11 */
    const double newton_thresh = pressure_r_error * pressure_r_error;
    double dtinv = 1. / dt;
    int nstep = 0, naccept = 0, info = 0;
    do {
        if (++nstep > 256) { s->last_newton = nstep - 1; return 2; } /* gas_layer_solve.cc:204 */
        if (gas_newton_step(s, dtinv, &info) < newton_thresh) naccept++;
        else naccept = 0;
        if (info) { s->last_newton = nstep; return 3; }             /* tri_solve.cc:29-32 */
    } while (naccept < 2);
    s->last_newton = nstep;
    return 0;
}

int vdo_gas_height(vdo_sim *s, double dt) { gas_update_height(s, dt); return 0; }

/* gas_layer.cc:36-46 -> gas_layer_solve.cc:215-222 */
int vdo_gas_update(vdo_sim *s, double t, double dt) {
    (void)t;
    memcpy(s->gp_old, s->gp, s->me * sizeof(double));
    memcpy(s->gtau_old, s->gtau, s->ml * sizeof(double));
    gas_update_height(s, dt);
    int rc = vdo_gas_newton(s, dt);
    if (rc) return rc;
    gas_liquid_fields(s);
    gas_apply_tau_bc(s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions                                                       */
/* ------------------------------------------------------------------------ */

/* fluid_2d_sio.cc:256-290: far-field ghosts, accumulated (quirk Q4) */
static void set_inviscid_boundaries(vdo_sim *s, double dt) {
    const int m = s->m, n = s->n, ml = s->ml, me = s->me;
    const vdo_params *P = &s->P;
    double prefac = s->dx * s->rhoinv * dt / (3 * M_PI);
    double *u = s->F[F_U], *v = s->F[F_V];
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++) {
        double du, dv, y = P->ay + (j + 0.5) * s->dy;
        vdo_simpson_velocity(me, P->ax, s->dx, P->x_sym, s->gp, P->bx + s->dx * 0.5, y, prefac, &du, &dv);
        AT(u, m, j) += du; AT(v, m, j) += dv;
        vdo_simpson_velocity(me, P->ax, s->dx, P->x_sym, s->gp, P->bx + s->dx * 1.5, y, prefac, &du, &dv);
        AT(u, m + 1, j) += du; AT(v, m + 1, j) += dv;
        if (!P->x_sym) {
            AT(u, -1, j) = -AT(u, m, j); AT(v, -1, j) = AT(v, m, j);
            AT(u, -2, j) = -AT(u, m + 1, j); AT(v, -2, j) = AT(v, m + 1, j);
        }
    }
#pragma omp parallel for schedule(static)
    for (int c = -2; c < m + 2; c++) {
        double du, dv, x = P->ax + (c + 0.5) * s->dx;
        vdo_simpson_velocity(me, P->ax, s->dx, P->x_sym, s->gp, x, P->by + s->dy * 0.5, prefac, &du, &dv);
        AT(u, c, n) += du; AT(v, c, n) += dv;
        vdo_simpson_velocity(me, P->ax, s->dx, P->x_sym, s->gp, x, P->by + s->dy * 1.5, prefac, &du, &dv);
        AT(u, c, n + 1) += du; AT(v, c, n + 1) += dv;
    }
}

/* fluid_2d_sio.cc:294-343 (order-dependent, quirk Q5; the ghost-row ux
 * scratch writes are dead and not stored) */
static void set_viscous_boundaries(vdo_sim *s) {
    const int m = s->m, n = s->n, ml = s->ml;
    double *u = s->F[F_U], *v = s->F[F_V];
    if (s->P.x_sym)
        for (int j = 0; j < n; j++) {
            AT(u, -1, j) = -AT(u, 0, j); AT(v, -1, j) = AT(v, 0, j);
            AT(u, -2, j) = -AT(u, 1, j); AT(v, -2, j) = AT(v, 1, j);
        }
    for (int c = -2; c < m + 2; c++) {
        double un_edge = s->muinv * s->gtau[c + 2];
        AT(u, c, -1) = AT(u, c, 0) - s->dy * un_edge;
        AT(u, c, -2) = AT(u, c, 1) - 3 * s->dy * un_edge;
    }
    for (int c = -1; c < m; c++) {
        double ux1 = 0.5 * s->xsp * (AT(u, c + 1, -1) - AT(u, c - 1, -1));
        double ux2 = 0.5 * s->xsp * (AT(u, c + 1, -2) - AT(u, c - 1, -2));
        AT(v, c, -1) = AT(v, c, 0) + s->dy * ux1;
        AT(v, c, -2) = AT(v, c, -1) + s->dy * ux2;
    }
}

/* fluid_2d_sio.cc:14-35 */
void vdo_init_fields(vdo_sim *s) {
    s->time = 0;
    size_t NP = (size_t)s->ml * (s->n + 4);
    memset(s->F[F_U], 0, NP * sizeof(double));
    memset(s->F[F_V], 0, NP * sizeof(double));
    memset(s->F[F_P], 0, NP * sizeof(double));
    gas_init_fields(s);
    set_inviscid_boundaries(s, 0);
    set_viscous_boundaries(s);
}

/* ------------------------------------------------------------------------ */
/* One time step                                                             */
/* ------------------------------------------------------------------------ */

static int mg_solve_sim(vdo_sim *s, vdo_mg *g, double *z, double t) {
    mg_initial_guess(g, z, t);
    int c = vdo_mg_solve(g, s->src, z, s->rel_tol, s->max_cycles);
    return c < 0 ? 1 : 0;
}

/* ---- the passes of fluid_2d::step_forward, one function each ---- */

#define STEP_LOCALS                                                                                 \
    const int m = s->m, n = s->n, ml = s->ml;                                                       \
    (void)m; (void)n; (void)ml;                                                                     \
    double *u = s->F[F_U], *v = s->F[F_V], *p = s->F[F_P];                                          \
    double *ux = s->F[F_UX], *uy = s->F[F_UY], *vx = s->F[F_VX], *vy = s->F[F_VY];                  \
    double *ud = s->F[F_UD], *vd = s->F[F_VD], *ul = s->F[F_UL], *vl = s->F[F_VL];                  \
    double *ur = s->F[F_UR], *vr = s->F[F_VR], *uu = s->F[F_UU], *vu = s->F[F_VU];                  \
    double *c0 = s->F[F_C0], *c1 = s->F[F_C1], *c2 = s->F[F_C2], *c3 = s->F[F_C3];                  \
    (void)u; (void)v; (void)p; (void)ux; (void)uy; (void)vx; (void)vy; (void)ud; (void)vd; (void)ul; \
    (void)vl; (void)ur; (void)vr; (void)uu; (void)vu; (void)c0; (void)c1; (void)c2; (void)c3

static double step_viscfac(const vdo_sim *s) {
    /* fluid_2d.cc:164 */
    return s->P.implicit_visc ? s->rhoinv * s->P.mu * 0.5 : s->rhoinv * s->P.mu;
}

/* K1, fluid_2d.cc:105-132: limited slopes and the normal half-step
 * extrapolation of u, v to the four faces */
static void pass_k1(vdo_sim *s, double dt) {
    STEP_LOCALS;
    const double dx = s->dx, dy = s->dy, xsp = s->xsp, ysp = s->ysp;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            AT(uy, i, j) = ysp * vdo_mono_diff(AT(u, i, j - 2), AT(u, i, j - 1), AT(u, i, j), AT(u, i, j + 1), AT(u, i, j + 2));
            AT(vy, i, j) = ysp * vdo_mono_diff(AT(v, i, j - 2), AT(v, i, j - 1), AT(v, i, j), AT(v, i, j + 1), AT(v, i, j + 2));
            AT(ux, i, j) = xsp * vdo_mono_diff(AT(u, i - 2, j), AT(u, i - 1, j), AT(u, i, j), AT(u, i + 1, j), AT(u, i + 2, j));
            AT(vx, i, j) = xsp * vdo_mono_diff(AT(v, i - 2, j), AT(v, i - 1, j), AT(v, i, j), AT(v, i + 1, j), AT(v, i + 2, j));
            double fu = AT(u, i, j), fv = AT(v, i, j);
            AT(ud, i, j) = fu - 0.5 * (dy + dt * fv) * AT(uy, i, j);
            AT(vd, i, j) = fv - 0.5 * (dy + dt * fv) * AT(vy, i, j);
            AT(ul, i, j) = fu - 0.5 * (dx + dt * fu) * AT(ux, i, j);
            AT(vl, i, j) = fv - 0.5 * (dx + dt * fu) * AT(vx, i, j);
            AT(ur, i, j) = fu + 0.5 * (dx - dt * fu) * AT(ux, i, j);
            AT(vr, i, j) = fv + 0.5 * (dx - dt * fu) * AT(vx, i, j);
            AT(uu, i, j) = fu + 0.5 * (dy - dt * fv) * AT(uy, i, j);
            AT(vu, i, j) = fv + 0.5 * (dy - dt * fv) * AT(vy, i, j);
        }
}

/* K2, fluid_2d.cc:136-150 (zero_edge_velocities :480-482): upwind choice at
 * every face, keeping the tangential-average form godunov_set_tang needs */
static void pass_k2(vdo_sim *s) {
    STEP_LOCALS;
    if (s->P.x_sym) for (int j = 0; j < n; j++) AT(ul, 0, j) = 0;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            if (i < m - 1) {
                vdo_godunov_set_tang(&AT(ur, i, j), &AT(ul, i + 1, j), &AT(vr, i, j), &AT(vl, i + 1, j));
                if (i == 0) AT(ur, -1, j) = AT(ul, 0, j);
            } else { AT(ul, m, j) = AT(ur, m - 1, j); AT(vl, m, j) = AT(vr, m - 1, j); }
            if (j < n - 1) {
                vdo_godunov_set_tang(&AT(vu, i, j), &AT(vd, i, j + 1), &AT(uu, i, j), &AT(ud, i, j + 1));
                if (j == 0) AT(vu, i, -1) = AT(vd, i, 0);
            } else { AT(vd, i, n) = AT(vu, i, n - 1); AT(ud, i, n) = AT(uu, i, n - 1); }
        }
}

/* K3, fluid_2d.cc:153-160: transverse (tangential) derivative terms */
static void pass_k3(vdo_sim *s) {
    STEP_LOCALS;
    const double xsp = s->xsp, ysp = s->ysp;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            AT(c0, i, j) = 0.5 * xsp * (AT(ul, i + 1, j) + AT(ul, i, j)) * (AT(ur, i, j) - AT(ur, i - 1, j));
            AT(c1, i, j) = 0.5 * xsp * (AT(ul, i + 1, j) + AT(ul, i, j)) * (AT(vl, i + 1, j) - AT(vl, i, j));
            AT(c2, i, j) = 0.5 * ysp * (AT(vd, i, j + 1) + AT(vd, i, j)) * (AT(ud, i, j + 1) - AT(ud, i, j));
            AT(c3, i, j) = 0.5 * ysp * (AT(vd, i, j + 1) + AT(vd, i, j)) * (AT(vu, i, j) - AT(vu, i, j - 1));
        }
}

/* K4, fluid_2d.cc:164-201 (centered_diff :412-417): full half-step face
 * predictor with the pressure gradient and the (explicit part of the)
 * viscous force; c0, c1 keep that force */
static void pass_k4(vdo_sim *s, double dt) {
    STEP_LOCALS;
    const double dx = s->dx, dy = s->dy, xsp = s->xsp, ysp = s->ysp, xxsp = s->xxsp, yysp = s->yysp;
    const double rhoinv = s->rhoinv, viscfac = step_viscfac(s);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double px = 0.5 * xsp * (AT(p, i + 1, j + 1) + AT(p, i + 1, j) - AT(p, i, j + 1) - AT(p, i, j));
            double py = 0.5 * ysp * (AT(p, i + 1, j + 1) - AT(p, i + 1, j) + AT(p, i, j + 1) - AT(p, i, j));
            double fu = AT(u, i, j), fv = AT(v, i, j);
            double uyy = yysp * (AT(u, i, j - 1) - 2 * fu + AT(u, i, j + 1));
            double vyy = yysp * (AT(v, i, j - 1) - 2 * fv + AT(v, i, j + 1));
            double uxx = xxsp * (AT(u, i - 1, j) - 2 * fu + AT(u, i + 1, j));
            double vxx = xxsp * (AT(v, i - 1, j) - 2 * fv + AT(v, i + 1, j));
            double f_x = -px * rhoinv + viscfac * (uxx + uyy);
            double f_y = -py * rhoinv + viscfac * (vxx + vyy);
            double fux = AT(ux, i, j), fuy = AT(uy, i, j), fvx = AT(vx, i, j), fvy = AT(vy, i, j);
            double fc0 = AT(c0, i, j), fc1 = AT(c1, i, j), fc2 = AT(c2, i, j), fc3 = AT(c3, i, j);
            AT(ud, i, j) = fu + 0.5 * (-dy * fuy + dt * (-fc0 - fv * fuy + f_x));
            AT(vd, i, j) = fv + 0.5 * (-dy * fvy + dt * (-fc1 - fv * fvy + f_y));
            AT(ul, i, j) = fu + 0.5 * (-dx * fux + dt * (-fu * fux - fc2 + f_x));
            AT(vl, i, j) = fv + 0.5 * (-dx * fvx + dt * (-fu * fvx - fc3 + f_y));
            AT(ur, i, j) = fu + 0.5 * (dx * fux + dt * (-fu * fux - fc2 + f_x));
            AT(vr, i, j) = fv + 0.5 * (dx * fvx + dt * (-fu * fvx - fc3 + f_y));
            AT(uu, i, j) = fu + 0.5 * (dy * fuy + dt * (-fc0 - fv * fuy + f_x));
            AT(vu, i, j) = fv + 0.5 * (dy * fvy + dt * (-fc1 - fv * fvy + f_y));
            AT(c0, i, j) = viscfac * (uxx + uyy);
            AT(c1, i, j) = viscfac * (vxx + vyy);
        }
}

/* K5, fluid_2d.cc:205-215: upwind choice at every face */
static void pass_k5(vdo_sim *s) {
    STEP_LOCALS;
    if (s->P.x_sym) for (int j = 0; j < n; j++) AT(ul, 0, j) = 0;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            if (i < m - 1) vdo_godunov_set(&AT(ur, i, j), &AT(ul, i + 1, j), &AT(vr, i, j), &AT(vl, i + 1, j));
            else { AT(ul, m, j) = AT(ur, m - 1, j); AT(vl, m, j) = AT(vr, m - 1, j); }
            if (j < n - 1) vdo_godunov_set(&AT(vu, i, j), &AT(vd, i, j + 1), &AT(uu, i, j), &AT(ud, i, j + 1));
            else { AT(vd, i, n) = AT(vu, i, n - 1); AT(ud, i, n) = AT(uu, i, n - 1); }
        }
}

/* K6, fluid_2d.cc:218-232: MAC right-hand side (face divergence) and the
 * Dirichlet data of the bottom face from the gas pressure (s_bc, quirk Q8) */
static void pass_k6(vdo_sim *s, double dt) {
    STEP_LOCALS;
    const double xsp = s->xsp, ysp = s->ysp, yysp = s->yysp;
    double *src = s->src, *s_bc = s->s_bc;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++)
            src[i + (size_t)m * j] = xsp * (AT(ul, i + 1, j) - AT(ul, i, j)) + ysp * (AT(vd, i, j + 1) - AT(vd, i, j));
    const double macfac = 0.125 * s->rhoinv * dt;
    for (int i = 0; i < m; i++)
        src[i] -= 2 * yysp * (s_bc[i] = macfac * (s->gp_old[i] + s->gp_old[i + 1] + s->gp[i] + s->gp[i + 1]));
}

/* K7, fluid_2d.cc:236-248: MAC correction of the face velocities */
static void pass_k7(vdo_sim *s) {
    STEP_LOCALS;
    const double xsp = s->xsp, ysp = s->ysp;
    const double *zm = s->zmac, *s_bc = s->s_bc;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            if (i > 0) AT(ul, i, j) -= (zm[i + (size_t)m * j] - zm[i - 1 + (size_t)m * j]) * xsp;
            if (j != 0) AT(vd, i, j) -= (zm[i + (size_t)m * j] - zm[i + (size_t)m * (j - 1)]) * ysp;
        }
    for (int i = 0; i < m; i++) AT(vd, i, 0) -= 2 * ysp * (zm[i] - s_bc[i]);
}

/* K8, fluid_2d.cc:253-261: advective update into c0, c1 */
static void pass_k8(vdo_sim *s, double dt) {
    STEP_LOCALS;
    const double xsp = s->xsp, ysp = s->ysp;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double fu = AT(u, i, j), fv = AT(v, i, j);
            double sx = xsp * (AT(ul, i + 1, j) + AT(ul, i, j)), sy = ysp * (AT(vd, i, j + 1) + AT(vd, i, j));
            AT(c0, i, j) = fu + dt * (-0.5 * (sx * (AT(ul, i + 1, j) - AT(ul, i, j)) + sy * (AT(ud, i, j + 1) - AT(ud, i, j))) + AT(c0, i, j));
            AT(c1, i, j) = fv + dt * (-0.5 * (sx * (AT(vl, i + 1, j) - AT(vl, i, j)) + sy * (AT(vd, i, j + 1) - AT(vd, i, j))) + AT(c1, i, j));
        }
}

/* Implicit viscosity right-hand side, fluid_2d.cc:272-324, component comp
 * (0: u from c0, 1: v from c1) into src */
static void pass_visco_rhs(vdo_sim *s, int comp, double dt) {
    STEP_LOCALS;
    const double dy = s->dy, xsp = s->xsp, viscfac = step_viscfac(s);
    const double vfx = viscfac * s->xxsp * dt, vfy = viscfac * s->yysp * dt;
    const double muinv = s->muinv;
    const int x_sym = s->P.x_sym;
    const double *w = comp ? v : u, *cc = comp ? c1 : c0;
    double *src = s->src;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double r = AT(cc, i, j);
            if (!x_sym && i == 0) r += vfx * AT(w, -1, j);
            if (i == m - 1) r += vfx * AT(w, m, j);
            if (j == n - 1) r += vfy * AT(w, i, n);
            if (j == 0) {
                if (!comp) {
                    double dudy = muinv * 0.5 * (s->gtau_old[i + 2] + s->gtau[i + 2]);
                    r -= vfy * dy * dudy;
                } else {
                    double dvdy = -0.5 * xsp * (AT(u, i + 1, 0) - AT(u, i - 1, 0));
                    r -= vfy * dy * dvdy;
                }
            }
            src[i + (size_t)m * j] = r;
        }
}

/* fluid_2d.cc:330-334: the two Helmholtz solutions become c0, c1 */
static void pass_visco_copy(vdo_sim *s) {
    STEP_LOCALS;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            AT(c0, i, j) = s->zvx[i + (size_t)m * j];
            AT(c1, i, j) = s->zvy[i + (size_t)m * j];
        }
}

/* K14, fluid_2d.cc:338-377: nodal weak divergence of (c0, c1) times rho/dt,
 * boundary flux terms and the Dirichlet bottom row from the gas pressure */
static void pass_k14(vdo_sim *s, double dt) {
    STEP_LOCALS;
    const int me = s->me, ne = s->ne, x_sym = s->P.x_sym;
    const double dx = s->dx, dy = s->dy, rdt = s->P.rho / dt;
    double *src = s->src;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ne; j++)
        for (int i = 0; i < me; i++) {
            double spx = 0, spy = 0;
            if (i > 0) { spx += AT(c0, i - 1, j - 1); spy += AT(c1, i - 1, j - 1); }
            if (i > 0 && j < n) { spx += AT(c0, i - 1, j); spy -= AT(c1, i - 1, j); }
            if (i < m) { spx -= AT(c0, i, j - 1); spy += AT(c1, i, j - 1); }
            if (i < m && j < n) { spx -= AT(c0, i, j); spy -= AT(c1, i, j); }
            src[i + (size_t)me * j] = 0.5 * (dy * spx + dx * spy) * rdt;
        }
    if (!x_sym)
        for (int j = 1; j < ne; j++) {
            if (j < n) src[(size_t)j * me] += 0.25 * dy * rdt * (AT(c0, 0, j) + AT(u, -1, j));
            src[(size_t)j * me] += 0.25 * dy * rdt * (AT(c0, 0, j - 1) + AT(u, -1, j - 1));
        }
    for (int j = 1; j < ne; j++) {
        if (j < n) src[m + (size_t)j * me] -= 0.25 * dy * rdt * (AT(c0, m - 1, j) + AT(u, m, j));
        src[m + (size_t)j * me] -= 0.25 * dy * rdt * (AT(c0, m - 1, j - 1) + AT(u, m, j - 1));
    }
    for (int i = 0; i < me; i++) {
        if (i > 0) src[(size_t)n * me + i] -= 0.25 * dx * rdt * (AT(c1, i - 1, n - 1) + AT(v, i - 1, n));
        if (i < m) src[(size_t)n * me + i] -= 0.25 * dx * rdt * (AT(c1, i, n - 1) + AT(v, i, n));
    }
    for (int i = 0; i < me; i++) src[i] = fem_acc(s->fem, &s->fem->L[0], i, 0) * s->gp[i];
}

/* K15, fluid_2d.cc:383-396: velocity update from the nodal pressure, and the
 * nodal pressure copied into the cell array (ghost column m and row n) */
static void pass_k15(vdo_sim *s, double dt) {
    STEP_LOCALS;
    const int me = s->me, ne = s->ne;
    const double xsp = s->xsp, ysp = s->ysp, rhoinv = s->rhoinv;
    const double *zf = s->zfem;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            const double *q = zf + i + (size_t)me * j;
            AT(u, i, j) = AT(c0, i, j) - 0.5 * dt * rhoinv * xsp * (q[me + 1] + q[1] - q[me] - q[0]);
            AT(v, i, j) = AT(c1, i, j) - 0.5 * dt * rhoinv * ysp * (q[me + 1] - q[1] + q[me] - q[0]);
        }
    for (int j = 0; j < ne; j++)
        for (int i = 0; i < me; i++) AT(p, i, j) = zf[i + (size_t)me * j];
}

static int step_solve(vdo_sim *s, int which, double t) {
    vdo_mg *g = which == VDO_KIND_MAC ? s->mac : which == VDO_KIND_FEM ? s->fem : which == VDO_KIND_VX ? s->vx : s->vy;
    double *z = which == VDO_KIND_MAC ? s->zmac : which == VDO_KIND_FEM ? s->zfem : which == VDO_KIND_VX ? s->zvx : s->zvy;
    return mg_solve_sim(s, g, z, t);
}

/* visco_impl.cc:21-39 (the coefficient part): re-set the hierarchy when the
 * coupling changed */
static void visco_setup(vdo_mg *g, double vfx, double vfy) {
    if (fabs(vfx - g->vfx) > 1e2 * DBL_EPSILON * g->vfx) {
        g->vfx = vfx; g->vfy = vfy;
        mg_set_coeffs(g, vfx, vfy);
    }
}

/* fluid_2d.cc:93-405 */
int vdo_step_forward(vdo_sim *s, double dt) {
    int rc;

    /* fluid_2d.cc:98 */
    update_tracers(s, dt);
    /* fluid_2d.cc:99-102 */
    if ((rc = vdo_gas_update(s, s->time, dt))) return rc == 2 ? 2 : 3;
    s->time += 0.5 * dt;

    pass_k1(s, dt);
    pass_k2(s);
    pass_k3(s);
    pass_k4(s, dt);
    pass_k5(s);
    pass_k6(s, dt);
    /* MAC solve, fluid_2d.cc:233 */
    if (step_solve(s, VDO_KIND_MAC, s->time)) return 1;
    pass_k7(s);
    s->time += 0.5 * dt;
    pass_k8(s, dt);

    /* fluid_2d.cc:267-269 -> :485-492 */
    set_inviscid_boundaries(s, dt);
    set_viscous_boundaries(s);

    if (s->P.implicit_visc) {
        /* fluid_2d.cc:272-335 */
        const double viscfac = step_viscfac(s);
        const double vfx = viscfac * s->xxsp * dt, vfy = viscfac * s->yysp * dt;
        for (int comp = 0; comp < 2; comp++) {
            pass_visco_rhs(s, comp, dt);
            vdo_mg *g = comp ? s->vy : s->vx;
            visco_setup(g, vfx, vfy);
            if (step_solve(s, comp ? VDO_KIND_VY : VDO_KIND_VX, s->time)) return 1;
        }
        pass_visco_copy(s);
    }

    pass_k14(s, dt);
    /* FEM solve, fluid_2d.cc:380 */
    if (step_solve(s, VDO_KIND_FEM, s->time)) return 1;
    pass_k15(s, dt);

    /* fluid_2d.cc:399-401 */
    set_viscous_boundaries(s);
    return 0;
}

/* ---- the passes one at a time (pinning tests) ---- */

int vdo_pass(vdo_sim *s, int pass, double dt) {
    switch (pass) {
    case VDO_PASS_K1: pass_k1(s, dt); return 0;
    case VDO_PASS_K2: pass_k2(s); return 0;
    case VDO_PASS_K3: pass_k3(s); return 0;
    case VDO_PASS_K4: pass_k4(s, dt); return 0;
    case VDO_PASS_K5: pass_k5(s); return 0;
    case VDO_PASS_K6: pass_k6(s, dt); return 0;
    case VDO_PASS_K7: pass_k7(s); return 0;
    case VDO_PASS_K8: pass_k8(s, dt); return 0;
    case VDO_PASS_INVISCID_BC: set_inviscid_boundaries(s, dt); return 0;
    case VDO_PASS_VISCOUS_BC: set_viscous_boundaries(s); return 0;
    case VDO_PASS_VISCO_RHS_U: if (!s->vx) return -1; pass_visco_rhs(s, 0, dt); return 0;
    case VDO_PASS_VISCO_RHS_V: if (!s->vy) return -1; pass_visco_rhs(s, 1, dt); return 0;
    case VDO_PASS_VISCO_COPY: if (!s->vx) return -1; pass_visco_copy(s); return 0;
    case VDO_PASS_K14: pass_k14(s, dt); return 0;
    case VDO_PASS_K15: pass_k15(s, dt); return 0;
    case VDO_PASS_GAS_LIQUID_FIELDS: gas_liquid_fields(s); return 0;
    case VDO_PASS_GAS_TAU_BC: gas_apply_tau_bc(s); return 0;
    case VDO_PASS_GAS_HX: gas_calculate_hx(s); return 0;
    case VDO_PASS_GAS_INIT: gas_init_fields(s); return 0;
    case VDO_PASS_TRACERS: update_tracers(s, dt); return 0;
    default: return -1;
    }
}

int vdo_initial_guess(vdo_sim *s, int which, double t) {
    vdo_mg *g = which == VDO_KIND_MAC ? s->mac : which == VDO_KIND_FEM ? s->fem : which == VDO_KIND_VX ? s->vx : s->vy;
    double *z = which == VDO_KIND_MAC ? s->zmac : which == VDO_KIND_FEM ? s->zfem : which == VDO_KIND_VX ? s->zvx : s->zvy;
    if (!g) return -1;
    mg_initial_guess(g, z, t);
    return 0;
}

int vdo_solve_v_cycle(vdo_sim *s, int which, double t) {
    if (which != VDO_KIND_MAC && which != VDO_KIND_FEM) return -1;
    if (step_solve(s, which, t)) return -1;
    return vdo_last_cycles(s, which);
}

int vdo_visco_setup_and_solve(vdo_sim *s, int which, double vfx, double vfy, double t) {
    if ((which != VDO_KIND_VX && which != VDO_KIND_VY) || !s->vx) return -1;
    visco_setup(which == VDO_KIND_VX ? s->vx : s->vy, vfx, vfy);
    if (step_solve(s, which, t)) return -1;
    return vdo_last_cycles(s, which);
}
