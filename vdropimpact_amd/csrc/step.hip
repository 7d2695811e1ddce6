// step.hip -- one time step of the drop-impact liquid solve on gfx950, and
// the C ABI (include/vdi.h).
//
// Data layout in HBM: every cell field is its own FP64 array (structure of
// arrays) with the reference's ghost padding, (m+4) x (n+4), cell (i,j) at
// (i+2) + (m+4)(j+2) -- the reference keeps the same fields in a 152-byte
// record per cell (fields.hh:7-19).  Node arrays (the FEM pressure) are
// (m+1) x (n+1).  The gas layer is a handful of 1D arrays.
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <algorithm>

#include "step.hpp"

namespace vdi {

static std::string g_last_error;

void set_error(const char *what, hipError_t e, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_last_error = buf;
}

#define IX(i, j) (((size_t)(j) + 2) * (size_t)g.ml + (size_t)((i) + 2))

// ------------------------------------------------------------------------
// Limiter and Godunov selection (fluid_2d.cc:419-476)
// ------------------------------------------------------------------------

__device__ __forceinline__ double delta_lim(double f0, double f1, double f2) {
    double drm = f1 - f0, drp = f2 - f1;
    return drm > 0 ? (drp > 0 ? 2 * fmin(drm, drp) : 0.0) : (drp > 0 ? 0.0 : -2 * fmax(drm, drp));
}
__device__ __forceinline__ double delta_f(double f0, double f1, double f2) {
    double drc = 0.5 * (f2 - f0);
    return drc > 0 ? fmin(drc, delta_lim(f0, f1, f2)) : fmax(drc, -delta_lim(f0, f1, f2));
}
__device__ __forceinline__ double mono_diff(double f0, double f1, double f2, double f3, double f4) {
    double tdrc = f3 - f1, s = (1 / 6.) * (4 * tdrc - delta_f(f0, f1, f2) - delta_f(f2, f3, f4)),
           t = delta_lim(f1, f2, f3);
    return tdrc > 0 ? fmin(s, t) : fmax(s, -t);
}
// fluid_2d.cc:455-461
__device__ __forceinline__ void godunov_set(double &u0, double &u1, double &v0, double &v1) {
    if (u0 > -u1) {
        if (u0 > 0) { u1 = u0; v1 = v0; return; }
    } else if (u1 < 0) return;
    u1 = 0;
    v1 = 0.5 * (v0 + v1);
}
// fluid_2d.cc:469-476
__device__ __forceinline__ void godunov_set_tang(double &u0, double &u1, double &v0, double &v1) {
    if (u0 > -u1) {
        if (u0 > 0) { u1 = u0; v1 = v0; return; }
    } else if (u1 < 0) { u0 = u1; return; }
    v1 = 0.5 * (v0 + v1);
    u0 = 0.5 * (u0 + u1);
    u1 = 0;
}

// Element-wise kernels: 256 threads as 64 x 4, tile (bx, by).
#define ELEM_IDX                                   \
    const int i = bx * kEX + (threadIdx.x & 63);   \
    const int j = by * kEY + (threadIdx.x >> 6);
#define REC_REFS                                   \
    const Geo &g = a.g;                            \
    const Fields &F = a.F;                         \
    (void)F;

// ------------------------------------------------------------------------
// Fused Godunov passes K1..K6 on LDS tiles
// ------------------------------------------------------------------------
//
// One workgroup owns kGX x kGY cells.  With T the tile and T+k its ring of
// width k (clipped to the domain):
//   K1 (slopes, first extrapolation) on T+2 from u,v on T+4;
//   post-K2 face states (tangential Godunov select, recomputed for every
//   face a cell needs) and K3 on T+1;
//   K4 (full predictor) on T+1;
//   K5 (normal Godunov select) and K6 (MAC RHS) on T.
// Only the final face states, the viscous force c0/c1 and src are written:
// about 10 words of traffic per cell instead of the ~75 of one pass per
// reference loop.  Boundary faces follow fluid_2d.cc:141-148 and :210-213.
constexpr int kGW = kGX + 8, kGH = kGY + 8;  // LDS frame of u, v: tile + 4 on each side
// Frames of the edge states (tile + 2, K1 on T+2) and of c0..c3 (tile + 1);
// sized to what each array holds so that three workgroups fit a CU's LDS.
constexpr int kEW = kGX + 4, kEH = kGY + 4, kCW = kGX + 2, kCH = kGY + 2;

struct GTile {
    int x0, y0;
    __device__ int L(int i, int j) const { return (j - y0 + 4) * kGW + (i - x0 + 4); }
    __device__ int E(int i, int j) const { return (j - y0 + 2) * kEW + (i - x0 + 2); }
    __device__ int C(int i, int j) const { return (j - y0 + 1) * kCW + (i - x0 + 1); }
};

// Tangential Godunov select on face between left state a and right state b
// (fluid_2d.cc:469-476); returns (u0', u1', v1').
__device__ __forceinline__ void tang3(double u0, double u1, double v0, double v1, double &ou0, double &ou1,
                                      double &ov1) {
    godunov_set_tang(u0, u1, v0, v1);
    ou0 = u0; ou1 = u1; ov1 = v1;
}

__device__ void godunov_body(const StepRec &a, int bx, int by) {
    const Geo &g = a.g;
    const Fields &F = a.F;
    const Gas &G = a.G;
    const double dt = a.dt, viscfac = a.viscfac, macfac = a.macfac;
    __shared__ double su[kGW * kGH], sv[kGW * kGH];
    // K1 edges, later K4 edges: ul, vl, ur, vr, ud, vd, uu, vu
    __shared__ double se[8][kEW * kEH];
    __shared__ double sc[4][kCW * kCH];         // c0..c3 of K3
    __shared__ double sp[(kGX + 3) * (kGY + 3)]; // nodal p on T+1 (+1 node)
    enum { UL, VL, UR, VR, UD, VD, UU, VU };
    const int tid = threadIdx.x;
    GTile t;
    t.x0 = bx * kGX;
    t.y0 = by * kGY;
    const int m = g.m, n = g.n;
    // u, v on T+4 within the padded array
    for (int k = tid; k < kGW * kGH; k += 256) {
        const int i = t.x0 - 4 + k % kGW, j = t.y0 - 4 + k / kGW;
        const bool in = i >= -2 && i < m + 2 && j >= -2 && j < n + 2;
        su[k] = in ? F.u[IX(i, j)] : 0.0;
        sv[k] = in ? F.v[IX(i, j)] : 0.0;
    }
    // p on nodes [x0-1, x0+kGX+1] x [y0-1, y0+kGY+1]
    for (int k = tid; k < (kGX + 3) * (kGY + 3); k += 256) {
        const int i = t.x0 - 1 + k % (kGX + 3), j = t.y0 - 1 + k / (kGX + 3);
        const bool in = i >= -2 && i < m + 2 && j >= -2 && j < n + 2;
        sp[k] = in ? F.p[IX(i, j)] : 0.0;
    }
    __syncthreads();
    // K1 on T+2 (fluid_2d.cc:105-132) + zero_edge_velocities
    const int ax0 = max(0, t.x0 - 2), ax1 = min(m, t.x0 + kGX + 2);
    const int ay0 = max(0, t.y0 - 2), ay1 = min(n, t.y0 + kGY + 2);
    for (int k = tid; k < (kGX + 4) * (kGY + 4); k += 256) {
        const int i = t.x0 - 2 + k % (kGX + 4), j = t.y0 - 2 + k / (kGX + 4);
        if (i < ax0 || i >= ax1 || j < ay0 || j >= ay1) continue;
        const int l = t.L(i, j), e = t.E(i, j);
        const double fu = su[l], fv = sv[l];
        const double uy = g.ysp * mono_diff(su[l - 2 * kGW], su[l - kGW], fu, su[l + kGW], su[l + 2 * kGW]);
        const double vy = g.ysp * mono_diff(sv[l - 2 * kGW], sv[l - kGW], fv, sv[l + kGW], sv[l + 2 * kGW]);
        const double ux = g.xsp * mono_diff(su[l - 2], su[l - 1], fu, su[l + 1], su[l + 2]);
        const double vx = g.xsp * mono_diff(sv[l - 2], sv[l - 1], fv, sv[l + 1], sv[l + 2]);
        se[UD][e] = fu - 0.5 * (g.dy + dt * fv) * uy;
        se[VD][e] = fv - 0.5 * (g.dy + dt * fv) * vy;
        se[UL][e] = (g.x_sym && i == 0) ? 0.0 : fu - 0.5 * (g.dx + dt * fu) * ux;
        se[VL][e] = fv - 0.5 * (g.dx + dt * fu) * vx;
        se[UR][e] = fu + 0.5 * (g.dx - dt * fu) * ux;
        se[VR][e] = fv + 0.5 * (g.dx - dt * fu) * vx;
        se[UU][e] = fu + 0.5 * (g.dy - dt * fv) * uy;
        se[VU][e] = fv + 0.5 * (g.dy - dt * fv) * vy;
    }
    __syncthreads();
    // post-K2 face states and K3 on T+1 (fluid_2d.cc:136-160)
    const int bx0 = max(0, t.x0 - 1), bx1 = min(m, t.x0 + kGX + 1);
    const int by0 = max(0, t.y0 - 1), by1 = min(n, t.y0 + kGY + 1);
    for (int k = tid; k < (kGX + 2) * (kGY + 2); k += 256) {
        const int i = t.x0 - 1 + k % (kGX + 2), j = t.y0 - 1 + k / (kGX + 2);
        if (i < bx0 || i >= bx1 || j < by0 || j >= by1) continue;
        const int l = t.E(i, j), lc = t.C(i, j);
        double ta, tb, tc;
        // x faces: left face (i-1/2) gives UR2(i-1), UL2(i), VL2(i); right face gives UR2(i), UL2(i+1), VL2(i+1)
        double urm, ul0, vl0, ur0, ulp, vlp;
        if (i == 0) { ul0 = se[UL][l]; vl0 = se[VL][l]; urm = ul0; }
        else { tang3(se[UR][l - 1], se[UL][l], se[VR][l - 1], se[VL][l], ta, tb, tc); urm = ta; ul0 = tb; vl0 = tc; }
        if (i == m - 1) { ur0 = se[UR][l]; ulp = se[UR][l]; vlp = se[VR][l]; }
        else { tang3(se[UR][l], se[UL][l + 1], se[VR][l], se[VL][l + 1], ta, tb, tc); ur0 = ta; ulp = tb; vlp = tc; }
        // y faces
        double vum, vd0, ud0, vu0, vdp, udp;
        if (j == 0) { vd0 = se[VD][l]; ud0 = se[UD][l]; vum = vd0; }
        else { tang3(se[VU][l - kEW], se[VD][l], se[UU][l - kEW], se[UD][l], ta, tb, tc); vum = ta; vd0 = tb; ud0 = tc; }
        if (j == n - 1) { vu0 = se[VU][l]; vdp = se[VU][l]; udp = se[UU][l]; }
        else { tang3(se[VU][l], se[VD][l + kEW], se[UU][l], se[UD][l + kEW], ta, tb, tc); vu0 = ta; vdp = tb; udp = tc; }
        const double sx = 0.5 * g.xsp * (ulp + ul0), sy = 0.5 * g.ysp * (vdp + vd0);
        sc[0][lc] = sx * (ur0 - urm);
        sc[1][lc] = sx * (vlp - vl0);
        sc[2][lc] = sy * (udp - ud0);
        sc[3][lc] = sy * (vu0 - vum);
    }
    __syncthreads();
    // K4 on T+1 (fluid_2d.cc:164-201); tile cells also store c0,c1
    for (int k = tid; k < (kGX + 2) * (kGY + 2); k += 256) {
        const int i = t.x0 - 1 + k % (kGX + 2), j = t.y0 - 1 + k / (kGX + 2);
        if (i < bx0 || i >= bx1 || j < by0 || j >= by1) continue;
        const int l = t.L(i, j), e = t.E(i, j), lc = t.C(i, j);
        const int q = (j - t.y0 + 1) * (kGX + 3) + (i - t.x0 + 1);
        const double px = 0.5 * g.xsp * (sp[q + kGX + 4] + sp[q + 1] - sp[q + kGX + 3] - sp[q]);
        const double py = 0.5 * g.ysp * (sp[q + kGX + 4] - sp[q + 1] + sp[q + kGX + 3] - sp[q]);
        const double fu = su[l], fv = sv[l];
        const double uyy = g.yysp * (su[l - kGW] - 2 * fu + su[l + kGW]);
        const double vyy = g.yysp * (sv[l - kGW] - 2 * fv + sv[l + kGW]);
        const double uxx = g.xxsp * (su[l - 1] - 2 * fu + su[l + 1]);
        const double vxx = g.xxsp * (sv[l - 1] - 2 * fv + sv[l + 1]);
        const double f_x = -px * g.rhoinv + viscfac * (uxx + uyy);
        const double f_y = -py * g.rhoinv + viscfac * (vxx + vyy);
        const double fuy = g.ysp * mono_diff(su[l - 2 * kGW], su[l - kGW], fu, su[l + kGW], su[l + 2 * kGW]);
        const double fvy = g.ysp * mono_diff(sv[l - 2 * kGW], sv[l - kGW], fv, sv[l + kGW], sv[l + 2 * kGW]);
        const double fux = g.xsp * mono_diff(su[l - 2], su[l - 1], fu, su[l + 1], su[l + 2]);
        const double fvx = g.xsp * mono_diff(sv[l - 2], sv[l - 1], fv, sv[l + 1], sv[l + 2]);
        const double fc0 = sc[0][lc], fc1 = sc[1][lc], fc2 = sc[2][lc], fc3 = sc[3][lc];
        const double dx = g.dx, dy = g.dy;
        se[UD][e] = fu + 0.5 * (-dy * fuy + dt * (-fc0 - fv * fuy + f_x));
        se[VD][e] = fv + 0.5 * (-dy * fvy + dt * (-fc1 - fv * fvy + f_y));
        se[UL][e] = (g.x_sym && i == 0) ? 0.0 : fu + 0.5 * (-dx * fux + dt * (-fu * fux - fc2 + f_x));
        se[VL][e] = fv + 0.5 * (-dx * fvx + dt * (-fu * fvx - fc3 + f_y));
        se[UR][e] = fu + 0.5 * (dx * fux + dt * (-fu * fux - fc2 + f_x));
        se[VR][e] = fv + 0.5 * (dx * fvx + dt * (-fu * fvx - fc3 + f_y));
        se[UU][e] = fu + 0.5 * (dy * fuy + dt * (-fc0 - fv * fuy + f_x));
        se[VU][e] = fv + 0.5 * (dy * fvy + dt * (-fc1 - fv * fvy + f_y));
        if (i >= t.x0 && i < t.x0 + kGX && j >= t.y0 && j < t.y0 + kGY) {
            F.c0[IX(i, j)] = viscfac * (uxx + uyy);
            F.c1[IX(i, j)] = viscfac * (vxx + vyy);
        }
    }
    __syncthreads();
    // K5 + K6 on T (fluid_2d.cc:205-232)
    {
        const int i = t.x0 + tid % kGX, j = t.y0 + tid / kGX;
        if (i >= m || j >= n) return;
        const int l = t.E(i, j);
        double a0, a1, b0, b1;
        // face i-1/2
        double ulL, vlL;
        if (i == 0) { ulL = se[UL][l]; vlL = se[VL][l]; }
        else { a0 = se[UR][l - 1]; a1 = se[UL][l]; b0 = se[VR][l - 1]; b1 = se[VL][l]; godunov_set(a0, a1, b0, b1); ulL = a1; vlL = b1; }
        // face i+1/2
        double ulR;
        if (i == m - 1) ulR = se[UR][l];
        else { a0 = se[UR][l]; a1 = se[UL][l + 1]; b0 = se[VR][l]; b1 = se[VL][l + 1]; godunov_set(a0, a1, b0, b1); ulR = a1; }
        // face j-1/2
        double vdB, udB;
        if (j == 0) { vdB = se[VD][l]; udB = se[UD][l]; }
        else { a0 = se[VU][l - kEW]; a1 = se[VD][l]; b0 = se[UU][l - kEW]; b1 = se[UD][l]; godunov_set(a0, a1, b0, b1); vdB = a1; udB = b1; }
        // face j+1/2
        double vdT;
        if (j == n - 1) vdT = se[VU][l];
        else { a0 = se[VU][l]; a1 = se[VD][l + kEW]; b0 = se[UU][l]; b1 = se[UD][l + kEW]; godunov_set(a0, a1, b0, b1); vdT = a1; }
        const size_t c = IX(i, j);
        F.ul[c] = ulL; F.vl[c] = vlL; F.vd[c] = vdB; F.ud[c] = udB;
        if (i == m - 1) { F.ul[IX(m, j)] = se[UR][l]; F.vl[IX(m, j)] = se[VR][l]; }
        if (j == n - 1) { F.vd[IX(i, n)] = se[VU][l]; F.ud[IX(i, n)] = se[UU][l]; }
        double s = g.xsp * (ulR - ulL) + g.ysp * (vdT - vdB);
        if (j == 0 && !a.defer_sbc) {
            const double sbc = macfac * (G.p_old[i] + G.p_old[i + 1] + G.p[i] + G.p[i + 1]);
            F.src[(size_t)m * n + i] = sbc;
            s -= 2 * g.yysp * sbc;
        }
        F.src[i + (size_t)m * j] = s;
    }
}

// K7 + K8 (fluid_2d.cc:236-261): the MAC gradient correction is applied to
// the four faces of the cell on the fly and the advective update is formed;
// the corrected faces are not needed afterwards, so they are not stored.
__device__ void correct_body(const StepRec &a, int bx, int by) {
    REC_REFS
    ELEM_IDX
    if (i >= g.m || j >= g.n) return;
    const double *z = a.zmac, dt = a.dt;
    const int m = g.m, n = g.n;
    const size_t q = i + (size_t)m * j, c = IX(i, j), r = IX(i + 1, j), t = IX(i, j + 1);
    const double *sbc = F.src + (size_t)m * n;
    double ulL = F.ul[c], ulR = F.ul[r], vdB = F.vd[c], vdT = F.vd[t];
    if (i > 0) ulL -= (z[q] - z[q - 1]) * g.xsp;
    if (i + 1 < m) ulR -= (z[q + 1] - z[q]) * g.xsp;
    if (j != 0) vdB -= (z[q] - z[q - m]) * g.ysp;
    else vdB -= 2 * g.ysp * (z[q] - sbc[i]);
    if (j + 1 < n) vdT -= (z[q + m] - z[q]) * g.ysp;
    const double sx = g.xsp * (ulR + ulL), sy = g.ysp * (vdT + vdB);
    F.c0[c] = F.u[c] + dt * (-0.5 * (sx * (ulR - ulL) + sy * (F.ud[t] - F.ud[c])) + F.c0[c]);
    F.c1[c] = F.v[c] + dt * (-0.5 * (sx * (F.vl[r] - F.vl[c]) + sy * (vdT - vdB)) + F.c1[c]);
}

// ------------------------------------------------------------------------
// Boundary conditions (fluid_2d_sio.cc:256-405)
// ------------------------------------------------------------------------

__device__ __forceinline__ void simpson_term(int x_sym, double xms, double xps, double ysq, double p,
                                             double &fsu, double &fsv) {
    double rm = xms * xms + ysq;
    rm = 1 / (rm * rm);
    if (x_sym) {
        double rp = xps * xps + ysq;
        rp = 1 / (rp * rp);
        fsu = p * (xms * rm + xps * rp);
        fsv = -p * ((xms * xms - ysq) * rm + (xps * xps - ysq) * rp);
    } else {
        fsu = p * xms * rm;
        fsv = -p * (xms * xms - ysq) * rm;
    }
}

// set_inviscid_boundaries: one wavefront per ghost line.  Waves [0,n) do the
// two right ghost cells of row j; waves [n, n+ml) the two top ghost cells of
// column c = w-n-2.  Composite Simpson over the gas pressure nodes, then the
// forward-Euler accumulation (fluid_2d_sio.cc:383-384, quirk Q4).
__device__ void simpson_body(const StepRec &a, int blk) {
    REC_REFS
    const double *gp = a.G.p, prefac = a.prefac;
    int w = (blk * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= g.n + g.ml) return;
    double x1, x2, y1, y2;
    bool right = w < g.n;
    int j = w, c = w - g.n - 2;
    if (right) {
        y1 = y2 = g.ay + (j + 0.5) * g.dy;
        x1 = g.bx + g.dx * 0.5;
        x2 = g.bx + g.dx * 1.5;
    } else {
        x1 = x2 = g.ax + (c + 0.5) * g.dx;
        y1 = g.by + g.dy * 0.5;
        y2 = g.by + g.dy * 1.5;
    }
    double su1 = 0, sv1 = 0, su2 = 0, sv2 = 0, ysq1 = y1 * y1, ysq2 = y2 * y2;
    for (int k = lane; k < g.me; k += 64) {
        double wgt = (k == 0 || k == g.me - 1) ? 1.0 : ((k & 1) ? 4.0 : 2.0);
        double s = k * g.dx, pk = gp[k], fsu, fsv;
        simpson_term(g.x_sym, x1 - g.ax - s, x1 + g.ax + s, ysq1, pk, fsu, fsv);
        su1 += wgt * fsu; sv1 += wgt * fsv;
        simpson_term(g.x_sym, x2 - g.ax - s, x2 + g.ax + s, ysq2, pk, fsu, fsv);
        su2 += wgt * fsu; sv2 += wgt * fsv;
    }
    for (int o = 32; o > 0; o >>= 1) {
        su1 += __shfl_xor(su1, o); sv1 += __shfl_xor(sv1, o);
        su2 += __shfl_xor(su2, o); sv2 += __shfl_xor(sv2, o);
    }
    if (lane) return;
    if (right) {
        size_t a = IX(g.m, j), b = IX(g.m + 1, j);
        F.u[a] += 2 * y1 * prefac * su1; F.v[a] += prefac * sv1;
        F.u[b] += 2 * y2 * prefac * su2; F.v[b] += prefac * sv2;
        if (!g.x_sym) {
            F.u[IX(-1, j)] = -F.u[a]; F.v[IX(-1, j)] = F.v[a];
            F.u[IX(-2, j)] = -F.u[b]; F.v[IX(-2, j)] = F.v[b];
        }
    } else {
        size_t a = IX(c, g.n), b = IX(c, g.n + 1);
        F.u[a] += 2 * y1 * prefac * su1; F.v[a] += prefac * sv1;
        F.u[b] += 2 * y2 * prefac * su2; F.v[b] += prefac * sv2;
    }
}

// set_viscous_boundaries (fluid_2d_sio.cc:294-343), phases in the
// reference's order (quirk Q5) with workgroup barriers between them.  The
// reference's ghost-row ux scratch stores are dead and are not made.
__device__ void viscous_bc_body(const StepRec &a) {
    REC_REFS
    const double *tau = a.G.tau;
    double *u = F.u, *v = F.v;
    const int T = blockDim.x;
    if (g.x_sym)
        for (int j = threadIdx.x; j < g.n; j += T) {
            u[IX(-1, j)] = -u[IX(0, j)]; v[IX(-1, j)] = v[IX(0, j)];
            u[IX(-2, j)] = -u[IX(1, j)]; v[IX(-2, j)] = v[IX(1, j)];
        }
    __syncthreads();
    for (int k = threadIdx.x; k < g.ml; k += T) {
        int c = k - 2;
        double un = g.muinv * tau[k];
        u[IX(c, -1)] = u[IX(c, 0)] - g.dy * un;
        u[IX(c, -2)] = u[IX(c, 1)] - 3 * g.dy * un;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < g.m + 1; k += T) {
        int c = k - 1;
        double ux1 = 0.5 * g.xsp * (u[IX(c + 1, -1)] - u[IX(c - 1, -1)]);
        double ux2 = 0.5 * g.xsp * (u[IX(c + 1, -2)] - u[IX(c - 1, -2)]);
        double v1 = v[IX(c, 0)] + g.dy * ux1;
        v[IX(c, -1)] = v1;
        v[IX(c, -2)] = v1 + g.dy * ux2;
    }
}

// ------------------------------------------------------------------------
// Implicit viscosity RHS (fluid_2d.cc:272-334) and FEM projection
// ------------------------------------------------------------------------

__device__ void visco_rhs_body(const StepRec &a, int bx, int by, int comp) {
    REC_REFS
    ELEM_IDX
    if (i >= g.m || j >= g.n) return;
    const Gas &G = a.G;
    const double vfx = a.vfx, vfy = a.vfy;
    const double *w = comp ? F.v : F.u;
    double r = comp ? F.c1[IX(i, j)] : F.c0[IX(i, j)];
    if (!g.x_sym && i == 0) r += vfx * w[IX(-1, j)];
    if (i == g.m - 1) r += vfx * w[IX(g.m, j)];
    if (j == g.n - 1) r += vfy * w[IX(i, g.n)];
    if (j == 0) {
        if (!comp) {
            double dudy = g.muinv * 0.5 * (G.tau_old[i + 2] + G.tau[i + 2]);
            r -= vfy * g.dy * dudy;
        } else {
            double dvdy = -0.5 * g.xsp * (F.u[IX(i + 1, 0)] - F.u[IX(i - 1, 0)]);
            r -= vfy * g.dy * dvdy;
        }
    }
    (comp ? F.src2 : F.src)[i + (size_t)g.m * j] = r;
}

__device__ void visco_copy_body(const StepRec &a, int bx, int by) {
    REC_REFS
    ELEM_IDX
    if (i >= g.m || j >= g.n) return;
    const double *zx = a.zvx, *zy = a.zvy;
    size_t q = i + (size_t)g.m * j;
    F.c0[IX(i, j)] = zx[q];
    F.c1[IX(i, j)] = zy[q];
}

// K14 (fluid_2d.cc:338-377): nodal weak divergence, boundary fluxes, and the
// Dirichlet bottom row src = a_cc p_gas.
__device__ void fem_rhs_body(const StepRec &a, int bx, int by) {
    REC_REFS
    ELEM_IDX
    if (i >= g.me || j >= g.ne) return;
    const double *gp = a.G.p, rdt = a.rdt, fm = a.fm;
    const int m = g.m, n = g.n;
    if (j == 0) {
        F.src[i] = (i > 0 && i < g.me - 1 ? 1.0 : 0.5) * 0.5 * fm * gp[i];
        return;
    }
    double spx = 0, spy = 0;
    if (i > 0) { spx += F.c0[IX(i - 1, j - 1)]; spy += F.c1[IX(i - 1, j - 1)]; }
    if (i > 0 && j < n) { spx += F.c0[IX(i - 1, j)]; spy -= F.c1[IX(i - 1, j)]; }
    if (i < m) { spx -= F.c0[IX(i, j - 1)]; spy += F.c1[IX(i, j - 1)]; }
    if (i < m && j < n) { spx -= F.c0[IX(i, j)]; spy -= F.c1[IX(i, j)]; }
    double s = 0.5 * (g.dy * spx + g.dx * spy) * rdt;
    if (!g.x_sym && i == 0) {
        if (j < n) s += 0.25 * g.dy * rdt * (F.c0[IX(0, j)] + F.u[IX(-1, j)]);
        s += 0.25 * g.dy * rdt * (F.c0[IX(0, j - 1)] + F.u[IX(-1, j - 1)]);
    }
    if (i == m) {
        if (j < n) s -= 0.25 * g.dy * rdt * (F.c0[IX(m - 1, j)] + F.u[IX(m, j)]);
        s -= 0.25 * g.dy * rdt * (F.c0[IX(m - 1, j - 1)] + F.u[IX(m, j - 1)]);
    }
    if (j == n) {
        if (i > 0) s -= 0.25 * g.dx * rdt * (F.c1[IX(i - 1, n - 1)] + F.v[IX(i - 1, n)]);
        if (i < m) s -= 0.25 * g.dx * rdt * (F.c1[IX(i, n - 1)] + F.v[IX(i, n)]);
    }
    F.src[i + (size_t)g.me * j] = s;
}

// K15 (fluid_2d.cc:383-396)
__device__ void update_body(const StepRec &a, int bx, int by) {
    REC_REFS
    ELEM_IDX
    if (i >= g.me || j >= g.ne) return;
    const double *zf = a.zfem, dt = a.dt;
    const int me = g.me;
    const double *q = zf + i + (size_t)me * j;
    if (i < g.m && j < g.n) {
        size_t c = IX(i, j);
        F.u[c] = F.c0[c] - 0.5 * dt * g.rhoinv * g.xsp * (q[me + 1] + q[1] - q[me] - q[0]);
        F.v[c] = F.c1[c] - 0.5 * dt * g.rhoinv * g.ysp * (q[me + 1] - q[1] + q[me] - q[0]);
    }
    F.p[IX(i, j)] = q[0];
}

// ------------------------------------------------------------------------
// Gas layer: one workgroup (gas_layer.cc:36-46, gas_layer_solve.cc:133-249)
// ------------------------------------------------------------------------

__device__ __forceinline__ double eno2(double p0, double p1, double p2, double p3) {
#pragma clang fp contract(off)
    return fabs(p0 - 2 * p1 + p2) > fabs(p1 - 2 * p2 + p3) ? 3 * p1 - 4 * p2 + p3 : p0 - p2;
}

// update_height (gas_layer_solve.cc:133-162) updates h in place while i
// sweeps upward, so cell i sees the new h below i and the old h from i on.
// gas_href is h_ref (gas_layer_solve.cc:176-180) as cell i sees it: HN holds
// new values below i, HO the old ones.  gas_cell is one cell of the sweep.
// Both fix their roundings (no contraction left to the compiler), so every
// caller computes the same bits.
__device__ __forceinline__ double gas_href(int k, int i, int m, int x_sym, const double *HN, const double *HO) {
#pragma clang fp contract(off)
    auto cur = [&](int q) { return q < i ? HN[q] : HO[q]; };
    if (k >= 0 && k < m) return cur(k);
    if (k >= m) return k == m ? 2 * cur(m - 1) - cur(m - 2) : 3 * cur(m - 1) - 2 * cur(m - 2);
    return x_sym ? cur(k == -1 ? 0 : 1) : (k == -1 ? 2 * cur(0) - cur(1) : 3 * cur(0) - 2 * cur(1));
}

__device__ __forceinline__ void gas_cell(int i, int m, int x_sym, double dt, double Vframe, const double *HN,
                                         const double *HO, double uc, double vc, double &hnew, double &htv) {
#pragma clang fp contract(off)
    auto h = [&](int k) { return gas_href(k, i, m, x_sym, HN, HO); };
    const double dhdx = uc > 0 ? eno2(h(i + 1), HO[i], h(i - 1), h(i - 2)) : -eno2(h(i - 1), HO[i], h(i + 1), h(i + 2));
    htv = -uc * dhdx + vc - Vframe;
    hnew = HO[i] + dt * htv;
}

// The sweep itself, by one lane.  HN may alias UC and HT VC: cell i reads its
// u_c and v_c before it overwrites them.
__device__ void gas_height_serial(int m, int x_sym, double dt, double Vframe, const double *HO, const double *UC,
                                  const double *VC, double *HN, double *HT) {
    for (int i = 0; i < m; i++) {
        const double uc = UC[i], vc = VC[i];
        double hn, ht;
        gas_cell(i, m, x_sym, dt, Vframe, HN, HO, uc, vc, hn, ht);
        HN[i] = hn;
        HT[i] = ht;
    }
}

// The same sweep by Jacobi passes over all cells at once: pass p computes
// every cell from pass p-1's values (the first from old h).  Cell i depends
// only on cells below it, so after p passes cells 0..p-1 hold the sweep's
// values exactly, and a pass that changes no bit has reached them everywhere:
// the result is the serial sweep's, bit for bit.  The map is contracting
// (the couplings are about 2 |u| dt / dx), so a few passes suffice instead of
// m dependent steps; after kGasJacobiPasses the serial sweep takes over.
// Returns the array holding new h (h_t in HT) and sets passes (0 when the
// serial sweep ran); all threads call it.
__device__ double *gas_height_jacobi(int m, int x_sym, double dt, double Vframe, const double *HO, const double *UC,
                                     const double *VC, double *H0, double *H1, double *HT, int &passes) {
    const int tid = threadIdx.x, T = blockDim.x;
    for (int i = tid; i < m; i += T) H0[i] = HO[i];
    __syncthreads();
    double *src = H0, *dst = H1;
    for (int pass = 0; pass < kGasJacobiPasses; pass++) {
        int changed = 0;
        for (int i = tid; i < m; i += T) {
            double hn, ht;
            gas_cell(i, m, x_sym, dt, Vframe, src, HO, UC[i], VC[i], hn, ht);
            changed |= __double_as_longlong(hn) != __double_as_longlong(src[i]);
            dst[i] = hn;
            HT[i] = ht;
        }
        if (!__syncthreads_or(changed)) {
            passes = pass + 1;
            return dst;
        }
        double *t = src;
        src = dst;
        dst = t;
    }
    if (tid == 0) gas_height_serial(m, x_sym, dt, Vframe, HO, UC, VC, H0, HT);
    __syncthreads();
    passes = 0;
    return H0;
}

// Deterministic workgroup sum; result broadcast to every thread.
__device__ double wg_sum(double x, double *red) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) red[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < nw; k++) s += red[k];
        red[32] = s;
    }
    __syncthreads();
    return red[32];
}

__device__ void gas_body(const StepRec &a, double *lds, double *red) {
    const Geo &g = a.g;
    const Fields &F = a.F;
    const Gas &G = a.G;
    const GasConst &K = a.K;
    const double dt = a.dt, Vframe = a.Vframe;
    const int m = g.m, me = g.me, ml = g.ml, T = blockDim.x, tid = threadIdx.x;
    const int x_sym = g.x_sym;

    // gas_layer.cc:39-40
    for (int k = tid; k < me; k += T) G.p_old[k] = G.p[k];
    for (int k = tid; k < ml; k += T) G.tau_old[k] = G.tau[k];

    // update_height: stage old h and the interface velocities in LDS
    double *HO = lds, *UC = lds + m, *VC = lds + 2 * m;
    for (int i = tid; i < m; i += T) {
        HO[i] = G.h[i];
        UC[i] = 0.25 * K.gxsp * (F.u[IX(i, 0)] + F.u[IX(i, -1)]);
        VC[i] = 0.5 * (F.v[IX(i, 0)] + F.v[IX(i, -1)]);
    }
    __syncthreads();
    // fbd += dt v_c does not take part in the in-place sweep
    for (int i = tid; i < m; i += T) G.fbd[i] += dt * VC[i];
    __syncthreads();
    double *HN = UC, *HT = VC;
    int passes = 0;
    if (m <= kGasJacobiM) {
        HT = lds + 5 * m;
        HN = gas_height_jacobi(m, x_sym, dt, Vframe, HO, UC, VC, lds + 3 * m, lds + 4 * m, HT, passes);
    } else {
        if (tid == 0) gas_height_serial(m, x_sym, dt, Vframe, HO, UC, VC, UC, VC);
        __syncthreads();
    }
    for (int i = tid; i < m; i += T) {
        G.h[i] = HN[i];
        G.ht[i] = HT[i];
    }
    // calculate_hx (gas_layer_solve.cc:58-62)
    for (int i = tid; i <= m; i += T) {
        double v;
        if (i >= 1 && i < m) v = K.gxsp * (HN[i] - HN[i - 1]);
        else if (i == 0) v = x_sym ? 0.0 : K.gxsp * (HN[1] - HN[0]);
        else v = K.gxsp * (HN[m - 1] - HN[m - 2]);
        G.hx[i] = v;
    }
    // update_gas_pressure: pg_old = pg
    for (int i = tid; i < me; i += T) G.pg_old[i] = G.pg[i];
    __syncthreads();

    // implicit_step / newton_step (gas_layer_solve.cc:67-129, :197-212)
    double *A = lds, *B = lds + me, *C = lds + 2 * me, *D = lds + 3 * me;
    const double mu = K.mug, alpha = K.alpha, Pamb = K.Pamb, xsp = K.gxsp, xxsp = K.gxxsp, dtinv = 1. / dt;
    const double pre = std::numeric_limits<double>::epsilon() * 1e8, thresh = pre * pre;
    int nstep = 0, naccept = 0, status = VDI_OK;
    while (true) {
        if (++nstep > 256) { status = VDI_ERR_NEWTON_NOCONV; nstep--; break; }
        double part = 0;
        for (int i = tid; i < me; i += T) {
            double a = 0, b = 1, c = 0, d;
            if (i == me - 1) d = -G.pg[me - 1];
            else if (i == 0 && !x_sym) d = -G.pg[0];
            else {
                const double *h = G.h, *ht = G.ht, *pg = G.pg;
                double h_ns = i > 0 ? 0.5 * (h[i - 1] + h[i]) : h[i];
                double ht_ns = i > 0 ? 0.5 * (ht[i - 1] + ht[i]) : ht[i];
                double pgi = pg[i], pgp = pg[i + 1], pgm = i > 0 ? pg[i - 1] : pg[i + 1];
                double px = i > 0 ? 0.5 * xsp * (pgp - pgm) : 0.0;
                double pxx = xxsp * (pgp - 2 * pgi + pgm);
                double hsq = h_ns * h_ns, hcu = hsq * h_ns, hxi = G.hx[i];
                double br = 3 * hsq * hxi * px + hcu * pxx - 12 * mu * alpha * h_ns * dtinv - 12 * mu * ht_ns;
                a = -xsp * alpha * hcu * px + pgi * (-1.5 * xsp * hsq * hxi + hcu * xxsp) +
                    Pamb * (-1.5 * xsp * hsq * hxi + hcu * xxsp);
                b = br - 2 * pgi * hcu * xxsp - 2 * Pamb * hcu * xxsp;
                c = xsp * alpha * hcu * px + pgi * (1.5 * xsp * hsq * hxi + hcu * xxsp) +
                    Pamb * (1.5 * xsp * hsq * hxi + hcu * xxsp);
                if (i == 0) { c += a; a = 0; }
                d = -(12 * mu * alpha * h_ns * dtinv * G.pg_old[i] + alpha * hcu * px * px + pgi * br +
                      Pamb * (-12 * mu * ht_ns + 3 * hsq * hxi * px + hcu * pxx));
                part += d * d;
            }
            A[i] = a; B[i] = b; C[i] = c; D[i] = d;
        }
        double ressq = wg_sum(part, red) / me;  // quirk Q7: divided by me
        // Tridiagonal solve without pivoting (the Newton matrix is diagonally
        // dominant), on rows scaled to b = 1: k steps of parallel cyclic
        // reduction (a step reads a, c, d of the two partner rows, the own row
        // staying in registers, and divides once) leave P = 2^k independent
        // systems of rows r, r + P, r + 2P, ... of at most about 16 rows, which
        // P threads then solve by Thomas elimination.  The last row is the
        // Dirichlet row pg = 0 (a = c = 0, b = 1), which a reduction step
        // leaves as it is, so only rows 0..me-2 are reduced.
        const int nr = me - 1;
        int P = 1;
        while (me > 16 * P && P < T) P <<= 1;
        double ra[kGasRows], rc[kGasRows], rd[kGasRows];
#pragma unroll
        for (int r = 0; r < kGasRows; r++) {
            const int i = tid + r * T;
            if (i >= me) continue;
            const double inv = 1.0 / B[i];
            ra[r] = A[i] * inv;
            rc[r] = C[i] * inv;
            rd[r] = D[i] * inv;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kGasRows; r++) {
            const int i = tid + r * T;
            if (i >= me) continue;
            A[i] = ra[r]; C[i] = rc[r]; D[i] = rd[r];
        }
        __syncthreads();
        for (int s = 1; s < P; s <<= 1) {
#pragma unroll
            for (int r = 0; r < kGasRows; r++) {
                const int i = tid + r * T;
                if (i >= nr) continue;
                double k1 = 0, k2 = 0, am = 0, cm = 0, dm = 0, ap = 0, cp = 0, dp = 0;
                if (i - s >= 0) { k1 = ra[r]; am = A[i - s]; cm = C[i - s]; dm = D[i - s]; }
                if (i + s < me) { k2 = rc[r]; ap = A[i + s]; cp = C[i + s]; dp = D[i + s]; }
                const double inv = 1.0 / (1.0 - cm * k1 - ap * k2);
                ra[r] = -am * k1 * inv;
                rc[r] = -cp * k2 * inv;
                rd[r] = (rd[r] - dm * k1 - dp * k2) * inv;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kGasRows; r++) {
                const int i = tid + r * T;
                if (i >= nr) continue;
                A[i] = ra[r]; C[i] = rc[r]; D[i] = rd[r];
            }
            __syncthreads();
        }
        // Thomas on system tid: row i couples to i - P (a) and i + P (c)
        if (tid < P) {
            double cq = 0, dq = 0;
            int last = tid;
            for (int i = tid; i < me; i += P) {
                const double a = A[i];
                const double inv = 1.0 / (1.0 - a * cq);
                cq = C[i] * inv;
                dq = (D[i] - a * dq) * inv;
                C[i] = cq;
                D[i] = dq;
                last = i;
            }
            double x = 0;
            for (int i = last; i >= tid; i -= P) {
                x = D[i] - C[i] * x;
                D[i] = x;
            }
        }
        __syncthreads();
        double bad = 0;
        for (int i = tid; i < me; i += T) {
            double x = D[i];
            if (!(x == x) || x - x != 0) bad = 1;
            G.pg[i] += x;
        }
        if (wg_sum(bad, red) != 0) { status = VDI_ERR_TRIDIAG; break; }
        if (ressq < thresh) naccept++;
        else naccept = 0;
        if (naccept >= 2) break;
    }
    __syncthreads();

    // calculate_liquid_fields (gas_layer_solve.cc:229-249)
    const double fac = 0.5 * K.sigma * xxsp, fac2 = K.sigma / K.R;
    const double *h = G.h, *pg = G.pg;
    for (int i = tid; i < me; i += T) {
        double v;
        if (i == 0) v = x_sym ? pg[0] + 2 * fac * (h[1] - h[0]) - fac2 : pg[0];
        else if (i == 1) v = x_sym ? pg[1] + fac * (h[2] - h[1]) - fac2 : pg[1];
        else if (i < m - 1) v = pg[i] + fac * (h[i + 1] - h[i] - h[i - 1] + h[i - 2]) - fac2;
        else v = pg[i];
        G.p[i] = v;
    }
    __syncthreads();
    for (int i = tid; i < m; i += T) G.tau[i] = -0.5 * xsp * mu * (G.p[i + 1] - G.p[i]) * h[i];
    __syncthreads();
    if (tid == 0) {
        // gas_layer.cc:25-30 (quirk Q1: tau[m], tau[m+1] are never written by the model)
        double *tau = G.tau;
        tau[0] = -tau[3]; tau[1] = -tau[2];
        tau[ml - 2] = -tau[ml - 3]; tau[ml - 1] = -tau[ml - 4];
        G.status[0] = status;
        G.status[1] = nstep;
        G.status[2] = passes;
    }
}

// gas_layer_solve.cc:32-53 + :229-249 (initial liquid fields, no tau BC: quirk Q6)
__global__ void __launch_bounds__(1024) k_gas_init(Geo g, Gas G, GasConst K) {
    const int m = g.m, me = g.me, T = blockDim.x, tid = threadIdx.x;
    const double dx = (g.bx - g.ax) / m;
    for (int i = tid; i < m; i += T) {
        G.fbd[i] = 0;
        double x = g.ax + (i + 0.5) * dx;
        G.h[i] = K.h0 + 0.5 / K.R * x * x;
        G.ht[i] = -K.V;
    }
    for (int i = tid; i < me; i += T) G.pg[i] = 0;
    __syncthreads();
    for (int i = tid; i <= m; i += T) {
        double v;
        if (i >= 1 && i < m) v = K.gxsp * (G.h[i] - G.h[i - 1]);
        else if (i == 0) v = g.x_sym ? 0.0 : K.gxsp * (G.h[1] - G.h[0]);
        else v = K.gxsp * (G.h[m - 1] - G.h[m - 2]);
        G.hx[i] = v;
    }
    const double fac = 0.5 * K.sigma * K.gxxsp, fac2 = K.sigma / K.R;
    const double *h = G.h, *pg = G.pg;
    for (int i = tid; i < me; i += T) {
        double v;
        if (i == 0) v = g.x_sym ? pg[0] + 2 * fac * (h[1] - h[0]) - fac2 : pg[0];
        else if (i == 1) v = g.x_sym ? pg[1] + fac * (h[2] - h[1]) - fac2 : pg[1];
        else if (i < m - 1) v = pg[i] + fac * (h[i + 1] - h[i] - h[i - 1] + h[i - 2]) - fac2;
        else v = pg[i];
        G.p[i] = v;
    }
    __syncthreads();
    for (int i = tid; i < m; i += T) G.tau[i] = -0.5 * K.gxsp * K.mug * (G.p[i + 1] - G.p[i]) * h[i];
}

#undef IX

// Gathers the four solve states and the gas-layer status for one readback.
__device__ void gather_body(const StepRec &a) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 4; k++)
        if (a.mgst[k]) a.rb_st[k] = *a.mgst[k];
    a.rb_gas[0] = a.G.status[0];
    a.rb_gas[1] = a.G.status[1];
    a.rb_gas[2] = a.G.status[2];
}

// Snapshot for a redo (u, v, p are the first three field arrays, then the
// gas arrays) and its restore; 2048 values per block.
__device__ void snapshot_body(const StepRec &a, int blk, bool restore) {
    const long n3 = 3 * a.np, tot = n3 + a.ng;
    for (long k = blk * 2048L + threadIdx.x; k < tot && k < (blk + 1) * 2048L; k += 256) {
        double *live = k < n3 ? a.F.u + k : a.gasbase + (k - n3);
        if (restore) *live = a.snap[k];
        else a.snap[k] = *live;
    }
}

// ------------------------------------------------------------------------
// Kernels: direct (one record, grid = tiles) and batched (records in device
// memory, a persistent grid walking the (record, tile) work list)
// ------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_godunov(StepRec a) { godunov_body(a, blockIdx.x, blockIdx.y); }
// The gas-pressure term of the bottom MAC source row (the j = 0 branch of
// godunov_body, same operations in the same order) when Godunov ran with
// defer_sbc, beside the gas update that produces p.
__global__ void __launch_bounds__(256) k_mac_bottom(StepRec a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const Geo &g = a.g;
    if (i >= g.m) return;
    const Gas &G = a.G;
    const size_t mn = (size_t)g.m * g.n;
    const double sbc = a.macfac * (G.p_old[i] + G.p_old[i + 1] + G.p[i] + G.p[i + 1]);
    a.F.src[mn + i] = sbc;
    double s = a.F.src[i];
    s -= 2 * g.yysp * sbc;
    a.F.src[i] = s;
}

__global__ void __launch_bounds__(256) k_correct_advect(StepRec a) { correct_body(a, blockIdx.x, blockIdx.y); }
__global__ void __launch_bounds__(256) k_simpson(StepRec a, Stamp *ts) {
    stamp_begin(ts);
    simpson_body(a, blockIdx.x);
    stamp_end(ts);
}
__global__ void __launch_bounds__(1024) k_viscous_bc(StepRec a) { viscous_bc_body(a); }
__global__ void __launch_bounds__(256) k_visco_rhs(StepRec a, int comp) {
    visco_rhs_body(a, blockIdx.x, blockIdx.y, comp);
}
__global__ void __launch_bounds__(256) k_visco_copy(StepRec a) { visco_copy_body(a, blockIdx.x, blockIdx.y); }
__global__ void __launch_bounds__(256) k_fem_rhs(StepRec a) { fem_rhs_body(a, blockIdx.x, blockIdx.y); }
__global__ void __launch_bounds__(256) k_update(StepRec a) { update_body(a, blockIdx.x, blockIdx.y); }
__global__ void __launch_bounds__(kGasThreads) k_gas(StepRec a, Stamp *ts) {
    extern __shared__ double lds[];
    __shared__ double red[33];
    stamp_begin(ts);
    gas_body(a, lds, red);
    stamp_end(ts);
}
__global__ void __launch_bounds__(256) k_snapshot(StepRec a, int restore) { snapshot_body(a, blockIdx.x, restore); }
__global__ void __launch_bounds__(64) k_gather_status(StepRec a) { gather_body(a); }

// fluid_2d::update_tracers (fluid_2d.cc:507-525): each tracer moves by dt times
// the bilinear interpolation of u, v at its position, from the cell whose
// centre lies at or below-left of it (indices truncated toward zero, then
// clamped to [-1, m-1] x [-1, n-1], so tracers near the edges extrapolate
// into the ghost cells).  The arithmetic is left uncontracted.
__global__ void __launch_bounds__(256) k_tracers(StepRec a, double *tm, int nt) {
#pragma clang fp contract(off)
    const Geo &g = a.g;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nt) return;
    double x = (tm[2 * k] - g.ax) * g.xsp - 0.5, y = (tm[2 * k + 1] - g.ay) * g.ysp - 0.5;
    int i = (int)x, j = (int)y;
    if (i < -1) i = -1;
    if (i > g.m - 1) i = g.m - 1;
    if (j < -1) j = -1;
    if (j > g.n - 1) j = g.n - 1;
    x -= i;
    y -= j;
    const size_t c = ((size_t)j + 2) * (size_t)g.ml + (size_t)(i + 2), r = c + 1, t = c + g.ml, tr = t + 1;
    const double *u = a.F.u, *v = a.F.v, dt = a.dt;
    tm[2 * k] += dt * ((1 - y) * (u[c] * (1 - x) + u[r] * x) + y * (u[t] * (1 - x) + u[tr] * x));
    tm[2 * k + 1] += dt * ((1 - y) * (v[c] * (1 - x) + v[r] * x) + y * (v[t] * (1 - x) + v[tr] * x));
}
__global__ void __launch_bounds__(256) k_fetch(const double *__restrict__ src, double *__restrict__ dst, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

int mapped_host_alloc(void **host, void **dev, size_t bytes) {
    VDI_HIP(hipHostMalloc(host, bytes, hipHostMallocMapped | hipHostMallocCoherent));
    memset(*host, 0, bytes);
    VDI_HIP(hipHostGetDevicePointer(dev, *host, 0));
    return VDI_OK;
}
int enqueue_fetch(const double *src, double *dst, int n, hipStream_t st) {
    if (n < 1) return VDI_ERR_ARG;
    hipLaunchKernelGGL(k_fetch, dim3(cdiv(n, 256)), dim3(256), 0, st, src, dst, n);
    VDI_LAUNCHED();
    return VDI_OK;
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

// Batched: tiles of kernel K for record a.
template <int K>
__device__ __forceinline__ void step_item(const StepRec &a, int tile, double *lds, double *red) {
    if constexpr (K == SK_GODUNOV) {
        const int nbx = (a.g.m + kGX - 1) / kGX;
        godunov_body(a, tile % nbx, tile / nbx);
    } else if constexpr (K == SK_CORRECT || K == SK_VISCO_RHS_U || K == SK_VISCO_RHS_V || K == SK_VISCO_COPY) {
        const int nbx = (a.g.m + kEX - 1) / kEX, bx = tile % nbx, by = tile / nbx;
        if constexpr (K == SK_CORRECT) correct_body(a, bx, by);
        else if constexpr (K == SK_VISCO_RHS_U) visco_rhs_body(a, bx, by, 0);
        else if constexpr (K == SK_VISCO_RHS_V) visco_rhs_body(a, bx, by, 1);
        else visco_copy_body(a, bx, by);
    } else if constexpr (K == SK_FEM_RHS || K == SK_UPDATE) {
        const int nbx = (a.g.me + kEX - 1) / kEX, bx = tile % nbx, by = tile / nbx;
        if constexpr (K == SK_FEM_RHS) fem_rhs_body(a, bx, by);
        else update_body(a, bx, by);
    } else if constexpr (K == SK_SIMPSON) {
        simpson_body(a, tile);
    } else if constexpr (K == SK_VISCOUS_BC) {
        viscous_bc_body(a);
    } else if constexpr (K == SK_GAS) {
        gas_body(a, lds, red);
    } else if constexpr (K == SK_SNAPSHOT || K == SK_RESTORE) {
        snapshot_body(a, tile, K == SK_RESTORE);
    } else if constexpr (K == SK_GATHER) {
        gather_body(a);
    }
}

template <int K, int NT>
__global__ void __launch_bounds__(NT) k_step_b(const StepRec *A, const WorkItem *W, int nw, BatchAux aux) {
    extern __shared__ double lds[];
    __shared__ double red[33];
    stamp_begin(aux.ts);
    for (int p = blockIdx.x; p < nw; p += gridDim.x) {
        const WorkItem w = W[p];
        const unsigned long long t0 = busy_begin(aux.busy);
        step_item<K>(A[w.rec], w.tile, lds, red);
        __syncthreads();
        busy_end(aux.busy, w.sim, t0);
    }
    stamp_end(aux.ts);
}

namespace stepk {

int threads(int k) { return (k == SK_VISCOUS_BC || k == SK_GAS) ? 1024 : 256; }

int tiles(int k, const Geo &g) {
    switch (k) {
    case SK_GODUNOV: return (int)(cdiv(g.m, kGX) * cdiv(g.n, kGY));
    case SK_CORRECT:
    case SK_VISCO_RHS_U:
    case SK_VISCO_RHS_V:
    case SK_VISCO_COPY: return (int)(cdiv(g.m, kEX) * cdiv(g.n, kEY));
    case SK_FEM_RHS:
    case SK_UPDATE: return (int)(cdiv(g.me, kEX) * cdiv(g.ne, kEY));
    case SK_SIMPSON: return (int)cdiv(g.n + g.ml, 4);
    case SK_SNAPSHOT:
    case SK_RESTORE: {
        const long np = (long)g.ml * (g.n + 4), ng = 5L * g.me + 2L * g.ml + 3L * g.m;
        return (int)cdiv(3 * np + ng, 2048);
    }
    default: return 1;
    }
}

size_t gas_lds(const Geo &g) {
    const size_t h = (g.m <= kGasJacobiM ? 6 : 3) * (size_t)g.m;  // height update
    return std::max<size_t>(h, 4 * (size_t)g.me) * sizeof(double);  // Newton rows
}

#define SK_CASES(X)                                                                                                \
    X(SK_GODUNOV, 256) X(SK_CORRECT, 256) X(SK_SIMPSON, 256) X(SK_VISCOUS_BC, 1024) X(SK_VISCO_RHS_U, 256)        \
    X(SK_VISCO_RHS_V, 256) X(SK_VISCO_COPY, 256) X(SK_FEM_RHS, 256) X(SK_UPDATE, 256) X(SK_GAS, 1024)             \
    X(SK_SNAPSHOT, 256) X(SK_RESTORE, 256) X(SK_GATHER, 256)

const void *kernel_b(int k) {
    switch (k) {
#define X(K, NT) \
    case K: return (const void *)k_step_b<K, NT>;
        SK_CASES(X)
#undef X
    default: return nullptr;
    }
}

int launch_b(int k, const StepRec *recs, const WorkItem *w, int nw, int grid, size_t lds, BatchAux aux, hipStream_t st) {
    switch (k) {
#define X(K, NT) \
    case K: hipLaunchKernelGGL((k_step_b<K, NT>), dim3(grid), dim3(NT), lds, st, recs, w, nw, aux); break;
        SK_CASES(X)
#undef X
    default: return VDI_ERR_ARG;
    }
    VDI_LAUNCHED();
    return VDI_OK;
}

}  // namespace stepk

}  // namespace vdi

// ------------------------------------------------------------------------
// Context and C ABI
// ------------------------------------------------------------------------

using namespace vdi;

namespace vdi {

static dim3 cells_grid(const Geo &g) { return dim3(cdiv(g.m, kEX), cdiv(g.n, kEY)); }
static dim3 nodes_grid(const Geo &g) { return dim3(cdiv(g.me, kEX), cdiv(g.ne, kEY)); }

double half_time(double t0, double dt) { return t0 + 0.5 * dt; }
double end_time(double t0, double dt) { return half_time(t0, dt) + 0.5 * dt; }

MG *step_solver(vdi_ctx *c, int k) {
    MG *gs[4] = {&c->mac, &c->vx, &c->vy, &c->fem};
    return gs[k];
}

bool solver_used(vdi_ctx *c, int k) { return c->P.implicit_visc || k == 0 || k == 3; }

StepRec step_record(vdi_ctx *c, double dt) {
    StepRec a{};
    const Geo &g = c->g;
    a.g = g;
    a.F = c->F;
    a.G = c->G;
    a.K = c->K;
    a.dt = dt;
    a.viscfac = c->P.implicit_visc ? g.rhoinv * c->P.mu * 0.5 : g.rhoinv * c->P.mu;
    a.macfac = 0.125 * g.rhoinv * dt;
    a.prefac = g.dx * g.rhoinv * dt / (3 * M_PI);
    a.vfx = a.viscfac * g.xxsp * dt;
    a.vfy = a.viscfac * g.yysp * dt;
    a.rdt = g.rho / dt;
    a.fm = c->fem.fe.fm;
    a.Vframe = c->Vframe;
    a.zmac = c->zmac;
    a.zvx = c->zvx;
    a.zvy = c->zvy;
    a.zfem = c->zfem;
    a.snap = c->snap.as<double>();
    a.gasbase = c->gasbuf.as<double>();
    a.np = (long)g.ml * (g.n + 4);
    a.ng = (long)(c->gasbuf.bytes / sizeof(double));
    for (int k = 0; k < 4; k++) a.mgst[k] = solver_used(c, k) ? step_solver(c, k)->state() : nullptr;
    vdi_ctx::Readback *rbd = c->rb_map;
    a.rb_st = rbd->st;
    a.rb_gas = rbd->gas;
    return a;
}

// visco_impl.cc:21-39: re-set the hierarchy when the coupling changed.
int visco_setup(MG &g, double vfx, double vfy) {
    if (std::fabs(vfx - g.vfx) > 1e2 * std::numeric_limits<double>::epsilon() * g.vfx) {
        g.vfx = vfx;
        g.vfy = vfy;
        return g.set_coeffs(vfx, vfy);
    }
    return VDI_OK;
}

}  // namespace vdi

extern "C" {

const char *vdi_status_string(int s) {
    switch (s) {
    case VDI_OK: return "OK";
    case VDI_ERR_MG_NOCONV: return "Multigrid routine failed";
    case VDI_ERR_NEWTON_NOCONV: return "Too many Newton steps";
    case VDI_ERR_TRIDIAG: return "Tridiagonal solve failed";
    case VDI_ERR_HIP: return "HIP runtime error";
    case VDI_ERR_ARG: return "Invalid argument";
    default: return "Unknown status";
    }
}

const char *vdi_last_error(void) { return g_last_error.c_str(); }

int vdi_create(const vdi_params *p, vdi_ctx **out) {
    if (!p || !out) return VDI_ERR_ARG;
    *out = nullptr;
    if (p->m < 4 || p->n < 4 || p->m + 1 > kGasThreads * kGasRows) return VDI_ERR_ARG;
    if (!(p->bx > p->ax) || !(p->by > p->ay) || !(p->rho > 0) || !(p->mu > 0)) return VDI_ERR_ARG;
    std::unique_ptr<vdi_ctx> c(new vdi_ctx());
    c->P = *p;
    Geo &g = c->g;
    g.m = p->m; g.n = p->n; g.ml = p->m + 4; g.me = p->m + 1; g.ne = p->n + 1; g.x_sym = p->x_sym;
    g.ax = p->ax; g.bx = p->bx; g.ay = p->ay; g.by = p->by;
    g.dx = (p->bx - p->ax) / p->m; g.dy = (p->by - p->ay) / p->n;
    g.xsp = 1. / g.dx; g.ysp = 1. / g.dy; g.xxsp = g.xsp * g.xsp; g.yysp = g.ysp * g.ysp;
    g.rho = p->rho; g.rhoinv = 1. / p->rho; g.mu = p->mu; g.muinv = 1. / p->mu;
    GasConst &K = c->K;
    K.alpha = p->alpha; K.mug = p->mug; K.Pamb = p->Pamb; K.h0 = p->h0; K.R = p->R; K.V = p->V;
    K.sigma = p->sigma;
    K.gxsp = 1. / ((p->bx - p->ax) / p->m);
    K.gxxsp = K.gxsp * K.gxsp;
    c->Vframe = p->V;
    VDI_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    VDI_HIP(hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
    VDI_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    VDI_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    VDI_HIP(hipEventCreateWithFlags(&c->ev_gfork, hipEventDisableTiming));
    VDI_HIP(hipEventCreateWithFlags(&c->ev_gjoin, hipEventDisableTiming));
    VDI_HIP(hipEventCreateWithFlags(&c->ev_god, hipEventDisableTiming));
    VDI_HIP(hipEventCreateWithFlags(&c->ev_bcjoin, hipEventDisableTiming));
    size_t NP = (size_t)g.ml * (g.n + 4);
    VDI_TRY(c->fields.alloc(19 * NP * sizeof(double)));
    double *f = c->fields.as<double>();
    double **dst[19] = {&c->F.u, &c->F.v, &c->F.p, &c->F.ux, &c->F.uy, &c->F.vx, &c->F.vy, &c->F.ud, &c->F.vd,
                        &c->F.ul, &c->F.vl, &c->F.ur, &c->F.vr, &c->F.uu, &c->F.vu, &c->F.c0, &c->F.c1,
                        &c->F.c2, &c->F.c3};
    for (int k = 0; k < 19; k++) *dst[k] = f + k * NP;
    VDI_TRY(c->srcbuf.alloc(((size_t)g.me * g.ne + (size_t)g.m * g.n) * sizeof(double)));
    c->F.src = c->srcbuf.as<double>();
    c->F.src2 = c->F.src + (size_t)g.me * g.ne;
    // gas arrays: p[me], tau[ml], p_old[me], tau_old[ml], h[m], hx[me], ht[m], fbd[m], pg[me], pg_old[me]
    size_t ng = 5 * (size_t)g.me + 2 * (size_t)g.ml + 3 * (size_t)g.m;
    VDI_TRY(c->gasbuf.alloc(ng * sizeof(double)));
    double *q = c->gasbuf.as<double>();
    c->G.p = q; q += g.me;
    c->G.tau = q; q += g.ml;
    c->G.p_old = q; q += g.me;
    c->G.tau_old = q; q += g.ml;
    c->G.h = q; q += g.m;
    c->G.hx = q; q += g.me;
    c->G.ht = q; q += g.m;
    c->G.fbd = q; q += g.m;
    c->G.pg = q; q += g.me;
    c->G.pg_old = q; q += g.me;
    VDI_TRY(c->statusbuf.alloc(4 * sizeof(int)));
    c->G.status = c->statusbuf.as<int>();
    VDI_HIP(hipHostMalloc((void **)&c->host_status, 4 * sizeof(int), hipHostMallocDefault));
    size_t mn = (size_t)g.m * g.n;
    VDI_TRY(c->zbuf.alloc((3 * mn + (size_t)g.me * g.ne) * sizeof(double)));
    double *z = c->zbuf.as<double>();
    c->zmac = z; c->zvx = z + mn; c->zvy = z + 2 * mn; c->zfem = z + 3 * mn;
    VDI_TRY(c->mac.init(VDI_MAC, g.m, g.n, g.xxsp, g.yysp, g.x_sym, c->zmac, c->F.src, c->st));
    VDI_TRY(c->fem.init(VDI_FEM, g.me, g.ne, g.dx, g.dy, g.x_sym, c->zfem, c->F.src, c->st));
    if (p->implicit_visc) {
        VDI_TRY(c->vx.init(VDI_VX, g.m, g.n, 0, 0, g.x_sym, c->zvx, c->F.src, c->st));
        VDI_TRY(c->vy.init(VDI_VY, g.m, g.n, 0, 0, g.x_sym, c->zvy, c->F.src2, c->st));
    }
    {
        MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
        for (int k = 0; k < 4; k++)
            if (gs[k]->nl)
                VDI_TRY(gs[k]->set_schedule(kStepNu[k][0], kStepNu[k][1], kStepNu[k][2], kStepNu[k][3]));
    }
    VDI_TRY(c->snap.alloc((3 * NP + ng) * sizeof(double)));
    VDI_TRY(c->coef_dev.alloc(8 * sizeof(double)));
    VDI_TRY(mapped_host_alloc((void **)&c->coef_host, (void **)&c->coef_map, 8 * sizeof(double)));
    VDI_TRY(mapped_host_alloc((void **)&c->rb, (void **)&c->rb_map, sizeof(vdi_ctx::Readback)));
    c->gas_lds = stepk::gas_lds(g);
    static size_t gas_attr = 0;
    if (c->gas_lds > 64 * 1024 && c->gas_lds > gas_attr) {
        VDI_HIP(hipFuncSetAttribute((const void *)k_gas, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->gas_lds));
        VDI_HIP(hipFuncSetAttribute(stepk::kernel_b(SK_GAS), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)c->gas_lds));
        gas_attr = c->gas_lds;
    }
    *out = c.release();
    return VDI_OK;
}

int vdi_destroy(vdi_ctx *ctx) {
    if (!ctx) return VDI_ERR_ARG;
    (void)hipStreamSynchronize(ctx->st);
    delete ctx;
    return VDI_OK;
}

int vdi_set_solver(vdi_ctx *c, int nu1, int nu2, double rel_tol, int max_cycles) {
    const bool keep = nu1 == 0 && nu2 == 0;
    if (!c || (!keep && (nu1 < 0 || nu2 < 0 || nu1 > kMaxNu || nu2 > kMaxNu)) || !(rel_tol > 0) || max_cycles < 1)
        return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->drop_graphs();
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    if (!keep) {
        // all four or none: a schedule one solver rejects restores the others
        int old[4][4];
        for (int k = 0; k < 4; k++) {
            old[k][0] = gs[k]->nu1;
            old[k][1] = gs[k]->nu2;
            old[k][2] = gs[k]->nu1s;
            old[k][3] = gs[k]->nu2s;
        }
        for (int k = 0; k < 4; k++) {
            if (!gs[k]->nl) continue;
            const int rc = gs[k]->set_schedule(nu1, nu2);
            if (rc != VDI_OK) {
                for (int q = 0; q < k; q++)
                    if (gs[q]->nl) (void)gs[q]->set_schedule(old[q][0], old[q][1], old[q][2], old[q][3]);
                return rc;
            }
        }
    }
    c->rel_tol = rel_tol;
    c->max_cycles = max_cycles;
    return VDI_OK;
}

int vdi_get_schedule(vdi_ctx *c, int which_solver, int *nu1, int *nu2) {
    if (!c || !nu1 || !nu2 || which_solver < VDI_MAC || which_solver > VDI_VY) return VDI_ERR_ARG;
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    if (!gs[which_solver]->nl) return VDI_ERR_ARG;
    *nu1 = gs[which_solver]->nu1;
    *nu2 = gs[which_solver]->nu2;
    return VDI_OK;
}

int vdi_set_schedule(vdi_ctx *c, int which_solver, int nu1, int nu2) {
    return vdi_set_schedule_levels(c, which_solver, nu1, nu2, nu1, nu2);
}

int vdi_set_schedule_levels(vdi_ctx *c, int which_solver, int nu1, int nu2, int nu1s, int nu2s) {
    if (!c || which_solver < VDI_MAC || which_solver > VDI_VY) return VDI_ERR_ARG;
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    MG &g = *gs[which_solver];
    if (!g.nl) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->drop_graphs();
    return g.set_schedule(nu1, nu2, nu1s, nu2s);
}

int vdi_get_schedule_levels(vdi_ctx *c, int which_solver, int out[4]) {
    if (!c || !out || which_solver < VDI_MAC || which_solver > VDI_VY) return VDI_ERR_ARG;
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    const MG &g = *gs[which_solver];
    if (!g.nl) return VDI_ERR_ARG;
    out[0] = g.nu1;
    out[1] = g.nu2;
    out[2] = g.nu1s;
    out[3] = g.nu2s;
    return VDI_OK;
}

int vdi_set_form(vdi_ctx *c, int form) {
    if (!c || (form != VDI_FORM_FP64 && form != VDI_FORM_MIXED)) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->form = form;
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    for (MG *g : gs) g->form = form;
    return VDI_OK;
}

}  // extern "C"

static int enqueue_viscous_bc(vdi_ctx *c, const StepRec &a, hipStream_t st = nullptr) {
    hipLaunchKernelGGL(k_viscous_bc, dim3(1), dim3(1024), 0, st ? st : c->st, a);
    VDI_LAUNCHED();
    return VDI_OK;
}

static int enqueue_inviscid_bc(vdi_ctx *c, const StepRec &a, hipStream_t st = nullptr) {
    const int waves = c->g.n + c->g.ml;
    // reads p once; reads and writes two ghost cells of u and v per line
    Stamp *ts = c->timer.slot(VDI_KCLASS_SIMPSON, 8.0 * (c->g.me + 8.0 * waves), -1, cdiv(waves, 4));
    hipLaunchKernelGGL(k_simpson, dim3(cdiv(waves, 4)), dim3(256), 0, st ? st : c->st, a, ts);
    VDI_LAUNCHED();
    return VDI_OK;
}

extern "C" int vdi_init_fields(vdi_ctx *c) {
    if (!c) return VDI_ERR_ARG;
    c->time = 0;
    size_t NP = (size_t)c->g.ml * (c->g.n + 4);
    VDI_HIP(hipMemsetAsync(c->F.u, 0, NP * sizeof(double), c->st));
    VDI_HIP(hipMemsetAsync(c->F.v, 0, NP * sizeof(double), c->st));
    VDI_HIP(hipMemsetAsync(c->F.p, 0, NP * sizeof(double), c->st));
    hipLaunchKernelGGL(k_gas_init, dim3(1), dim3(1024), 0, c->st, c->g, c->G, c->K);
    VDI_LAUNCHED();
    const StepRec a = step_record(c, 0.0);
    VDI_TRY(enqueue_inviscid_bc(c, a));
    VDI_TRY(enqueue_viscous_bc(c, a));
    VDI_HIP(hipStreamSynchronize(c->st));
    return VDI_OK;
}

static int enqueue_gas(vdi_ctx *c, const StepRec &a, hipStream_t st = nullptr) {
    const Geo &g = c->g;
    // reads u,v rows 0,-1, h, ht, fbd, pg, p, tau; writes p_old, tau_old, h, ht, fbd, hx, pg, pg_old, p, tau
    Stamp *ts = c->timer.slot(VDI_KCLASS_GAS, 8.0 * (4.0 * g.m + 3.0 * g.m + 3.0 * g.me + 2.0 * g.ml + 3.0 * g.m +
                                                      4.0 * g.me + 2.0 * g.ml));
    hipLaunchKernelGGL(k_gas, dim3(1), dim3(kGasThreads), c->gas_lds, st ? st : c->st, a, ts);
    VDI_LAUNCHED();
    return VDI_OK;
}

static int gas_update(vdi_ctx *c, const StepRec &a) {
    VDI_TRY(enqueue_gas(c, a));
    VDI_HIP(hipMemcpyAsync(c->host_status, c->G.status, 3 * sizeof(int), hipMemcpyDeviceToHost, c->st));
    VDI_HIP(hipStreamSynchronize(c->st));
    c->last_newton = c->host_status[1];
    c->last_height_passes = c->host_status[2];
    return c->host_status[0];
}

extern "C" int vdi_gas_layer_update(vdi_ctx *c, double t, double dt) {
    (void)t;
    if (!c || !(dt > 0)) return VDI_ERR_ARG;
    return gas_update(c, step_record(c, dt));
}

static int solve_mg(vdi_ctx *c, MG &g, double t, int *cyc) {
    VDI_TRY(g.initial_guess(t));
    return g.solve(c->rel_tol, c->max_cycles, cyc);
}

extern "C" int vdi_mg_solve_v_cycle(vdi_ctx *c, int which, double t, int *vcycles) {
    if (!c || (which != VDI_MAC && which != VDI_FEM)) return VDI_ERR_ARG;
    return solve_mg(c, which == VDI_MAC ? c->mac : c->fem, t, vcycles);
}

extern "C" int vdi_visco_setup_and_solve(vdi_ctx *c, int which, double vfx, double vfy, double t, int *vcycles) {
    if (!c || !c->P.implicit_visc || (which != VDI_VX && which != VDI_VY)) return VDI_ERR_ARG;
    if (!(vfx >= 0) || !(vfy >= 0)) return VDI_ERR_ARG;
    MG &g = which == VDI_VX ? c->vx : c->vy;
    VDI_TRY(visco_setup(g, vfx, vfy));
    // Both reference solvers read the shared src (visco_impl.cc:13); inside
    // the step the v solve has a right-hand side buffer of its own (so that
    // the two solves can run side by side), which this call fills from src.
    if (which == VDI_VY)
        VDI_HIP(hipMemcpyAsync(c->F.src2, c->F.src, (size_t)c->g.m * c->g.n * sizeof(double),
                               hipMemcpyDeviceToDevice, c->st));
    return solve_mg(c, g, t, vcycles);
}

// How a solve inside the step body is carried out.
enum SolveMode {
    SOLVE_DIRECT,  // initial guess + solve with host synchronisation
    SOLVE_AGAIN,   // solve from the current z (redo of a step), no extrapolation
    SOLVE_FIXED    // device-coefficient extrapolation + fixed budget, no host synchronisation
};

static int body_solve(vdi_ctx *c, int k, double t, SolveMode mode) {
    MG &g = *step_solver(c, k);
    // the learned guess replaces the MAC system's extrapolated guess when it
    // lowers the initial residual (mode 2), or records training pairs (mode 1)
    LearnedGuess *L = (k == 0 && c->guess) ? c->guess : nullptr;
    const int lmode = guess_mode(L);
    if (mode == SOLVE_FIXED) {
        VDI_TRY(g.enqueue_extrapolate(c->coef_dev.as<double>() + 2 * k));
        if (lmode == 2) VDI_TRY(guess_infer(*L, g, c->st));
        VDI_TRY(g.enqueue_fixed(c->rel_tol, c->budget[k]));
        if (lmode == 2) VDI_TRY(guess_after_solve(*L, g, c->st));
        return VDI_OK;
    }
    int cyc = 0;
    if (mode == SOLVE_DIRECT) {
        VDI_TRY(g.initial_guess(t));
        if (lmode == 1) VDI_TRY(guess_capture_pre(*L, g, c->st));
        if (lmode == 2) VDI_TRY(guess_infer(*L, g, c->st));
    }
    VDI_TRY(g.solve(c->rel_tol, c->max_cycles, &cyc));
    if (mode == SOLVE_DIRECT && lmode) VDI_TRY(guess_after_solve(*L, g, c->st));
    // a redo continues from the captured step's iterates: the whole solve
    // took those cycles plus the redo's
    c->record_cycles(k, (mode == SOLVE_AGAIN ? c->redo_base[k] : 0) + cyc);
    return VDI_OK;
}

// fluid_2d::step_forward (fluid_2d.cc:93-405) from the Godunov passes on; the
// gas layer has been enqueued already.  Host times: t0 at the start of the
// step; the clock advances in two half steps, as fluid_2d.cc:102 and :251 do,
// so that solve times match bit for bit.
static int enqueue_body(vdi_ctx *c, const StepRec &a, double t0, SolveMode mode, bool gas_branch = false) {
    const Geo &g = c->g;
    const dim3 gc = cells_grid(g), gn = nodes_grid(g);
    const double dt = a.dt;
    if (gas_branch) {
        // The gas update (st2) and the Godunov passes (st) share no data but
        // the gas pressure of the bottom MAC source row, which is added once
        // both are done.  The boundary kernels that follow on st2 (far-field
        // ghosts from the gas pressure, then the viscous ghosts) wait for
        // Godunov, which reads the old ghosts, and run beside the MAC solve
        // and the advective update: those read and write no ghost cell of u
        // or v (only interior u, v, the faces, c0/c1, z and src).
        VDI_HIP(hipEventRecord(c->ev_gfork, c->st));
        VDI_HIP(hipStreamWaitEvent(c->st2, c->ev_gfork, 0));
        VDI_TRY(enqueue_gas(c, a, c->st2));
        VDI_HIP(hipEventRecord(c->ev_gjoin, c->st2));
        StepRec ad = a;
        ad.defer_sbc = 1;
        hipLaunchKernelGGL(k_godunov, dim3(cdiv(g.m, kGX), cdiv(g.n, kGY)), dim3(256), 0, c->st, ad);
        VDI_LAUNCHED();
        VDI_HIP(hipEventRecord(c->ev_god, c->st));
        VDI_HIP(hipStreamWaitEvent(c->st2, c->ev_god, 0));
        VDI_TRY(enqueue_inviscid_bc(c, a, c->st2));
        VDI_TRY(enqueue_viscous_bc(c, a, c->st2));
        VDI_HIP(hipEventRecord(c->ev_bcjoin, c->st2));
        VDI_HIP(hipStreamWaitEvent(c->st, c->ev_gjoin, 0));
        hipLaunchKernelGGL(k_mac_bottom, dim3(cdiv(g.m, 256)), dim3(256), 0, c->st, a);
        VDI_LAUNCHED();
    } else {
        hipLaunchKernelGGL(k_godunov, dim3(cdiv(g.m, kGX), cdiv(g.n, kGY)), dim3(256), 0, c->st, a);
        VDI_LAUNCHED();
    }
    VDI_TRY(body_solve(c, 0, half_time(t0, dt), mode));
    hipLaunchKernelGGL(k_correct_advect, gc, dim3(256), 0, c->st, a);
    VDI_LAUNCHED();
    if (gas_branch) {
        VDI_HIP(hipStreamWaitEvent(c->st, c->ev_bcjoin, 0));
    } else {
        VDI_TRY(enqueue_inviscid_bc(c, a));
        VDI_TRY(enqueue_viscous_bc(c, a));
    }
    const double t1 = end_time(t0, dt);
    if (c->P.implicit_visc) {
        hipLaunchKernelGGL(k_visco_rhs, gc, dim3(256), 0, c->st, a, 0);
        VDI_LAUNCHED();
        VDI_TRY(visco_setup(c->vx, a.vfx, a.vfy));
        VDI_TRY(visco_setup(c->vy, a.vfx, a.vfy));
        if (mode == SOLVE_FIXED && c->par_visco) {
            // The u and v solves are independent (separate right-hand sides
            // and solutions): in a captured step the v branch (its right-hand
            // side and solve) runs on st2 beside the u solve on st.
            VDI_HIP(hipEventRecord(c->ev_fork, c->st));
            VDI_HIP(hipStreamWaitEvent(c->st2, c->ev_fork, 0));
            hipLaunchKernelGGL(k_visco_rhs, gc, dim3(256), 0, c->st2, a, 1);
            VDI_LAUNCHED();
            c->vy.st = c->st2;
            const int rc = body_solve(c, 2, t1, mode);
            c->vy.st = c->st;
            VDI_TRY(rc);
            VDI_TRY(body_solve(c, 1, t1, mode));
            VDI_HIP(hipEventRecord(c->ev_join, c->st2));
            VDI_HIP(hipStreamWaitEvent(c->st, c->ev_join, 0));
        } else {
            VDI_TRY(body_solve(c, 1, t1, mode));
            hipLaunchKernelGGL(k_visco_rhs, gc, dim3(256), 0, c->st, a, 1);
            VDI_LAUNCHED();
            VDI_TRY(body_solve(c, 2, t1, mode));
        }
        hipLaunchKernelGGL(k_visco_copy, gc, dim3(256), 0, c->st, a);
        VDI_LAUNCHED();
    }
    hipLaunchKernelGGL(k_fem_rhs, gn, dim3(256), 0, c->st, a);
    VDI_LAUNCHED();
    VDI_TRY(body_solve(c, 3, t1, mode));
    hipLaunchKernelGGL(k_update, gn, dim3(256), 0, c->st, a);
    VDI_LAUNCHED();
    VDI_TRY(enqueue_viscous_bc(c, a));
    return VDI_OK;
}

// One step with host synchronisation after the gas layer and inside every
// solve (the path of the first steps, and of a redo).
static int step_direct(vdi_ctx *c, double dt, SolveMode mode) {
    const double t0 = c->time;
    const StepRec a = step_record(c, dt);
    if (c->timer.cls >= 0) VDI_TRY(c->timer.reset(c->timer.reserved, c->timer.cap, c->st));
    c->timer.next = c->timer.reserved;
    VDI_TRY(gas_update(c, a));
    VDI_TRY(enqueue_body(c, a, t0, mode));
    c->time = end_time(t0, dt);
    if (c->timer.cls >= 0) VDI_TRY(c->timer.harvest(c->timer.reserved, c->timer.next, c->st));
    return VDI_OK;
}

static bool graph_ready(vdi_ctx *c) {
    if (!c->use_graphs || guess_mode(c->guess) == 1) return false;
    for (int k = 0; k < 4; k++) {
        if (!solver_used(c, k)) continue;
        if (step_solver(c, k)->scount < 2 || c->budget[k] < 1) return false;
    }
    return true;
}

static int capture_step(vdi_ctx *c, double dt, vdi_ctx::GraphEntry &e) {
    const StepRec a = step_record(c, dt);
    const Geo &g = c->g;
    hipGraph_t graph;
    e.slot_a = c->timer.next;
    // The viscous couplings of this time step (a new coarse inverse when they
    // change) are set up before capture, so that nothing inside the capture
    // allocates or copies; the graph keeps the inverses it was recorded with.
    if (c->P.implicit_visc) {
        VDI_TRY(visco_setup(c->vx, a.vfx, a.vfy));
        VDI_TRY(visco_setup(c->vy, a.vfx, a.vfy));
        e.vfx = c->vx.vfx;
        e.vfy = c->vx.vfy;
        e.vinv[0] = c->vx.coarse_factors_in_use();
        e.vinv[1] = c->vy.coarse_factors_in_use();
    }
    VDI_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
    int rc = VDI_OK;
    do {
        // extrapolation coefficients of this step, written by the host into
        // mapped memory before each replay
        if ((rc = enqueue_fetch(c->coef_map, c->coef_dev.as<double>(), 8, c->st)) != VDI_OK) break;
        // the slots this graph will use are cleared at its start; their
        // count is known only after capture, so clear the rest of the pool
        if (c->timer.cls >= 0 && (rc = c->timer.reset(e.slot_a, c->timer.cap, c->st)) != VDI_OK) break;
        hipLaunchKernelGGL(k_snapshot, dim3(stepk::tiles(SK_SNAPSHOT, g)), dim3(256), 0, c->st, a, 0);
        if (c->par_visco) {
            if ((rc = enqueue_body(c, a, c->time, SOLVE_FIXED, true)) != VDI_OK) break;
        } else {
            if ((rc = enqueue_gas(c, a)) != VDI_OK) break;
            if ((rc = enqueue_body(c, a, c->time, SOLVE_FIXED)) != VDI_OK) break;
        }
        // the solve states and the gas status, written into mapped host memory
        hipLaunchKernelGGL(k_gather_status, dim3(1), dim3(64), 0, c->st, a);
    } while (0);
    hipError_t ec = hipStreamEndCapture(c->st, &graph);
    if (rc != VDI_OK) return rc;
    VDI_HIP(ec);
    hipError_t ei = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    VDI_HIP(ei);
    e.slot_b = c->timer.next;
    c->timer.reserved = c->timer.next;
    return VDI_OK;
}

namespace vdi {
int redo_step(vdi_ctx *c, double dt, double t0) {
    const StepRec a = step_record(c, dt);
    hipLaunchKernelGGL(k_snapshot, dim3(stepk::tiles(SK_RESTORE, c->g)), dim3(256), 0, c->st, a, 1);
    VDI_LAUNCHED();
    c->redo_steps++;
    c->time = t0;
    return step_direct(c, dt, SOLVE_AGAIN);
}
}  // namespace vdi

// A captured step: every solve gets its predicted cycle budget (the cycles
// the previous solve of that system needed) and convergence is only checked
// at the end.  If a solve fell short, the state is restored from the
// snapshot the graph took and the step is run again on the direct path.
static int step_graph(vdi_ctx *c, double dt) {
    const double t0 = c->time;
    // host half of the initial guesses (MAC at t+dt/2, the others at t+dt)
    const double ts[4] = {half_time(t0, dt), end_time(t0, dt), end_time(t0, dt), end_time(t0, dt)};
    for (int k = 0; k < 4; k++) {
        if (!solver_used(c, k)) { c->coef_host[2 * k] = 1; c->coef_host[2 * k + 1] = 0; continue; }
        double bb = 1, cc = 0;
        if (step_solver(c, k)->guess_coeffs(ts[k], &bb, &cc) != 1) { bb = 1; cc = 0; }
        c->coef_host[2 * k] = bb;
        c->coef_host[2 * k + 1] = cc;
    }
    vdi_ctx::GraphEntry *e = nullptr;
    for (auto &x : c->graphs)
        if (x.dt == dt && x.form == c->form && std::equal(x.budget, x.budget + 4, c->budget)) e = &x;
    if (!e) {
        if ((int)c->graphs.size() >= vdi_ctx::kMaxGraphs) {
            // bounded cache: destroy the least recently replayed graph
            auto lru = std::min_element(c->graphs.begin(), c->graphs.end(),
                                        [](const vdi_ctx::GraphEntry &x, const vdi_ctx::GraphEntry &y) {
                                            return x.used < y.used;
                                        });
            (void)hipGraphExecDestroy(lru->exec);
            c->graphs.erase(lru);
        }
        vdi_ctx::GraphEntry ne{};
        std::copy(c->budget, c->budget + 4, ne.budget);
        ne.dt = dt;
        ne.form = c->form;
        VDI_TRY(capture_step(c, dt, ne));
        c->graphs.push_back(ne);
        e = &c->graphs.back();
    } else if (c->P.implicit_visc && (c->vx.vfx != e->vfx || c->vy.vfx != e->vfx)) {
        // another time step was set up since this graph was recorded: the
        // host-side couplings and inverses go back to the graph's (a redo on
        // the direct path then finds them consistent)
        c->vx.adopt_coeffs(e->vfx, e->vfy, e->vinv[0]);
        c->vy.adopt_coeffs(e->vfx, e->vfy, e->vinv[1]);
    }
    e->used = ++c->graph_clock;
    VDI_HIP(hipGraphLaunch(e->exec, c->st));
    VDI_HIP(hipStreamSynchronize(c->st));
    if (c->timer.cls >= 0) VDI_TRY(c->timer.harvest(e->slot_a, e->slot_b, c->st));
    c->graph_steps++;
    const vdi_ctx::Readback &r = *c->rb;
    c->last_newton = r.gas[1];
    c->last_height_passes = r.gas[2];
    if (r.gas[0] != VDI_OK) return r.gas[0];
    bool ok = true;
    for (int k = 0; k < 4; k++) {
        if (!solver_used(c, k)) continue;
        MG &g = *step_solver(c, k);
        const MGState &st = r.st[k];
        g.total_cycles += st.cycles;
        c->redo_base[k] = st.cycles;
        if (st.done == 1 || st.done == 3) {
            g.last_cycles = st.cycles;
            g.solves++;
            c->record_cycles(k, st.cycles);
        } else {
            ok = false;
        }
    }
    if (ok) {
        c->time = end_time(t0, dt);
        return VDI_OK;
    }
    // restore the state the step started from and solve again from the
    // partially converged iterates
    return redo_step(c, dt, t0);
}

extern "C" {

int vdi_step_forward(vdi_ctx *c, double dt) {
    if (!c || !(dt > 0)) return VDI_ERR_ARG;
    // fluid_2d.cc:98: the tracers move first, with the velocity at the start
    // of the step (outside the captured step, so a redo does not move them
    // twice)
    if (c->ntrace > 0) {
        hipLaunchKernelGGL(k_tracers, dim3(cdiv(c->ntrace, 256)), dim3(256), 0, c->st, step_record(c, dt),
                           c->tracers.as<double>(), c->ntrace);
        VDI_LAUNCHED();
    }
    int rc = graph_ready(c) ? step_graph(c, dt) : step_direct(c, dt, SOLVE_DIRECT);
    if (rc == VDI_OK) c->steps++;
    return rc;
}

int vdi_set_tracers(vdi_ctx *c, const double *pos, int ntrace) {
    if (!c || ntrace < 0 || (ntrace > 0 && !pos)) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    if (ntrace == 0) {
        c->tracers.release();
        c->ntrace = 0;
        return VDI_OK;
    }
    VDI_TRY(c->tracers.alloc(2 * sizeof(double) * (size_t)ntrace));
    VDI_HIP(hipMemcpy(c->tracers.p, pos, 2 * sizeof(double) * (size_t)ntrace, hipMemcpyHostToDevice));
    c->ntrace = ntrace;
    return VDI_OK;
}

int vdi_get_tracers(vdi_ctx *c, double *pos, int ntrace) {
    if (!c || ntrace != c->ntrace || (ntrace > 0 && !pos)) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    if (ntrace > 0) VDI_HIP(hipMemcpy(pos, c->tracers.p, 2 * sizeof(double) * (size_t)ntrace, hipMemcpyDeviceToHost));
    return VDI_OK;
}

int vdi_run_pass(vdi_ctx *c, int which, double dt, int reps, double *ms) {
    if (!c || !(dt > 0) || reps < 1 || (which != VDI_PASS_GODUNOV && which != VDI_PASS_CORRECT_ADVECT))
        return VDI_ERR_ARG;
    const StepRec a = step_record(c, dt);
    const Geo &g = c->g;
    hipEvent_t e0, e1;
    VDI_HIP(hipEventCreate(&e0));
    VDI_HIP(hipEventCreate(&e1));
    VDI_HIP(hipEventRecord(e0, c->st));
    for (int r = 0; r < reps; r++) {
        if (which == VDI_PASS_GODUNOV)
            hipLaunchKernelGGL(k_godunov, dim3(cdiv(g.m, kGX), cdiv(g.n, kGY)), dim3(256), 0, c->st, a);
        else
            hipLaunchKernelGGL(k_correct_advect, cells_grid(g), dim3(256), 0, c->st, a);
    }
    VDI_LAUNCHED();
    VDI_HIP(hipEventRecord(e1, c->st));
    VDI_HIP(hipEventSynchronize(e1));
    float t = 0;
    VDI_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return VDI_OK;
}

static double *array_ptr(vdi_ctx *c, int id, long *len) {
    long NP = (long)c->g.ml * (c->g.n + 4), mn = (long)c->g.m * c->g.n, nn = (long)c->g.me * c->g.ne;
    *len = 0;
    switch (id) {
    case VDI_U: *len = NP; return c->F.u;
    case VDI_V: *len = NP; return c->F.v;
    case VDI_P: *len = NP; return c->F.p;
    case VDI_UL: *len = NP; return c->F.ul;
    case VDI_VL: *len = NP; return c->F.vl;
    case VDI_UD: *len = NP; return c->F.ud;
    case VDI_VD: *len = NP; return c->F.vd;
    case VDI_C0: *len = NP; return c->F.c0;
    case VDI_C1: *len = NP; return c->F.c1;
    case VDI_SRC: *len = nn; return c->F.src;
    case VDI_GP: *len = c->g.me; return c->G.p;
    case VDI_GTAU: *len = c->g.ml; return c->G.tau;
    case VDI_GP_OLD: *len = c->g.me; return c->G.p_old;
    case VDI_GTAU_OLD: *len = c->g.ml; return c->G.tau_old;
    case VDI_H: *len = c->g.m; return c->G.h;
    case VDI_HX: *len = c->g.me; return c->G.hx;
    case VDI_HT: *len = c->g.m; return c->G.ht;
    case VDI_FBD: *len = c->g.m; return c->G.fbd;
    case VDI_PG: *len = c->g.me; return c->G.pg;
    case VDI_PG_OLD: *len = c->g.me; return c->G.pg_old;
    case VDI_Z_MAC: *len = mn; return c->zmac;
    case VDI_Z_FEM: *len = nn; return c->zfem;
    case VDI_Z_VX: *len = c->P.implicit_visc ? mn : 0; return c->P.implicit_visc ? c->zvx : nullptr;
    case VDI_Z_VY: *len = c->P.implicit_visc ? mn : 0; return c->P.implicit_visc ? c->zvy : nullptr;
    default: return nullptr;
    }
}

long vdi_array_len(vdi_ctx *c, int id) {
    if (!c) return -1;
    long len;
    array_ptr(c, id, &len);
    return len;
}

int vdi_get_array(vdi_ctx *c, int id, double *host, long len) {
    if (!c || !host) return VDI_ERR_ARG;
    long n;
    double *p = array_ptr(c, id, &n);
    if (!p || n != len) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    VDI_HIP(hipMemcpy(host, p, n * sizeof(double), hipMemcpyDeviceToHost));
    return VDI_OK;
}

int vdi_set_array(vdi_ctx *c, int id, const double *host, long len) {
    if (!c || !host) return VDI_ERR_ARG;
    long n;
    double *p = array_ptr(c, id, &n);
    if (!p || n != len) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    VDI_HIP(hipMemcpy(p, host, n * sizeof(double), hipMemcpyHostToDevice));
    return VDI_OK;
}

int vdi_get_time(vdi_ctx *c, double *t) {
    if (!c || !t) return VDI_ERR_ARG;
    *t = c->time;
    return VDI_OK;
}

int vdi_set_time(vdi_ctx *c, double t) {
    if (!c) return VDI_ERR_ARG;
    c->time = t;
    return VDI_OK;
}

int vdi_get_stats(vdi_ctx *c, vdi_stats *s) {
    if (!c || !s) return VDI_ERR_ARG;
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    for (int k = 0; k < 4; k++) {
        s->last_cycles[k] = gs[k]->last_cycles;
        s->total_cycles[k] = gs[k]->total_cycles;
        s->solves[k] = gs[k]->solves;
    }
    s->last_newton = c->last_newton;
    s->steps = c->steps;
    s->time = c->time;
    s->last_height_passes = c->last_height_passes;
    return VDI_OK;
}

int vdi_reset_stats(vdi_ctx *c) {
    if (!c) return VDI_ERR_ARG;
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    for (MG *g : gs) { g->total_cycles = 0; g->solves = 0; }
    c->steps = 0;
    return VDI_OK;
}

int vdi_synchronize(vdi_ctx *c) {
    if (!c) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    return VDI_OK;
}

int vdi_time_kernels(vdi_ctx *c, int kclass, int max_launches) {
    if (!c || max_launches < 0 || kclass < -1 || kclass > VDI_KCLASS_SFEM_UP) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->drop_graphs();  // captured graphs own slots of the previous pool
    VDI_TRY(c->timer.setup(kclass, max_launches));
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    for (MG *g : gs) g->timer = kclass < 0 ? nullptr : &c->timer;
    return VDI_OK;
}

// Totals of the harvested launches of a timer.  Launches that returned at
// once (issued after their solve had already converged) move no data; they
// are left out.
static void timer_totals(const KernelTimer &tm, long *launches, double *total_ms, double *alg_bytes) {
    const std::vector<float> &ms = tm.dur_ms;
    std::vector<float> sorted = ms;
    std::sort(sorted.begin(), sorted.end());
    const float med = sorted.empty() ? 0.f : sorted[sorted.size() / 2];
    long n = 0;
    double t = 0, b = 0;
    for (size_t k = 0; k < ms.size(); k++) {
        if (ms[k] < 0.2f * med) continue;
        n++;
        t += ms[k];
        b += tm.bytes[k];
    }
    *launches = n;
    *total_ms = t;
    *alg_bytes = b;
}

int vdi_kernel_times(vdi_ctx *c, long *launches, double *total_ms, double *alg_bytes) {
    if (!c || !launches || !total_ms || !alg_bytes) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    timer_totals(c->timer, launches, total_ms, alg_bytes);
    return VDI_OK;
}

int vdi_get_counters(vdi_ctx *c, long *graph_steps, long *redo_steps) {
    if (!c || !graph_steps || !redo_steps) return VDI_ERR_ARG;
    *graph_steps = c->graph_steps;
    *redo_steps = c->redo_steps;
    return VDI_OK;
}

int vdi_get_budgets(vdi_ctx *c, int *out) {
    if (!c || !out) return VDI_ERR_ARG;
    for (int k = 0; k < 4; k++) out[k] = c->budget[k];
    return VDI_OK;
}

int vdi_set_graphs(vdi_ctx *c, int enable) {
    if (!c) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->use_graphs = enable ? 1 : 0;
    if (!enable) c->drop_graphs();
    return VDI_OK;
}

int vdi_set_concurrency(vdi_ctx *c, int enable) {
    if (!c) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->par_visco = enable ? 1 : 0;
    c->drop_graphs();
    return VDI_OK;
}

long vdi_state_len(vdi_ctx *c) {
    if (!c) return -1;
    const Geo &g = c->g;
    return 1 + 3L * g.ml * (g.n + 4) + 1 + g.me + g.m + g.m + g.m + g.me;
}

int vdi_download_state(vdi_ctx *c, double *host, long len) {
    if (!c || !host || len != vdi_state_len(c)) return VDI_ERR_ARG;
    const Geo &g = c->g;
    long NP = (long)g.ml * (g.n + 4);
    std::vector<double> u(NP), v(NP), p(NP);
    VDI_TRY(vdi_get_array(c, VDI_U, u.data(), NP));
    VDI_TRY(vdi_get_array(c, VDI_V, v.data(), NP));
    VDI_TRY(vdi_get_array(c, VDI_P, p.data(), NP));
    double *o = host;
    *o++ = c->time;
    for (long k = 0; k < NP; k++) { *o++ = u[k]; *o++ = v[k]; *o++ = p[k]; }
    *o++ = c->Vframe;
    std::vector<double> tau(g.ml);
    VDI_HIP(hipMemcpy(o, c->G.p, g.me * sizeof(double), hipMemcpyDeviceToHost)); o += g.me;
    VDI_HIP(hipMemcpy(tau.data(), c->G.tau, g.ml * sizeof(double), hipMemcpyDeviceToHost));
    memcpy(o, tau.data() + 2, g.m * sizeof(double)); o += g.m;
    VDI_HIP(hipMemcpy(o, c->G.h, g.m * sizeof(double), hipMemcpyDeviceToHost)); o += g.m;
    VDI_HIP(hipMemcpy(o, c->G.ht, g.m * sizeof(double), hipMemcpyDeviceToHost)); o += g.m;
    VDI_HIP(hipMemcpy(o, c->G.pg, g.me * sizeof(double), hipMemcpyDeviceToHost)); o += g.me;
    return VDI_OK;
}

int vdi_upload_state(vdi_ctx *c, const double *host, long len) {
    if (!c || !host || len != vdi_state_len(c)) return VDI_ERR_ARG;
    const Geo &g = c->g;
    long NP = (long)g.ml * (g.n + 4);
    std::vector<double> u(NP), v(NP), p(NP);
    const double *o = host;
    c->time = *o++;
    for (long k = 0; k < NP; k++) { u[k] = *o++; v[k] = *o++; p[k] = *o++; }
    c->Vframe = *o++;
    VDI_TRY(vdi_set_array(c, VDI_U, u.data(), NP));
    VDI_TRY(vdi_set_array(c, VDI_V, v.data(), NP));
    VDI_TRY(vdi_set_array(c, VDI_P, p.data(), NP));
    // gas_layer_solve.cc:313-324: p, tau[2..m+1] + tau BC, h + hx, ht, pg
    std::vector<double> tau(g.ml, 0.0), h(o + g.me, o + g.me + g.m), hx(g.me);
    VDI_HIP(hipMemcpy(c->G.p, o, g.me * sizeof(double), hipMemcpyHostToDevice)); o += g.me;
    memcpy(tau.data() + 2, o, g.m * sizeof(double)); o += g.m;
    tau[0] = -tau[3]; tau[1] = -tau[2]; tau[g.ml - 2] = -tau[g.ml - 3]; tau[g.ml - 1] = -tau[g.ml - 4];
    VDI_HIP(hipMemcpy(c->G.tau, tau.data(), g.ml * sizeof(double), hipMemcpyHostToDevice));
    VDI_HIP(hipMemcpy(c->G.h, o, g.m * sizeof(double), hipMemcpyHostToDevice)); o += g.m;
    for (int i = 1; i < g.m; i++) hx[i] = c->K.gxsp * (h[i] - h[i - 1]);
    hx[0] = g.x_sym ? 0 : hx[1];
    hx[g.m] = hx[g.m - 1];
    VDI_HIP(hipMemcpy(c->G.hx, hx.data(), g.me * sizeof(double), hipMemcpyHostToDevice));
    VDI_HIP(hipMemcpy(c->G.ht, o, g.m * sizeof(double), hipMemcpyHostToDevice)); o += g.m;
    VDI_HIP(hipMemcpy(c->G.pg, o, g.me * sizeof(double), hipMemcpyHostToDevice)); o += g.me;
    return VDI_OK;
}

// ---- standalone linear systems ----

}  // extern "C"

struct vdi_mg {
    MG g;
    hipStream_t st = nullptr;
    KernelTimer timer;  // vdi_mg_time_kernels
    ~vdi_mg() {
        if (st) (void)hipStreamDestroy(st);
    }
};

extern "C" {

int vdi_device_info(int *info, int n) {
    if (!info || n < 1) return VDI_ERR_ARG;
    const hipDeviceAttribute_t attrs[] = {
        hipDeviceAttributeMaxSharedMemoryPerBlock, hipDeviceAttributeMultiprocessorCount,
        hipDeviceAttributeMaxThreadsPerBlock,      hipDeviceAttributeWarpSize,
        hipDeviceAttributeClockRate,               hipDeviceAttributeMemoryClockRate,
        hipDeviceAttributeMemoryBusWidth,          hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
        hipDeviceAttributeL2CacheSize};
    const int na = (int)(sizeof attrs / sizeof attrs[0]);
    for (int k = 0; k < n; k++) {
        info[k] = -1;
        if (k < na) VDI_HIP(hipDeviceGetAttribute(&info[k], attrs[k], 0));
    }
    return VDI_OK;
}

int vdi_copy_probe(long n, int reps, double *ms) {
    if (n <= 0 || reps <= 0 || !ms) return VDI_ERR_ARG;
    DevBuf a, b;
    VDI_TRY(a.alloc(n * sizeof(double)));
    VDI_TRY(b.alloc(n * sizeof(double)));
    hipEvent_t e0, e1;
    VDI_HIP(hipEventCreate(&e0));
    VDI_HIP(hipEventCreate(&e1));
    const unsigned nb = (unsigned)std::min<long>(4096, (n + 255) / 256);
    VDI_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, nullptr, a.as<double>(), b.as<double>(), n);
    VDI_LAUNCHED();
    VDI_HIP(hipEventRecord(e1, nullptr));
    VDI_HIP(hipEventSynchronize(e1));
    float t = 0;
    VDI_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return VDI_OK;
}

int vdi_mg_create(int kind, int m, int n, double cx, double cy, int x_sym, vdi_mg **out) {
    if (!out || kind < VDI_MAC || kind > VDI_VY || m < 2 || n < 2) return VDI_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<vdi_mg> h(new vdi_mg());
    VDI_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    VDI_TRY(h->g.init(kind, m, n, cx, cy, x_sym, nullptr, nullptr, h->st));
    *out = h.release();
    return VDI_OK;
}

int vdi_mg_destroy(vdi_mg *h) {
    if (!h) return VDI_ERR_ARG;
    (void)hipStreamSynchronize(h->st);
    delete h;
    return VDI_OK;
}

int vdi_mg_set_schedule(vdi_mg *h, int nu1, int nu2) {
    if (!h) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    return h->g.set_schedule(nu1, nu2);
}

int vdi_mg_set_form(vdi_mg *h, int form) {
    if (!h || (form != VDI_FORM_FP64 && form != VDI_FORM_MIXED)) return VDI_ERR_ARG;
    h->g.form = form;
    return VDI_OK;
}

int vdi_mg_levels(vdi_mg *h) { return h ? h->g.nl : -1; }

int vdi_mg_set_b(vdi_mg *h, const double *host, long len) {
    if (!h || !host || len != h->g.unknowns()) return VDI_ERR_ARG;
    VDI_HIP(hipMemcpyAsync(h->g.b(), host, len * sizeof(double), hipMemcpyHostToDevice, h->st));
    VDI_HIP(hipStreamSynchronize(h->st));
    return VDI_OK;
}

int vdi_mg_set_z(vdi_mg *h, const double *host, long len) {
    if (!h || !host || len != h->g.unknowns()) return VDI_ERR_ARG;
    VDI_HIP(hipMemcpyAsync(h->g.z(), host, len * sizeof(double), hipMemcpyHostToDevice, h->st));
    VDI_HIP(hipStreamSynchronize(h->st));
    return VDI_OK;
}

int vdi_mg_get_z(vdi_mg *h, double *host, long len) {
    if (!h || !host || len != h->g.unknowns()) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    VDI_HIP(hipMemcpy(host, h->g.z(), len * sizeof(double), hipMemcpyDeviceToHost));
    return VDI_OK;
}

int vdi_mg_apply(vdi_mg *h, const double *z, double *out, long len) {
    if (!h || !z || !out || len != h->g.unknowns()) return VDI_ERR_ARG;
    DevBuf dz, dout;
    VDI_TRY(dz.alloc(len * sizeof(double)));
    VDI_TRY(dout.alloc(len * sizeof(double)));
    VDI_HIP(hipMemcpy(dz.p, z, len * sizeof(double), hipMemcpyHostToDevice));
    VDI_TRY(h->g.apply(dz.as<double>(), dout.as<double>()));
    VDI_HIP(hipStreamSynchronize(h->st));
    VDI_HIP(hipMemcpy(out, dout.p, len * sizeof(double), hipMemcpyDeviceToHost));
    return VDI_OK;
}

int vdi_mg_solve(vdi_mg *h, double rel_tol, int max_cycles, int *vcycles) {
    if (!h || !(rel_tol > 0) || max_cycles < 0) return VDI_ERR_ARG;
    return h->g.solve(rel_tol, max_cycles, vcycles);
}

int vdi_mg_vcycles(vdi_mg *h, int k) {
    if (!h || k < 0) return VDI_ERR_ARG;
    return h->g.vcycles(k);
}

int vdi_mg_device_arrays(vdi_mg *h, double **z_dev, double **b_dev) {
    if (!h || !z_dev || !b_dev) return VDI_ERR_ARG;
    *z_dev = h->g.z();
    *b_dev = h->g.b();
    return VDI_OK;
}

int vdi_mg_solve_fixed(vdi_mg *h, double rel_tol, int budget, int *done, int *vcycles) {
    if (!h || !done || !vcycles || !(rel_tol > 0) || budget < 0) return VDI_ERR_ARG;
    VDI_TRY(h->g.enqueue_fixed(rel_tol, budget));
    MGState st;
    VDI_TRY(h->g.read_state(&st));
    *done = st.done;
    *vcycles = st.cycles;
    return VDI_OK;
}

int vdi_mg_residual(vdi_mg *h, double *rnorm, double *bnorm) {
    if (!h || !rnorm || !bnorm) return VDI_ERR_ARG;
    return h->g.residual(rnorm, bnorm);
}

int vdi_mg_cycle_bytes(vdi_mg *h, double *bytes) {
    if (!h || !bytes) return VDI_ERR_ARG;
    *bytes = h->g.cycle_bytes();
    return VDI_OK;
}

int vdi_mg_synchronize(vdi_mg *h) {
    if (!h) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    return VDI_OK;
}

int vdi_tile_shape(int shape, int *tx, int *ty) {
    if (shape < 0 || shape >= kNumTileShapes || !tx || !ty) return VDI_ERR_ARG;
    *tx = kTileShapes[shape].tx;
    *ty = kTileShapes[shape].ty;
    return VDI_OK;
}

int vdi_mg_time_kernels(vdi_mg *h, int kclass, int max_launches) {
    if (!h || max_launches < 0 || kclass < -1 || kclass > VDI_KCLASS_SFEM_UP) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    VDI_TRY(h->timer.setup(kclass, max_launches));
    h->g.timer = kclass < 0 ? nullptr : &h->timer;
    if (kclass >= 0) VDI_TRY(h->timer.reset(0, h->timer.cap, h->st));
    h->timer.next = 0;
    return VDI_OK;
}

int vdi_mg_kernel_times(vdi_mg *h, long *launches, double *total_ms, double *alg_bytes) {
    if (!h || !launches || !total_ms || !alg_bytes) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    if (h->timer.cls >= 0) {
        VDI_TRY(h->timer.harvest(0, h->timer.next, h->st));
        VDI_TRY(h->timer.reset(0, h->timer.cap, h->st));
        h->timer.next = 0;
    }
    timer_totals(h->timer, launches, total_ms, alg_bytes);
    return VDI_OK;
}

int vdi_tile_stream(int shape) {
    if (shape < 0 || shape >= kNumTileShapes) return -1;  // documented: negative when out of range
    return kTileShapes[shape].stream;
}

int vdi_mg_set_tile(vdi_mg *h, int prec, int which, int level, int shape) {
    if (!h) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    return h->g.set_tile(prec, which, level, shape);
}

int vdi_mg_autotune(vdi_mg *h, int reps) {
    if (!h || reps < 1) return VDI_ERR_ARG;
    return h->g.autotune(reps);
}

int vdi_autotune(vdi_ctx *c, int reps) {
    if (!c || reps < 1) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->drop_graphs();
    for (int k = 0; k < 4; k++)
        if (solver_used(c, k)) VDI_TRY(step_solver(c, k)->autotune(reps));
    return VDI_OK;
}

int vdi_get_tiles(vdi_ctx *c, int which_solver, int *out, int n) {
    if (!c || !out || which_solver < VDI_MAC || which_solver > VDI_VY) return VDI_ERR_ARG;
    MG *gs[4] = {&c->mac, &c->fem, &c->vx, &c->vy};
    MG &g = *gs[which_solver];
    if (!g.nl) return VDI_ERR_ARG;
    int q = 0;
    for (int pr = 0; pr < 2; pr++)
        for (int w = 0; w < 2; w++)
            for (int l = 0; l < g.lc; l++) {
                if (q < n) out[q] = g.tile[pr][w][l];
                q++;
            }
    if (q < n) out[q] = g.tile_mx;
    return q + 1;
}

}  // extern "C"
