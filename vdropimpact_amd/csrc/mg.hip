// mg.hip -- multigrid kernels and driver (see mg.hpp for the scheme).
#include "mg.hpp"

#include <cmath>
#include <cstring>
#include <limits>

namespace vdi {

// ------------------------------------------------------------------------
// Cell-centred operator (mgs_mac.hh:37-50, visco_impl.hh:32-53)
// ------------------------------------------------------------------------

__device__ __forceinline__ double bc_diag_factor(int bc) {
    return bc == BC_DF ? 2.0 : (bc == BC_DC ? 1.0 : 0.0);
}
__device__ __forceinline__ double bc_ghost_factor(int bc) {
    return bc == BC_DF ? -1.0 : (bc == BC_DC ? 0.0 : 1.0);
}

__device__ __forceinline__ double cc_diag(const CCOp &op, const Level &L, int i, int j) {
    double d = 0;
    d += (i > 0 ? 1.0 : bc_diag_factor(op.bcl)) * L.cx;
    d += (i < L.m - 1 ? 1.0 : bc_diag_factor(op.bcr)) * L.cx;
    d += (j > 0 ? 1.0 : bc_diag_factor(op.bcb)) * L.cy;
    d += (j < L.n - 1 ? 1.0 : bc_diag_factor(op.bct)) * L.cy;
    return op.sigma - op.s * d;
}

__device__ __forceinline__ double cc_off(const Level &L, const double *z, int i, int j) {
    const double *w = z + i + (size_t)L.m * j;
    double sx = 0, sy = 0;
    if (i > 0) sx += w[-1];
    if (i < L.m - 1) sx += w[1];
    if (j > 0) sy += w[-L.m];
    if (j < L.n - 1) sy += w[L.m];
    return L.cx * sx + L.cy * sy;
}

__device__ __forceinline__ double cc_Az(const CCOp &op, const Level &L, const double *z, int i, int j) {
    return op.s * cc_off(L, z, i, j) + cc_diag(op, L, i, j) * z[i + (size_t)L.m * j];
}

// ------------------------------------------------------------------------
// Nodal FEM operator (mgs_fem.hh:42-56, mgs_fem.cc:34-50)
// ------------------------------------------------------------------------

__device__ __forceinline__ double fem_acc(const FemOp &op, const Level &L, int i, int j) {
    return (i > 0 && i < L.m - 1 ? 1.0 : 0.5) * (j > 0 && j < L.n - 1 ? 1.0 : 0.5) * op.fm;
}

__device__ __forceinline__ double fem_off(const FemOp &op, const Level &L, const double *z, int i, int j) {
    const int M = L.m;
    const double *w = z + i + (size_t)M * j, *wm = w - M;
    if (j == 0) return 0;
    if (j == L.n - 1) {
        if (i == 0) return op.fc * wm[1] + op.hey * wm[0] + op.hex * w[1];
        if (i == M - 1) return op.fc * wm[-1] + op.hey * wm[0] + op.hex * w[-1];
        return op.fc * (wm[-1] + wm[1]) + op.fey * wm[0] + op.hex * (w[-1] + w[1]);
    }
    const double *wp = w + M;
    if (i == 0) return op.fc * (wm[1] + wp[1]) + op.hey * (wm[0] + wp[0]) + op.fex * w[1];
    if (i == M - 1) return op.fc * (wm[-1] + wp[-1]) + op.hey * (wm[0] + wp[0]) + op.fex * w[-1];
    return op.fc * (wm[-1] + wm[1] + wp[-1] + wp[1]) + op.fey * (wm[0] + wp[0]) + op.fex * (w[-1] + w[1]);
}

__device__ __forceinline__ double fem_Az(const FemOp &op, const Level &L, const double *z, int i, int j) {
    return fem_acc(op, L, i, j) * z[i + (size_t)L.m * j] + fem_off(op, L, z, i, j);
}

// ------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------

// Red-black half sweep: points with (i+j)&1 == color.
__global__ void k_cc_smooth(Level L, CCOp op, int color) {
    int ii = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (j >= L.n) return;
    int i = 2 * ii + ((j + color) & 1);
    if (i >= L.m) return;
    size_t ij = i + (size_t)L.m * j;
    L.z[ij] = (L.b[ij] - op.s * cc_off(L, L.z, i, j)) / cc_diag(op, L, i, j);
}

// Residual on the four children, averaged into the coarse RHS; clears the
// coarse solution.
__global__ void k_cc_restrict(Level F, Level C, CCOp op) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= C.m || J >= C.n) return;
    int i = 2 * I, j = 2 * J;
    size_t f = i + (size_t)F.m * j;
    double r = F.b[f] - cc_Az(op, F, F.z, i, j);
    r += F.b[f + 1] - cc_Az(op, F, F.z, i + 1, j);
    r += F.b[f + F.m] - cc_Az(op, F, F.z, i, j + 1);
    r += F.b[f + F.m + 1] - cc_Az(op, F, F.z, i + 1, j + 1);
    size_t c = I + (size_t)C.m * J;
    C.b[c] = 0.25 * r;
    C.z[c] = 0;
}

// Bilinear prolongation of the coarse correction, added to the fine z.
__global__ void k_cc_prolong(Level F, Level C, CCOp op) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= F.m || j >= F.n) return;
    int I = i >> 1, J = j >> 1;
    int I2 = (i & 1) ? I + 1 : I - 1, J2 = (j & 1) ? J + 1 : J - 1;
    double fx = 1, fy = 1;
    if (I2 < 0) { fx = bc_ghost_factor(op.bcl); I2 = I; }
    else if (I2 >= C.m) { fx = bc_ghost_factor(op.bcr); I2 = I; }
    if (J2 < 0) { fy = bc_ghost_factor(op.bcb); J2 = J; }
    else if (J2 >= C.n) { fy = bc_ghost_factor(op.bct); J2 = J; }
    const double *e = C.z;
    double e00 = e[I + (size_t)C.m * J], e10 = fx * e[I2 + (size_t)C.m * J];
    double e01 = fy * e[I + (size_t)C.m * J2], e11 = fx * fy * e[I2 + (size_t)C.m * J2];
    F.z[i + (size_t)F.m * j] += 0.5625 * e00 + 0.1875 * (e10 + e01) + 0.0625 * e11;
}

// Coarsest level: ncs red-black sweeps in one workgroup.
__global__ void k_cc_coarse(Level L, CCOp op, int ncs) {
    const int N = L.m * L.n;
    for (int k = 0; k < ncs; k++)
        for (int c = 0; c < 2; c++) {
            for (int ij = threadIdx.x; ij < N; ij += blockDim.x) {
                int i = ij % L.m, j = ij / L.m;
                if (((i + j) & 1) != c) continue;
                L.z[ij] = (L.b[ij] - op.s * cc_off(L, L.z, i, j)) / cc_diag(op, L, i, j);
            }
            __syncthreads();
        }
}

// FEM four-colour sweep for colour (ci,cj); row 0 is never updated.
__global__ void k_fem_smooth(Level L, FemOp op, int ci, int cj) {
    int ii = blockIdx.x * blockDim.x + threadIdx.x;
    int jj = blockIdx.y * blockDim.y + threadIdx.y;
    int i = 2 * ii + ci, j = 2 * jj + cj;
    if (i >= L.m || j >= L.n || j == 0) return;
    size_t ij = i + (size_t)L.m * j;
    L.z[ij] = (L.b[ij] - fem_off(op, L, L.z, i, j)) / fem_acc(op, L, i, j);
}

// Full-weighting restriction of the FEM residual; coarse row 0 gets 0.
__global__ void k_fem_restrict(Level F, Level C, FemOp op) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= C.m || J >= C.n) return;
    double v = 0;
    if (J > 0) {
        for (int b2 = -1; b2 <= 1; b2++) {
            int jf = 2 * J + b2;
            if (jf < 0 || jf >= F.n) continue;
            double wy = b2 ? 0.5 : 1.0, sx = 0;
            for (int a2 = -1; a2 <= 1; a2++) {
                int i_f = 2 * I + a2;
                if (i_f < 0 || i_f >= F.m) continue;
                double r = F.b[i_f + (size_t)F.m * jf] - fem_Az(op, F, F.z, i_f, jf);
                sx += (a2 ? 0.5 : 1.0) * r;
            }
            v += wy * sx;
        }
    }
    size_t c = I + (size_t)C.m * J;
    C.b[c] = v;
    C.z[c] = 0;
}

__global__ void k_fem_prolong(Level F, Level C) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= F.m || j >= F.n || j == 0) return;
    int I = i >> 1, J = j >> 1;
    const double *q = C.z + I + (size_t)C.m * J;
    double e;
    if (!(i & 1) && !(j & 1)) e = q[0];
    else if (!(j & 1)) e = 0.5 * (q[0] + q[1]);
    else if (!(i & 1)) e = 0.5 * (q[0] + q[C.m]);
    else e = 0.25 * (q[0] + q[1] + q[C.m] + q[C.m + 1]);
    F.z[i + (size_t)F.m * j] += e;
}

__global__ void k_fem_coarse(Level L, FemOp op, int ncs) {
    const int N = L.m * L.n;
    for (int k = 0; k < ncs; k++)
        for (int c = 0; c < 4; c++) {
            int ci = c & 1, cj = c >> 1;
            for (int ij = threadIdx.x; ij < N; ij += blockDim.x) {
                int i = ij % L.m, j = ij / L.m;
                if ((i & 1) != ci || (j & 1) != cj || j == 0) continue;
                L.z[ij] = (L.b[ij] - fem_off(op, L, L.z, i, j)) / fem_acc(op, L, i, j);
            }
            __syncthreads();
        }
}

__global__ void k_fem_dirichlet(Level L, FemOp op) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.m) return;
    L.z[i] = L.b[i] / fem_acc(op, L, i, 0);
}

// Block sum of two doubles over 256 threads (fixed order: deterministic).
__device__ __forceinline__ void block_sum2(double &a, double &b, double *sh) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { sh[2 * w] = a; sh[2 * w + 1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0; b = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) { a += sh[2 * k]; b += sh[2 * k + 1]; }
    }
}

template <bool FEM>
__global__ void __launch_bounds__(256) k_norms(Level L, CCOp cop, FemOp fop, double *partial) {
    __shared__ double sh[16];
    double rr = 0, bb = 0;
    const long N = (long)L.m * L.n;
    for (long ij = blockIdx.x * (long)blockDim.x + threadIdx.x; ij < N; ij += (long)gridDim.x * blockDim.x) {
        int i = (int)(ij % L.m), j = (int)(ij / L.m);
        double az = FEM ? fem_Az(fop, L, L.z, i, j) : cc_Az(cop, L, L.z, i, j);
        double bv = L.b[ij], r = bv - az;
        rr += r * r;
        bb += bv * bv;
    }
    block_sum2(rr, bb, sh);
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = rr; partial[2 * blockIdx.x + 1] = bb; }
}

__global__ void __launch_bounds__(256) k_sum_partials(const double *partial, int nblk, double *out) {
    __shared__ double sh[16];
    double a = 0, b = 0;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) { a += partial[2 * k]; b += partial[2 * k + 1]; }
    block_sum2(a, b, sh);
    if (threadIdx.x == 0) { out[0] = a; out[1] = b; }
}

template <bool FEM>
__global__ void k_apply(Level L, CCOp cop, FemOp fop, double *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= L.m || j >= L.n) return;
    out[i + (size_t)L.m * j] = FEM ? fem_Az(fop, L, L.z, i, j) : cc_Az(cop, L, L.z, i, j);
}

// mgs_common.cc:37-44: z <- b z - c z_prev, z_prev <- old z.
__global__ void k_extrapolate(double *z, double *zp, long N, double bb, double c) {
    long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (ij >= N) return;
    double zz = z[ij];
    z[ij] = bb * zz - c * zp[ij];
    zp[ij] = zz;
}

// ------------------------------------------------------------------------
// Driver
// ------------------------------------------------------------------------

static const dim3 kBlk2(64, 4);

int MG::init(int kind_, int m, int n, double cx, double cy, int x_sym, double *z0, double *b0,
             hipStream_t stream) {
    kind = kind_;
    fem = kind == VDI_FEM;
    st = stream;
    if (m < 2 || n < 2) return VDI_ERR_ARG;
    if (kind == VDI_MAC) {
        cc = CCOp{0.0, 1.0, BC_N, BC_N, BC_DF, BC_N};
    } else if (kind == VDI_VX || kind == VDI_VY) {
        int bl = x_sym ? (kind == VDI_VX ? BC_DF : BC_N) : BC_DC;
        cc = CCOp{1.0, -1.0, bl, BC_DC, BC_N, BC_DC};
        vfx = cx; vfy = cy;
    } else if (kind == VDI_FEM) {
        double dxdy = cx / cy, dydx = cy / cx;
        fe.fm = 4. / 3. * (dxdy + dydx);
        fe.fey = 1. / 3. * (-2 * dxdy + dydx); fe.hey = 0.5 * fe.fey;
        fe.fex = 1. / 3. * (-2 * dydx + dxdy); fe.hex = 0.5 * fe.fex;
        fe.fc = -1. / 6. * (dxdy + dydx);
    } else {
        return VDI_ERR_ARG;
    }
    // level sizes: halve while cell counts are even and stay >= 2
    lm.clear(); ln.clear();
    int mc = fem ? m - 1 : m, nc = fem ? n - 1 : n;
    for (;;) {
        lm.push_back(fem ? mc + 1 : mc);
        ln.push_back(fem ? nc + 1 : nc);
        if ((int)lm.size() >= 24) break;
        if (mc % 2 || nc % 2 || mc / 2 < 2 || nc / 2 < 2) break;
        mc /= 2; nc /= 2;
    }
    nl = (int)lm.size();
    size_t coarse = 0;
    for (int l = 1; l < nl; l++) coarse += (size_t)lm[l] * ln[l];
    VDI_TRY(store_.alloc(2 * coarse * sizeof(double)));
    long N0 = (long)m * n;
    if (!z0) { VDI_TRY(own_z_.alloc(N0 * sizeof(double))); z0 = own_z_.as<double>(); }
    if (!b0) { VDI_TRY(own_b_.alloc(N0 * sizeof(double))); b0 = own_b_.as<double>(); }
    VDI_TRY(zprev_.alloc(N0 * sizeof(double)));
    L.assign(nl, Level{});
    double *p = store_.as<double>();
    for (int l = 0; l < nl; l++) {
        L[l].m = lm[l];
        L[l].n = ln[l];
        if (l == 0) { L[l].z = z0; L[l].b = b0; }
        else {
            size_t Nl = (size_t)lm[l] * ln[l];
            L[l].z = p; p += Nl;
            L[l].b = p; p += Nl;
        }
    }
    set_coeffs(cx, cy);
    nblk_norm_ = (int)std::min<long>(1024, (N0 + 255) / 256);
    VDI_TRY(partial_.alloc(2 * nblk_norm_ * sizeof(double)));
    VDI_TRY(norms_.alloc(2 * sizeof(double)));
    if (!host_norms_) VDI_HIP(hipHostMalloc((void **)&host_norms_, 2 * sizeof(double), hipHostMallocDefault));
    return VDI_OK;
}

MG::~MG() {
    if (host_norms_) (void)hipHostFree(host_norms_);
}

void MG::set_coeffs(double cx, double cy) {
    for (int l = 0; l < nl; l++) {
        L[l].cx = cx;
        L[l].cy = cy;
        cx *= 0.25;
        cy *= 0.25;
    }
}

int MG::sweep(int l, int count) {
    const Level &Lv = L[l];
    for (int k = 0; k < count; k++) {
        if (!fem) {
            dim3 g(cdiv((Lv.m + 1) / 2, kBlk2.x), cdiv(Lv.n, kBlk2.y));
            for (int c = 0; c < 2; c++) {
                hipLaunchKernelGGL(k_cc_smooth, g, kBlk2, 0, st, Lv, cc, c);
                VDI_LAUNCHED();
            }
        } else {
            dim3 g(cdiv((Lv.m + 1) / 2, kBlk2.x), cdiv((Lv.n + 1) / 2, kBlk2.y));
            for (int c = 0; c < 4; c++) {
                hipLaunchKernelGGL(k_fem_smooth, g, kBlk2, 0, st, Lv, fe, c & 1, c >> 1);
                VDI_LAUNCHED();
            }
        }
    }
    return VDI_OK;
}

int MG::vcycle_level(int l) {
    const Level &F = L[l];
    if (l == nl - 1) {
        if (!fem) hipLaunchKernelGGL(k_cc_coarse, dim3(1), dim3(256), 0, st, F, cc, ncs);
        else hipLaunchKernelGGL(k_fem_coarse, dim3(1), dim3(256), 0, st, F, fe, ncs);
        VDI_LAUNCHED();
        return VDI_OK;
    }
    const Level &C = L[l + 1];
    VDI_TRY(sweep(l, nu1));
    dim3 gc(cdiv(C.m, kBlk2.x), cdiv(C.n, kBlk2.y));
    if (!fem) hipLaunchKernelGGL(k_cc_restrict, gc, kBlk2, 0, st, F, C, cc);
    else hipLaunchKernelGGL(k_fem_restrict, gc, kBlk2, 0, st, F, C, fe);
    VDI_LAUNCHED();
    VDI_TRY(vcycle_level(l + 1));
    dim3 gf(cdiv(F.m, kBlk2.x), cdiv(F.n, kBlk2.y));
    if (!fem) hipLaunchKernelGGL(k_cc_prolong, gf, kBlk2, 0, st, F, C, cc);
    else hipLaunchKernelGGL(k_fem_prolong, gf, kBlk2, 0, st, F, C);
    VDI_LAUNCHED();
    VDI_TRY(sweep(l, nu2));
    return VDI_OK;
}

int MG::vcycle() { return vcycle_level(0); }

int MG::enforce_dirichlet() {
    if (!fem) return VDI_OK;
    hipLaunchKernelGGL(k_fem_dirichlet, dim3(cdiv(L[0].m, 256)), dim3(256), 0, st, L[0], fe);
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::enqueue_norms() {
    if (fem) hipLaunchKernelGGL(k_norms<true>, dim3(nblk_norm_), dim3(256), 0, st, L[0], cc, fe, partial_.as<double>());
    else hipLaunchKernelGGL(k_norms<false>, dim3(nblk_norm_), dim3(256), 0, st, L[0], cc, fe, partial_.as<double>());
    VDI_LAUNCHED();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial_.as<double>(), nblk_norm_, norms_.as<double>());
    VDI_LAUNCHED();
    return VDI_OK;
}

int MG::residual(double *rnorm, double *bnorm) {
    VDI_TRY(enqueue_norms());
    VDI_HIP(hipMemcpyAsync(host_norms_, norms_.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    VDI_HIP(hipStreamSynchronize(st));
    *rnorm = std::sqrt(host_norms_[0]);
    *bnorm = std::sqrt(host_norms_[1]);
    return VDI_OK;
}

int MG::vcycles(int k) {
    VDI_TRY(enforce_dirichlet());
    for (int c = 0; c < k; c++) VDI_TRY(vcycle());
    return VDI_OK;
}

int MG::solve(double tol, int max_cycles, int *cycles) {
    VDI_TRY(enforce_dirichlet());
    double rn, bn;
    int c = 0;
    for (;; c++) {
        VDI_TRY(residual(&rn, &bn));
        if (bn == 0) {
            VDI_HIP(hipMemsetAsync(L[0].z, 0, sizeof(double) * unknowns(), st));
            break;
        }
        if (!(rn == rn) || c > max_cycles) {
            last_cycles = c;
            if (cycles) *cycles = c;
            return VDI_ERR_MG_NOCONV;
        }
        if (rn <= tol * bn) break;
        if (c == max_cycles) {
            last_cycles = c;
            if (cycles) *cycles = c;
            return VDI_ERR_MG_NOCONV;
        }
        VDI_TRY(vcycle());
    }
    last_cycles = c;
    total_cycles += c;
    solves++;
    if (cycles) *cycles = c;
    return VDI_OK;
}

int MG::initial_guess(double t) {
    long N = unknowns();
    if (scount < 2) {
        if (scount == 1) {
            VDI_HIP(hipMemcpyAsync(zprev_.p, L[0].z, N * sizeof(double), hipMemcpyDeviceToDevice, st));
            t0 = t;
        } else {
            t_prev = t;
        }
        scount++;
        return VDI_OK;
    }
    if (std::fabs(t0 - t_prev) < std::numeric_limits<double>::epsilon() * 10) {
        fputs("Multigrid routine called for duplicate times\n", stderr);
        return VDI_OK;
    }
    double a = 1. / (t0 - t_prev), bb = (t - t_prev) * a, c = (t - t0) * a;
    hipLaunchKernelGGL(k_extrapolate, dim3(cdiv(N, 256)), dim3(256), 0, st, L[0].z, zprev_.as<double>(), N, bb, c);
    VDI_LAUNCHED();
    t_prev = t0;
    t0 = t;
    return VDI_OK;
}

int MG::apply(const double *z, double *out) {
    Level Lz = L[0];
    Lz.z = const_cast<double *>(z);
    dim3 g(cdiv(Lz.m, kBlk2.x), cdiv(Lz.n, kBlk2.y));
    if (fem) hipLaunchKernelGGL(k_apply<true>, g, kBlk2, 0, st, Lz, cc, fe, out);
    else hipLaunchKernelGGL(k_apply<false>, g, kBlk2, 0, st, Lz, cc, fe, out);
    VDI_LAUNCHED();
    return VDI_OK;
}

}  // namespace vdi
