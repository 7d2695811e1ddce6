// nn.hip -- the U-Net of nn.hpp: BF16 MFMA convolutions, the backward
// pass, Adam, and the learned initial guess of the MAC solve.
//
// Convolution (3x3, zero padding) as an implicit GEMM on
// v_mfma_f32_16x16x16_bf16: a workgroup owns an 8 x 8 pixel tile and stages
// the 10 x 10 input halo (all input channels) in LDS; each of its four waves
// computes 16 pixels x all output channels, one 16 x 16 FP32 accumulator
// block per 16 output channels, stepping K over (tap, 16 input channels).
// Operand layout of the 16x16x16 instruction (lane i, element j):
//   A[i % 16][4 (i / 16) + j],  B[4 (i / 16) + j][i % 16],  D[4 (i / 16) + j][i % 16].
// The backward pass for the inputs is the same kernel with the weights
// flipped in space and transposed in channels; the weight gradient is a
// second MFMA kernel (K = pixels) writing per-split partials that a third
// kernel sums in a fixed order.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "step.hpp"
#include "nn.hpp"

namespace vdi {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(((unsigned)h) << 16); }

// ------------------------------------------------------------------------
// Convolution (forward, and backward for the inputs)
// ------------------------------------------------------------------------

struct ConvArgs {
    const bf16_t *x;
    int xs, xo;        // input channel stride / first channel
    const bf16_t *w;   // [Cout][9][Cin]
    const float *bias; // [Cout] or null
    bf16_t *y;         // BF16 output (or null)
    int ys, yo;
    float *yf;         // FP32 output (or null), += when acc
    int yfs, yfo;
    int H, W, Cin, Cout, act, acc;
};

template <int NOB>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
    extern __shared__ bf16_t xl[];  // 10 x 10 x Cin
    const int tilesx = a.W >> 3;
    const int x0 = (blockIdx.x % tilesx) * 8, y0 = (blockIdx.x / tilesx) * 8;
    const int tid = threadIdx.x;
    const int c4n = a.Cin >> 2;
    for (int idx = tid; idx < 100 * c4n; idx += 256) {
        const int q = idx / c4n, c4 = idx - q * c4n;
        const int hy = q / 10, hx = q - hy * 10;
        const int gy = y0 - 1 + hy, gx = x0 - 1 + hx;
        uint2 v = make_uint2(0u, 0u);
        if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
            v = *reinterpret_cast<const uint2 *>(a.x + ((size_t)gy * a.W + gx) * a.xs + a.xo + 4 * c4);
        *reinterpret_cast<uint2 *>(xl + q * a.Cin + 4 * c4) = v;
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63, r = lane & 15, kq = lane >> 4;
    const int py = 2 * wave + (r >> 3), px = r & 7;
    f32x4 acc[NOB];
#pragma unroll
    for (int ob = 0; ob < NOB; ob++) acc[ob] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int cbn = a.Cin >> 4;
    const size_t wob = (size_t)16 * 9 * a.Cin;
    for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3, dx = tap - 3 * dy;
        const bf16_t *arow = xl + ((py + dy) * 10 + px + dx) * a.Cin + 4 * kq;
        const bf16_t *wrow = a.w + ((size_t)r * 9 + tap) * a.Cin + 4 * kq;
        for (int cb = 0; cb < cbn; cb++) {
            const bf16x4 A = *reinterpret_cast<const bf16x4 *>(arow + cb * 16);
#pragma unroll
            for (int ob = 0; ob < NOB; ob++) {
                const bf16x4 B = *reinterpret_cast<const bf16x4 *>(wrow + ob * wob + cb * 16);
                acc[ob] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, acc[ob], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int ob = 0; ob < NOB; ob++) {
        const int co = ob * 16 + r;
        const float bv = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int p = 4 * kq + j;
            const int gy = y0 + 2 * wave + (p >> 3), gx = x0 + (p & 7);
            float v = acc[ob][j] + bv;
            if (a.act) v = v > 0.f ? v : 0.1f * v;
            const size_t pp = (size_t)gy * a.W + gx;
            if (a.y) a.y[pp * a.ys + a.yo + co] = f2bf(v);
            if (a.yf) {
                float *d = a.yf + pp * a.yfs + a.yfo + co;
                *d = a.acc ? *d + v : v;
            }
        }
    }
}

// ------------------------------------------------------------------------
// Weight gradient: dW[co][tap][ci] = sum_p D[p][co] X[p + off(tap)][ci],
// db[co] = sum_p D[p][co].  Workgroup = (64 output channels) x (64 input
// channels) x (a range of 8 x 8 pixel chunks); wave w owns output block w.
// ------------------------------------------------------------------------

struct WgArgs {
    const bf16_t *d;   // [P][Cout] (gradient at the convolution output)
    const bf16_t *x;   // input slice
    int xs, xo;
    int H, W, Cin, Cout;
    int gco, gci, nsplit;
    float *part;       // [nsplit][Cout*9*Cin + Cout]
};

__global__ void __launch_bounds__(256) k_wgrad(WgArgs a) {
    __shared__ bf16_t dl[64 * 64];
    __shared__ bf16_t xl[100 * 64];
    const int ng = a.gco * a.gci;
    const int g = blockIdx.x % ng, s = blockIdx.x / ng;
    const int co_g = g % a.gco, ci_g = g / a.gco;
    const int co_base = co_g * 64, ci_base = ci_g * 64;
    const int nco = min(4, (a.Cout - co_base) >> 4), nci = min(4, (a.Cin - ci_base) >> 4);
    const int tilesx = a.W >> 3, total = tilesx * (a.H >> 3);
    const int c_a = (int)((long)s * total / a.nsplit), c_b = (int)((long)(s + 1) * total / a.nsplit);
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, r = lane & 15, kq = lane >> 4;
    f32x4 acc[4][9];
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int t = 0; t < 9; t++) acc[cb][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (int ch = c_a; ch < c_b; ch++) {
        const int y0 = (ch / tilesx) * 8, x0 = (ch % tilesx) * 8;
        for (int idx = tid; idx < 64 * 16; idx += 256) {
            const int p = idx >> 4, c4 = idx & 15, co = co_base + 4 * c4;
            uint2 v = make_uint2(0u, 0u);
            if (co < a.Cout)
                v = *reinterpret_cast<const uint2 *>(a.d + ((size_t)(y0 + (p >> 3)) * a.W + x0 + (p & 7)) * a.Cout + co);
            *reinterpret_cast<uint2 *>(dl + p * 64 + 4 * c4) = v;
        }
        for (int idx = tid; idx < 100 * 16; idx += 256) {
            const int q = idx >> 4, c4 = idx & 15, ci = ci_base + 4 * c4;
            const int hy = q / 10, hx = q - hy * 10, gy = y0 - 1 + hy, gx = x0 - 1 + hx;
            uint2 v = make_uint2(0u, 0u);
            if (ci < a.Cin && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
                v = *reinterpret_cast<const uint2 *>(a.x + ((size_t)gy * a.W + gx) * a.xs + a.xo + ci);
            *reinterpret_cast<uint2 *>(xl + q * 64 + 4 * c4) = v;
        }
        __syncthreads();
        if (ci_g == 0 && tid < 64 && co_base + tid < a.Cout)
            for (int p = 0; p < 64; p++) bsum += bf2f(dl[p * 64 + tid]);
        if (w < nco) {
            for (int ks = 0; ks < 4; ks++) {
                bf16x4 A;
#pragma unroll
                for (int j = 0; j < 4; j++) A[j] = (short)dl[(ks * 16 + 4 * kq + j) * 64 + w * 16 + r];
#pragma unroll
                for (int t = 0; t < 9; t++) {
                    const int dy = t / 3, dx = t - 3 * dy;
#pragma unroll
                    for (int cb = 0; cb < 4; cb++) {
                        if (cb >= nci) break;
                        bf16x4 B;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const int k = ks * 16 + 4 * kq + j;
                            B[j] = (short)xl[(((k >> 3) + dy) * 10 + (k & 7) + dx) * 64 + cb * 16 + r];
                        }
                        acc[cb][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, acc[cb][t], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
    }
    const size_t L = (size_t)a.Cout * 9 * a.Cin + a.Cout;
    float *P = a.part + (size_t)s * L;
    if (w < nco) {
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            if (cb >= nci) break;
#pragma unroll
            for (int t = 0; t < 9; t++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int co = co_base + w * 16 + 4 * kq + j, ci = ci_base + cb * 16 + r;
                    P[((size_t)co * 9 + t) * a.Cin + ci] = acc[cb][t][j];
                }
        }
    }
    if (ci_g == 0 && tid < 64 && co_base + tid < a.Cout) P[(size_t)a.Cout * 9 * a.Cin + co_base + tid] = bsum;
}

// grad[q] += sum_s part[s][q], s in order.
__global__ void k_wreduce(const float *part, long L, int nsplit, float *grad) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= L) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s += part[(size_t)k * L + q];
    grad[q] += s;
}

// Wt[ci][tap][co] = W[co][8 - tap][ci]
__global__ void k_wflip(const bf16_t *w, bf16_t *wt, int Cin, int Cout) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= (long)Cin * 9 * Cout) return;
    const int co = (int)(q % Cout), t = (int)((q / Cout) % 9), ci = (int)(q / ((long)Cout * 9));
    wt[q] = w[((size_t)co * 9 + (8 - t)) * Cin + ci];
}

// ------------------------------------------------------------------------
// Pooling, upsampling, activation gradient, loss, Adam
// ------------------------------------------------------------------------

// 2x2 average: out (level l+1, Ho x Wo) from in (level l)
__global__ void k_pool(const bf16_t *x, int xs, int xo, bf16_t *y, int ys, int yo, int C, int Ho, int Wo) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= (long)Ho * Wo * C) return;
    const int c = (int)(q % C), p = (int)(q / C), yy = p / Wo, xx = p - yy * Wo, Wi = 2 * Wo;
    const size_t a = (size_t)(2 * yy) * Wi + 2 * xx;
    const float s = bf2f(x[a * xs + xo + c]) + bf2f(x[(a + 1) * xs + xo + c]) + bf2f(x[(a + Wi) * xs + xo + c]) +
                    bf2f(x[(a + Wi + 1) * xs + xo + c]);
    y[(size_t)p * ys + yo + c] = f2bf(0.25f * s);
}
__global__ void k_pool_bwd(const float *gy, int gys, int gyo, float *gx, int gxs, int gxo, int C, int Ho, int Wo) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= (long)4 * Ho * Wo * C) return;
    const int c = (int)(q % C), p = (int)(q / C), Wi = 2 * Wo, yy = p / Wi, xx = p - yy * Wi;
    const size_t par = (size_t)(yy >> 1) * Wo + (xx >> 1);
    gx[(size_t)p * gxs + gxo + c] += 0.25f * gy[par * gys + gyo + c];
}
// nearest x2: out (level l, Ho x Wo) from in (level l+1)
__global__ void k_up(const bf16_t *x, int xs, int xo, bf16_t *y, int ys, int yo, int C, int Ho, int Wo) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= (long)Ho * Wo * C) return;
    const int c = (int)(q % C), p = (int)(q / C), yy = p / Wo, xx = p - yy * Wo;
    y[(size_t)p * ys + yo + c] = x[((size_t)(yy >> 1) * (Wo >> 1) + (xx >> 1)) * xs + xo + c];
}
__global__ void k_up_bwd(const float *gy, int gys, int gyo, float *gx, int gxs, int gxo, int C, int Hi, int Wi) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= (long)Hi * Wi * C) return;
    const int c = (int)(q % C), p = (int)(q / C), yy = p / Wi, xx = p - yy * Wi, Wo = 2 * Wi;
    const size_t a = (size_t)(2 * yy) * Wo + 2 * xx;
    const float s = gy[a * gys + gyo + c] + gy[(a + 1) * gys + gyo + c] + gy[(a + Wo) * gys + gyo + c] +
                    gy[(a + Wo + 1) * gys + gyo + c];
    gx[(size_t)p * gxs + gxo + c] += s;
}
// D[p][c] = bf16(G[p][c] * act'(y))
__global__ void k_actgrad(const float *g, int gs, int go, const bf16_t *y, int ys, int yo, int C, long P, int act,
                          bf16_t *d) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= P * C) return;
    const int c = (int)(q % C);
    const long p = q / C;
    float v = g[p * gs + go + c];
    if (act && !(bf2f(y[p * ys + yo + c]) > 0.f)) v *= 0.1f;
    d[q] = f2bf(v);
}
// loss = mean (y0 - t)^2: G_y[p][0] = 2 (y0 - t) / P, other channels 0;
// block partials of the squared error.
__global__ void k_loss(const float *yf, const float *t, float *gy, long P, float *part) {
    __shared__ float sh[4];
    const long p = blockIdx.x * 256L + threadIdx.x;
    float e2 = 0.f;
    if (p < P) {
        const float e = yf[p * 16] - t[p];
        e2 = e * e;
        gy[p * 16] = 2.f * e / (float)P;
        for (int c = 1; c < 16; c++) gy[p * 16 + c] = 0.f;
    }
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ void k_adam(float *w, float *g, float *m, float *v, bf16_t *w16, long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q >= n) return;
    const float gg = g[q];
    const float mm = b1 * m[q] + (1.f - b1) * gg, vv = b2 * v[q] + (1.f - b2) * gg * gg;
    m[q] = mm;
    v[q] = vv;
    const float wn = w[q] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    w[q] = wn;
    w16[q] = f2bf(wn);
    g[q] = 0.f;
}
__global__ void k_to_bf16(const float *w, bf16_t *w16, long n) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q < n) w16[q] = f2bf(w[q]);
}
__global__ void k_zero_f(float *p, long n) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q < n) p[q] = 0.f;
}
// x0[p] = (in[0][p], ..., in[kNNInputs-1][p], 0, ..., 0) in BF16 (planar inputs)
__global__ void k_expand(const float *in, bf16_t *x0, long P) {
    const long p = blockIdx.x * 256L + threadIdx.x;
    if (p >= P) return;
    for (int c = 0; c < 16; c++) x0[p * 16 + c] = c < kNNInputs ? f2bf(in[c * P + p]) : (bf16_t)0;
}

static unsigned g1(long n) { return (unsigned)((n + 255) / 256); }

// ------------------------------------------------------------------------
// UNet host side
// ------------------------------------------------------------------------

static unsigned long long splitmix(unsigned long long &s) {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int UNet::init(int H_, int W_, hipStream_t stream, unsigned long long seed) {
    // the deepest level (1/8 resolution) must hold whole 8 x 8 tiles
    if (H_ % 64 || W_ % 64 || H_ < 64 || W_ < 64) return VDI_ERR_ARG;
    H = H_;
    W = W_;
    st = stream;
    static bool attr = false;
    if (!attr) {
        const void *ks[] = {(const void *)k_conv<1>, (const void *)k_conv<2>, (const void *)k_conv<3>,
                            (const void *)k_conv<6>, (const void *)k_conv<8>, (const void *)k_conv<12>,
                            (const void *)k_conv<18>};
        for (const void *k : ks)
            VDI_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 100 * 288 * (int)sizeof(bf16_t)));
        attr = true;
    }
    const int C[NBUF] = {16, 16, 48, 16, 32, 128, 32, 96, 288, 96, 192, 192, 96, 96, 32, 32, 16, 16, 16};
    const int Lv[NBUF] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    for (int b = 0; b < NBUF; b++) { bufC[b] = C[b]; bufL[b] = Lv[b]; }
    ops.clear();
    long wo = 0;
    auto conv = [&](Slice in, Slice out, int level, int act) {
        Op o{CONV, in, out, level, act, wo};
        wo += (long)out.C * 9 * in.C + out.C;
        ops.push_back(o);
    };
    auto pool = [&](Slice in, Slice out, int level) { ops.push_back(Op{POOL, in, out, level, 0, 0}); };
    auto up = [&](Slice in, Slice out, int level) { ops.push_back(Op{UP, in, out, level, 0, 0}); };
    conv({X0, 0, 16}, {E1A, 0, 16}, 0, 1);
    conv({E1A, 0, 16}, {CAT1, 32, 16}, 0, 1);
    pool({CAT1, 32, 16}, {P1, 0, 16}, 0);
    conv({P1, 0, 16}, {E2A, 0, 32}, 1, 1);
    conv({E2A, 0, 32}, {CAT2, 96, 32}, 1, 1);
    pool({CAT2, 96, 32}, {P2, 0, 32}, 1);
    conv({P2, 0, 32}, {E3A, 0, 96}, 2, 1);
    conv({E3A, 0, 96}, {CAT3, 192, 96}, 2, 1);
    pool({CAT3, 192, 96}, {P3, 0, 96}, 2);
    conv({P3, 0, 96}, {B1, 0, 192}, 3, 1);
    conv({B1, 0, 192}, {B2, 0, 192}, 3, 1);
    up({B2, 0, 192}, {CAT3, 0, 192}, 2);
    conv({CAT3, 0, 288}, {D3A, 0, 96}, 2, 1);
    conv({D3A, 0, 96}, {D3B, 0, 96}, 2, 1);
    up({D3B, 0, 96}, {CAT2, 0, 96}, 1);
    conv({CAT2, 0, 128}, {D2A, 0, 32}, 1, 1);
    conv({D2A, 0, 32}, {D2B, 0, 32}, 1, 1);
    up({D2B, 0, 32}, {CAT1, 0, 32}, 0);
    conv({CAT1, 0, 48}, {D1A, 0, 16}, 0, 1);
    conv({D1A, 0, 16}, {D1B, 0, 16}, 0, 1);
    conv({D1B, 0, 16}, {Y, 0, 16}, 0, 0);
    nparams = wo;
    // activations (BF16) and their gradients (FP32)
    size_t na = 0;
    std::vector<size_t> off(NBUF);
    for (int b = 0; b < NBUF; b++) { off[b] = na; na += pixels(bufL[b]) * bufC[b]; }
    VDI_TRY(acts_.alloc(na * sizeof(bf16_t)));
    VDI_TRY(grads_.alloc(na * sizeof(float)));
    for (int b = 0; b < NBUF; b++) {
        act_[b] = acts_.as<bf16_t>() + off[b];
        gact_[b] = grads_.as<float>() + off[b];
    }
    VDI_TRY(yf_.alloc(pixels(0) * 16 * sizeof(float)));
    VDI_TRY(w32_.alloc(nparams * sizeof(float)));
    VDI_TRY(w16_.alloc(nparams * sizeof(bf16_t)));
    VDI_TRY(g32_.alloc(nparams * sizeof(float)));
    VDI_TRY(m_.alloc(nparams * sizeof(float)));
    VDI_TRY(v_.alloc(nparams * sizeof(float)));
    long maxw = 0;
    size_t maxd = 0;
    for (const Op &o : ops)
        if (o.kind == CONV) {
            maxw = std::max(maxw, (long)o.out.C * 9 * o.in.C);
            maxd = std::max(maxd, pixels(o.level) * o.out.C);
        }
    VDI_TRY(wflip_.alloc(maxw * sizeof(bf16_t)));
    VDI_TRY(dbuf_.alloc(maxd * sizeof(bf16_t)));
    VDI_TRY(lossbuf_.alloc((pixels(0) / 256 + 2) * sizeof(float)));
    // He initialisation (uniform with variance 2 / fan_in), biases zero
    std::vector<float> w(nparams, 0.f);
    unsigned long long s = seed;
    for (const Op &o : ops) {
        if (o.kind != CONV) continue;
        const long nw = (long)o.out.C * 9 * o.in.C;
        const float lim = std::sqrt(6.f / (9.f * o.in.C)) * (o.act ? 1.f : 0.1f);
        for (long k = 0; k < nw; k++) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            w[o.woff + k] = (float)((2 * u - 1) * lim);
        }
    }
    VDI_HIP(hipMemcpy(w32_.p, w.data(), nparams * sizeof(float), hipMemcpyHostToDevice));
    VDI_TRY(sync_bf16());
    adam_t_ = 0;
    return VDI_OK;
}

int UNet::sync_bf16() {
    hipLaunchKernelGGL(k_to_bf16, dim3(g1(nparams)), dim3(256), 0, st, w32_.as<float>(), w16_.as<bf16_t>(), nparams);
    VDI_LAUNCHED();
    return VDI_OK;
}

int UNet::zero_grads() {
    hipLaunchKernelGGL(k_zero_f, dim3(g1(nparams)), dim3(256), 0, st, g32_.as<float>(), nparams);
    VDI_LAUNCHED();
    return VDI_OK;
}

static int conv_launch(const ConvArgs &a, hipStream_t st) {
    const int nob = a.Cout / 16;
    const size_t lds = (size_t)100 * a.Cin * sizeof(bf16_t);
    const dim3 g((a.W / 8) * (a.H / 8));
    switch (nob) {
    case 1: hipLaunchKernelGGL(k_conv<1>, g, dim3(256), lds, st, a); break;
    case 2: hipLaunchKernelGGL(k_conv<2>, g, dim3(256), lds, st, a); break;
    case 3: hipLaunchKernelGGL(k_conv<3>, g, dim3(256), lds, st, a); break;
    case 6: hipLaunchKernelGGL(k_conv<6>, g, dim3(256), lds, st, a); break;
    case 8: hipLaunchKernelGGL(k_conv<8>, g, dim3(256), lds, st, a); break;
    case 12: hipLaunchKernelGGL(k_conv<12>, g, dim3(256), lds, st, a); break;
    case 18: hipLaunchKernelGGL(k_conv<18>, g, dim3(256), lds, st, a); break;
    default: return VDI_ERR_ARG;
    }
    VDI_LAUNCHED();
    return VDI_OK;
}

int UNet::conv_fwd(const Op &op, const bf16_t *w, const bf16_t *x, int xs, int xo, int Cin, int Cout, bf16_t *y,
                   int ys, int yo, float *yf, int yfs, int yfo, int act, int acc, const float *bias) {
    ConvArgs a{};
    a.x = x; a.xs = xs; a.xo = xo;
    a.w = w;
    a.bias = bias;
    a.y = y; a.ys = ys; a.yo = yo;
    a.yf = yf; a.yfs = yfs; a.yfo = yfo;
    a.H = H >> op.level;
    a.W = W >> op.level;
    a.Cin = Cin;
    a.Cout = Cout;
    a.act = act;
    a.acc = acc;
    return conv_launch(a, st);
}

int UNet::forward() {
    const bf16_t *w16 = w16_.as<bf16_t>();
    const float *w32 = w32_.as<float>();
    for (const Op &o : ops) {
        const int H2 = H >> o.level, W2 = W >> o.level;
        if (o.kind == CONV) {
            const bool head = o.out.buf == Y;
            VDI_TRY(conv_fwd(o, w16 + o.woff, act_[o.in.buf], bufC[o.in.buf], o.in.off, o.in.C, o.out.C,
                             head ? nullptr : act_[o.out.buf], bufC[o.out.buf], o.out.off, head ? y() : nullptr, 16, 0,
                             o.act, 0, w32 + o.woff + (long)o.out.C * 9 * o.in.C));
        } else if (o.kind == POOL) {
            const int Ho = H2 / 2, Wo = W2 / 2;
            hipLaunchKernelGGL(k_pool, dim3(g1((long)Ho * Wo * o.in.C)), dim3(256), 0, st, act_[o.in.buf],
                               bufC[o.in.buf], o.in.off, act_[o.out.buf], bufC[o.out.buf], o.out.off, o.in.C, Ho, Wo);
            VDI_LAUNCHED();
        } else {
            hipLaunchKernelGGL(k_up, dim3(g1((long)H2 * W2 * o.in.C)), dim3(256), 0, st, act_[o.in.buf],
                               bufC[o.in.buf], o.in.off, act_[o.out.buf], bufC[o.out.buf], o.out.off, o.in.C, H2, W2);
            VDI_LAUNCHED();
        }
    }
    return VDI_OK;
}

int UNet::backward(const float *target) {
    const long P0 = (long)pixels(0);
    // gradients of every activation start at zero
    size_t na = grads_.bytes / sizeof(float);
    hipLaunchKernelGGL(k_zero_f, dim3(g1((long)na)), dim3(256), 0, st, grads_.as<float>(), (long)na);
    VDI_LAUNCHED();
    hipLaunchKernelGGL(k_loss, dim3(g1(P0)), dim3(256), 0, st, (const float *)y(), target, gact_[Y], P0,
                       lossbuf_.as<float>());
    VDI_LAUNCHED();
    const bf16_t *w16 = w16_.as<bf16_t>();
    bf16_t *D = dbuf_.as<bf16_t>();
    for (int k = (int)ops.size() - 1; k >= 0; k--) {
        const Op &o = ops[k];
        const int H2 = H >> o.level, W2 = W >> o.level;
        const long P = (long)H2 * W2;
        if (o.kind == CONV) {
            const int Cin = o.in.C, Cout = o.out.C;
            // D = G_out * act'(out)
            hipLaunchKernelGGL(k_actgrad, dim3(g1(P * Cout)), dim3(256), 0, st, (const float *)gact_[o.out.buf],
                               bufC[o.out.buf], o.out.off, (const bf16_t *)act_[o.out.buf], bufC[o.out.buf],
                               o.out.off, Cout, P, o.act, D);
            VDI_LAUNCHED();
            // weight and bias gradient
            WgArgs wa{};
            wa.d = D;
            wa.x = act_[o.in.buf];
            wa.xs = bufC[o.in.buf];
            wa.xo = o.in.off;
            wa.H = H2;
            wa.W = W2;
            wa.Cin = Cin;
            wa.Cout = Cout;
            wa.gco = (Cout + 63) / 64;
            wa.gci = (Cin + 63) / 64;
            const int chunks = (H2 / 8) * (W2 / 8);
            wa.nsplit = std::max(1, std::min(chunks, 512 / (wa.gco * wa.gci)));
            const size_t L = (size_t)Cout * 9 * Cin + Cout;
            if (part_cap_ < wa.nsplit * L) {
                VDI_HIP(hipStreamSynchronize(st));
                VDI_TRY(part_.alloc(wa.nsplit * L * sizeof(float)));
                part_cap_ = wa.nsplit * L;
            }
            wa.part = part_.as<float>();
            hipLaunchKernelGGL(k_wgrad, dim3(wa.gco * wa.gci * wa.nsplit), dim3(256), 0, st, wa);
            VDI_LAUNCHED();
            hipLaunchKernelGGL(k_wreduce, dim3(g1((long)L)), dim3(256), 0, st, (const float *)part_.as<float>(),
                               (long)L, wa.nsplit, g32_.as<float>() + o.woff);
            VDI_LAUNCHED();
            // gradient of the input (not needed for the network input)
            if (o.in.buf != X0) {
                hipLaunchKernelGGL(k_wflip, dim3(g1((long)Cin * 9 * Cout)), dim3(256), 0, st, w16 + o.woff,
                                   wflip_.as<bf16_t>(), Cin, Cout);
                VDI_LAUNCHED();
                ConvArgs a{};
                a.x = D; a.xs = Cout; a.xo = 0;
                a.w = wflip_.as<bf16_t>();
                a.bias = nullptr;
                a.y = nullptr;
                a.yf = gact_[o.in.buf]; a.yfs = bufC[o.in.buf]; a.yfo = o.in.off;
                a.H = H2; a.W = W2; a.Cin = Cout; a.Cout = Cin; a.act = 0; a.acc = 1;
                VDI_TRY(conv_launch(a, st));
            }
        } else if (o.kind == POOL) {
            const int Ho = H2 / 2, Wo = W2 / 2;
            hipLaunchKernelGGL(k_pool_bwd, dim3(g1((long)4 * Ho * Wo * o.in.C)), dim3(256), 0, st,
                               (const float *)gact_[o.out.buf], bufC[o.out.buf], o.out.off, gact_[o.in.buf],
                               bufC[o.in.buf], o.in.off, o.in.C, Ho, Wo);
            VDI_LAUNCHED();
        } else {
            const int Hi = H2 / 2, Wi = W2 / 2;
            hipLaunchKernelGGL(k_up_bwd, dim3(g1((long)Hi * Wi * o.in.C)), dim3(256), 0, st,
                               (const float *)gact_[o.out.buf], bufC[o.out.buf], o.out.off, gact_[o.in.buf],
                               bufC[o.in.buf], o.in.off, o.in.C, Hi, Wi);
            VDI_LAUNCHED();
        }
    }
    return VDI_OK;
}

int UNet::adam(float lr, float b1, float b2, float eps) {
    adam_t_++;
    const float bc1 = 1.f - std::pow(b1, (float)adam_t_), bc2 = 1.f - std::pow(b2, (float)adam_t_);
    hipLaunchKernelGGL(k_adam, dim3(g1(nparams)), dim3(256), 0, st, w32_.as<float>(), g32_.as<float>(),
                       m_.as<float>(), v_.as<float>(), w16_.as<bf16_t>(), nparams, lr, b1, b2, eps, bc1, bc2);
    VDI_LAUNCHED();
    return VDI_OK;
}

double UNet::last_loss() {
    const long nb = g1((long)pixels(0));
    std::vector<float> h(nb);
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    if (hipMemcpy(h.data(), lossbuf_.p, nb * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    double s = 0;
    for (float v : h) s += v;
    return s / (double)pixels(0);
}

// ------------------------------------------------------------------------
// Learned initial guess of the MAC solve (FP64 grid side)
// ------------------------------------------------------------------------

constexpr int kGuessBlocks = 256;

__global__ void __launch_bounds__(256) k_g_sumsq(const double *a, int N, double *part) {
    __shared__ double sh[4];
    double s = 0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < N; k += gridDim.x * 256) s += a[k] * a[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// stats[0] = RMS(r0), stats[1] = RMS(z_ex), stats[2] = ||r0||^2 (sums in
// block order by one thread)
__global__ void k_g_scales(const double *pr, const double *pz, int nb, int N, double *stats) {
    if (threadIdx.x != 0) return;
    double r = 0, z = 0;
    for (int k = 0; k < nb; k++) { r += pr[k]; z += pz[k]; }
    stats[0] = r > 0 ? sqrt(r / N) : 1.0;
    stats[1] = z > 0 ? sqrt(z / N) : 1.0;
    stats[2] = r;
}

// Network inputs of the current guess: r0 / RMS(r0), z_ex / RMS(z_ex) and
// the previous solve's correction in target units, d_prev xxsp / (RMS(r0) dscale).
__global__ void k_g_prep(const double *r0, const double *zex, const double *dprev, const double *stats, int N,
                         double xxsp, double dscale, bf16_t *x0, float *slot) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= N) return;
    const float a = (float)(r0[p] / stats[0]), b = (float)(zex[p] / stats[1]);
    const float c = (float)(dprev[p] * xxsp / (stats[0] * dscale));
    if (x0) {
        x0[(size_t)p * 16] = f2bf(a);
        x0[(size_t)p * 16 + 1] = f2bf(b);
        x0[(size_t)p * 16 + 2] = f2bf(c);
        for (int k = 3; k < 16; k++) x0[(size_t)p * 16 + k] = 0;
    }
    if (slot) {
        slot[p] = a;
        slot[N + p] = b;
        slot[2 * N + p] = c;
    }
}

// After a MAC solve: d_prev = z - z_ex, and (capture) the scaled target.
__global__ void k_g_after(const double *z, const double *zex, const double *stats, int N, double xxsp, double *dprev,
                          float *t) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= N) return;
    const double d = z[p] - zex[p];
    dprev[p] = d;
    if (t) t[p] = (float)(d * xxsp / stats[0]);
}

__global__ void k_g_cand(const double *zex, const float *yf, const double *stats, int N, double scale, double *z) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < N) z[p] = zex[p] + (double)yf[(size_t)p * 16] * stats[0] * scale;
}

// Keeps the network's guess only if its residual is below the extrapolated
// guess's (the fallback rule).
__global__ void k_g_select(const double *pr1, int nb, const double *stats, int *flag, long long *counters) {
    if (threadIdx.x != 0) return;
    double r1 = 0;
    for (int k = 0; k < nb; k++) r1 += pr1[k];
    const int ok = r1 < stats[2];
    *flag = ok;
    counters[ok ? 0 : 1] += 1;
}

__global__ void k_g_apply(const int *flag, const double *zex, double *z, int N) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < N && !*flag) z[p] = zex[p];
}

__global__ void k_g_scale_f(float *a, long n, float s) {
    const long q = blockIdx.x * 256L + threadIdx.x;
    if (q < n) a[q] *= s;
}

__global__ void __launch_bounds__(256) k_g_sumsq_f(const float *a, long n, double *part) {
    __shared__ double sh[4];
    double s = 0;
    for (long k = blockIdx.x * 256L + threadIdx.x; k < n; k += (long)gridDim.x * 256) s += (double)a[k] * a[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

int LearnedGuess::init(int m_, int n_, double xxsp_, hipStream_t st, unsigned long long seed, int cap) {
    m = m_;
    n = n_;
    xxsp = xxsp_;
    VDI_TRY(net.init(n, m, st, seed));
    const size_t N = (size_t)m * n;
    VDI_TRY(zex_.alloc(N * sizeof(double)));
    VDI_TRY(r0_.alloc(N * sizeof(double)));
    VDI_TRY(part_.alloc(3 * kGuessBlocks * sizeof(double)));
    VDI_TRY(stats_.alloc(8 * sizeof(double)));
    VDI_TRY(counters_.alloc(2 * sizeof(long long) + 2 * sizeof(int)));
    VDI_TRY(dprev_.alloc(N * sizeof(double)));
    capacity = std::max(0, cap);
    if (capacity) VDI_TRY(data_.alloc((size_t)capacity * kSlot * N * sizeof(float)));
    return VDI_OK;
}

// Input features of the current extrapolated guess (z) of the MAC system:
// z_ex copy, r0 = b - A z_ex, RMS scales.
static int guess_features(LearnedGuess &L, MG &g, hipStream_t st) {
    const int N = L.m * L.n;
    double *part = L.part_.as<double>();
    VDI_HIP(hipMemcpyAsync(L.zex_.p, g.z(), N * sizeof(double), hipMemcpyDeviceToDevice, st));
    VDI_TRY(g.residual_out(L.zex_.as<double>(), L.r0_.as<double>(), part, kGuessBlocks));
    hipLaunchKernelGGL(k_g_sumsq, dim3(kGuessBlocks), dim3(256), 0, st, (const double *)L.zex_.as<double>(), N,
                       part + kGuessBlocks);
    hipLaunchKernelGGL(k_g_scales, dim3(1), dim3(64), 0, st, (const double *)part, (const double *)(part + kGuessBlocks),
                       kGuessBlocks, N, L.stats_.as<double>());
    VDI_LAUNCHED();
    return VDI_OK;
}

int guess_capture_pre(LearnedGuess &L, MG &g, hipStream_t st) {
    VDI_TRY(guess_features(L, g, st));
    if (!L.capacity) return VDI_OK;
    const int N = L.m * L.n;
    float *slot = L.data_.as<float>() + (size_t)L.next * kSlot * N;
    hipLaunchKernelGGL(k_g_prep, dim3(g1(N)), dim3(256), 0, st, (const double *)L.r0_.as<double>(),
                       (const double *)L.zex_.as<double>(), (const double *)L.dprev_.as<double>(),
                       (const double *)L.stats_.as<double>(), N, L.xxsp, 1.0, (bf16_t *)nullptr, slot);
    VDI_LAUNCHED();
    return VDI_OK;
}

int guess_after_solve(LearnedGuess &L, MG &g, hipStream_t st) {
    const int N = L.m * L.n;
    float *t = nullptr;
    if (L.mode == 1 && L.capacity) t = L.data_.as<float>() + ((size_t)L.next * kSlot + 3) * N;
    hipLaunchKernelGGL(k_g_after, dim3(g1(N)), dim3(256), 0, st, (const double *)g.z(),
                       (const double *)L.zex_.as<double>(), (const double *)L.stats_.as<double>(), N, L.xxsp,
                       L.dprev_.as<double>(), t);
    VDI_LAUNCHED();
    if (t) {
        L.next = (L.next + 1) % L.capacity;
        L.count = std::min(L.count + 1, L.capacity);
    }
    return VDI_OK;
}

int guess_infer(LearnedGuess &L, MG &g, hipStream_t st) {
    const int N = L.m * L.n;
    VDI_TRY(guess_features(L, g, st));
    hipLaunchKernelGGL(k_g_prep, dim3(g1(N)), dim3(256), 0, st, (const double *)L.r0_.as<double>(),
                       (const double *)L.zex_.as<double>(), (const double *)L.dprev_.as<double>(),
                       (const double *)L.stats_.as<double>(), N, L.xxsp, L.kappa, L.net.x0(), (float *)nullptr);
    VDI_LAUNCHED();
    VDI_TRY(L.net.forward());
    hipLaunchKernelGGL(k_g_cand, dim3(g1(N)), dim3(256), 0, st, (const double *)L.zex_.as<double>(),
                       (const float *)L.net.y(), (const double *)L.stats_.as<double>(), N, L.kappa / L.xxsp, g.z());
    VDI_LAUNCHED();
    double *part = L.part_.as<double>() + 2 * kGuessBlocks;
    VDI_TRY(g.residual_out(g.z(), nullptr, part, kGuessBlocks));
    long long *cnt = L.counters_.as<long long>();
    int *flag = reinterpret_cast<int *>(cnt + 2);
    hipLaunchKernelGGL(k_g_select, dim3(1), dim3(64), 0, st, (const double *)part, kGuessBlocks,
                       (const double *)L.stats_.as<double>(), flag, cnt);
    hipLaunchKernelGGL(k_g_apply, dim3(g1(N)), dim3(256), 0, st, (const int *)flag, (const double *)L.zex_.as<double>(),
                       g.z(), N);
    VDI_LAUNCHED();
    return VDI_OK;
}

void guess_destroy(LearnedGuess *L) { delete L; }
int guess_mode(const LearnedGuess *L) { return L ? L->mode : 0; }

// Trains on the captured pairs: targets are first scaled by their dataset
// RMS kappa (once), then epochs of per-sample forward/backward with an
// Adam step every `batch` samples, in an order drawn from SplitMix64.
static int guess_train(LearnedGuess &L, int epochs, int batch, double lr, unsigned long long seed, double *loss) {
    if (L.count < 1 || batch < 1 || epochs < 1) return VDI_ERR_ARG;
    UNet &net = L.net;
    hipStream_t st = net.st;
    const long N = (long)L.m * L.n;
    if (L.kappa == 1.0) {
        DevBuf part;
        VDI_TRY(part.alloc(1024 * sizeof(double)));
        double tot = 0;
        std::vector<double> h(1024);
        for (int k = 0; k < L.count; k++) {
            hipLaunchKernelGGL(k_g_sumsq_f, dim3(1024), dim3(256), 0, st,
                               (const float *)(L.data_.as<float>() + ((size_t)k * kSlot + 3) * N), N, part.as<double>());
            VDI_LAUNCHED();
            VDI_HIP(hipMemcpyAsync(h.data(), part.p, 1024 * sizeof(double), hipMemcpyDeviceToHost, st));
            VDI_HIP(hipStreamSynchronize(st));
            for (double v : h) tot += v;
        }
        const double kap = std::sqrt(tot / ((double)N * L.count));
        L.kappa = kap > 0 ? kap : 1.0;
        for (int k = 0; k < L.count; k++) {
            // the previous correction (input 2) and the target share the scale
            hipLaunchKernelGGL(k_g_scale_f, dim3(g1(2 * N)), dim3(256), 0, st,
                               L.data_.as<float>() + ((size_t)k * kSlot + 2) * N, 2 * N, (float)(1.0 / L.kappa));
            VDI_LAUNCHED();
        }
    }
    VDI_TRY(net.zero_grads());
    std::vector<int> order(L.count);
    unsigned long long s = seed;
    // Adam on this small data set can leave a good basin late in training.
    // The weights at the start of each epoch are saved; the mean of the
    // per-sample losses seen during the epoch rates them, and the best rated
    // weights are restored at the end (their rating is reported).
    DevBuf snap[2];
    VDI_TRY(snap[0].alloc((size_t)net.nparams * sizeof(float)));
    VDI_TRY(snap[1].alloc((size_t)net.nparams * sizeof(float)));
    const size_t wbytes = (size_t)net.nparams * sizeof(float);
    int ib = 0;  // snap[ib]: best so far; snap[1 - ib]: start of this epoch
    double best_loss = -1;
    for (int e = 0; e < epochs; e++) {
        VDI_HIP(hipMemcpyAsync(snap[1 - ib].p, net.params(), wbytes, hipMemcpyDeviceToDevice, st));
        for (int k = 0; k < L.count; k++) order[k] = k;
        for (int k = L.count - 1; k > 0; k--) std::swap(order[k], order[splitmix(s) % (unsigned long long)(k + 1)]);
        double acc = 0;
        for (int q = 0; q < L.count; q++) {
            const float *slot = L.data_.as<float>() + (size_t)order[q] * kSlot * N;
            hipLaunchKernelGGL(k_expand, dim3(g1(N)), dim3(256), 0, st, slot, net.x0(), N);
            VDI_LAUNCHED();
            VDI_TRY(net.forward());
            VDI_TRY(net.backward(slot + 3 * N));
            acc += net.last_loss();
            if ((q + 1) % batch == 0 || q == L.count - 1) VDI_TRY(net.adam((float)lr, 0.9f, 0.999f, 1e-8f));
        }
        const double el = acc / L.count;
        if (!(el >= 0)) return VDI_ERR_ARG;  // NaN loss: the data or the rate is unusable
        if (best_loss < 0 || el < best_loss) {
            best_loss = el;
            ib = 1 - ib;
        }
    }
    VDI_HIP(hipMemcpyAsync(net.params(), snap[ib].p, wbytes, hipMemcpyDeviceToDevice, st));
    VDI_TRY(net.sync_bf16());
    VDI_HIP(hipStreamSynchronize(st));
    if (loss) *loss = best_loss;
    return VDI_OK;
}

}  // namespace vdi

using namespace vdi;

struct vdi_nn {
    UNet net;
    hipStream_t st = nullptr;
    DevBuf x2, t;
    ~vdi_nn() {
        if (st) (void)hipStreamSynchronize(st);
        if (st) (void)hipStreamDestroy(st);
    }
};

extern "C" {

int vdi_guess_attach(vdi_ctx *c, int capacity, unsigned long long seed) {
    if (!c || capacity < 0) return VDI_ERR_ARG;
    if (c->g.m % 64 || c->g.n % 64) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->drop_graphs();
    std::unique_ptr<LearnedGuess> L(new LearnedGuess());
    VDI_TRY(L->init(c->g.m, c->g.n, c->g.xxsp, c->st, seed, capacity));
    if (c->guess) guess_destroy(c->guess);
    c->guess = L.release();
    return VDI_OK;
}

int vdi_guess_mode(vdi_ctx *c, int mode) {
    if (!c || !c->guess || mode < 0 || mode > 2) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    c->drop_graphs();
    c->guess->mode = mode;
    return VDI_OK;
}

int vdi_guess_train(vdi_ctx *c, int epochs, int batch, double lr, unsigned long long seed, double *final_loss) {
    if (!c || !c->guess) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    return guess_train(*c->guess, epochs, batch, lr, seed, final_loss);
}

int vdi_guess_stats(vdi_ctx *c, long *pairs, long *accepted, long *rejected, double *kappa) {
    if (!c || !c->guess || !pairs || !accepted || !rejected || !kappa) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(c->st));
    long long h[2];
    VDI_HIP(hipMemcpy(h, c->guess->counters_.p, sizeof h, hipMemcpyDeviceToHost));
    *pairs = c->guess->count;
    *accepted = (long)h[0];
    *rejected = (long)h[1];
    *kappa = c->guess->kappa;
    return VDI_OK;
}

int vdi_guess_copy(vdi_ctx *dst, vdi_ctx *src) {
    if (!dst || !src || !dst->guess || !src->guess) return VDI_ERR_ARG;
    if (dst->guess->net.nparams != src->guess->net.nparams) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(src->st));
    VDI_HIP(hipStreamSynchronize(dst->st));
    VDI_HIP(hipMemcpy(dst->guess->net.params(), src->guess->net.params(), src->guess->net.nparams * sizeof(float),
                      hipMemcpyDeviceToDevice));
    VDI_TRY(dst->guess->net.sync_bf16());
    dst->guess->kappa = src->guess->kappa;
    dst->drop_graphs();
    return VDI_OK;
}

int vdi_nn_create(int H, int W, unsigned long long seed, vdi_nn **out) {
    if (!out) return VDI_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<vdi_nn> h(new vdi_nn());
    VDI_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    VDI_TRY(h->net.init(H, W, h->st, seed));
    VDI_TRY(h->x2.alloc(kNNInputs * (size_t)H * W * sizeof(float)));
    VDI_TRY(h->t.alloc((size_t)H * W * sizeof(float)));
    *out = h.release();
    return VDI_OK;
}

int vdi_nn_destroy(vdi_nn *h) {
    if (!h) return VDI_ERR_ARG;
    delete h;
    return VDI_OK;
}

long vdi_nn_nparams(vdi_nn *h) { return h ? h->net.nparams : -1; }

int vdi_nn_get_params(vdi_nn *h, float *out, long n) {
    if (!h || !out || n != h->net.nparams) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    VDI_HIP(hipMemcpy(out, h->net.params(), n * sizeof(float), hipMemcpyDeviceToHost));
    return VDI_OK;
}

int vdi_nn_set_params(vdi_nn *h, const float *in, long n) {
    if (!h || !in || n != h->net.nparams) return VDI_ERR_ARG;
    VDI_HIP(hipStreamSynchronize(h->st));
    VDI_HIP(hipMemcpy(h->net.params(), in, n * sizeof(float), hipMemcpyHostToDevice));
    return h->net.sync_bf16();
}

// x2: the input channels, planar [kNNInputs][H*W]; y: channel 0 of the head.
int vdi_nn_forward(vdi_nn *h, const float *x2, float *y) {
    if (!h || !x2 || !y) return VDI_ERR_ARG;
    const long P = (long)h->net.pixels(0);
    VDI_HIP(hipMemcpy(h->x2.p, x2, kNNInputs * P * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_expand, dim3(g1(P)), dim3(256), 0, h->st, (const float *)h->x2.as<float>(), h->net.x0(), P);
    VDI_LAUNCHED();
    VDI_TRY(h->net.forward());
    VDI_HIP(hipStreamSynchronize(h->st));
    std::vector<float> yf(16 * P);
    VDI_HIP(hipMemcpy(yf.data(), h->net.y(), 16 * P * sizeof(float), hipMemcpyDeviceToHost));
    for (long p = 0; p < P; p++) y[p] = yf[16 * p];
    return VDI_OK;
}

// Gradient of loss = mean_p (y0 - t)^2 with respect to every parameter.
int vdi_nn_gradient(vdi_nn *h, const float *x2, const float *t, float *grad, double *loss) {
    if (!h || !x2 || !t || !grad || !loss) return VDI_ERR_ARG;
    const long P = (long)h->net.pixels(0);
    VDI_HIP(hipMemcpy(h->x2.p, x2, kNNInputs * P * sizeof(float), hipMemcpyHostToDevice));
    VDI_HIP(hipMemcpy(h->t.p, t, P * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_expand, dim3(g1(P)), dim3(256), 0, h->st, (const float *)h->x2.as<float>(), h->net.x0(), P);
    VDI_LAUNCHED();
    VDI_TRY(h->net.zero_grads());
    VDI_TRY(h->net.forward());
    VDI_TRY(h->net.backward(h->t.as<float>()));
    *loss = h->net.last_loss();
    VDI_HIP(hipMemcpy(grad, h->net.grads(), h->net.nparams * sizeof(float), hipMemcpyDeviceToHost));
    return VDI_OK;
}

}  // extern "C"
