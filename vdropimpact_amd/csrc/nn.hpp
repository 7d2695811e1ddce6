// nn.hpp -- learned initial guess for the MAC projection (no counterpart in
// the reference, whose guess is the time extrapolation of mgs_common.cc:17-50).
//
// A 4-level convolutional U-Net (about 1.0M parameters) maps three input
// channels on the m x n MAC grid -- the residual r0 = b - A z_ex of the
// extrapolated guess z_ex, z_ex itself (each scaled by its RMS), and the
// previous step's correction in target units -- to the correction delta with
// A delta = r0, i.e. z = z_ex + delta.  Activations are BF16 NHWC tensors,
// m and n are multiples of 64 (whole 8 x 8 tiles at 1/8 resolution);
// every convolution is a 3x3 implicit GEMM on MFMA (bf16 inputs, FP32
// accumulation); training keeps FP32 master weights, gradients and Adam
// moments.  All kernels are hand-written (nn.hip).
//
// Layers (level l has (n >> l) x (m >> l) pixels; channels in brackets):
//   enc1  x0[16] -> e1a[16] -> s1[16]           pool -> p1 (level 1)
//   enc2  p1 -> e2a[32] -> s2[32]                pool -> p2 (level 2)
//   enc3  p2 -> e3a[96] -> s3[96]                pool -> p3 (level 3)
//   bott  p3 -> b1[192] -> b2[192]               up   -> cat3[0:192]
//   dec3  cat3 = [up(b2) | s3] [288] -> d3a[96] -> d3b[96]   up -> cat2[0:96]
//   dec2  cat2 = [up(d3b) | s2] [128] -> d2a[32] -> d2b[32]  up -> cat1[0:32]
//   dec1  cat1 = [up(d2b) | s1] [48] -> d1a[16] -> d1b[16]
//   head  d1b -> y[16] (no activation; channel 0 is the prediction)
// LeakyReLU(0.1) after every convolution but the head; 2x2 average pooling;
// nearest-neighbour upsampling.
#pragma once

#include "mg.hpp"

namespace vdi {

typedef unsigned short bf16_t;

// Network inputs (channels 0..2 of the 16-channel input tensor) and the
// floats per recorded training pair: the three inputs and the target.
constexpr int kNNInputs = 3;
constexpr int kSlot = kNNInputs + 1;

class UNet {
  public:
    // Tensors: activation buffers (BF16, NHWC, C channels) at a level.
    enum Buf {
        X0, E1A, CAT1, P1, E2A, CAT2, P2, E3A, CAT3, P3, B1, B2, D3A, D3B, D2A, D2B, D1A, D1B, Y, NBUF
    };
    struct Slice {
        int buf, off, C;
    };
    enum OpKind { CONV, POOL, UP };
    struct Op {
        OpKind kind;
        Slice in, out;
        int level;   // level of the output (CONV, UP) or input (POOL)
        int act;     // CONV: LeakyReLU
        long woff;   // CONV: offset of W [Cout][9][Cin] in the parameter vector; bias follows
    };

    int H = 0, W = 0;  // level-0 pixels (rows = n, columns = m)
    std::vector<Op> ops;
    int bufC[NBUF] = {}, bufL[NBUF] = {};
    long nparams = 0;
    hipStream_t st = nullptr;

    int init(int H, int W, hipStream_t stream, unsigned long long seed);
    // Forward pass on x0 (BF16 [P0][16]); y (FP32 [P0][16]) gets the head output.
    int forward();
    // Backward pass for loss = mean over pixels of (y0 - t)^2, t FP32 [P0];
    // accumulates parameter gradients; loss partials for loss().
    int backward(const float *target);
    // Adam step (bias-corrected), then zero gradients.
    int adam(float lr, float beta1, float beta2, float eps);
    int zero_grads();
    bf16_t *x0() const { return act_[X0]; }
    float *y() const { return yf_.as<float>(); }
    float *params() const { return w32_.as<float>(); }
    float *grads() const { return g32_.as<float>(); }
    // Copies FP32 master weights to the BF16 copies the convolutions read.
    int sync_bf16();
    double last_loss();  // synchronous
    size_t pixels(int level) const { return (size_t)(H >> level) * (W >> level); }

  private:
    DevBuf acts_, grads_, yf_, gy_;
    bf16_t *act_[NBUF] = {};
    float *gact_[NBUF] = {};
    DevBuf w32_, w16_, wflip_, g32_, m_, v_, part_, dbuf_, lossbuf_;
    size_t part_cap_ = 0;
    long adam_t_ = 0;
    int conv_fwd(const Op &op, const bf16_t *w, const bf16_t *x, int xs, int xo, int Cin, int Cout, bf16_t *y,
                 int ys, int yo, float *yf, int yfs, int yfo, int act, int acc, const float *bias);
};

// The learned guess attached to one context's MAC solve.
struct LearnedGuess {
    UNet net;
    int m = 0, n = 0;
    double xxsp = 1;
    double kappa = 1;   // dataset RMS of the scaled target
    DevBuf zex_, r0_, dprev_, part_, stats_, cand_;
    // dataset: per pair the kNNInputs input planes and the target (FP32),
    // capacity pairs
    DevBuf data_;
    int capacity = 0, count = 0, next = 0;
    int mode = 0;       // 0 off, 1 capture, 2 infer
    long used = 0, fallbacks = 0;
    DevBuf counters_;   // device: [0] accepted, [1] rejected
    int init(int m, int n, double xxsp, hipStream_t st, unsigned long long seed, int capacity);
};

}  // namespace vdi
