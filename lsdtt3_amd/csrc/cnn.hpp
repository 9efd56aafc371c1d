// cnn.hpp -- the learned Newton warm start: a 2-D encoder-decoder CNN
// (U-Net) predicting the hillslope step's elevation change from the
// post-incision surface, trained and run on the device.
//
// No counterpart in the reference (the reference always starts Newton from
// z^n).  Architecture: `levels` resolution levels, `base` channels doubling
// per level, two 3x3 convolutions (+ bias, ReLU) per level on the way down
// and up, 2x2 average pooling, nearest 2x upsampling with skip
// concatenation ([upsampled coarse, skip] channel order), and a 1x1 FP32
// head.  Activations and weights are BF16 (NHWC, channels innermost), every
// 3x3 convolution is an implicit GEMM on MFMA (v_mfma_f32_16x16x32_bf16) with
// FP32 accumulation; the head, gradients and Adam state are FP32.
//
// Backward: the ReLU gradient is taken where the stored BF16 output is > 0;
// the masked output gradient is rounded to BF16 and feeds three MFMA /
// vector kernels: input gradient (the forward kernel on flipped weights),
// weight gradient (pixels as the reduction dimension, LDS-transposed tiles)
// and bias gradient.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"

namespace lsdtt_gpu {

struct ConvLayer {
    int cin, cin_pad, cout;  // cin_pad: multiple of 8 (zero channels)
    int kpad;                // 9 * cin_pad rounded up to 32
    size_t woff, boff;       // offsets of W [cout][9][cin_pad] and bias in the parameter vector
};

class UNet {
public:
    // base a multiple of 32 and base * 2^(levels-1) <= 256; levels >= 2
    UNet(int in_channels, int base, int levels, uint64_t seed);

    int in_channels() const { return in_ch_; }
    int base() const { return base_; }
    int levels() const { return levels_; }
    size_t num_params() const { return nparam_; }
    // spatial sizes must be divisible by 2^(levels-1), and the pixel count
    // of the coarsest level by 16
    bool shape_ok(int H, int W) const;

    // x: FP32 NHWC [N][H][W][in_channels] (device); y: FP32 [N][H][W]
    // (device, may be null: the output stays in out_device())
    void forward(const float* x, int N, int H, int W, float* y, hipStream_t s);
    const float* out_device() const { return out.p; }
    // MSE over the pixels with loss_mask != 0 (uint8 [N][H][W], null = all)
    // of the last forward(); adds the FP32 parameter gradient to grads;
    // returns the loss
    double backward(const float* target, const uint8_t* loss_mask, hipStream_t s);
    // Adam on the FP32 master weights, then the BF16 working copy
    void adam_step(double lr, double beta1, double beta2, double eps, hipStream_t s);
    void zero_grad(hipStream_t s);

    void get_params(float* host) const;
    void set_params(const float* host);
    void get_grads(float* host) const;

    // multiply-adds x 2 of the 3x3 convolutions of one forward pass
    double conv_flops(int N, int H, int W) const;
    // time of the last forward's convolution stack (ms) when timing is on
    bool time_convs = false;
    double conv_ms = 0.0;

private:
    int in_ch_, base_, levels_;
    size_t nparam_ = 0;
    std::vector<ConvLayer> enc_a, enc_b, dec_a, dec_b;  // per level
    size_t head_w = 0, head_b = 0;
    dbuf<float> params, grads, adam_m, adam_v;
    dbuf<__bf16> wbf;             // BF16 weights, per layer [cout][kpad]
    dbuf<__bf16> wflip;           // flipped weights [cin_pad][9 * cout] (backward-data)
    std::vector<size_t> wbf_off;  // per conv layer (enc_a..., enc_b..., dec_a..., dec_b...)
    long adam_t = 0;

    // activations of the last forward (BF16 NHWC)
    int N_ = 0, H_ = 0, W_ = 0;
    std::vector<dbuf<__bf16>> act;
    std::vector<dbuf<float>> gskip;  // per level: skip-connection gradients
    dbuf<float> out, dout;           // head output; its gradient
    dbuf<float> g0, g1, g2;          // FP32 gradient scratch
    dbuf<__bf16> gb;                 // BF16 masked gradient
    dbuf<double> loss_sum;
    hbuf<double> hpin;

    ConvLayer make_layer(int cin, int cout);
    void convert_weights(hipStream_t s);
    // input [upsample2(u) (cs channels), x] (u null, cs 0: x only)
    void conv(const ConvLayer& L, size_t wo, const __bf16* u, int cs, const __bf16* x, int N, int H, int W, __bf16* y,
              hipStream_t s);
    void conv_backward(const ConvLayer& L, size_t wo, const __bf16* u, int cs, const __bf16* x, const __bf16* y,
                       int N, int H, int W, const float* g_out, float* g_in, hipStream_t s);
};

// One 3x3 convolution of the network on host arrays, through the same
// kernels: x [N][H][W][cin] and w [cout][9][cin] are rounded to BF16, y
// [N][H][W][cout] receives the BF16 output (+ bias, ReLU if relu) as float.
// With gy (FP32 output gradient) also the weight gradient gw [cout][9][cin],
// bias gradient gb [cout] and, when gx is not null (cin % 32 == 0), the
// input gradient gx [N][H][W][cin]; the masked gradient is rounded to BF16
// first, as in training.
void conv3x3_layer(const float* x, const float* w, const float* b, int N, int H, int W, int cin, int cout, int relu,
                   float* y, const float* gy, float* gx, float* gw, float* gb);

}  // namespace lsdtt_gpu
