// cnn.hip -- U-Net forward, backward and Adam on MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

#include "cnn.hpp"

namespace lsdtt_gpu {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kB = 256;

inline int nblk(long long n) { return static_cast<int>((n + kB - 1) / kB); }

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 a;
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = __bf16(0.0f);
    return a;
}

// ---------------------------------------------------------------------------
// 3x3 convolution (stride 1, zero padding) as an implicit GEMM on
// v_mfma_f32_16x16x32_bf16: C[pixel][cout] = sum_k A[pixel][k] B[k][cout],
// k = tap * cin + ci (tap = 3 (dy + 1) + (dx + 1)), A the im2col rows of X
// (NHWC, cin a multiple of 8, so the 8 consecutive k of one lane share a
// tap), B the weights stored [cout][kpad].  A wavefront computes 16 pixels
// x 32 output channels (two MFMA tiles sharing the A operand); the four
// wavefronts of a workgroup take 64 consecutive pixels of the same 32
// channels.  Operand layouts (gfx950): lane l supplies A[l % 16][8 (l / 16)
// + i] and B[8 (l / 16) + i][l % 16], i = 0..7, and receives
// C[4 (l / 16) + i][l % 16], i = 0..3.
// ---------------------------------------------------------------------------
template <bool RELU, bool OUT_BF16>
__global__ void __launch_bounds__(kB) k_conv(const __bf16* __restrict__ X, const __bf16* __restrict__ Wt,
                                             const float* __restrict__ bias, void* __restrict__ Y, int N, int H,
                                             int W, int cin, int cout, int kpad)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long P = (long long)N * H * W;
    const long long p0 = ((long long)blockIdx.x * 4 + wave) * 16;
    if (p0 >= P) return;  // P is a multiple of 16: whole wavefronts retire
    const int co0 = blockIdx.y * 32;
    const long long pa = p0 + (lane & 15);
    const int x = static_cast<int>(pa % W);
    const long long t = pa / W;
    const int y = static_cast<int>(t % H);
    const int n = static_cast<int>(t / H);
    const int kg = lane >> 4;
    const __bf16* w0 = Wt + (size_t)(co0 + (lane & 15)) * kpad;
    const __bf16* w1 = w0 + (size_t)16 * kpad;
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < kpad; kc += 32) {
        const int k0 = kc + 8 * kg;
        const int tap = k0 / cin, ci0 = k0 - tap * cin;
        bf16x8 a = zero8();
        if (tap < 9) {
            const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                a = *reinterpret_cast<const bf16x8*>(X + ((((size_t)n * H + yy) * W + xx) * cin + ci0));
        }
        const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(w0 + k0);
        const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(w1 + k0);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
    }
    const int c0 = co0 + (lane & 15), c1 = c0 + 16;
    const float bias0 = bias ? bias[c0] : 0.f, bias1 = bias ? bias[c1] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long long p = p0 + 4 * kg + i;
        float v0 = acc0[i] + bias0, v1 = acc1[i] + bias1;
        if (RELU) {
            v0 = fmaxf(v0, 0.f);
            v1 = fmaxf(v1, 0.f);
        }
        if (OUT_BF16) {
            __bf16* yb = static_cast<__bf16*>(Y);
            yb[(size_t)p * cout + c0] = __bf16(v0);
            yb[(size_t)p * cout + c1] = __bf16(v1);
        } else {
            float* yf = static_cast<float*>(Y);
            yf[(size_t)p * cout + c0] = v0;
            yf[(size_t)p * cout + c1] = v1;
        }
    }
}

// ---------------------------------------------------------------------------
// 3x3 convolution for cin and cout multiples of 32: a workgroup computes a
// block of RB rows x 16 pixels of one image for 32 WC output channels.  Its
// four wavefronts are WR = 4 / WC along the rows (R output rows each) by WC
// along the channels (32 each, two 16-channel MFMA groups).  Per 32-channel
// chunk of the input the block plus a one-pixel halo ((RB + 2) x 18 pixels x
// 32 channels, pixel stride 40 elements so that 16-byte reads of 16
// consecutive pixels fall in distinct LDS banks) is staged in LDS; the next
// chunk is loaded into registers while this one is used, and the chunk's
// weights (9 taps x 2 groups of B operands) are held in registers.
//
// Operand reuse is what the MFMA rate needs: for one tap column dx, the A
// fragment of staged row s (16 pixels x 32 channels, one 16-byte LDS read
// per lane) serves output rows s, s - 1 and s - 2 through taps dy = 0, 1, 2,
// and each of the two channel groups: 3 (R + 2) LDS reads per chunk feed
// 18 R MFMAs, and each B operand, loaded once per chunk, feeds R rows (8;
// 6 or 4 for 32-channel blocks, whose taller staged block needs more
// prefetch registers, 4 where all nine taps' weights stay in registers).  The B operands are held one tap column at a time (the next
// column's load in flight), which keeps the registers within two
// wavefronts per SIMD.
//
// The input may be the decoder's concatenation [upsample2(U), X]: channels
// below cs come from U (cs channels at half resolution, pixel (y/2, x/2)),
// the rest from X (cin - cs channels); cs is a multiple of 32 (0: X only).
// ---------------------------------------------------------------------------
constexpr int kTile = 16;                // output block width (pixels): one MFMA's rows
constexpr int kHalo = kTile + 2;         // staged width
constexpr int kPixStride = 40;           // LDS elements per staged pixel (32 + 8 pad)

template <int WC, int R>
struct ConvRowsShape {
    static constexpr int WR = 4 / WC;                 // wavefronts along the rows
    static constexpr int RB = WR * R;                 // block rows
    static constexpr int SR = RB + 2;                 // staged rows
    static constexpr int UNITS = SR * kHalo * 4;      // 16-byte units staged per chunk
    static constexpr int PF = (UNITS + kB - 1) / kB;  // units per thread
};

template <bool RELU, bool OUT_BF16, int WC, int R, bool ONE>
__global__ void __launch_bounds__(kB, 2) k_conv_tile(const __bf16* __restrict__ U, int cs,
                                                     const __bf16* __restrict__ X, const __bf16* __restrict__ Wt,
                                                     const float* __restrict__ bias, void* __restrict__ Y, int N,
                                                     int H, int W, int cin, int cout, int kpad, int tiles_x,
                                                     int tiles_y)
{
    using S = ConvRowsShape<WC, R>;
    __shared__ __align__(16) __bf16 tile[S::SR * kHalo * kPixStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = wave % S::WR, wc = wave / S::WR;
    const int ntiles = tiles_x * tiles_y * N;
    const int nchunk = cin / 32;
    const int co0 = blockIdx.y * (32 * WC) + 32 * wc;  // this wavefront's first output channel
    const int col = lane & 15, kg = lane >> 4;
    const __bf16* wrow = Wt + (size_t)(co0 + col) * kpad + 8 * kg;
    const int ck = cin - cs, Hh = H >> 1, Wh = W >> 1;
    const int srow0 = wr * R;  // staged row of this wavefront's first output row (minus the halo)
    // persistent: the block walks the pixel blocks t = blockIdx.x + k
    // gridDim.x, and the 32-channel chunks of each; the input of the next
    // (block, chunk) is loaded into registers while this one computes
    bf16x8 pf[S::PF];
    auto fetch = [&](int t, int ci0) {
        const int tx0 = (t % tiles_x) * kTile;
        const int rest = t / tiles_x;
        const int ty0 = (rest % tiles_y) * S::RB;
        const int n = rest / tiles_y;
        const bool from_u = ci0 < cs;
#pragma unroll
        for (int j = 0; j < S::PF; ++j) {
            const int u = threadIdx.x + kB * j;
            bf16x8 v = zero8();
            if (u < S::UNITS) {
                const int pix = u >> 2, c8 = (u & 3) * 8;
                const int yy = ty0 - 1 + pix / kHalo, xx = tx0 - 1 + pix % kHalo;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    const __bf16* src =
                        from_u ? U + ((((size_t)n * Hh + (yy >> 1)) * Wh + (xx >> 1)) * cs + ci0 + c8)
                               : X + ((((size_t)n * H + yy) * W + xx) * ck + (ci0 - cs) + c8);
                    v = *reinterpret_cast<const bf16x8*>(src);
                }
            }
            pf[j] = v;
        }
    };
    // B operands of one tap column (3 taps x 2 channel groups), double
    // buffered across the columns of a chunk
    auto load_column = [&](bf16x8 (&bw)[3][2], int ci0, int dx) {
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const int tap = 3 * dy + dx;
            bw[dy][0] = *reinterpret_cast<const bf16x8*>(wrow + tap * cin + ci0);
            bw[dy][1] = *reinterpret_cast<const bf16x8*>(wrow + (size_t)16 * kpad + tap * cin + ci0);
        }
    };
    float bs[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) bs[g] = bias ? bias[co0 + 16 * g + col] : 0.f;
    int t = blockIdx.x;
    if (t >= ntiles) return;
    fetch(t, 0);
    bf16x8 bA[3][2], bB[3][2];
    // a single input chunk: the weights of all nine taps stay in registers
    // for every pixel block the block computes
    bf16x8 b0[3][2], b1[3][2], b2[3][2];
    if (ONE) {
        load_column(b0, 0, 0);
        load_column(b1, 0, 1);
        load_column(b2, 0, 2);
    }
    for (; t < ntiles; t += gridDim.x) {
        floatx4 acc[R][2];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            acc[r][0] = floatx4{0.f, 0.f, 0.f, 0.f};
            acc[r][1] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
        for (int c = 0; c < nchunk; ++c) {
            __syncthreads();  // the previous chunk's reads are done
#pragma unroll
            for (int j = 0; j < S::PF; ++j) {
                const int u = threadIdx.x + kB * j;
                if (u < S::UNITS) *reinterpret_cast<bf16x8*>(&tile[(u >> 2) * kPixStride + (u & 3) * 8]) = pf[j];
            }
            if (!ONE) load_column(bA, 32 * c, 0);
            if (c + 1 < nchunk) fetch(t, 32 * (c + 1));
            else if (t + (int)gridDim.x < ntiles) fetch(t + gridDim.x, 0);
            __syncthreads();
            auto column = [&](const bf16x8 (&bw)[3][2], int dx) {
#pragma unroll
                for (int s = 0; s < R + 2; ++s) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(
                        &tile[((srow0 + s) * kHalo + col + dx) * kPixStride + 8 * kg]);
#pragma unroll
                    for (int dy = 0; dy < 3; ++dy) {
                        const int r = s - dy;
                        if (r < 0 || r >= R) continue;
                        acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[dy][0], acc[r][0], 0, 0, 0);
                        acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[dy][1], acc[r][1], 0, 0, 0);
                    }
                }
            };
            if (ONE) {
                column(b0, 0);
                column(b1, 1);
                column(b2, 2);
            } else {
                load_column(bB, 32 * c, 1);
                column(bA, 0);
                load_column(bA, 32 * c, 2);
                column(bB, 1);
                column(bA, 2);
            }
        }
        // lane (col, kg) holds, for output row r, the pixels 4 kg + i of
        // that row and output channels co0 + 16 g + col
        const int tx0 = (t % tiles_x) * kTile;
        const int rest = t / tiles_x;
        const int ty0 = (rest % tiles_y) * S::RB;
        const int n = rest / tiles_y;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int yy = ty0 + srow0 + r;
            if (yy >= H) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xx = tx0 + 4 * kg + i;
                if (xx >= W) continue;
                const size_t p = ((size_t)n * H + yy) * W + xx;
#pragma unroll
                for (int g = 0; g < 2; ++g) {
                    float v = acc[r][g][i] + bs[g];
                    if (RELU) v = fmaxf(v, 0.f);
                    if (OUT_BF16) static_cast<__bf16*>(Y)[p * cout + co0 + 16 * g + col] = __bf16(v);
                    else static_cast<float*>(Y)[p * cout + co0 + 16 * g + col] = v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient of a 3x3 convolution: dW[co][tap][ci] += sum over pixels p
// of dY[p][co] X[p + off(tap)][ci].  A workgroup owns 32 output x 32 input
// channels of one tap over a chunk of pixels.  Per iteration it stages 128
// pixels of both operands transposed in LDS ([channel][pixel]); wavefront w
// reduces pixels 32w..32w+31 into the whole 32 x 32 block (four MFMA tiles),
// so the pixel index is the MFMA's k dimension.  The four wavefronts' sums
// are combined in LDS and added to dW atomically (FP32).  The input may be
// the decoder's concatenation [upsample2(U), X] (cs channels from U, a
// multiple of 32; cs = 0: X only).
// ---------------------------------------------------------------------------
constexpr int kWgPix = 128;
constexpr int kWgChunk = 8192;

__global__ void __launch_bounds__(kB) k_wgrad(const __bf16* __restrict__ dY, const __bf16* __restrict__ U, int cs,
                                              const __bf16* __restrict__ X, float* __restrict__ dW, int N, int H,
                                              int W, int cin_pad, int cout)
{
    __shared__ __align__(16) __bf16 At[32][kWgPix + 8];  // [co][pixel]
    __shared__ __align__(16) __bf16 Bt[32][kWgPix + 8];  // [ci][pixel]
    __shared__ float red[4][32][33];
    const long long P = (long long)N * H * W;
    const int co0 = blockIdx.x * 32;
    const int nci = (cin_pad + 31) / 32;
    const int tap = blockIdx.y / nci, ci0 = (blockIdx.y % nci) * 32;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const long long pbeg = (long long)blockIdx.z * kWgChunk;
    const long long pend = std::min<long long>(P, pbeg + kWgChunk);
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int q = t >> 1, h16 = (t & 1) * 16;  // loader: pixel, channel half
    floatx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (long long ps = pbeg; ps < pend; ps += kWgPix) {
        const long long p = ps + q;
        bf16x8 a0 = zero8(), a1 = zero8(), b0 = zero8(), b1 = zero8();
        if (p < pend) {
            const __bf16* dys = dY + (size_t)p * cout + co0 + h16;
            a0 = *reinterpret_cast<const bf16x8*>(dys);
            a1 = *reinterpret_cast<const bf16x8*>(dys + 8);
            const int x = static_cast<int>(p % W);
            const long long tt = p / W;
            const int y = static_cast<int>(tt % H);
            const int n = static_cast<int>(tt / H);
            const int yy = y + dy, xx = x + dx;
            const int c = ci0 + h16;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W && c < cin_pad) {
                // channels below cs: the upsampled half-resolution input U
                const __bf16* xs = c < cs ? U + ((((size_t)n * (H >> 1) + (yy >> 1)) * (W >> 1) + (xx >> 1)) * cs + c)
                                          : X + ((((size_t)n * H + yy) * W + xx) * (cin_pad - cs) + (c - cs));
                b0 = *reinterpret_cast<const bf16x8*>(xs);
                if (c + 8 < cin_pad) b1 = *reinterpret_cast<const bf16x8*>(xs + 8);
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            At[h16 + j][q] = a0[j];
            At[h16 + 8 + j][q] = a1[j];
            Bt[h16 + j][q] = b0[j];
            Bt[h16 + 8 + j][q] = b1[j];
        }
        __syncthreads();
        const int kcol = 32 * wave + 8 * (lane >> 4);
        bf16x8 av[2], bv[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            av[h] = *reinterpret_cast<const bf16x8*>(&At[h * 16 + (lane & 15)][kcol]);
            bv[h] = *reinterpret_cast<const bf16x8*>(&Bt[h * 16 + (lane & 15)][kcol]);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[wave][a * 16 + 4 * (lane >> 4) + i][b * 16 + (lane & 15)] = acc[a][b][i];
    __syncthreads();
    for (int e = t; e < 1024; e += kB) {
        const int row = e >> 5, col = e & 31;
        if (ci0 + col < cin_pad) {
            const float s = (red[0][row][col] + red[1][row][col]) + (red[2][row][col] + red[3][row][col]);
            atomicAdd(&dW[((size_t)(co0 + row) * 9 + tap) * cin_pad + ci0 + col], s);
        }
    }
}

// out[c] += sum over pixels of (d ? d[p] : 1) X[p][c]; C divides kB
__global__ void __launch_bounds__(kB) k_chan_sum(const __bf16* __restrict__ X, const float* __restrict__ d,
                                                 long long P, int C, float* __restrict__ out)
{
    __shared__ float sh[kB];
    const int c = threadIdx.x % C, lanes = kB / C, sub = threadIdx.x / C;
    float s = 0.f;
    for (long long p = (long long)blockIdx.x * lanes + sub; p < P; p += (long long)gridDim.x * lanes)
        s += (d ? d[p] : 1.f) * static_cast<float>(X[(size_t)p * C + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    if (sub == 0) {
        for (int k = 1; k < lanes; ++k) s += sh[k * C + c];
        atomicAdd(&out[c], s);
    }
}

__global__ void __launch_bounds__(kB) k_sum(const float* __restrict__ d, long long P, float* __restrict__ out)
{
    __shared__ float sh[kB];
    float s = 0.f;
    for (long long p = blockIdx.x * (long long)kB + threadIdx.x; p < P; p += (long long)gridDim.x * kB) s += d[p];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}

// FP32 input [P][cin] -> BF16 [P][cin_pad]
__global__ void k_pack_input(const float* __restrict__ x, long long P, int cin, int cin_pad, __bf16* __restrict__ y)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= P * cin_pad) return;
    const long long p = i / cin_pad;
    const int c = static_cast<int>(i - p * cin_pad);
    y[i] = __bf16(c < cin ? x[p * cin + c] : 0.f);
}

// 2x2 average pooling, BF16 NHWC, C a multiple of 8: a thread owns 8
// consecutive channels of one output pixel (16-byte loads and store)
__global__ void k_pool(const __bf16* __restrict__ x, int N, int H, int W, int C, __bf16* __restrict__ y)
{
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const long long total = (long long)N * Ho * Wo * C8;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = static_cast<int>(i % C8) * 8;
    long long t = i / C8;
    const int xo = static_cast<int>(t % Wo);
    t /= Wo;
    const int yo = static_cast<int>(t % Ho);
    const int n = static_cast<int>(t / Ho);
    const __bf16* b = x + (((size_t)n * H + 2 * yo) * W + 2 * xo) * C + c;
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(b);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(b + C);
    const bf16x8 a2 = *reinterpret_cast<const bf16x8*>(b + (size_t)W * C);
    const bf16x8 a3 = *reinterpret_cast<const bf16x8*>(b + (size_t)W * C + C);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float s = (static_cast<float>(a0[k]) + static_cast<float>(a1[k])) +
                        (static_cast<float>(a2[k]) + static_cast<float>(a3[k]));
        o[k] = __bf16(0.25f * s);
    }
    *reinterpret_cast<bf16x8*>(y + (((size_t)n * Ho + yo) * Wo + xo) * C + c) = o;
}

// g_x[y][x] += g_y[y / 2][x / 2] / 4 (FP32, fine grid)
__global__ void k_pool_bwd(const float* __restrict__ gy, int N, int H, int W, int C, float* __restrict__ gx)
{
    const long long total = (long long)N * H * W * C;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = static_cast<int>(i % C);
    long long t = i / C;
    const int x = static_cast<int>(t % W);
    t /= W;
    const int y = static_cast<int>(t % H);
    const int n = static_cast<int>(t / H);
    gx[i] += 0.25f * gy[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

// concat backward, coarse part: g_c[y][x] = sum of the 4 fine gradients
__global__ void k_up_bwd_coarse(const float* __restrict__ gu, int N, int H, int W, int Cs, int Ck,
                                float* __restrict__ gc)
{
    const int C = Cs + Ck, Hc = H / 2, Wc = W / 2;
    const long long total = (long long)N * Hc * Wc * Cs;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = static_cast<int>(i % Cs);
    long long t = i / Cs;
    const int x = static_cast<int>(t % Wc);
    t /= Wc;
    const int y = static_cast<int>(t % Hc);
    const int n = static_cast<int>(t / Hc);
    const float* b = gu + (((size_t)n * H + 2 * y) * W + 2 * x) * C + c;
    gc[i] = (b[0] + b[C]) + (b[(size_t)W * C] + b[(size_t)W * C + C]);
}

// concat backward, skip part
__global__ void k_up_bwd_skip(const float* __restrict__ gu, long long P, int Cs, int Ck, float* __restrict__ gk)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= P * Ck) return;
    const long long p = i / Ck;
    const int c = static_cast<int>(i - p * Ck);
    gk[i] = gu[p * (Cs + Ck) + Cs + c];
}

// gb = bf16(g * (y > 0)): the ReLU gradient at the stored output
__global__ void k_mask_grad(const float* __restrict__ g, const __bf16* __restrict__ y, long long n,
                            __bf16* __restrict__ gb)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) gb[i] = __bf16(static_cast<float>(y[i]) > 0.f ? g[i] : 0.f);
}

// head: out[p] = sum_c w[c] x[p][c] + b (FP32)
// C a multiple of 8: the pixel's channels are read as 16-byte vectors, all
// issued before the sum, which adds them in channel order
__global__ void k_head(const __bf16* __restrict__ x, long long P, int C, const float* __restrict__ w,
                       const float* __restrict__ b, float* __restrict__ out)
{
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= P) return;
    float s = b[0];
    const bf16x8* xv = reinterpret_cast<const bf16x8*>(x + p * C);
    int c = 0;
    for (; c + 32 <= C; c += 32) {
        bf16x8 v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = xv[c / 8 + j];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < 8; ++k) s += w[c + 8 * j + k] * static_cast<float>(v[j][k]);
    }
    for (; c < C; c += 8) {
        const bf16x8 v = xv[c / 8];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += w[c + k] * static_cast<float>(v[k]);
    }
    out[p] = s;
}

// loss: err = out - target on masked pixels (0 elsewhere); FP64 sums of
// err^2 and of the mask
__global__ void __launch_bounds__(kB) k_loss(const float* __restrict__ out, const float* __restrict__ target,
                                             const uint8_t* __restrict__ mask, long long P, float* __restrict__ err,
                                             double* __restrict__ sums)
{
    __shared__ double s0[kB], s1[kB];
    double a = 0.0, c = 0.0;
    for (long long p = blockIdx.x * (long long)kB + threadIdx.x; p < P; p += (long long)gridDim.x * kB) {
        const bool m = mask ? mask[p] != 0 : true;
        const float e = m ? out[p] - target[p] : 0.f;
        err[p] = e;
        a += (double)e * e;
        c += m ? 1.0 : 0.0;
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] += s1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], s0[0]);
        atomicAdd(&sums[1], s1[0]);
    }
}

// d(loss)/d(out) = 2 err / count, in place
__global__ void k_dout(float* __restrict__ err, const double* __restrict__ sums, long long P)
{
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float scale = sums[1] > 0.0 ? static_cast<float>(2.0 / sums[1]) : 0.f;
    err[p] *= scale;
}

// head input gradient: g[p][c] = d[p] w[c]
__global__ void k_head_gx(const float* __restrict__ d, const float* __restrict__ w, long long P, int C,
                          float* __restrict__ g)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= P * C) return;
    g[i] = d[i / C] * w[i % C];
}

// wflip[ci][tap' * cout + co] = W[co][(8 - tap') * cin_pad + ci]: the
// input gradient is the forward convolution of dY with these weights
__global__ void k_flip(const __bf16* __restrict__ w, int cout, int cin_pad, int kpad, __bf16* __restrict__ wf)
{
    const long long total = (long long)cin_pad * 9 * cout;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int co = static_cast<int>(i % cout);
    const long long t = i / cout;
    const int tp = static_cast<int>(t % 9);
    const int ci = static_cast<int>(t / 9);
    wf[i] = w[(size_t)co * kpad + (8 - tp) * cin_pad + ci];
}

// FP32 parameters [cout][9 * cin_pad] -> BF16 [cout][kpad] (zero tail)
__global__ void k_to_bf16(const float* __restrict__ p, int cout, int k9, int kpad, __bf16* __restrict__ w)
{
    const long long total = (long long)cout * kpad;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = static_cast<int>(i % kpad);
    const int co = static_cast<int>(i / kpad);
    w[i] = __bf16(k < k9 ? p[(size_t)co * k9 + k] : 0.f);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_zero(float* __restrict__ a, long long n)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.f;
}

// stored activations: 0 the packed input; per level l the two encoder
// outputs, the pooled encoder output, (unused slot) and the two decoder
// outputs; the decoder's concatenated input is never stored (its
// convolutions read the upsampled coarse tensor and the skip in place)
inline int S_X0() { return 0; }
inline int S_EA(int l) { return 1 + 6 * l; }
inline int S_EB(int l) { return 2 + 6 * l; }
inline int S_P(int l) { return 3 + 6 * l; }
inline int S_DA(int l) { return 5 + 6 * l; }
inline int S_DB(int l) { return 6 + 6 * l; }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
ConvLayer UNet::make_layer(int cin, int cout)
{
    ConvLayer L;
    L.cin = cin;
    L.cin_pad = (cin + 7) / 8 * 8;
    L.cout = cout;
    L.kpad = (9 * L.cin_pad + 31) / 32 * 32;
    L.woff = nparam_;
    nparam_ += (size_t)cout * 9 * L.cin_pad;
    L.boff = nparam_;
    nparam_ += cout;
    return L;
}

UNet::UNet(int in_channels, int base, int levels, uint64_t seed) : in_ch_(in_channels), base_(base), levels_(levels)
{
    if (levels < 2 || levels > 6 || base < 32 || base % 32 != 0 || (base << (levels - 1)) > 256 || in_channels < 1 ||
        in_channels > 32)
        throw std::invalid_argument(
            "UNet: need 2 <= levels <= 6, base a multiple of 32 with base * 2^(levels-1) <= 256, 1..32 input channels");
    for (int l = 0; l < levels; ++l) {
        const int c = base << l;
        enc_a.push_back(make_layer(l == 0 ? in_channels : (base << (l - 1)), c));
        enc_b.push_back(make_layer(c, c));
    }
    for (int l = 0; l < levels - 1; ++l) {
        const int c = base << l;
        dec_a.push_back(make_layer((base << (l + 1)) + c, c));
        dec_b.push_back(make_layer(c, c));
    }
    head_w = nparam_;
    nparam_ += base;
    head_b = nparam_;
    nparam_ += 1;
    // He-normal weights (fan_in = 9 cin), zero biases; padded input
    // channels stay zero (their gradient is zero, so Adam keeps them there)
    std::vector<float> h(nparam_, 0.f);
    std::mt19937_64 rng(seed);
    auto init = [&](const ConvLayer& L) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * L.cin)));
        for (int co = 0; co < L.cout; ++co)
            for (int t = 0; t < 9; ++t)
                for (int ci = 0; ci < L.cin; ++ci) h[L.woff + ((size_t)co * 9 + t) * L.cin_pad + ci] = nd(rng);
    };
    for (auto* v : {&enc_a, &enc_b, &dec_a, &dec_b})
        for (auto& L : *v) init(L);
    std::normal_distribution<float> nh(0.f, std::sqrt(1.f / base));
    for (int c = 0; c < base; ++c) h[head_w + c] = nh(rng);
    params.reserve(nparam_);
    grads.reserve(nparam_);
    adam_m.reserve(nparam_);
    adam_v.reserve(nparam_);
    LSDTT_HIP(hipMemcpy(params.p, h.data(), sizeof(float) * nparam_, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemset(grads.p, 0, sizeof(float) * nparam_));
    LSDTT_HIP(hipMemset(adam_m.p, 0, sizeof(float) * nparam_));
    LSDTT_HIP(hipMemset(adam_v.p, 0, sizeof(float) * nparam_));
    size_t off = 0, maxflip = 0;
    for (auto* v : {&enc_a, &enc_b, &dec_a, &dec_b})
        for (auto& L : *v) {
            wbf_off.push_back(off);
            off += (size_t)L.cout * L.kpad;
            maxflip = std::max(maxflip, (size_t)L.cin_pad * 9 * L.cout);
        }
    wbf.reserve(off);
    wflip.reserve(maxflip);
    loss_sum.reserve(2);
    hpin.reserve(2);
    act.resize(6 * levels + 1);
    gskip.resize(levels);
    convert_weights(nullptr);
    LSDTT_HIP(hipDeviceSynchronize());
}

bool UNet::shape_ok(int H, int W) const
{
    const int f = 1 << (levels_ - 1);
    return H > 0 && W > 0 && H % f == 0 && W % f == 0 && ((long long)(H / f) * (W / f)) % 16 == 0;
}

double UNet::conv_flops(int N, int H, int W) const
{
    double f = 0.0;
    for (int l = 0; l < levels_; ++l) {
        const double px = (double)N * (H >> l) * (W >> l);
        f += 18.0 * px * ((double)enc_a[l].cin * enc_a[l].cout + (double)enc_b[l].cin * enc_b[l].cout);
        if (l < levels_ - 1)
            f += 18.0 * px * ((double)dec_a[l].cin * dec_a[l].cout + (double)dec_b[l].cin * dec_b[l].cout);
    }
    return f;
}

void UNet::convert_weights(hipStream_t s)
{
    size_t i = 0;
    for (auto* v : {&enc_a, &enc_b, &dec_a, &dec_b})
        for (auto& L : *v) {
            k_to_bf16<<<nblk((long long)L.cout * L.kpad), kB, 0, s>>>(params.p + L.woff, L.cout, 9 * L.cin_pad,
                                                                      L.kpad, wbf.p + wbf_off[i]);
            ++i;
        }
    LSDTT_LAUNCH_CHECK();
}

void UNet::get_params(float* host) const
{
    LSDTT_HIP(hipMemcpy(host, params.p, sizeof(float) * nparam_, hipMemcpyDeviceToHost));
}

void UNet::set_params(const float* host)
{
    LSDTT_HIP(hipMemcpy(params.p, host, sizeof(float) * nparam_, hipMemcpyHostToDevice));
    convert_weights(nullptr);
    LSDTT_HIP(hipDeviceSynchronize());
}

void UNet::get_grads(float* host) const
{
    LSDTT_HIP(hipMemcpy(host, grads.p, sizeof(float) * nparam_, hipMemcpyDeviceToHost));
}

namespace {
// blocks of `k` (256 threads) that the device holds at once
template <class K>
int resident_conv_blocks(K k)
{
    int dev = 0, cus = 0, per = 0;
    LSDTT_HIP(hipGetDevice(&dev));
    LSDTT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    LSDTT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, kB, 0));
    return std::max(1, per) * std::max(1, cus);
}

// y = conv3x3([upsample2(u), x]) (+ bias, ReLU), u with cs channels (cs = 0:
// no u): the LDS block kernel when cin is a multiple of 32 (64-channel
// output blocks when cout allows), the direct kernel otherwise (the
// 8-channel input layer, which has no u)
template <bool RELU, bool OUT_BF16>
void launch_conv(const __bf16* u, int cs, const __bf16* x, const __bf16* w, const float* bias, void* y, int N, int H,
                 int W, int cin, int cout, int kpad, hipStream_t s)
{
    if (cin % 32 == 0 && cs % 32 == 0 && cout % 32 == 0) {
        // the widest channel block that divides cout: 128, 64 or 32 channels
        // over 8, 16 or 32 rows
        const int wc = cout % 128 == 0 ? 4 : (cout % 64 == 0 ? 2 : 1);
        const bool one = cin == 32;
        const int rb = wc == 4 ? ConvRowsShape<4, 8>::RB
                               : (wc == 2 ? ConvRowsShape<2, 8>::RB
                                          : (one ? ConvRowsShape<1, 4>::RB : ConvRowsShape<1, 6>::RB));
        const int tx = (W + kTile - 1) / kTile, ty = (H + rb - 1) / rb;
        const long long nt = (long long)tx * ty * N;
        if (nt >= (1LL << 31)) throw std::invalid_argument("conv3x3: too many pixel blocks");
        const int gy = cout / (32 * wc);
        // one input chunk: persistent blocks (what the device holds at once,
        // shared among the channel blocks), each loading the next pixel
        // block while it computes this one; several chunks: a block per
        // pixel block (the chunk pipeline hides the loads)
        auto launch = [&](auto k) {
            const long long gx = one ? std::max(1LL, std::min(nt, (long long)resident_conv_blocks(k) / gy)) : nt;
            k<<<dim3((unsigned)gx, (unsigned)gy), kB, 0, s>>>(u, cs, x, w, bias, y, N, H, W, cin, cout, kpad, tx, ty);
        };
        if (one) {
            if (wc == 4) launch(k_conv_tile<RELU, OUT_BF16, 4, 8, true>);
            else if (wc == 2) launch(k_conv_tile<RELU, OUT_BF16, 2, 8, true>);
            else launch(k_conv_tile<RELU, OUT_BF16, 1, 4, true>);
        } else {
            if (wc == 4) launch(k_conv_tile<RELU, OUT_BF16, 4, 8, false>);
            else if (wc == 2) launch(k_conv_tile<RELU, OUT_BF16, 2, 8, false>);
            else launch(k_conv_tile<RELU, OUT_BF16, 1, 6, false>);
        }
    } else {
        if (cs != 0) throw std::logic_error("conv3x3: an upsampled input needs cin and cs multiples of 32");
        const long long P = (long long)N * H * W;
        dim3 g((unsigned)((P + 63) / 64), (unsigned)(cout / 32));
        k_conv<RELU, OUT_BF16><<<g, kB, 0, s>>>(x, w, bias, y, N, H, W, cin, cout, kpad);
    }
    LSDTT_LAUNCH_CHECK();
}
}  // namespace

void UNet::conv(const ConvLayer& L, size_t wo, const __bf16* u, int cs, const __bf16* x, int N, int H, int W,
                __bf16* y, hipStream_t s)
{
    launch_conv<true, true>(u, cs, x, wbf.p + wo, params.p + L.boff, y, N, H, W, L.cin_pad, L.cout, L.kpad, s);
}

void UNet::forward(const float* x, int N, int H, int W, float* y, hipStream_t s)
{
    if (N < 1 || !shape_ok(H, W))
        throw std::invalid_argument("UNet: tile sides must be divisible by 2^(levels-1), coarsest pixel count by 16");
    // the concatenated level-0 tensor (3 base channels per pixel) and every
    // flat index of it stay below 2^31 elements per pixel row arithmetic
    if ((long long)N * H * W * 3 * base_ >= (1LL << 31))
        throw std::invalid_argument("UNet: batch too large (N * H * W * 3 * base must stay below 2^31)");
    N_ = N;
    H_ = H;
    W_ = W;
    const int L = levels_;
    auto px = [&](int l) { return (long long)N * (H >> l) * (W >> l); };
    act[S_X0()].reserve(px(0) * enc_a[0].cin_pad);
    for (int l = 0; l < L; ++l) {
        const int c = base_ << l;
        act[S_EA(l)].reserve(px(l) * c);
        act[S_EB(l)].reserve(px(l) * c);
        if (l < L - 1) {
            act[S_P(l)].reserve(px(l + 1) * c);
            act[S_DA(l)].reserve(px(l) * c);
            act[S_DB(l)].reserve(px(l) * c);
        }
    }
    out.reserve(px(0));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (time_convs) {
        LSDTT_HIP(hipEventCreate(&e0));
        LSDTT_HIP(hipEventCreate(&e1));
    }
    k_pack_input<<<nblk(px(0) * enc_a[0].cin_pad), kB, 0, s>>>(x, px(0), in_ch_, enc_a[0].cin_pad, act[S_X0()].p);
    if (time_convs) LSDTT_HIP(hipEventRecord(e0, s));
    for (int l = 0; l < L; ++l) {
        const int Hl = H >> l, Wl = W >> l;
        const __bf16* in = l == 0 ? act[S_X0()].p : act[S_P(l - 1)].p;
        conv(enc_a[l], wbf_off[l], nullptr, 0, in, N, Hl, Wl, act[S_EA(l)].p, s);
        conv(enc_b[l], wbf_off[L + l], nullptr, 0, act[S_EA(l)].p, N, Hl, Wl, act[S_EB(l)].p, s);
        if (l < L - 1)
            k_pool<<<nblk(px(l + 1) * (base_ << l) / 8), kB, 0, s>>>(act[S_EB(l)].p, N, Hl, Wl, base_ << l,
                                                                 act[S_P(l)].p);
    }
    for (int l = L - 2; l >= 0; --l) {
        const int Hl = H >> l, Wl = W >> l, cs = base_ << (l + 1);
        const __bf16* coarse = (l == L - 2) ? act[S_EB(L - 1)].p : act[S_DB(l + 1)].p;
        // input [upsample2(coarse), e_b[l]], read in place by the convolution
        conv(dec_a[l], wbf_off[2 * L + l], coarse, cs, act[S_EB(l)].p, N, Hl, Wl, act[S_DA(l)].p, s);
        conv(dec_b[l], wbf_off[3 * L - 1 + l], nullptr, 0, act[S_DA(l)].p, N, Hl, Wl, act[S_DB(l)].p, s);
    }
    if (time_convs) LSDTT_HIP(hipEventRecord(e1, s));
    k_head<<<nblk(px(0)), kB, 0, s>>>(act[S_DB(0)].p, px(0), base_, params.p + head_w, params.p + head_b, out.p);
    LSDTT_LAUNCH_CHECK();
    if (y) LSDTT_HIP(hipMemcpyAsync(y, out.p, sizeof(float) * px(0), hipMemcpyDeviceToDevice, s));
    if (time_convs) {
        LSDTT_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LSDTT_HIP(hipEventElapsedTime(&ms, e0, e1));
        conv_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
}

void UNet::zero_grad(hipStream_t s)
{
    k_zero<<<nblk((long long)nparam_), kB, 0, s>>>(grads.p, (long long)nparam_);
    LSDTT_LAUNCH_CHECK();
}

// one convolution's backward: g_out (FP32, gradient at the post-ReLU output
// y) -> weight and bias gradients, and g_in (FP32, gradient at x) unless null
void UNet::conv_backward(const ConvLayer& L, size_t wo, const __bf16* u, int cs, const __bf16* x, const __bf16* y,
                         int N, int H, int W, const float* g_out, float* g_in, hipStream_t s)
{
    const long long P = (long long)N * H * W;
    k_mask_grad<<<nblk(P * L.cout), kB, 0, s>>>(g_out, y, P * L.cout, gb.p);
    dim3 gw((unsigned)(L.cout / 32), (unsigned)(9 * ((L.cin_pad + 31) / 32)),
            (unsigned)((P + kWgChunk - 1) / kWgChunk));
    k_wgrad<<<gw, kB, 0, s>>>(gb.p, u, cs, x, grads.p + L.woff, N, H, W, L.cin_pad, L.cout);
    k_chan_sum<<<std::min(nblk(P * L.cout), 1024), kB, 0, s>>>(gb.p, nullptr, P, L.cout, grads.p + L.boff);
    if (g_in) {
        if (L.cin_pad % 32 != 0) throw std::logic_error("UNet: input gradient needs cin_pad % 32 == 0");
        k_flip<<<nblk((long long)L.cin_pad * 9 * L.cout), kB, 0, s>>>(wbf.p + wo, L.cout, L.cin_pad, L.kpad,
                                                                      wflip.p);
        launch_conv<false, false>(nullptr, 0, gb.p, wflip.p, nullptr, g_in, N, H, W, L.cout, L.cin_pad, 9 * L.cout, s);
    }
    LSDTT_LAUNCH_CHECK();
}

double UNet::backward(const float* target, const uint8_t* loss_mask, hipStream_t s)
{
    if (N_ == 0) throw std::logic_error("UNet: backward before forward");
    const int N = N_, H = H_, W = W_, L = levels_;
    auto px = [&](int l) { return (long long)N * (H >> l) * (W >> l); };
    size_t big = 0;
    for (int l = 0; l < L; ++l) big = std::max(big, (size_t)px(l) * (3 * (base_ << l)));
    g0.reserve(big);
    g1.reserve(big);
    g2.reserve(big);
    gb.reserve(big);
    dout.reserve(px(0));
    for (int l = 0; l < L - 1; ++l) gskip[l].reserve(px(l) * (base_ << l));
    LSDTT_HIP(hipMemsetAsync(loss_sum.p, 0, sizeof(double) * 2, s));
    k_loss<<<std::min(nblk(px(0)), 2048), kB, 0, s>>>(out.p, target, loss_mask, px(0), dout.p, loss_sum.p);
    k_dout<<<nblk(px(0)), kB, 0, s>>>(dout.p, loss_sum.p, px(0));
    k_chan_sum<<<std::min(nblk(px(0) * base_), 1024), kB, 0, s>>>(act[S_DB(0)].p, dout.p, px(0), base_,
                                                                   grads.p + head_w);
    k_sum<<<std::min(nblk(px(0)), 1024), kB, 0, s>>>(dout.p, px(0), grads.p + head_b);
    k_head_gx<<<nblk(px(0) * base_), kB, 0, s>>>(dout.p, params.p + head_w, px(0), base_, g0.p);
    LSDTT_LAUNCH_CHECK();
    // decoder, in reverse order of the forward (fine level first):
    // g0 holds the gradient at d_b[l]
    for (int l = 0; l <= L - 2; ++l) {
        const int Hl = H >> l, Wl = W >> l, c = base_ << l, cs = base_ << (l + 1);
        const __bf16* coarse = (l == L - 2) ? act[S_EB(L - 1)].p : act[S_DB(l + 1)].p;
        conv_backward(dec_b[l], wbf_off[3 * L - 1 + l], nullptr, 0, act[S_DA(l)].p, act[S_DB(l)].p, N, Hl, Wl, g0.p,
                      g1.p, s);
        conv_backward(dec_a[l], wbf_off[2 * L + l], coarse, cs, act[S_EB(l)].p, act[S_DA(l)].p, N, Hl, Wl, g1.p,
                      g2.p, s);
        k_up_bwd_skip<<<nblk(px(l) * c), kB, 0, s>>>(g2.p, px(l), cs, c, gskip[l].p);
        // the coarse part is the gradient at d_b[l + 1] (e_b[L - 1] at the bottom)
        k_up_bwd_coarse<<<nblk(px(l + 1) * cs), kB, 0, s>>>(g2.p, N, Hl, Wl, cs, c, g0.p);
        LSDTT_LAUNCH_CHECK();
    }
    // encoder, coarse to fine: gcur holds the full gradient at e_b[l]
    const float* gcur = g0.p;
    for (int l = L - 1; l >= 0; --l) {
        const int Hl = H >> l, Wl = W >> l;
        conv_backward(enc_b[l], wbf_off[L + l], nullptr, 0, act[S_EA(l)].p, act[S_EB(l)].p, N, Hl, Wl, gcur, g1.p,
                      s);
        if (l == 0) {
            conv_backward(enc_a[0], wbf_off[0], nullptr, 0, act[S_X0()].p, act[S_EA(0)].p, N, Hl, Wl, g1.p, nullptr,
                          s);
        } else {
            // gradient at the pooled e_b[l - 1], then added to its skip gradient
            conv_backward(enc_a[l], wbf_off[l], nullptr, 0, act[S_P(l - 1)].p, act[S_EA(l)].p, N, Hl, Wl, g1.p, g2.p,
                          s);
            const int c = base_ << (l - 1);
            k_pool_bwd<<<nblk(px(l - 1) * c), kB, 0, s>>>(g2.p, N, H >> (l - 1), W >> (l - 1), c, gskip[l - 1].p);
            LSDTT_LAUNCH_CHECK();
            gcur = gskip[l - 1].p;
        }
    }
    LSDTT_HIP(hipMemcpyAsync(hpin.p, loss_sum.p, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    return hpin.p[1] > 0.0 ? hpin.p[0] / hpin.p[1] : 0.0;
}

void conv3x3_layer(const float* x, const float* w, const float* b, int N, int H, int W, int cin, int cout, int relu,
                   float* y, const float* gy, float* gx, float* gw, float* gb)
{
    if (N < 1 || H < 1 || W < 1 || cin < 1 || cout < 32 || cout % 32 != 0 || ((long long)N * H * W) % 16 != 0)
        throw std::invalid_argument("conv3x3 layer: cout a multiple of 32, N*H*W a multiple of 16");
    if (gx && cin % 32 != 0) throw std::invalid_argument("conv3x3 layer: the input gradient needs cin % 32 == 0");
    const int cin_pad = (cin + 7) / 8 * 8, kpad = (9 * cin_pad + 31) / 32 * 32;
    const long long P = (long long)N * H * W;
    dbuf<float> fx, fw, fb, fgy, fg, fgx;
    dbuf<__bf16> bx, bw, by, bgm, bflip;
    fx.reserve(P * cin);
    fw.reserve((size_t)cout * 9 * cin_pad);
    fb.reserve(cout);
    bx.reserve(P * cin_pad);
    bw.reserve((size_t)cout * kpad);
    by.reserve(P * cout);
    // weights arrive as [cout][9][cin]; stored with the channel padding
    std::vector<float> wp((size_t)cout * 9 * cin_pad, 0.f);
    for (int co = 0; co < cout; ++co)
        for (int t = 0; t < 9; ++t)
            for (int ci = 0; ci < cin; ++ci) wp[((size_t)co * 9 + t) * cin_pad + ci] = w[((size_t)co * 9 + t) * cin + ci];
    LSDTT_HIP(hipMemcpy(fx.p, x, sizeof(float) * P * cin, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemcpy(fw.p, wp.data(), sizeof(float) * wp.size(), hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemcpy(fb.p, b, sizeof(float) * cout, hipMemcpyHostToDevice));
    k_pack_input<<<nblk(P * cin_pad), kB>>>(fx.p, P, cin, cin_pad, bx.p);
    k_to_bf16<<<nblk((long long)cout * kpad), kB>>>(fw.p, cout, 9 * cin_pad, kpad, bw.p);
    if (relu) launch_conv<true, true>(nullptr, 0, bx.p, bw.p, fb.p, by.p, N, H, W, cin_pad, cout, kpad, nullptr);
    else launch_conv<false, true>(nullptr, 0, bx.p, bw.p, fb.p, by.p, N, H, W, cin_pad, cout, kpad, nullptr);
    {
        std::vector<__bf16> hy((size_t)P * cout);
        LSDTT_HIP(hipMemcpy(hy.data(), by.p, sizeof(__bf16) * hy.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < hy.size(); ++i) y[i] = static_cast<float>(hy[i]);
    }
    if (!gy) return;
    fgy.reserve(P * cout);
    bgm.reserve(P * cout);
    fg.reserve((size_t)cout * 9 * cin_pad + cout);
    LSDTT_HIP(hipMemcpy(fgy.p, gy, sizeof(float) * P * cout, hipMemcpyHostToDevice));
    LSDTT_HIP(hipMemset(fg.p, 0, sizeof(float) * ((size_t)cout * 9 * cin_pad + cout)));
    // without the ReLU every output passes its gradient
    if (relu) k_mask_grad<<<nblk(P * cout), kB>>>(fgy.p, by.p, P * cout, bgm.p);
    else k_pack_input<<<nblk(P * cout), kB>>>(fgy.p, P * cout, 1, 1, bgm.p);
    dim3 gw2((unsigned)(cout / 32), (unsigned)(9 * ((cin_pad + 31) / 32)), (unsigned)((P + kWgChunk - 1) / kWgChunk));
    k_wgrad<<<gw2, kB>>>(bgm.p, nullptr, 0, bx.p, fg.p, N, H, W, cin_pad, cout);
    k_chan_sum<<<std::min(nblk(P * cout), 1024), kB>>>(bgm.p, nullptr, P, cout, fg.p + (size_t)cout * 9 * cin_pad);
    if (gx) {
        fgx.reserve(P * cin_pad);
        bflip.reserve((size_t)cin_pad * 9 * cout);
        k_flip<<<nblk((long long)cin_pad * 9 * cout), kB>>>(bw.p, cout, cin_pad, kpad, bflip.p);
        launch_conv<false, false>(nullptr, 0, bgm.p, bflip.p, nullptr, fgx.p, N, H, W, cout, cin_pad, 9 * cout,
                                  nullptr);
    }
    LSDTT_LAUNCH_CHECK();
    LSDTT_HIP(hipDeviceSynchronize());
    std::vector<float> hg((size_t)cout * 9 * cin_pad + cout);
    LSDTT_HIP(hipMemcpy(hg.data(), fg.p, sizeof(float) * hg.size(), hipMemcpyDeviceToHost));
    for (int co = 0; co < cout; ++co)
        for (int t = 0; t < 9; ++t)
            for (int ci = 0; ci < cin; ++ci) gw[((size_t)co * 9 + t) * cin + ci] = hg[((size_t)co * 9 + t) * cin_pad + ci];
    for (int co = 0; co < cout; ++co) gb[co] = hg[(size_t)cout * 9 * cin_pad + co];
    if (gx) LSDTT_HIP(hipMemcpy(gx, fgx.p, sizeof(float) * P * cin, hipMemcpyDeviceToHost));
}

void UNet::adam_step(double lr, double beta1, double beta2, double eps, hipStream_t s)
{
    ++adam_t;
    const double c1 = 1.0 - std::pow(beta1, (double)adam_t), c2 = 1.0 - std::pow(beta2, (double)adam_t);
    k_adam<<<nblk((long long)nparam_), kB, 0, s>>>(params.p, grads.p, adam_m.p, adam_v.p, (long long)nparam_,
                                                   (float)lr, (float)beta1, (float)beta2, (float)eps, (float)c1,
                                                   (float)c2);
    LSDTT_LAUNCH_CHECK();
    convert_weights(s);
}

}  // namespace lsdtt_gpu
