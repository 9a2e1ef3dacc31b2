// denoise.hip — the learned denoiser: a 3-level convolutional U-Net
// (16 / 32 / 64 channels, 3x3 convolutions, skip concatenation) in BF16 on
// the matrix cores with FP32 accumulation, its training (forward, backward,
// Adam) and its inference on a frame of the stochastic sampling mode, with a
// per-tile fallback to the traced frame. No counterpart in the reference,
// whose screen-space roughness blur (rough_blur.h) it stands in for.
//
// Layout: activations NHWC in bf16 with the channel count padded to a
// multiple of 16; weights kept in fp32 (master copy, Adam moments) and
// copied to bf16 in the two orders the convolutions read: [tap][cout][cin]
// for the forward pass and [8 - tap][cin][cout] for the data gradient, which
// makes the data gradient the same kernel as the forward convolution.
//
// Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16: a wave
// computes 16 consecutive pixels of one row against all output channels,
// K = 9 taps x input channels, one 16-byte load of 8 channels per lane and
// operand. The weight gradient is the GEMM over pixels (K = pixels) per
// (tap, 16 input, 16 output channel) tile, summed over pixel chunks in a
// fixed order, so training steps are deterministic.
#include "rgrt_internal.h"

#include <cmath>
#include <cstring>
#include <random>

namespace rgrt {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                           // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }
__device__ __forceinline__ float h2f(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }
__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

constexpr int kThreads = 256;
constexpr int kConvStrips = 4; // 16-pixel strips a convolution wave computes (64 pixels of one row)

inline uint32_t grid_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// Convolution 3x3, stride 1, zero padding 1: out = act(conv(in, wt) + bias)
// in [N][H][W][KIN], wt [9][NOUT][KIN], out [N][H][W][NOUT]
// With C0 > 0 the input is the concatenation the decoder convolves (skip
// connection): channels [0, C0) are the nearest 2x upsampling of
// low [N][H/2][W/2][C0], channels [C0, KIN) come from in [N][H][W][KIN - C0];
// the concatenated tensor is never stored.
// ---------------------------------------------------------------------------
template <int KIN, int NOUT, int C0 = 0>
__global__ __launch_bounds__(kThreads) void k_conv(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wt,
                                                   const float* __restrict__ bias, uint16_t* __restrict__ out, int N, int H,
                                                   int W, int relu, const uint16_t* __restrict__ low)
{
    static_assert(KIN % 16 == 0 && NOUT % 16 == 0, "channel counts are multiples of 16");
    static_assert(C0 % 8 == 0 && C0 < KIN, "a load of 8 channels comes from one source");
    constexpr int C1 = KIN - C0;
    constexpr int KT = 9 * KIN;
    constexpr int KSTEPS = (KT + 31) / 32;
    constexpr int NT = NOUT / 16;
    constexpr int SPW = kConvStrips; // 16-pixel strips of one row per wave, sharing the weight operands
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const int stripsPerRow = (W + 15) / 16;
    const int groupsPerRow = (stripsPerRow + SPW - 1) / SPW;
    if(wave >= (long)N * H * groupsPerRow) return; // whole waves only
    const int sg = (int)(wave % groupsPerRow);
    const long row = wave / groupsPerRow;
    const int y = (int)(row % H);
    const int n = (int)(row / H);
    const int m = lane & 15, g = lane >> 4;
    const int x0 = sg * SPW * 16 + m; // pixel of this lane in the group's first strip
    f32x4 acc[SPW][NT];
#pragma unroll
    for(int q = 0; q < SPW; ++q)
#pragma unroll
        for(int j = 0; j < NT; ++j) acc[q][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u});
    for(int s = 0; s < KSTEPS; ++s) {
        const int k0 = 32 * s + 8 * g;
        const int tap = k0 / KIN;
        const int c0 = k0 - tap * KIN;
        const bool kin = k0 < KT;
        const int yy = y + tap / 3 - 1;
        const bool rowIn = kin && yy >= 0 && yy < H;
        bf16x8 b[NT];
#pragma unroll
        for(int j = 0; j < NT; ++j) {
            b[j] = zero8;
            if(kin) b[j] = __builtin_bit_cast(bf16x8, ld_global16(wt + ((long)tap * NOUT + 16 * j + m) * KIN + c0));
        }
        bf16x8 a[SPW];
#pragma unroll
        for(int q = 0; q < SPW; ++q) {
            a[q] = zero8;
            const int xx = x0 + 16 * q + tap % 3 - 1;
            if(rowIn && xx >= 0 && xx < W) {
                if(C0 > 0 && c0 < C0)
                    a[q] = __builtin_bit_cast(bf16x8,
                                              ld_global16(low + (((long)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C0 + c0));
                else
                    a[q] = __builtin_bit_cast(bf16x8, ld_global16(in + (((long)n * H + yy) * W + xx) * C1 + (c0 - C0)));
            }
        }
#pragma unroll
        for(int q = 0; q < SPW; ++q)
#pragma unroll
            for(int j = 0; j < NT; ++j) acc[q][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b[j], acc[q][j], 0, 0, 0);
    }
    // lane holds D[row 4g + i][column m] of each strip: pixel 4g + i of the
    // strip, channel 16j + m
#pragma unroll
    for(int j = 0; j < NT; ++j) {
        const int co = 16 * j + m;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for(int q = 0; q < SPW; ++q)
#pragma unroll
            for(int i = 0; i < 4; ++i) {
                const int px = (sg * SPW + q) * 16 + 4 * g + i;
                if(px < W) {
                    float v = acc[q][j][i] + bv;
                    if(relu) v = fmaxf(v, 0.0f);
                    out[(((long)n * H + y) * W + px) * NOUT + co] = f2bf(v);
                }
            }
    }
}

// Weight gradient of one layer: dW[tap][ci][co] = sum_p x[p + off(tap)][ci] * dz[p][co].
// A wave computes one (tap, 16 ci, 16 co) tile over `chunk` pixels into
// partial[chunkIndex][tile][16][16].
template <int CIN, int COUT>
__global__ __launch_bounds__(kThreads) void k_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dz,
                                                    float* __restrict__ partial, int N, int H, int W, int chunk,
                                                    int numChunks)
{
    constexpr int CT = CIN / 16, OT = COUT / 16, TILES = 9 * CT * OT;
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if(wave >= (long)TILES * numChunks) return;
    const int tile = (int)(wave % TILES);
    const int ch = (int)(wave / TILES);
    const int tap = tile / (CT * OT);
    const int rem = tile - tap * CT * OT;
    const int ct = rem / OT, ot = rem - (rem / OT) * OT;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int m = lane & 15, g = lane >> 4;
    const long P = (long)N * H * W;
    const long p0 = (long)ch * chunk;
    const long p1 = p0 + chunk < P ? p0 + chunk : P;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for(long pb = p0; pb < p1; pb += 32) {
        uint16_t av[8], bv[8];
#pragma unroll
        for(int j = 0; j < 8; ++j) {
            const long p = pb + 8 * g + j;
            av[j] = 0;
            bv[j] = 0;
            if(p < p1) {
                const int px = (int)(p % W);
                const long r = p / W;
                const int py = (int)(r % H);
                const int pn = (int)(r / H);
                const int yy = py + dy, xx = px + dx;
                if(yy >= 0 && yy < H && xx >= 0 && xx < W) av[j] = x[(((long)pn * H + yy) * W + xx) * CIN + ct * 16 + m];
                bv[j] = dz[p * COUT + ot * 16 + m];
            }
        }
        bf16x8 a, b;
#pragma unroll
        for(int j = 0; j < 8; ++j) {
            a[j] = __builtin_bit_cast(__bf16, av[j]);
            b[j] = __builtin_bit_cast(__bf16, bv[j]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    float* out = partial + ((long)ch * TILES + tile) * 256;
#pragma unroll
    for(int i = 0; i < 4; ++i) out[(4 * g + i) * 16 + m] = acc[i];
}

// dW (fp32, [9][CIN][COUT]) = sum over chunks of the partial tiles, in chunk order.
__global__ void k_wgrad_reduce(const float* __restrict__ partial, int numChunks, int cin, int cout,
                               float* __restrict__ dw)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = 9 * cin * cout;
    if(e >= total) return;
    const int tap = e / (cin * cout);
    const int r = e - tap * cin * cout;
    const int ci = r / cout, co = r - (r / cout) * cout;
    const int CT = cin / 16, OT = cout / 16, TILES = 9 * CT * OT;
    const int tile = (tap * CT + ci / 16) * OT + co / 16;
    const int idx = (ci % 16) * 16 + co % 16;
    float s = 0.0f;
    for(int c = 0; c < numChunks; ++c) s += partial[((long)c * TILES + tile) * 256 + idx];
    dw[e] = s;
}

// db[co] = sum_p dz[p][co]: one block per channel, fixed summation order.
__global__ void k_bias_grad(const uint16_t* __restrict__ dz, long P, int C, float* __restrict__ db)
{
    __shared__ float part[kThreads];
    const int co = blockIdx.x;
    float s = 0.0f;
    for(long p = threadIdx.x; p < P; p += kThreads) s += bf2f(dz[p * C + co]);
    part[threadIdx.x] = s;
    __syncthreads();
    for(int off = kThreads / 2; off > 0; off >>= 1) {
        if((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if(threadIdx.x == 0) db[co] = part[0];
}

// ---------------------------------------------------------------------------
// Elementwise parts of the network
// ---------------------------------------------------------------------------
// 2x2 average pooling: in [N][H][W][C] -> out [N][H/2][W/2][C]
__global__ void k_pool(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int N, int H, int W, int C)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = H / 2, w = W / 2;
    if(e >= (long)N * h * w * C) return;
    const int c = (int)(e % C);
    const long q = e / C;
    const int x = (int)(q % w);
    const long r = q / w;
    const int y = (int)(r % h), n = (int)(r / h);
    const long b = (((long)n * H + 2 * y) * W + 2 * x) * C + c;
    const float s = (bf2f(in[b]) + bf2f(in[b + C])) + (bf2f(in[b + (long)W * C]) + bf2f(in[b + (long)W * C + C]));
    out[e] = f2bf(s * 0.25f);
}

// out [N][H][W][C0 + C1]: channels < C0 from low [N][H/2][W/2][C0] (nearest
// upsampling), the rest from skip [N][H][W][C1].
__global__ void k_up_concat(const uint16_t* __restrict__ low, const uint16_t* __restrict__ skip, uint16_t* __restrict__ out,
                            int N, int H, int W, int C0, int C1)
{
    const int C = C0 + C1;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= (long)N * H * W * C) return;
    const int c = (int)(e % C);
    const long q = e / C;
    const int x = (int)(q % W);
    const long r = q / W;
    const int y = (int)(r % H), n = (int)(r / H);
    out[e] = c < C0 ? low[(((long)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C0 + c] : skip[q * C1 + (c - C0)];
}

// Backward of k_up_concat: dlow = sum of the four fine pixels' first C0
// channels; dskip = the remaining channels.
__global__ void k_split_down(const uint16_t* __restrict__ dcat, uint16_t* __restrict__ dlow, uint16_t* __restrict__ dskip,
                             int N, int H, int W, int C0, int C1)
{
    const int C = C0 + C1;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h = H / 2, w = W / 2;
    const long nLow = (long)N * h * w * C0;
    if(e < nLow) {
        const int c = (int)(e % C0);
        const long q = e / C0;
        const int x = (int)(q % w);
        const long r = q / w;
        const int y = (int)(r % h), n = (int)(r / h);
        const long b = (((long)n * H + 2 * y) * W + 2 * x) * C + c;
        const float s = (bf2f(dcat[b]) + bf2f(dcat[b + C])) + (bf2f(dcat[b + (long)W * C]) + bf2f(dcat[b + (long)W * C + C]));
        dlow[e] = f2bf(s);
        return;
    }
    const long f = e - nLow;
    if(f >= (long)N * H * W * C1) return;
    const int c = (int)(f % C1);
    const long q = f / C1;
    dskip[f] = dcat[q * C + C0 + c];
}

// Backward of k_pool plus the skip path's gradient: out = dp[y/2][x/2] / 4 + extra.
__global__ void k_pool_back_add(const uint16_t* __restrict__ dp, const uint16_t* __restrict__ extra, uint16_t* __restrict__ out,
                                int N, int H, int W, int C)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= (long)N * H * W * C) return;
    const int c = (int)(e % C);
    const long q = e / C;
    const int x = (int)(q % W);
    const long r = q / W;
    const int y = (int)(r % H), n = (int)(r / H);
    const float v = bf2f(dp[(((long)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]) * 0.25f + bf2f(extra[e]);
    out[e] = f2bf(v);
}

// dz = dy where the layer's output y (after ReLU) is positive.
__global__ void k_relu_back(const uint16_t* __restrict__ dy, const uint16_t* __restrict__ y, uint16_t* __restrict__ dz, long n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(e < n) dz[e] = bf2f(y[e]) > 0.0f ? dy[e] : (uint16_t)0;
}

// Mean squared error of y = noisy + o[:3] against the target, and its
// gradient with respect to o (channels 3..15 zero). One partial sum per block.
__global__ void k_loss(const uint16_t* __restrict__ o, const uint16_t* __restrict__ x0, const float* __restrict__ target,
                       long P, uint16_t* __restrict__ dz, float* __restrict__ lossPart)
{
    __shared__ float part[kThreads];
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if(p < P) {
        const float scale = 2.0f / (3.0f * (float)P);
        for(int c = 0; c < 16; ++c) {
            float gval = 0.0f;
            if(c < 3) {
                const float yv = bf2f(x0[p * 16 + c]) + bf2f(o[p * 16 + c]);
                const float d = yv - target[p * 3 + c];
                l += d * d;
                gval = d * scale;
            }
            dz[p * 16 + c] = f2bf(gval);
        }
    }
    part[threadIdx.x] = l;
    __syncthreads();
    for(int off = kThreads / 2; off > 0; off >>= 1) {
        if((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if(threadIdx.x == 0) lossPart[blockIdx.x] = part[0];
}

__global__ void k_sum(const float* __restrict__ in, int n, float scale, float* __restrict__ out)
{
    __shared__ float part[kThreads];
    float s = 0.0f;
    for(int i = threadIdx.x; i < n; i += kThreads) s += in[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for(int off = kThreads / 2; off > 0; off >>= 1) {
        if((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if(threadIdx.x == 0) *out = part[0] * scale;
}

// Adam on the fp32 parameters (bias-corrected step size precomputed on the host).
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g, long n,
                       float stepSize, float b1, float b2, float eps)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= n) return;
    const float gv = g[e];
    const float mv = b1 * m[e] + (1.0f - b1) * gv;
    const float vv = b2 * v[e] + (1.0f - b2) * gv * gv;
    m[e] = mv;
    v[e] = vv;
    w[e] = w[e] - stepSize * mv / (sqrtf(vv) + eps);
}

// bf16 weight copies of one layer from the fp32 master [9][cin][cout]:
// forward [tap][cout][cin], data gradient [8 - tap][cin][cout].
__global__ void k_prep_weights(const float* __restrict__ w, int cin, int cout, uint16_t* __restrict__ wf,
                               uint16_t* __restrict__ wb)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= 9 * cin * cout) return;
    const int tap = e / (cin * cout);
    const int r = e - tap * cin * cout;
    const int ci = r / cout, co = r - (r / cout) * cout;
    const uint16_t b = f2bf(w[e]);
    wf[((long)tap * cout + co) * cin + ci] = b;
    wb[((long)(8 - tap) * cin + ci) * cout + co] = b;
}

// Network input from a frame of the stochastic mode: noisy RGB (BASE),
// albedo and roughness (FEATURES), normal (NORMAL), depth as log(1 + t) / 10.
__global__ void k_pack(const uint16_t* __restrict__ base, const uint16_t* __restrict__ normal, const uint16_t* __restrict__ feat,
                       int H, int W, uint16_t* __restrict__ x0)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= (long)H * W) return;
    float v[16];
    for(int c = 0; c < 16; ++c) v[c] = 0.0f;
    for(int c = 0; c < 3; ++c) {
        v[c] = h2f(base[4 * p + c]);
        v[3 + c] = h2f(feat[8 * p + c]);
        v[6 + c] = h2f(normal[4 * p + c]);
    }
    v[9] = log1pf(fminf(fmaxf(h2f(feat[8 * p + 4]), 0.0f), 10000.0f)) / 10.0f;
    v[10] = h2f(feat[8 * p + 3]);
    for(int c = 0; c < 16; ++c) x0[p * 16 + c] = f2bf(v[c]);
}

// target (fp32 RGB) [+]= weight * BASE.rgb of a frame
__global__ void k_target_add(const uint16_t* __restrict__ base, long P, float weight, int accumulate, float* __restrict__ target)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= P) return;
    for(int c = 0; c < 3; ++c) {
        const float v = weight * h2f(base[4 * p + c]);
        target[p * 3 + c] = accumulate ? target[p * 3 + c] + v : v;
    }
}

// Per 16x16 tile: the denoiser's change of the displayed colour (the noisy
// RGB and the denoised RGB both clamped to [0, 1], the range the post pass
// keeps) against the frame's noise level. A tile whose mean squared change
// exceeds a * (mean variance of the pixel means) + b keeps the traced frame.
__global__ void k_tile_check(const uint16_t* __restrict__ o, const uint16_t* __restrict__ base,
                             const uint16_t* __restrict__ feat, int H, int W, float a, float b,
                             uint8_t* __restrict__ fallback)
{
    __shared__ float se[kThreads], sv[kThreads];
    const int tx = blockIdx.x, ty = blockIdx.y;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int x = tx * 16 + lx, y = ty * 16 + ly;
    float e = 0.0f, v = 0.0f, cnt = 0.0f;
    if(x < W && y < H) {
        const long p = (long)y * W + x;
        for(int c = 0; c < 3; ++c) {
            const float x = h2f(base[4 * p + c]);
            const float y = x + bf2f(o[p * 16 + c]);
            const float d = fminf(fmaxf(y, 0.0f), 1.0f) - fminf(fmaxf(x, 0.0f), 1.0f);
            e += d * d;
        }
        e /= 3.0f;
        v = h2f(feat[8 * p + 5]);
        cnt = 1.0f;
    }
    se[threadIdx.x] = e;
    sv[threadIdx.x] = v;
    __syncthreads();
    for(int off = kThreads / 2; off > 0; off >>= 1) {
        if((int)threadIdx.x < off) {
            se[threadIdx.x] += se[threadIdx.x + off];
            sv[threadIdx.x] += sv[threadIdx.x + off];
        }
        __syncthreads();
    }
    (void)cnt;
    if(threadIdx.x == 0) {
        const int tw = min(16, W - tx * 16), th = min(16, H - ty * 16);
        const float n = (float)(tw * th);
        fallback[ty * gridDim.x + tx] = (se[0] / n > a * (sv[0] / n) + b) ? 1 : 0;
    }
}

// BASE.rgb = noisy + o outside fallback tiles (alpha 0, as the stochastic
// mode leaves it); fallback tiles keep the traced frame.
__global__ void k_denoise_out(const uint16_t* __restrict__ o, const uint8_t* __restrict__ fallback, int H, int W,
                              uint16_t* __restrict__ base)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= (long)H * W) return;
    const int x = (int)(p % W), y = (int)(p / W);
    const int tilesX = (W + 15) / 16;
    if(fallback[(y / 16) * tilesX + x / 16]) return;
    for(int c = 0; c < 3; ++c) base[4 * p + c] = f2h(h2f(base[4 * p + c]) + bf2f(o[p * 16 + c]));
    base[4 * p + 3] = 0;
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
template <int KIN, int NOUT, int C0 = 0>
void conv_t(hipStream_t s, const uint16_t* in, const uint16_t* wt, const float* bias, uint16_t* out, int N, int H, int W,
            int relu, const uint16_t* low = nullptr)
{
    const long waves = (long)N * H * (((W + 15) / 16 + kConvStrips - 1) / kConvStrips);
    k_conv<KIN, NOUT, C0><<<grid_for(waves, kThreads / 64), kThreads, 0, s>>>(in, wt, bias, out, N, H, W, relu, low);
}

void conv(hipStream_t s, int kin, int nout, const uint16_t* in, const uint16_t* wt, const float* bias, uint16_t* out, int N,
          int H, int W, int relu)
{
#define RGRT_CONV(a, b)                                                                                                      \
    if(kin == a && nout == b) return conv_t<a, b>(s, in, wt, bias, out, N, H, W, relu);
    RGRT_CONV(16, 16)
    RGRT_CONV(16, 32)
    RGRT_CONV(32, 32)
    RGRT_CONV(32, 64)
    RGRT_CONV(64, 64)
    RGRT_CONV(96, 32)
    RGRT_CONV(48, 16)
    RGRT_CONV(16, 48)
    RGRT_CONV(32, 96)
    RGRT_CONV(64, 32)
    RGRT_CONV(32, 16)
#undef RGRT_CONV
    throw std::runtime_error("denoiser: no convolution for these channel counts");
}

constexpr int kWgradChunk = 2048;

template <int CIN, int COUT>
void wgrad_t(hipStream_t s, const uint16_t* x, const uint16_t* dz, float* partial, int N, int H, int W, int numChunks)
{
    constexpr int TILES = 9 * (CIN / 16) * (COUT / 16);
    k_wgrad<CIN, COUT><<<grid_for((long)TILES * numChunks, kThreads / 64), kThreads, 0, s>>>(x, dz, partial, N, H, W,
                                                                                              kWgradChunk, numChunks);
}

void wgrad(hipStream_t s, int cin, int cout, const uint16_t* x, const uint16_t* dz, float* partial, float* dw, int N, int H,
           int W)
{
    const long P = (long)N * H * W;
    const int numChunks = (int)((P + kWgradChunk - 1) / kWgradChunk);
#define RGRT_WGRAD(a, b)                                                                                                     \
    if(cin == a && cout == b) {                                                                                              \
        wgrad_t<a, b>(s, x, dz, partial, N, H, W, numChunks);                                                               \
        k_wgrad_reduce<<<grid_for(9 * a * b, kThreads), kThreads, 0, s>>>(partial, numChunks, a, b, dw);                     \
        return;                                                                                                              \
    }
    RGRT_WGRAD(16, 16)
    RGRT_WGRAD(16, 32)
    RGRT_WGRAD(32, 32)
    RGRT_WGRAD(32, 64)
    RGRT_WGRAD(64, 64)
    RGRT_WGRAD(96, 32)
    RGRT_WGRAD(48, 16)
#undef RGRT_WGRAD
    throw std::runtime_error("denoiser: no weight gradient for these channel counts");
}

// The network: 11 layers (input padded from 11 to 16 channels, output from 3 to 16).
struct LayerDef {
    int cin, cout, level, relu;
};
constexpr LayerDef kLayers[Denoiser::kNumLayers] = {
    {16, 16, 0, 1}, {16, 16, 0, 1},               // enc1
    {16, 32, 1, 1}, {32, 32, 1, 1},               // enc2 (after pooling)
    {32, 64, 2, 1}, {64, 64, 2, 1},               // bottleneck
    {96, 32, 1, 1}, {32, 32, 1, 1},               // dec2 (upsampled bottleneck + enc2)
    {48, 16, 0, 1}, {16, 16, 0, 1},               // dec1 (upsampled dec2 + enc1)
    {16, 16, 0, 0},                                // output (3 channels used)
};

} // namespace

// ---------------------------------------------------------------------------
// Denoiser
// ---------------------------------------------------------------------------
Denoiser::Denoiser(uint32_t seed)
{
    size_t off = 0;
    for(int l = 0; l < kNumLayers; ++l) {
        wOff[l] = off;
        off += 9ull * kLayers[l].cin * kLayers[l].cout;
        bOff[l] = off;
        off += kLayers[l].cout;
    }
    numParams = off;
    size_t bfOff = 0;
    for(int l = 0; l < kNumLayers; ++l) {
        wbfOff[l] = bfOff;
        bfOff += 9ull * kLayers[l].cin * kLayers[l].cout;
    }
    check(hipMalloc(&params, numParams * 4));
    check(hipMalloc(&grads, numParams * 4));
    check(hipMalloc(&adamM, numParams * 4));
    check(hipMalloc(&adamV, numParams * 4));
    check(hipMalloc(&wFwd, bfOff * 2));
    check(hipMalloc(&wBwd, bfOff * 2));
    check(hipMemset(adamM, 0, numParams * 4));
    check(hipMemset(adamV, 0, numParams * 4));
    check(hipDeviceSynchronize()); // the fills are done before any stream's work reads them
    // He initialisation of the weights over the real fan-in (11 input
    // channels in the first layer), zero biases; padded output channels of
    // the last layer stay zero (their gradient is zero too).
    std::vector<float> h(numParams, 0.0f);
    std::mt19937 rng(seed);
    for(int l = 0; l < kNumLayers; ++l) {
        const int cin = kLayers[l].cin, cout = kLayers[l].cout;
        const int realIn = l == 0 ? 11 : cin;
        const int realOut = l == kNumLayers - 1 ? 3 : cout;
        const float scale = (l == kNumLayers - 1 ? 0.1f : 1.0f) * std::sqrt(2.0f / (9.0f * realIn));
        std::normal_distribution<float> nd(0.0f, scale);
        for(int t = 0; t < 9; ++t)
            for(int ci = 0; ci < cin; ++ci)
                for(int co = 0; co < cout; ++co) {
                    const float v = nd(rng);
                    h[wOff[l] + ((size_t)t * cin + ci) * cout + co] = (ci < realIn && co < realOut) ? v : 0.0f;
                }
    }
    check(hipMemcpy(params, h.data(), numParams * 4, hipMemcpyHostToDevice));
    stepCount = 0;
}

Denoiser::~Denoiser()
{
    for(void* p: {(void*)params, (void*)grads, (void*)adamM, (void*)adamV, (void*)wFwd, (void*)wBwd, (void*)ws, (void*)dsX,
                  (void*)dsT})
        if(p) (void)hipFree(p);
}

void Denoiser::prep_weights(hipStream_t s)
{
    for(int l = 0; l < kNumLayers; ++l) {
        const int n = 9 * kLayers[l].cin * kLayers[l].cout;
        k_prep_weights<<<grid_for(n, kThreads), kThreads, 0, s>>>(params + wOff[l], kLayers[l].cin, kLayers[l].cout,
                                                                  wFwd + wbfOff[l], wBwd + wbfOff[l]);
    }
    weightsReady = true;
}

void Denoiser::set_params(hipStream_t s, const float* host)
{
    check(hipMemcpyAsync(params, host, numParams * 4, hipMemcpyHostToDevice, s));
    check(hipMemsetAsync(adamM, 0, numParams * 4, s));
    check(hipMemsetAsync(adamV, 0, numParams * 4, s));
    check(hipStreamSynchronize(s));
    stepCount = 0;
    weightsReady = false;
    paramsLoaded = true;
}

void Denoiser::get_params(hipStream_t s, float* host)
{
    check(hipStreamSynchronize(s));
    check(hipMemcpy(host, params, numParams * 4, hipMemcpyDeviceToHost));
}

// Workspace of the forward and backward tensors for N images of H x W.
void Denoiser::reserve(int N, int H, int W, bool training)
{
    if(H % 4 || W % 4) throw std::runtime_error("denoiser: image sizes must be multiples of 4");
    const size_t p0 = (size_t)N * H * W, p1 = p0 / 4, p2 = p0 / 16;
    // channel counts of every tensor, in the order of the Buf enum
    const size_t sizes[kNumBufs] = {
        p0 * 16, p0 * 16, p0 * 16, p1 * 16, p1 * 32, p1 * 32, p2 * 32, p2 * 64, p2 * 64, p1 * 96, p1 * 32, p1 * 32,
        p0 * 48, p0 * 16, p0 * 16, p0 * 16,
        // backward
        p0 * 16, p0 * 16, p0 * 48, p1 * 32, p0 * 16, p1 * 96, p2 * 64, p1 * 32, p2 * 64, p2 * 32, p1 * 32, p1 * 16,
        p0 * 16, p0 * 16, p1 * 32};
    const int used = training ? kNumBufs : kNumFwdBufs;
    size_t total = 0;
    for(int b = 0; b < used; ++b) total += (sizes[b] + 127) & ~(size_t)127;
    const size_t partialFloats = training ? (size_t)((p0 + kWgradChunk - 1) / kWgradChunk) * 9 * 6 * 4 * 256 : 0;
    const size_t bytes = total * 2 + partialFloats * 4 + 4096 * 4 + (size_t)((H + 15) / 16) * ((W + 15) / 16) + 256;
    if(bytes > wsBytes) {
        if(ws) check(hipFree(ws));
        ws = nullptr;
        check(hipMalloc(&ws, bytes));
        wsBytes = bytes;
    }
    char* p = reinterpret_cast<char*>(ws);
    for(int b = 0; b < kNumBufs; ++b) buf[b] = nullptr;
    for(int b = 0; b < used; ++b) {
        buf[b] = reinterpret_cast<uint16_t*>(p);
        p += ((sizes[b] + 127) & ~(size_t)127) * 2;
    }
    partial = training ? reinterpret_cast<float*>(p) : nullptr;
    p += partialFloats * 4;
    lossPart = reinterpret_cast<float*>(p);
    p += 4096 * 4;
    fallback = reinterpret_cast<uint8_t*>(p);
    wsN = N;
    wsH = H;
    wsW = W;
    wsTraining = training;
}

void Denoiser::forward(hipStream_t s, int N, int H, int W, bool keepConcat)
{
    if(!weightsReady) prep_weights(s);
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
    auto w = [&](int l) { return wFwd + wbfOff[l]; };
    auto b = [&](int l) { return params + bOff[l]; };
    conv(s, 16, 16, buf[X0], w(0), b(0), buf[A1], N, H, W, 1);
    conv(s, 16, 16, buf[A1], w(1), b(1), buf[S1], N, H, W, 1);
    k_pool<<<grid_for((long)N * H1 * W1 * 16, kThreads), kThreads, 0, s>>>(buf[S1], buf[P1], N, H, W, 16);
    conv(s, 16, 32, buf[P1], w(2), b(2), buf[A2], N, H1, W1, 1);
    conv(s, 32, 32, buf[A2], w(3), b(3), buf[S2], N, H1, W1, 1);
    k_pool<<<grid_for((long)N * H2 * W2 * 32, kThreads), kThreads, 0, s>>>(buf[S2], buf[P2], N, H1, W1, 32);
    conv(s, 32, 64, buf[P2], w(4), b(4), buf[A3], N, H2, W2, 1);
    conv(s, 64, 64, buf[A3], w(5), b(5), buf[B3], N, H2, W2, 1);
    // the decoder convolutions read up(b3) ++ s2 and up(u2) ++ s1 directly;
    // the concatenations are stored only for the backward pass
    if(keepConcat)
        k_up_concat<<<grid_for((long)N * H1 * W1 * 96, kThreads), kThreads, 0, s>>>(buf[B3], buf[S2], buf[C2], N, H1, W1, 64, 32);
    conv_t<96, 32, 64>(s, buf[S2], w(6), b(6), buf[A4], N, H1, W1, 1, buf[B3]);
    conv(s, 32, 32, buf[A4], w(7), b(7), buf[U2], N, H1, W1, 1);
    if(keepConcat)
        k_up_concat<<<grid_for((long)N * H * W * 48, kThreads), kThreads, 0, s>>>(buf[U2], buf[S1], buf[C1], N, H, W, 32, 16);
    conv_t<48, 16, 32>(s, buf[S1], w(8), b(8), buf[A5], N, H, W, 1, buf[U2]);
    conv(s, 16, 16, buf[A5], w(9), b(9), buf[U1], N, H, W, 1);
    conv(s, 16, 16, buf[U1], w(10), b(10), buf[OUT], N, H, W, 0);
}

// Loss (mean squared error of noisy + o against target) and the gradients
// of every parameter into grads; the loss lands in lossPart[4095].
void Denoiser::backward(hipStream_t s, const float* target, int N, int H, int W)
{
    const long P0 = (long)N * H * W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
    const long P1 = P0 / 4, P2 = P0 / 16;
    auto wb = [&](int l) { return wBwd + wbfOff[l]; };
    auto gw = [&](int l) { return grads + wOff[l]; };
    auto gb = [&](int l) { return grads + bOff[l]; };
    const int lossBlocks = (int)grid_for(P0, kThreads);
    if(lossBlocks > 4095) throw std::runtime_error("denoiser: training batch too large for the loss reduction");
    k_loss<<<lossBlocks, kThreads, 0, s>>>(buf[OUT], buf[X0], target, P0, buf[DZO], lossPart);
    k_sum<<<1, kThreads, 0, s>>>(lossPart, lossBlocks, 1.0f / (3.0f * (float)P0), lossPart + 4095);
    auto relu_back = [&](uint16_t* dy, uint16_t* y, uint16_t* dz, long n) {
        k_relu_back<<<grid_for(n, kThreads), kThreads, 0, s>>>(dy, y, dz, n);
    };
    auto layer_grads = [&](int l, const uint16_t* x, const uint16_t* dz, int n, int h, int w) {
        wgrad(s, kLayers[l].cin, kLayers[l].cout, x, dz, partial, gw(l), n, h, w);
        k_bias_grad<<<kLayers[l].cout, kThreads, 0, s>>>(dz, (long)n * h * w, kLayers[l].cout, gb(l));
    };
    // output layer (no ReLU): dz = dL/do
    layer_grads(10, buf[U1], buf[DZO], N, H, W);
    conv(s, 16, 16, buf[DZO], wb(10), nullptr, buf[GA], N, H, W, 0); // d u1
    // d1b
    relu_back(buf[GA], buf[U1], buf[GB], P0 * 16);
    layer_grads(9, buf[A5], buf[GB], N, H, W);
    conv(s, 16, 16, buf[GB], wb(9), nullptr, buf[GA], N, H, W, 0); // d a5
    // d1a (input c1 = up(u2) ++ s1)
    relu_back(buf[GA], buf[A5], buf[GB], P0 * 16);
    layer_grads(8, buf[C1], buf[GB], N, H, W);
    conv(s, 16, 48, buf[GB], wb(8), nullptr, buf[GC1], N, H, W, 0); // d c1
    k_split_down<<<grid_for(P1 * 32 + P0 * 16, kThreads), kThreads, 0, s>>>(buf[GC1], buf[GU2], buf[GS1X], N, H, W, 32, 16);
    // d2b
    relu_back(buf[GU2], buf[U2], buf[GU2], P1 * 32);
    layer_grads(7, buf[A4], buf[GU2], N, H1, W1);
    conv(s, 32, 32, buf[GU2], wb(7), nullptr, buf[GL1A], N, H1, W1, 0); // d a4
    // d2a (input c2 = up(b3) ++ s2)
    relu_back(buf[GL1A], buf[A4], buf[GL1A], P1 * 32);
    layer_grads(6, buf[C2], buf[GL1A], N, H1, W1);
    conv(s, 32, 96, buf[GL1A], wb(6), nullptr, buf[GC2], N, H1, W1, 0); // d c2
    k_split_down<<<grid_for(P2 * 64 + P1 * 32, kThreads), kThreads, 0, s>>>(buf[GC2], buf[GB3], buf[GS2X], N, H1, W1, 64, 32);
    // bb
    relu_back(buf[GB3], buf[B3], buf[GB3], P2 * 64);
    layer_grads(5, buf[A3], buf[GB3], N, H2, W2);
    conv(s, 64, 64, buf[GB3], wb(5), nullptr, buf[GL2A], N, H2, W2, 0); // d a3
    // ba (input p2)
    relu_back(buf[GL2A], buf[A3], buf[GL2A], P2 * 64);
    layer_grads(4, buf[P2], buf[GL2A], N, H2, W2);
    conv(s, 64, 32, buf[GL2A], wb(4), nullptr, buf[GP2], N, H2, W2, 0); // d p2
    k_pool_back_add<<<grid_for(P1 * 32, kThreads), kThreads, 0, s>>>(buf[GP2], buf[GS2X], buf[GS2], N, H1, W1, 32); // d s2
    // e2b
    relu_back(buf[GS2], buf[S2], buf[GS2], P1 * 32);
    layer_grads(3, buf[A2], buf[GS2], N, H1, W1);
    conv(s, 32, 32, buf[GS2], wb(3), nullptr, buf[GL1A], N, H1, W1, 0); // d a2
    // e2a (input p1)
    relu_back(buf[GL1A], buf[A2], buf[GL1A], P1 * 32);
    layer_grads(2, buf[P1], buf[GL1A], N, H1, W1);
    conv(s, 32, 16, buf[GL1A], wb(2), nullptr, buf[GP1], N, H1, W1, 0); // d p1
    k_pool_back_add<<<grid_for(P0 * 16, kThreads), kThreads, 0, s>>>(buf[GP1], buf[GS1X], buf[GS1], N, H, W, 16); // d s1
    // e1b
    relu_back(buf[GS1], buf[S1], buf[GS1], P0 * 16);
    layer_grads(1, buf[A1], buf[GS1], N, H, W);
    conv(s, 16, 16, buf[GS1], wb(1), nullptr, buf[GA], N, H, W, 0); // d a1
    // e1a (input x0; no data gradient needed)
    relu_back(buf[GA], buf[A1], buf[GB], P0 * 16);
    layer_grads(0, buf[X0], buf[GB], N, H, W);
}

void Denoiser::adam_step(hipStream_t s, float lr)
{
    ++stepCount;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float stepSize = lr * std::sqrt(1.0f - std::pow(b2, (float)stepCount)) / (1.0f - std::pow(b1, (float)stepCount));
    k_adam<<<grid_for(numParams, kThreads), kThreads, 0, s>>>(params, adamM, adamV, grads, (long)numParams, stepSize, b1, b2, eps);
    prep_weights(s);
}

float Denoiser::loss_and_grads(hipStream_t s, const uint16_t* x0Host, const float* targetHost, int N, int H, int W,
                               float* gradHost)
{
    reserve(N, H, W, true);
    const size_t P = (size_t)N * H * W;
    float* tgt = nullptr;
    check(hipMalloc(&tgt, P * 12));
    try {
        check(hipMemcpyAsync(buf[X0], x0Host, P * 32, hipMemcpyHostToDevice, s));
        check(hipMemcpyAsync(tgt, targetHost, P * 12, hipMemcpyHostToDevice, s));
        forward(s, N, H, W, true);
        backward(s, tgt, N, H, W);
        float loss = 0.0f;
        check(hipMemcpyAsync(&loss, lossPart + 4095, 4, hipMemcpyDeviceToHost, s));
        if(gradHost) check(hipMemcpyAsync(gradHost, grads, numParams * 4, hipMemcpyDeviceToHost, s));
        check(hipStreamSynchronize(s));
        (void)hipFree(tgt);
        return loss;
    }
    catch(...) {
        (void)hipFree(tgt);
        throw;
    }
}

void Denoiser::infer(hipStream_t s, const uint16_t* x0Host, int N, int H, int W, uint16_t* outHost)
{
    if(!ws || wsN != N || wsH != H || wsW != W) {
        check(hipStreamSynchronize(s));
        reserve(N, H, W, false);
    }
    check(hipMemcpyAsync(buf[X0], x0Host, (size_t)N * H * W * 32, hipMemcpyHostToDevice, s));
    forward(s, N, H, W, false);
    read_output(s, outHost);
}

void Denoiser::read_output(hipStream_t s, uint16_t* host)
{
    check(hipStreamSynchronize(s));
    check(hipMemcpy(host, buf[OUT], (size_t)wsN * wsH * wsW * 32, hipMemcpyDeviceToHost));
}

void Denoiser::dataset_reserve(uint32_t count, int H, int W)
{
    if(ds && (count != dsCount || H != dsH || W != dsW)) {
        (void)hipFree(dsX);
        (void)hipFree(dsT);
        dsX = nullptr;
        dsT = nullptr;
    }
    if(!dsX) {
        check(hipMalloc(&dsX, (size_t)count * H * W * 32));
        check(hipMalloc(&dsT, (size_t)count * H * W * 12));
    }
    ds = true;
    dsCount = count;
    dsH = H;
    dsW = W;
}

void Denoiser::dataset_add_input(hipStream_t s, const Images& img, uint32_t slot)
{
    if(!ds || slot >= dsCount || (int)img.w != dsW || (int)img.h != dsH)
        throw std::runtime_error("denoiser: dataset slot or frame size does not match the dataset");
    k_pack<<<grid_for((long)dsH * dsW, kThreads), kThreads, 0, s>>>(img.base, img.normal, img.features, dsH, dsW,
                                                                     dsX + (size_t)slot * dsH * dsW * 16);
}

void Denoiser::dataset_add_target(hipStream_t s, const Images& img, uint32_t slot, float weight, bool accumulate)
{
    if(!ds || slot >= dsCount || (int)img.w != dsW || (int)img.h != dsH)
        throw std::runtime_error("denoiser: dataset slot or frame size does not match the dataset");
    k_target_add<<<grid_for((long)dsH * dsW, kThreads), kThreads, 0, s>>>(img.base, (long)dsH * dsW, weight, accumulate ? 1 : 0,
                                                                           dsT + (size_t)slot * dsH * dsW * 3);
}

void Denoiser::dataset_read(hipStream_t s, uint32_t first, uint32_t count, int H, int W, uint16_t* x0Host, float* targetHost)
{
    if(!ds || first > dsCount || count > dsCount - first) throw std::runtime_error("denoiser: dataset slots out of range");
    if(H != dsH || W != dsW) throw std::runtime_error("denoiser: dataset frame size differs");
    const size_t per = (size_t)dsH * dsW;
    check(hipStreamSynchronize(s));
    if(x0Host) check(hipMemcpy(x0Host, dsX + first * per * 16, count * per * 32, hipMemcpyDeviceToHost));
    if(targetHost) check(hipMemcpy(targetHost, dsT + first * per * 3, count * per * 12, hipMemcpyDeviceToHost));
}

float Denoiser::train(hipStream_t s, uint32_t steps, uint32_t batch, float lr, uint32_t seed, uint32_t used)
{
    if(!ds || used == 0 || used > dsCount) throw std::runtime_error("denoiser: no training data");
    if(batch == 0 || batch > used) throw std::runtime_error("denoiser: batch size");
    reserve((int)batch, dsH, dsW, true);
    const size_t per = (size_t)dsH * dsW;
    float* tgt = nullptr;
    check(hipMalloc(&tgt, (size_t)batch * per * 12));
    std::mt19937 rng(seed);
    std::uniform_int_distribution<uint32_t> pick(0, used - 1);
    double acc = 0.0;
    int accN = 0;
    try {
        for(uint32_t it = 0; it < steps; ++it) {
            for(uint32_t b = 0; b < batch; ++b) {
                const uint32_t k = pick(rng);
                check(hipMemcpyAsync(buf[X0] + b * per * 16, dsX + k * per * 16, per * 32, hipMemcpyDeviceToDevice, s));
                check(hipMemcpyAsync(tgt + b * per * 3, dsT + k * per * 3, per * 12, hipMemcpyDeviceToDevice, s));
            }
            forward(s, (int)batch, dsH, dsW, true);
            backward(s, tgt, (int)batch, dsH, dsW);
            adam_step(s, lr);
            if(it + 16 >= steps) { // mean loss of the last steps
                float l = 0.0f;
                check(hipMemcpyAsync(&l, lossPart + 4095, 4, hipMemcpyDeviceToHost, s));
                check(hipStreamSynchronize(s));
                acc += l;
                ++accN;
            }
        }
        check(hipStreamSynchronize(s));
        (void)hipFree(tgt);
    }
    catch(...) {
        (void)hipFree(tgt);
        throw;
    }
    paramsLoaded = true;
    return accN ? (float)(acc / accN) : 0.0f;
}

uint32_t Denoiser::denoise(hipStream_t s, Images& img, float fallbackA, float fallbackB)
{
    const int H = (int)img.h, W = (int)img.w;
    if(!ws || wsN != 1 || wsH != H || wsW != W) reserve(1, H, W, false);
    k_pack<<<grid_for((long)H * W, kThreads), kThreads, 0, s>>>(img.base, img.normal, img.features, H, W, buf[X0]);
    forward(s, 1, H, W, false);
    const dim3 tiles((W + 15) / 16, (H + 15) / 16);
    k_tile_check<<<tiles, kThreads, 0, s>>>(buf[OUT], img.base, img.features, H, W, fallbackA, fallbackB, fallback);
    k_denoise_out<<<grid_for((long)H * W, kThreads), kThreads, 0, s>>>(buf[OUT], fallback, H, W, img.base);
    return tiles.x * tiles.y;
}

uint32_t Denoiser::count_fallback(hipStream_t s, uint32_t tiles)
{
    std::vector<uint8_t> f(tiles);
    check(hipStreamSynchronize(s));
    check(hipMemcpy(f.data(), fallback, tiles, hipMemcpyDeviceToHost));
    uint32_t n = 0;
    for(uint8_t v: f) n += v;
    return n;
}

} // namespace rgrt
