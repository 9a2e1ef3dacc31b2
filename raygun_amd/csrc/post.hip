// post.hip — the compute post passes of Raytracer::doRaytracing
// (raygun/render/raytracer.cpp:106-140) on the HIP core.
//
// k_rough_post fuses rough_prepare (resources/shaders/rough_prepare.comp:28-54),
// the ten (rough_blur_h, rough_blur_v) pairs (rough_blur.h:23-40,
// raytracer.cpp:116-121) and postprocess (postprocess.comp:28-50) into one
// launch: a 32x32 output tile is computed from a 52x52 region (10-pixel
// halo, one pixel per H or V pass), the two colour ping-pong images live in
// LDS as fp16 exactly as the reference's rgba16f images hold them between
// passes, and transitions are stored R8_SNORM. k_fxaa is FXAA (the quality
// variant, Raygun's preset 39 settings, fxaa.comp:29-36) written from the
// published algorithm, with an emulated bilinear clamp-to-edge sampler
// (8-bit sub-texel weights).
#pragma clang fp contract(off)

#include "rgrt_internal.h"

namespace rgrt {

namespace {

constexpr int kTile = 32;
constexpr int kHalo = 10;
constexpr int kRegion = kTile + 2 * kHalo; // 52
constexpr int kRegionPix = kRegion * kRegion;
// 1024 threads per tile: the block's 54 KB of LDS allow two blocks per CU, so
// the block size sets the waves that hide the LDS and barrier latency of the
// 20 passes (measured at 1080p: 0.303 ms with 256, 0.199 with 512, 0.155 with
// 1024 threads).
constexpr int kPostThreads = 1024;

__device__ __forceinline__ float h2f(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }
__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

struct Px4 {
    float x, y, z, w;
};

__device__ __forceinline__ Px4 load16(const uint16_t* img, int W, int H, int x, int y)
{
    if(x < 0 || y < 0 || x >= W || y >= H) return {0.0f, 0.0f, 0.0f, 0.0f};
    const ushort4 v = reinterpret_cast<const ushort4*>(img)[(size_t)y * W + x];
    return {h2f(v.x), h2f(v.y), h2f(v.z), h2f(v.w)};
}

__device__ __forceinline__ float dist4(Px4 a, Px4 b)
{
    const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    return sqrtf(((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3);
}

__device__ __forceinline__ int8_t to_snorm8(float f)
{
    const float c = clampf(f, -1.0f, 1.0f);
    return (int8_t)__builtin_rintf(c * 127.0f);
}
__device__ __forceinline__ float from_snorm8(int8_t v) { return fmaxf((float)v / 127.0f, -1.0f); }

// One 32x32 output tile at (tx0, ty0) of rough prepare, the 20 blur passes
// and the composite; A, B, T are the block's LDS region arrays.
__device__ __forceinline__ void rough_post_tile(const Images& img, float4 fade, int tx0, int ty0, ushort4* A, ushort4* B, float* T)
{
    const int W = (int)img.w, H = (int)img.h;
    const int rx0 = tx0 - kHalo, ry0 = ty0 - kHalo;

    // rough_prepare over the region
    for(int q = threadIdx.x; q < kRegionPix; q += kPostThreads) {
        const int x = rx0 + q % kRegion, y = ry0 + q / kRegion;
        if(x < 0 || y < 0 || x >= W || y >= H) {
            A[q] = make_ushort4(0, 0, 0, 0);
            B[q] = make_ushort4(0, 0, 0, 0);
            T[q] = 0.0f;
            continue;
        }
        const Px4 r = load16(img.rough, W, H, x, y);
        const Px4 ru = load16(img.rough, W, H, x, y + 1);
        const Px4 rd = load16(img.rough, W, H, x, y - 1);
        const Px4 rl = load16(img.rough, W, H, x + 1, y);
        const Px4 rr = load16(img.rough, W, H, x - 1, y);
        const Px4 n = load16(img.normal, W, H, x, y);
        const Px4 nu = load16(img.normal, W, H, x, y + 1);
        const Px4 nd = load16(img.normal, W, H, x, y - 1);
        const Px4 nl = load16(img.normal, W, H, x + 1, y);
        const Px4 nr = load16(img.normal, W, H, x - 1, y);
        const float uf = fminf(ru.w, r.w) * clampf(1.0f - dist4(nu, n) * 10.0f, 0.0f, 1.0f);
        const float df = fminf(rd.w, r.w) * clampf(1.0f - dist4(nd, n) * 10.0f, 0.0f, 1.0f);
        const float lf = fminf(rl.w, r.w) * clampf(1.0f - dist4(nl, n) * 10.0f, 0.0f, 1.0f);
        const float rf = fminf(rr.w, r.w) * clampf(1.0f - dist4(nr, n) * 10.0f, 0.0f, 1.0f);
        const int8_t tq = to_snorm8(fminf(fminf(fminf(uf, df), lf), rf));
        T[q] = from_snorm8(tq);
        const ushort4 rb = reinterpret_cast<const ushort4*>(img.rough)[(size_t)y * W + x];
        A[q] = rb;
        B[q] = rb;
        const int lx = x - tx0, ly = y - ty0;
        if(lx >= 0 && ly >= 0 && lx < kTile && ly < kTile) img.transitions[(size_t)y * W + x] = tq;
    }
    __syncthreads();

    // 10 x (H: A -> B, V: B -> A). Region-border pixels read a missing
    // neighbour as 0; that error stays inside the halo (one pixel per pass).
    // A pass only computes the pixels the remaining passes still read: the
    // tile grown by one pixel in x per H pass after it and one pixel in y per
    // V pass after it (these sets shrink from pass to pass, so every pixel a
    // pass computes was computed by all earlier passes too).
    const int tcol = threadIdx.x % 32, trow = threadIdx.x / 32;
    for(int pass = 0; pass < 20; ++pass) {
        const bool horiz = (pass & 1) == 0;
        const ushort4* in = horiz ? A : B;
        ushort4* out = horiz ? B : A;
        const int ox = horiz ? 1 : 0, oy = horiz ? 0 : 1;
        const int hAfter = 9 - pass / 2;                     // H passes after this one
        const int vAfter = horiz ? 10 - pass / 2 : 9 - pass / 2; // V passes after this one
        const int x0 = kHalo - hAfter, x1 = kHalo + kTile + hAfter;
        const int y0 = kHalo - vAfter, y1 = kHalo + kTile + vAfter;
        for(int ly = y0 + trow; ly < y1; ly += kPostThreads / 32)
            for(int lx = x0 + tcol; lx < x1; lx += 32) {
            const int q = ly * kRegion + lx;
            const float t = T[q];
            if(t < 0.001f) continue;
            const ushort4 c = in[q];
            const int q1x = lx + ox, q1y = ly + oy, q2x = lx - ox, q2y = ly - oy;
            const ushort4 z = make_ushort4(0, 0, 0, 0);
            const ushort4 c1 = (q1x < kRegion && q1y < kRegion) ? in[q1y * kRegion + q1x] : z;
            const ushort4 c2 = (q2x >= 0 && q2y >= 0) ? in[q2y * kRegion + q2x] : z;
            const float w0 = (1.0f - t) - t;
            ushort4 o;
            o.x = f2h(((h2f(c.x) * w0) + (h2f(c1.x) * t)) + (h2f(c2.x) * t));
            o.y = f2h(((h2f(c.y) * w0) + (h2f(c1.y) * t)) + (h2f(c2.y) * t));
            o.z = f2h(((h2f(c.z) * w0) + (h2f(c1.z) * t)) + (h2f(c2.z) * t));
            o.w = c.w;
            out[q] = o;
        }
        __syncthreads();
    }

    // write colorsA / colorsB and the composite (postprocess.comp:30-38)
    for(int q = threadIdx.x; q < kTile * kTile; q += kPostThreads) {
        const int x = tx0 + q % kTile, y = ty0 + q / kTile;
        if(x >= W || y >= H) continue;
        const int rq = (y - ry0) * kRegion + (x - rx0);
        const size_t pix = (size_t)y * W + x;
        const ushort4 a = A[rq];
        reinterpret_cast<ushort4*>(img.colorsA)[pix] = a;
        reinterpret_cast<ushort4*>(img.colorsB)[pix] = B[rq];
        const ushort4 bb = reinterpret_cast<const ushort4*>(img.base)[pix];
        const float ba = h2f(bb.w);
        const float c0 = clampf(mixf(h2f(bb.x), h2f(a.x), ba), 0.0f, 1.0f);
        const float c1 = clampf(mixf(h2f(bb.y), h2f(a.y), ba), 0.0f, 1.0f);
        const float c2 = clampf(mixf(h2f(bb.z), h2f(a.z), ba), 0.0f, 1.0f);
        const float luma = ((c0 * 0.299f) + (c1 * 0.587f)) + (c2 * 0.114f);
        ushort4 f;
        f.x = f2h(mixf(c0, fade.x, fade.w));
        f.y = f2h(mixf(c1, fade.y, fade.w));
        f.z = f2h(mixf(c2, fade.z, fade.w));
        f.w = f2h(luma);
        reinterpret_cast<ushort4*>(img.final_)[pix] = f;
    }
}

__global__ __launch_bounds__(kPostThreads) void k_rough_post(Images img, float4 fade)
{
    __shared__ ushort4 A[kRegionPix];
    __shared__ ushort4 B[kRegionPix];
    __shared__ float T[kRegionPix];
    rough_post_tile(img, fade, blockIdx.x * kTile, blockIdx.y * kTile, A, B, T);
}

// The post passes of every scene of a batch in one launch: block b is tile
// b - start[s] of scene s (row-major over its 32x32 tiles).
__global__ __launch_bounds__(kPostThreads) void k_rough_post_multi(const ShadeParams* __restrict__ ps,
                                                                   const uint32_t* __restrict__ start, int numScenes)
{
    __shared__ ushort4 A[kRegionPix];
    __shared__ ushort4 B[kRegionPix];
    __shared__ float T[kRegionPix];
    const int s = block_scene(start, numScenes, blockIdx.x);
    const ShadeParams& p = ps[s];
    const uint32_t t = blockIdx.x - start[s];
    const uint32_t tilesX = (p.img.w + kTile - 1) / kTile;
    const float4 fade = make_float4(p.ubo.fadeColor[0], p.ubo.fadeColor[1], p.ubo.fadeColor[2], p.ubo.fadeColor[3]);
    rough_post_tile(p.img, fade, (int)(t % tilesX) * kTile, (int)(t / tilesX) * kTile, A, B, T);
}

// postprocess.comp:41-49 (ubo.showAlpha): every image shows its alpha.
__global__ void k_show_alpha(Images img)
{
    const size_t npix = (size_t)img.w * img.h;
    const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if(p >= npix) return;
    uint16_t* imgs[6] = {img.final_, img.normal, img.rough, img.base, img.colorsA, img.colorsB};
    for(int k = 0; k < 6; ++k) {
        ushort4 v = reinterpret_cast<ushort4*>(imgs[k])[p];
        v.x = v.y = v.z = v.w;
        reinterpret_cast<ushort4*>(imgs[k])[p] = v;
    }
    img.transitions[p] = to_snorm8(1.0f); // .aaaa of a one-channel image reads alpha = 1
}

// ---------------------------------------------------------------------------
// FXAA (Lottes, "FXAA" white paper, NVIDIA 2009, and the FXAA 3.11 quality
// algorithm as published): Raygun runs it with preset 39, subpixel quality
// 1.0, edge threshold 0.063, edge threshold minimum 0.0312 and the green
// channel as luma (fxaa.comp:29-36; knobs fxaa.h:178-216, 487-501).
//
// Per pixel:
//  1. contrast gate: luma of the pixel and its 4 neighbours; a pixel whose
//     luma range is below max(0.0312, 0.063 * brightest) keeps its colour;
//  2. orientation: the second differences across rows (weights 1-2-1 over
//     the three columns) against those across columns; the larger says
//     whether the edge runs horizontally or vertically;
//  3. side: of the two neighbours across the edge, the one with the larger
//     luma step; the edge lies half a pixel toward it, and its luma and the
//     pixel's average to the edge's reference luma;
//  4. walk: from that half-pixel point, along the edge in both directions
//     with the preset's growing strides, sampling (bilinearly, so the two
//     rows either side of the edge are averaged) until the luma leaves the
//     reference by a quarter of the step found in 3;
//  5. blend: if the nearer end turns the opposite way from the pixel, the
//     pixel is moved across the edge by 0.5 - (distance to the nearer end)
//     / (length of the span); a subpixel term from the 3x3 neighbourhood's
//     mean luma (smoothstep of its contrast, squared) is the lower bound;
//     the colour is resampled there, and alpha carries the pixel's luma.
// Distances along the edge are kept as exact texel counts (the strides'
// running sums), so the span ratio of 5 has no accumulated rounding.
// ---------------------------------------------------------------------------
struct Tex {
    const uint16_t* p;
    int W, H;
};

// textureLod / textureLodOffset on the linear clamp-to-edge sampler
// (raygun/compute/compute_system.cpp:76-85), fp32 filtering of the fp16
// texels with weights rounded to 1/256.
__device__ __forceinline__ Px4 tex_sample(const Tex& t, float px, float py, int ox, int oy)
{
    const float u = (px * (float)t.W - 0.5f) + (float)ox;
    const float v = (py * (float)t.H - 0.5f) + (float)oy;
    const float fu = floorf(u), fv = floorf(v);
    const float au = __builtin_rintf((u - fu) * 256.0f) / 256.0f;
    const float av = __builtin_rintf((v - fv) * 256.0f) / 256.0f;
    int x0 = (int)fu, y0 = (int)fv;
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = min(max(x0, 0), t.W - 1);
    x1 = min(max(x1, 0), t.W - 1);
    y0 = min(max(y0, 0), t.H - 1);
    y1 = min(max(y1, 0), t.H - 1);
    const Px4 t00 = load16(t.p, t.W, t.H, x0, y0), t10 = load16(t.p, t.W, t.H, x1, y0);
    const Px4 t01 = load16(t.p, t.W, t.H, x0, y1), t11 = load16(t.p, t.W, t.H, x1, y1);
    Px4 r;
    r.x = mixf(mixf(t00.x, t10.x, au), mixf(t01.x, t11.x, au), av);
    r.y = mixf(mixf(t00.y, t10.y, au), mixf(t01.y, t11.y, au), av);
    r.z = mixf(mixf(t00.z, t10.z, au), mixf(t01.z, t11.z, au), av);
    r.w = mixf(mixf(t00.w, t10.w, au), mixf(t01.w, t11.w, au), av);
    return r;
}

// Running sums of preset 39's strides (1, 1, 1, 1, 1, 1.5, 2, 2, 2, 2, 4, 8):
// the distance in texels from the half-pixel start after each stride.
__constant__ float kWalkReach[12] = {1.0f, 2.0f, 3.0f, 4.0f, 5.0f, 6.5f, 8.5f, 10.5f, 12.5f, 14.5f, 18.5f, 26.5f};
constexpr int kWalkProbes = 11; // samples per direction; the last stride moves without sampling

struct EdgeEnd {
    float reach; // texels from the start point
    float delta; // luma there minus the reference luma (of the last sample taken)
};

// One direction of step 4: (su, sv) is the start point, (du, dv) one texel
// along the edge in texture coordinates, signed for the direction.
__device__ __forceinline__ EdgeEnd walk_edge(const Tex& t, float su, float sv, float du, float dv, float refLuma, float limit)
{
    EdgeEnd end{kWalkReach[kWalkProbes], 0.0f};
    for(int k = 0; k < kWalkProbes; ++k) {
        const float r = kWalkReach[k];
        const float d = tex_sample(t, su + du * r, sv + dv * r, 0, 0).y - refLuma;
        end.delta = d;
        if(fabsf(d) >= limit) {
            end.reach = r;
            break;
        }
    }
    return end;
}

__device__ __forceinline__ float second_diff(float a, float b, float c) { return fabsf((a + c) - 2.0f * b); }

__device__ __forceinline__ void fxaa_pixel(const uint16_t* __restrict__ src, uint16_t* __restrict__ out, int W, int H, int x, int y)
{
    if(x >= W || y >= H) return;
    const Tex t{src, W, H};
    const float texelU = 1.0f / (float)W, texelV = 1.0f / (float)H;
    const float u0 = ((float)x + 0.5f) / (float)W, v0 = ((float)y + 0.5f) / (float)H;
    ushort4* dst = reinterpret_cast<ushort4*>(out) + (size_t)y * W + x;

    // 1. contrast gate
    const Px4 centre = tex_sample(t, u0, v0, 0, 0);
    const float lC = centre.y;
    const float lUp = tex_sample(t, u0, v0, 0, -1).y, lDown = tex_sample(t, u0, v0, 0, 1).y;
    const float lLeft = tex_sample(t, u0, v0, -1, 0).y, lRight = tex_sample(t, u0, v0, 1, 0).y;
    const float brightest = fmaxf(fmaxf(fmaxf(lUp, lDown), fmaxf(lLeft, lRight)), lC);
    const float darkest = fminf(fminf(fminf(lUp, lDown), fminf(lLeft, lRight)), lC);
    const float contrast = brightest - darkest;
    if(contrast < fmaxf(0.0312f, brightest * 0.063f)) {
        *dst = make_ushort4(f2h(centre.x), f2h(centre.y), f2h(centre.z), f2h(centre.w));
        return;
    }

    // 2. orientation
    const float lUL = tex_sample(t, u0, v0, -1, -1).y, lUR = tex_sample(t, u0, v0, 1, -1).y;
    const float lDL = tex_sample(t, u0, v0, -1, 1).y, lDR = tex_sample(t, u0, v0, 1, 1).y;
    const float acrossRows =
        2.0f * second_diff(lUp, lC, lDown) + (second_diff(lUL, lLeft, lDL) + second_diff(lUR, lRight, lDR));
    const float acrossCols =
        2.0f * second_diff(lLeft, lC, lRight) + (second_diff(lUL, lUp, lUR) + second_diff(lDL, lDown, lDR));
    const bool edgeRunsAlongX = acrossRows >= acrossCols;

    // subpixel lower bound of step 5
    const float ringMean = (2.0f * ((lUp + lDown) + (lLeft + lRight)) + ((lUL + lUR) + (lDL + lDR))) / 12.0f;
    float sub = clampf(fabsf(ringMean - lC) / contrast, 0.0f, 1.0f);
    sub = (sub * sub) * (3.0f - 2.0f * sub);
    const float subShift = (sub * sub) * 1.0f;

    // 3. side
    const float lNeg = edgeRunsAlongX ? lUp : lLeft, lPos = edgeRunsAlongX ? lDown : lRight;
    const float stepNeg = lNeg - lC, stepPos = lPos - lC;
    const bool towardNeg = fabsf(stepNeg) >= fabsf(stepPos);
    const float texelAcross = edgeRunsAlongX ? texelV : texelU;
    const float across = towardNeg ? -texelAcross : texelAcross;
    const float refLuma = (lC + (towardNeg ? lNeg : lPos)) * 0.5f;
    const float limit = fmaxf(fabsf(stepNeg), fabsf(stepPos)) * 0.25f;
    const bool centreBelow = (lC - refLuma) < 0.0f;

    // 4. walk both ways from the half-pixel point on the edge
    const float su = edgeRunsAlongX ? u0 : u0 + across * 0.5f;
    const float sv = edgeRunsAlongX ? v0 + across * 0.5f : v0;
    const float du = edgeRunsAlongX ? texelU : 0.0f, dv = edgeRunsAlongX ? 0.0f : texelV;
    const EdgeEnd back = walk_edge(t, su, sv, -du, -dv, refLuma, limit);
    const EdgeEnd fwd = walk_edge(t, su, sv, du, dv, refLuma, limit);

    // 5. blend
    const bool backNearer = back.reach < fwd.reach;
    const EdgeEnd near = backNearer ? back : fwd;
    const bool turnsAway = (near.delta < 0.0f) != centreBelow;
    const float spanShift = turnsAway ? 0.5f - near.reach / (back.reach + fwd.reach) : 0.0f;
    const float shift = fmaxf(spanShift, subShift);
    const float fu = edgeRunsAlongX ? u0 : u0 + shift * across;
    const float fv = edgeRunsAlongX ? v0 + shift * across : v0;
    const Px4 c = tex_sample(t, fu, fv, 0, 0);
    *dst = make_ushort4(f2h(c.x), f2h(c.y), f2h(c.z), f2h(lC));
}

__global__ __launch_bounds__(256) void k_fxaa(const uint16_t* __restrict__ src, uint16_t* __restrict__ out, int W, int H)
{
    fxaa_pixel(src, out, W, H, blockIdx.x * 16 + (threadIdx.x & 15), blockIdx.y * 16 + (threadIdx.x >> 4));
}

// FXAA of every scene of a batch that asks for it in one launch: block b is
// 16x16 tile b - start[s] of scene s (scenes without FXAA have no tiles).
__global__ __launch_bounds__(256) void k_fxaa_multi(const ShadeParams* __restrict__ ps, const uint32_t* __restrict__ start,
                                                    int numScenes)
{
    const int s = block_scene(start, numScenes, blockIdx.x);
    const ShadeParams& p = ps[s];
    const uint32_t t = blockIdx.x - start[s];
    const uint32_t tilesX = (p.img.w + 15) / 16;
    fxaa_pixel(p.img.final_, p.img.base, (int)p.img.w, (int)p.img.h, (int)(t % tilesX) * 16 + (threadIdx.x & 15),
               (int)(t / tilesX) * 16 + (threadIdx.x >> 4));
}

// ---------------------------------------------------------------------------
// Readback in a display format (the swapchain blit of
// raygun/render/render_system.cpp:118-147 converts the rgba16f result to the
// surface format with nearest filtering, i.e. per pixel)
// ---------------------------------------------------------------------------
// Vulkan's float -> unorm8 conversion: clamp to [0, 1] (NaN to 0), then the
// nearest of the 256 levels (x * 255 rounded half to even).
__device__ __forceinline__ uint32_t unorm8(float x)
{
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint32_t)rintf(x * 255.0f);
}

// sRGB transfer function (IEC 61966-2-1) of a linear value in [0, 1].
__device__ __forceinline__ float srgb_encode(float l)
{
    l = fminf(fmaxf(l, 0.0f), 1.0f);
    return l <= 0.0031308f ? 12.92f * l : 1.055f * powf(l, 1.0f / 2.4f) - 0.055f;
}

__global__ void k_to_rgba8(const uint16_t* __restrict__ src, uint32_t* __restrict__ dst, size_t npix, int bgra, int srgb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= npix) return;
    const ushort4 v = reinterpret_cast<const ushort4*>(src)[i];
    float r = h2f(v.x), g = h2f(v.y), b = h2f(v.z);
    const float a = h2f(v.w); // alpha is never sRGB-encoded
    if(srgb) {
        r = srgb_encode(r);
        g = srgb_encode(g);
        b = srgb_encode(b);
    }
    const uint32_t c0 = unorm8(bgra ? b : r), c2 = unorm8(bgra ? r : b);
    dst[i] = c0 | (unorm8(g) << 8) | (c2 << 16) | (unorm8(a) << 24);
}

} // namespace

void launch_to_rgba8(hipStream_t s, const uint16_t* src, uint8_t* dst, size_t npix, int bgra, int srgb)
{
    if(npix)
        k_to_rgba8<<<(unsigned)((npix + 255) / 256), 256, 0, s>>>(src, reinterpret_cast<uint32_t*>(dst), npix, bgra, srgb);
}

// Post passes: rough + composite into `final_`, optional showAlpha, and
// FXAA from `final_` into `base`. `roughEnd` (may be null) is recorded
// after the fused rough/composite kernel.
// Post passes of a batch: `start` (device) holds numScenes + 1 block starts
// for the rough/composite tiles, then numScenes + 1 for the FXAA tiles;
// blocks are the totals (host copies of the last entries).
void launch_post_multi(hipStream_t s, const ShadeParams* ps, const uint32_t* start, int numScenes, uint32_t roughBlocks,
                       uint32_t fxaaBlocks)
{
    if(roughBlocks) k_rough_post_multi<<<roughBlocks, kPostThreads, 0, s>>>(ps, start, numScenes);
    if(fxaaBlocks) k_fxaa_multi<<<fxaaBlocks, 256, 0, s>>>(ps, start + numScenes + 1, numScenes);
}

uint32_t post_tiles(uint32_t w, uint32_t h) { return ((w + kTile - 1) / kTile) * ((h + kTile - 1) / kTile); }
uint32_t fxaa_tiles(uint32_t w, uint32_t h) { return ((w + 15) / 16) * ((h + 15) / 16); }

void launch_post(hipStream_t s, const Images& img, const rgrt_ubo& ubo, int useFxaa, hipEvent_t roughEnd)
{
    const dim3 grid((img.w + kTile - 1) / kTile, (img.h + kTile - 1) / kTile);
    const float4 fade = make_float4(ubo.fadeColor[0], ubo.fadeColor[1], ubo.fadeColor[2], ubo.fadeColor[3]);
    k_rough_post<<<grid, kPostThreads, 0, s>>>(img, fade);
    if(roughEnd) (void)hipEventRecord(roughEnd, s);
    if(ubo.showAlpha) {
        const size_t npix = (size_t)img.w * img.h;
        k_show_alpha<<<(unsigned)((npix + 255) / 256), 256, 0, s>>>(img);
    }
    if(useFxaa) {
        const dim3 g2((img.w + 15) / 16, (img.h + 15) / 16);
        k_fxaa<<<g2, 256, 0, s>>>(img.final_, img.base, (int)img.w, (int)img.h);
    }
}

} // namespace rgrt
