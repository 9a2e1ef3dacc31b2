// wavefront.hip — raygen and the closest-hit / miss stages as a wavefront.
//
// The reference traces one recursive ray tree per pixel inside
// traceRaysKHR: raygen (resources/shaders/raygen.h:69-115) calls
// traceRayEXT, the closest-hit shader (closesthit.rchit:93-268) calls it
// again for shadow, reflection and refraction rays, and every call shares
// one payload (payload.h:29-39) by reference. Here each pixel keeps that
// payload, its sample accumulators and a stack of suspended closest-hit
// frames in HBM. A pass of the shading kernel takes one traced ray per
// pixel, runs the closest-hit or miss stage for it, resumes suspended
// frames until one of them issues the next ray (or the pixel's samples are
// done) and appends that ray to the next queue. Appends are compacted per
// wave with a ballot and one atomic. The traversal pass in between is
// trace.hip.
//
// Frame stages (where a suspended closest-hit continues):
//   G_SHADOW   after the direct shadow ray          closesthit.rchit:193-197
//   G_REFL     after the reflection ray             :215-220
//   G_REFR     after the refraction ray             :232-236 / :244-249
//   SI_TRACE   shadow-internal, after SHADOW_TRACE  :133-138
//   SI_SKY     shadow-internal, after the sky ray   :144-145
//   ST_INTERNAL shadow-trace, after SHADOW_INTERNAL :158-159
// A ray with cull mask 0 (:144) cannot hit and a refraction ray with a zero
// direction (refract() at total internal reflection) is defined here not to
// hit either; both go straight to their miss shader without a queue entry.
#pragma clang fp contract(off)

#include "rgrt_internal.h"
#include "traverse.h"

namespace rgrt {

namespace {

constexpr int RT_GENERIC = 0, RT_SHADOW_TRACE = 1, RT_SHADOW_INTERNAL = 2;
constexpr int kShadeBlock = 256;

// Minimum waves per SIMD the shading kernels are compiled for (the register
// budget is then 512 / waves VGPRs; what does not fit spills to scratch).
#ifndef RGRT_SHADE_WAVES
#define RGRT_SHADE_WAVES 1
#endif
#ifndef RGRT_DRAIN_WAVES
#define RGRT_DRAIN_WAVES 1
#endif
// The shading variant for one sample per slot fits in 128 VGPRs (4 waves);
// compiled for 5 waves it spills 29 VGPRs and is slightly faster beside
// other frames' traversal (config 2 at 1080p, four frames in flight:
// 5.36 against 5.40-5.43 ms per frame).
#ifndef RGRT_SHADE1_WAVES
#define RGRT_SHADE1_WAVES 5
#endif
// The drain variant for one sample per slot (no next-sample start) fits in
// 143 VGPRs (3 waves); compiled for 4 waves it spills 12 VGPRs and runs
// faster (config 2 at 1080p: 1.58-1.60 against 1.69-1.82 ms). At 5 waves it
// spills 55 inside the traversal loop: 5.60 against 5.34-5.36 ms per
// frame with four frames in flight.
#ifndef RGRT_DRAIN1_WAVES
#define RGRT_DRAIN1_WAVES 4
#endif

enum Stage : int { G_SHADOW = 1, G_REFL = 2, G_REFR = 3, SI_TRACE = 4, SI_SKY = 5, ST_INTERNAL = 6 };
enum Step : int { RET = 0, EMIT_FRAME = 1, EMIT = 2, DONE = 3 };

__constant__ float kAAOffsets[9][8][2] = { // raygen.h:37-67
    {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}},
    {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}},
    {{0.25f, 0.25f}, {-0.25f, -0.25f}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}},
    {{-0.125f, -0.375f}, {0.375f, -0.125f}, {-0.375f, 0.125f}, {0.125f, 0.375f}, {0, 0}, {0, 0}, {0, 0}, {0, 0}},
    {{-0.125f, -0.375f}, {0.375f, -0.125f}, {-0.375f, 0.125f}, {0.125f, 0.375f}, {0, 0}, {0, 0}, {0, 0}, {0, 0}},
    {{0.0625f, -0.1875f}, {-0.0625f, 0.1875f}, {0.3125f, 0.0625f}, {-0.1875f, -0.3125f}, {-0.3125f, 0.3125f},
     {-0.4375f, -0.0625f}, {0.1875f, 0.4375f}, {0.4375f, -0.4375f}},
    {{0.0625f, -0.1875f}, {-0.0625f, 0.1875f}, {0.3125f, 0.0625f}, {-0.1875f, -0.3125f}, {-0.3125f, 0.3125f},
     {-0.4375f, -0.0625f}, {0.1875f, 0.4375f}, {0.4375f, -0.4375f}},
    {{0.0625f, -0.1875f}, {-0.0625f, 0.1875f}, {0.3125f, 0.0625f}, {-0.1875f, -0.3125f}, {-0.3125f, 0.3125f},
     {-0.4375f, -0.0625f}, {0.1875f, 0.4375f}, {0.4375f, -0.4375f}},
    {{0.0625f, -0.1875f}, {-0.0625f, 0.1875f}, {0.3125f, 0.0625f}, {-0.1875f, -0.3125f}, {-0.3125f, 0.3125f},
     {-0.4375f, -0.0625f}, {0.1875f, 0.4375f}, {0.4375f, -0.4375f}},
};

struct Payload {
    F3 hitValue;
    float reflectContribution;
    F3 normal;
    float rv[4];
    float depth, curIOR, refDepth;
    int rayType, recDepth;
};

// A closest-hit invocation in registers. Stored while suspended are only
// stage, frontFacing, mat, t, origin, dir, vn, gridRef, shadowColor,
// reflectColor and reflectDepth; the material's fields, the grid effect
// and baseColor are recomputed from those when it resumes (load_frame).
struct Frame {
    int stage, frontFacing;
    uint32_t mat;  // material index (clamped to the table)
    float gridRef; // the payload's refDepth at the hit (grid effect input)
    F3 origin, dir, vn;
    float t;
    F3 diffuse;
    float transparency;
    F3 specular;
    float reflectivity, roughness, ior;
    int rayConsumption;
    float emission;
    F3 baseColor, shadowColor, reflectColor;
    float reflectDepth;
};

struct OutRay {
    F3 o;
    float tmin;
    F3 d;
    float tmax;
    uint32_t mask;
    uint32_t miss1;
};

struct Ctx {
    const ShadeParams& p;
    F3 L;
    int maxRec;
    uint32_t key; // sample key of the ray being shaded (stochastic sampling)
};

// ---------------------------------------------------------------------------
// State in HBM
// ---------------------------------------------------------------------------
// Between passes a pixel keeps only the payload fields that a later pass
// reads: hitValue (the input of a shadow ray's hit), depth (read after a
// nested trace that may not set it), curIOR, refDepth, recDepth and
// rayType. normal, rv and reflectContribution are written by the outermost
// closest-hit (or the primary ray's miss) in the same pass as the sample's
// accumulation, after any earlier write, so they start at their initial
// zeros when a pass loads the state. Record: hitValue, depth | curIOR,
// refDepth, recDepth, sp | rayType << 8 | sample << 10.
__device__ __forceinline__ void load_state(const ShadeParams& p, uint32_t pix, Payload& P, int& sp, int& sample)
{
    const float4* b = reinterpret_cast<const float4*>(p.st.payload) + (size_t)pix * 2;
    const float4 a = b[0], c = b[1];
    P.hitValue = f3(a.x, a.y, a.z);
    P.depth = a.w;
    P.curIOR = c.x;
    P.refDepth = c.y;
    P.recDepth = __float_as_int(c.z);
    const uint32_t m = __float_as_uint(c.w);
    sp = (int)(m & 0xffu);
    P.rayType = (int)((m >> 8) & 3u);
    sample = (int)(m >> 10);
    P.normal = f3(0.0f);
    P.rv[0] = P.rv[1] = P.rv[2] = P.rv[3] = 0.0f;
    P.reflectContribution = 0.0f;
}

__device__ __forceinline__ void store_state(const ShadeParams& p, uint32_t pix, const Payload& P, int sp, int sample)
{
    float4* b = reinterpret_cast<float4*>(p.st.payload) + (size_t)pix * 2;
    const uint32_t m = (uint32_t)sp | ((uint32_t)P.rayType << 8) | ((uint32_t)sample << 10);
    b[0] = make_float4(P.hitValue.x, P.hitValue.y, P.hitValue.z, P.depth);
    b[1] = make_float4(P.curIOR, P.refDepth, __int_as_float(P.recDepth), __uint_as_float(m));
}

// Frame record (kFrameStride words) in five 16-byte parts: [0] stage |
// frontFacing << 3 | mat << 4, t, origin.xy; [1] origin.z, dir; [2] vn,
// gridRef; [3] shadowColor, reflectDepth; [4] reflectColor, 0.
// Only the parts a later resume reads and that changed are written: a frame
// suspended by the closest-hit stage (`fresh`) writes [0..2]; [3] once a
// stage after it reads shadowColor or reflectDepth (every stage but G_SHADOW,
// whose resume sets shadowColor, reflectColor and reflectDepth before reading
// them, and ST_INTERNAL, which reads nothing but the stage); [4] for G_REFR,
// the only stage whose resume reads reflectColor. A resumed frame's [1], [2]
// are unchanged in memory.
__device__ __forceinline__ void store_frame(const ShadeParams& p, uint32_t pix, int level, const Frame& F, bool fresh)
{
    float4* b = reinterpret_cast<float4*>(p.st.frames + ((size_t)level * p.st.frameSlots + pix) * kFrameStride);
    const uint32_t w0 = ((uint32_t)F.stage & 7u) | (((uint32_t)F.frontFacing & 1u) << 3) | (F.mat << 4);
    b[0] = make_float4(__uint_as_float(w0), F.t, F.origin.x, F.origin.y);
    if(fresh) {
        b[1] = make_float4(F.origin.z, F.dir.x, F.dir.y, F.dir.z);
        b[2] = make_float4(F.vn.x, F.vn.y, F.vn.z, F.gridRef);
    }
    if(F.stage != G_SHADOW && F.stage != ST_INTERNAL)
        b[3] = make_float4(F.shadowColor.x, F.shadowColor.y, F.shadowColor.z, F.reflectDepth);
    if(F.stage == G_REFR) b[4] = make_float4(F.reflectColor.x, F.reflectColor.y, F.reflectColor.z, 0.0f);
}

__device__ __forceinline__ void init_payload(Payload& P) // raygen.h:92-100
{
    P.hitValue = f3(0.0f);
    P.normal = f3(0.0f);
    P.rv[0] = P.rv[1] = P.rv[2] = P.rv[3] = 0.0f;
    P.depth = 0.0f;
    P.refDepth = 0.0f;
    P.curIOR = 1.0f;
    P.rayType = RT_GENERIC;
    P.recDepth = 0;
    P.reflectContribution = 0.0f;
}

// raygen.h:78-90: origin = viewInverse * (0,0,0,1); direction =
// viewInverse * (normalize((projInverse * (d, 1, 1)).xyz), 0).
// With stochastic sampling the pixel offset is uniform in [-0.5, 0.5)^2
// from the sample key instead of the SSAA table.
__device__ __forceinline__ void primary_ray(const rgrt_ubo& u, uint32_t px, uint32_t py, uint32_t W, uint32_t H, int sample,
                                            uint32_t sampling, uint32_t seed, OutRay& r)
{
    const float* vi = u.viewInverse;
    const float* pi = u.projInverse;
    const int n = u.numSamples;
    const int row = n < 8 ? n : 8;
    float offx = kAAOffsets[row][sample & 7][0];
    float offy = kAAOffsets[row][sample & 7][1];
    if(sampling == RGRT_SAMPLING_STOCHASTIC) {
        const uint32_t key = sample_key(seed, py * W + px, (uint32_t)sample);
        offx = u01(hash32(key ^ 0x68bc21ebu)) - 0.5f;
        offy = u01(hash32(key ^ 0x02e5be93u)) - 0.5f;
    }
    const float pcx = ((float)px + 0.5f) + offx;
    const float pcy = ((float)py + 0.5f) + offy;
    const float ux = pcx / (float)W, uy = pcy / (float)H;
    const float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    float tg[3];
    for(int k = 0; k < 3; ++k) tg[k] = ((pi[k] * dx + pi[4 + k] * dy) + pi[8 + k] * 1.0f) + pi[12 + k] * 1.0f;
    const F3 dn = normalize(f3(tg[0], tg[1], tg[2]));
    r.d = f3(((vi[0] * dn.x + vi[4] * dn.y) + vi[8] * dn.z) + vi[12] * 0.0f,
             ((vi[1] * dn.x + vi[5] * dn.y) + vi[9] * dn.z) + vi[13] * 0.0f,
             ((vi[2] * dn.x + vi[6] * dn.y) + vi[10] * dn.z) + vi[14] * 0.0f);
    r.o = f3(((vi[0] * 0.0f + vi[4] * 0.0f) + vi[8] * 0.0f) + vi[12] * 1.0f,
             ((vi[1] * 0.0f + vi[5] * 0.0f) + vi[9] * 0.0f) + vi[13] * 1.0f,
             ((vi[2] * 0.0f + vi[6] * 0.0f) + vi[10] * 0.0f) + vi[14] * 1.0f);
    r.tmin = 0.001f;
    r.tmax = 10000.0f;
    r.mask = 0xffu;
    r.miss1 = 0u;
}

// ---------------------------------------------------------------------------
// Miss stages
// ---------------------------------------------------------------------------
__device__ F3 sky(F3 worldDir, F3 L) // miss.rmiss:38-74 with the constants of :78
{
    const F3 sunTone = f3(1.0f, 0.6f, 0.05f), skyTone = f3(0.2f, 0.4f, 0.8f), scatterTone = f3(1.0f, 0.3f, 0.0f);
    const float scatterFactor = 1.3f, powFactor = 80.0f;
    const bool zero = worldDir.x == 0.0f && worldDir.y == 0.0f && worldDir.z == 0.0f;
    const F3 rayDir = zero ? f3(0.0f) : normalize(worldDir);
    const float y = fabsf(worldDir.y + 1.5f) / 3.0f;
    float sun = 1.0f - distance(rayDir, normalize(-L));
    sun = clampf(sun, 0.0f, 2.0f);
    float glow = sun;
    glow = clampf(glow, 0.0f, 1.0f);
    sun = powf(sun, powFactor);
    sun *= 1000.0f;
    sun = clampf(sun, 0.0f, 16.0f);
    glow = powf(glow, 6.0f) * 1.0f;
    glow = powf(glow, y);
    glow = clampf(glow, 0.0f, 1.0f);
    sun *= powf(y * y, (float)(1.0 / 1.65));
    glow *= powf(y * y, (float)(1.0 / 2.0));
    sun += glow;
    const F3 sunColor = sunTone * sun;
    const float atmosphere = sqrtf(1.0f - y);
    float scatter = powf(4.0f - L.y, (float)(1.0 / 15.0));
    scatter = 1.0f - clampf(scatter, 0.8f, 1.0f);
    const F3 scatterColor = mix(f3(1.0f), scatterTone * 1.5f, scatter);
    const F3 skyScatter = mix(skyTone, scatterColor, atmosphere / scatterFactor);
    return sunColor + skyScatter;
}

__device__ __forceinline__ void run_miss(const Ctx& c, Payload& P, F3 dir, uint32_t miss1)
{
    if(miss1) { // shadowMiss.rmiss:30-34
        P.hitValue = f3(1.0f);
        return;
    }
    const F3 res = sky(dir, c.L); // miss.rmiss:76-83
    P.hitValue = res;
    P.rv[0] = res.x;
    P.rv[1] = res.y;
    P.rv[2] = res.z;
    P.rv[3] = 0.0f;
    P.depth = 10000.0f;
}

// ---------------------------------------------------------------------------
// Closest-hit stage
// ---------------------------------------------------------------------------
__device__ __forceinline__ void set_ray(OutRay& r, F3 o, float tmin, F3 d, float tmax, uint32_t mask, uint32_t miss1)
{
    r.o = o;
    r.tmin = tmin;
    r.d = d;
    r.tmax = tmax;
    r.mask = mask;
    r.miss1 = miss1;
}

// Final colour of a generic hit (closesthit.rchit:253-267). The 0/0 weight
// of a material that neither reflects nor refracts is taken as 0.
__device__ void generic_finish(const Frame& F, Payload& P, F3 refractColor)
{
    F3 baseColor = F.baseColor * F.shadowColor;
    baseColor = baseColor + F.emission * F.diffuse;
    const float totalContrib = fmaxf(F.transparency, F.reflectivity);
    const float wsum = F.transparency + F.reflectivity;
    const float w = wsum == 0.0f ? 0.0f : F.reflectivity / wsum;
    const F3 roughCol = mix(refractColor, F.reflectColor, w);
    P.hitValue = mix(baseColor, roughCol, totalContrib);
    if(P.recDepth == 0) {
        P.hitValue = baseColor;
        P.normal = F.vn;
        P.rv[0] = roughCol.x;
        P.rv[1] = roughCol.y;
        P.rv[2] = roughCol.z;
        P.rv[3] = fminf((F.reflectDepth / 50.0f) * F.roughness, F.roughness / 2.1f);
        P.reflectContribution = totalContrib;
    }
    P.depth = F.t;
}

// Refraction part (closesthit.rchit:223-251).
__device__ int generic_refraction(const Ctx& c, Frame& F, Payload& P, OutRay& out)
{
    if(P.recDepth < c.maxRec && F.transparency > 0.0f) {
        F3 dir;
        if(F.frontFacing) {
            const float eta = P.curIOR / F.ior;
            dir = refract(F.dir, F.vn, eta);
            P.recDepth++;
            P.curIOR = F.ior;
        }
        else {
            const float eta = F.ior / 1.0f;
            dir = refract(F.dir, F.vn, eta);
            P.recDepth++;
            P.curIOR = 1.0f;
        }
        set_ray(out, F.origin, 0.01f, dir, 1000.0f, 0xffu, 0u);
        F.stage = G_REFR;
        return EMIT_FRAME;
    }
    generic_finish(F, P, f3(1.0f));
    return RET;
}

// Reflection part (closesthit.rchit:206-221).
__device__ int generic_reflection(const Ctx& c, Frame& F, Payload& P, OutRay& out)
{
    F.reflectColor = f3(1.0f);
    F.reflectDepth = 0.0f;
    if(P.recDepth < c.maxRec && F.reflectivity > 0.0f) {
        F3 dir = reflect(F.dir, F.vn);
        if(c.p.sampling == RGRT_SAMPLING_STOCHASTIC && F.roughness > 0.0f)
            dir = glossy_dir(c.key, F.origin, dir, F.vn, F.roughness);
        P.recDepth += F.rayConsumption;
        P.refDepth += F.t;
        set_ray(out, F.origin, 0.01f, dir, 1000.0f, 0xffu, 0u);
        F.stage = G_REFL;
        return EMIT_FRAME;
    }
    return generic_refraction(c, F, P, out);
}

__device__ void grid_effect(float refDepth, float hitT, Frame& F, F3 pos) // closesthit.rchit:74-91
{
    const float aa = ((refDepth + hitT) + 8.0f) / 30.0f;
    const float aa2 = aa / 2.0f;
    float minmod = fminf(fabsf(glsl_mod((pos.x + 1000.0f) * 10.0f + aa2, 20.0f) - aa2),
                         fabsf(glsl_mod((pos.z + 1000.0f) * 10.0f + aa2, 20.0f) - aa2));
    if(minmod < aa2) {
        minmod -= aa2 - powf(aa, 2.0f) / 3.0f;
        minmod *= 3.0f / powf(aa, 2.0f);
        const float m = mixf(aa / 10.0f, 1.0f, minmod);
        F.diffuse = F.diffuse * m;
        F.specular = F.specular * m;
        F.reflectivity *= m;
    }
    if(glsl_mod((pos.x + 1000.0f) * 5.0f, 20.0f) < 10.0f && glsl_mod((pos.z + 1000.0f) * 5.0f, 20.0f) < 10.0f) {
        F.reflectivity *= 1.5f;
    }
}

// The material's fields of a frame (closesthit.rchit:103-110), then the grid
// effect (:113) for materials with effectId 1.
__device__ __forceinline__ void frame_material(const ShadeParams& p, Frame& F)
{
    const rgrt_material& m = p.mats[F.mat];
    F.diffuse = f3(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
    F.transparency = m.transparency;
    F.specular = f3(m.specular[0], m.specular[1], m.specular[2]);
    F.reflectivity = m.reflectivity;
    F.roughness = m.roughness;
    F.ior = m.ior;
    F.rayConsumption = (int)m.rayConsumption;
    F.emission = m.emission;
    if(m.effectId == 1u) grid_effect(F.gridRef, F.t, F, F.origin);
}

// A suspended frame back in registers: the stored fields, then what the
// closest-hit stage had derived from them (material, grid effect and
// baseColor = max(dot(-L, vn), 0.2) * diffuse, closesthit.rchit:174-175),
// with the same operations on the same values.
__device__ __forceinline__ void load_frame(const Ctx& c, uint32_t pix, int level, Frame& F)
{
    const ShadeParams& p = c.p;
    const float4* b = reinterpret_cast<const float4*>(p.st.frames + ((size_t)level * p.st.frameSlots + pix) * kFrameStride);
    // all five parts in one go (reading only the written ones would wait for
    // the stage first); parts store_frame skipped hold values that are set
    // before they are read
    const float4 a0 = b[0], a1 = b[1], a2 = b[2], a3 = b[3], a4 = b[4];
    const uint32_t w0 = __float_as_uint(a0.x);
    F.stage = (int)(w0 & 7u);
    F.frontFacing = (int)((w0 >> 3) & 1u);
    F.mat = min(w0 >> 4, p.numMaterials ? p.numMaterials - 1 : 0u); // as stored; the clamp only guards the table
    F.t = a0.y;
    F.origin = f3(a0.z, a0.w, a1.x);
    F.dir = f3(a1.y, a1.z, a1.w);
    F.vn = f3(a2.x, a2.y, a2.z);
    F.gridRef = a2.w;
    F.shadowColor = f3(a3.x, a3.y, a3.z);
    F.reflectDepth = a3.w;
    F.reflectColor = f3(a4.x, a4.y, a4.z);
    frame_material(p, F);
    const float dp = fmaxf(dot(-c.L, F.vn), 0.2f);
    F.baseColor = dp * F.diffuse;
}

// The shading record of triangle (indexBufferOffset + 3 prim) / 3 when one
// gathered with this table entry's vertex offset is there (false otherwise:
// the caller gathers the vertices itself).
__device__ __forceinline__ bool shade_record(const ShadeParams& p, const rgrt_instance_offsets& table, uint32_t prim, F3& n0,
                                             F3& n1, F3& n2, uint32_t& matIndex)
{
    if(table.indexBufferOffset % 3u != 0u || table.vertexBufferOffset == kNoShadeRecord) return false;
    const uint64_t g = (uint64_t)(table.indexBufferOffset / 3u) + prim;
    if(g >= p.numIndices / 3) return false;
    const float4* q = reinterpret_cast<const float4*>(p.srec + g);
    const float4 a = q[0], b = q[1], c = q[2];
    if(__float_as_uint(b.w) != table.vertexBufferOffset) return false;
    n0 = f3(a.x, a.y, a.z);
    n1 = f3(b.x, b.y, b.z);
    n2 = f3(c.x, c.y, c.z);
    matIndex = __float_as_uint(a.w);
    return true;
}

// closesthit.rchit:93-204 up to the first nested trace.
// Material of a hit (closesthit.rchit:95-110: offset table of the instance's
// customIndex, vertex 0's matIndex), with the gathers clamped to the buffers.
__device__ const rgrt_material& hit_material(const ShadeParams& p, uint32_t instPos, uint32_t prim)
{
    const InstanceShade& is = p.inst[min(instPos, p.numInstances - 1)];
    const rgrt_instance_offsets table = p.inst[min(is.customIndex, p.numInstances - 1)].offsets;
    uint32_t matIndex;
    F3 n0, n1, n2;
    if(!shade_record(p, table, prim, n0, n1, n2, matIndex)) {
        const uint64_t lastIndex = p.numIndices ? p.numIndices - 1 : 0;
        const uint64_t lastVertex = p.numVertices ? p.numVertices - 1 : 0;
        const uint64_t ib0 = (uint64_t)table.indexBufferOffset + 3ull * prim;
        const uint32_t i0 = p.ib[min(ib0, lastIndex)];
        const uint64_t vi0 = (uint64_t)table.vertexBufferOffset + i0;
        matIndex = p.vb[min(vi0, lastVertex)].matIndex;
    }
    const uint32_t mi = table.materialBufferOffset + matIndex;
    return p.mats[p.numMaterials ? min(mi, p.numMaterials - 1) : 0u];
}

__device__ int closest_hit(const Ctx& c, Payload& P, uint32_t instPos, uint32_t prim, float hitT, float u, float v, F3 O, F3 D,
                           Frame& F, OutRay& out)
{
    const ShadeParams& p = c.p;
    // Gathers are clamped to the buffers: an inconsistent offset table reads
    // wrong data (as the reference would) but never outside an allocation.
    const InstanceShade& is = p.inst[min(instPos, p.numInstances - 1)];
    const rgrt_instance_offsets table = p.inst[min(is.customIndex, p.numInstances - 1)].offsets; // offset table[customIndex]

    const float b0 = (1.0f - u) - v, b1 = u, b2 = v;
    // the three vertex normals and vertex 0's material index: from the
    // triangle's shading record, or gathered through the index buffer
    F3 n0, n1, n2;
    uint32_t matIndex;
    if(!shade_record(p, table, prim, n0, n1, n2, matIndex)) {
        const uint64_t lastIndex = p.numIndices ? p.numIndices - 1 : 0;
        const uint64_t lastVertex = p.numVertices ? p.numVertices - 1 : 0;
        const uint64_t ib0 = (uint64_t)table.indexBufferOffset + 3ull * prim;
        const uint32_t i0 = p.ib[min(ib0 + 0, lastIndex)];
        const uint32_t i1 = p.ib[min(ib0 + 1, lastIndex)];
        const uint32_t i2 = p.ib[min(ib0 + 2, lastIndex)];
        const rgrt_vertex& v0 = p.vb[min((uint64_t)table.vertexBufferOffset + i0, lastVertex)];
        const rgrt_vertex& v1 = p.vb[min((uint64_t)table.vertexBufferOffset + i1, lastVertex)];
        const rgrt_vertex& v2 = p.vb[min((uint64_t)table.vertexBufferOffset + i2, lastVertex)];
        n0 = f3(v0.normal[0], v0.normal[1], v0.normal[2]);
        n1 = f3(v1.normal[0], v1.normal[1], v1.normal[2]);
        n2 = f3(v2.normal[0], v2.normal[1], v2.normal[2]);
        matIndex = v0.matIndex;
    }
    const uint32_t mi = table.materialBufferOffset + matIndex;
    F.mat = p.numMaterials ? min(mi, p.numMaterials - 1) : 0u;
    F.frontFacing = 0; // set below for a generic hit

    const float tmin = 0.01f, tmax = 1000.0f;
    const F3 origin = O + D * hitT;
    const F3 vn = (n0 * b0 + n1 * b1) + n2 * b2;
    const float* w = is.o2w;
    const F3 wn = f3((w[0] * vn.x + w[1] * vn.y) + w[2] * vn.z, (w[4] * vn.x + w[5] * vn.y) + w[6] * vn.z,
                     (w[8] * vn.x + w[9] * vn.y) + w[10] * vn.z);
    F3 vnW = normalize(wn);

    F.origin = origin;
    F.dir = D;
    F.t = hitT;
    F.gridRef = P.refDepth;
    frame_material(p, F); // the material's fields and the grid effect

    if(P.rayType == RT_SHADOW_INTERNAL) { // :125-152
        const float thicknessModulation = clampf((hitT * (1.0f - F.transparency)) * 10.0f, 0.0f, 1.0f);
        const F3 shadowCol =
            P.hitValue - mix(f3(0.0f), normalize(f3(1.1f) - F.diffuse) + f3(0.1f), thicknessModulation);
        if(P.recDepth < c.maxRec) {
            P.rayType = RT_SHADOW_TRACE;
            P.recDepth++;
            set_ray(out, origin, tmin, D, tmax, 0xffu, 0u);
            F.shadowColor = shadowCol;
            F.vn = vnW;
            F.stage = SI_TRACE;
            return EMIT_FRAME;
        }
        P.hitValue = shadowCol * f3(0.4f);
        return RET;
    }
    if(P.rayType == RT_SHADOW_TRACE) { // :153-166
        if(F.transparency > 0.0f) {
            if(P.recDepth < c.maxRec) {
                P.rayType = RT_SHADOW_INTERNAL;
                P.recDepth++;
                set_ray(out, origin, tmin, D, tmax, 0xffu, 1u);
                F.stage = ST_INTERNAL;
                return EMIT_FRAME;
            }
        }
        else {
            P.hitValue = P.hitValue * mix(f3(0.4f), f3(0.8f), clampf(logf(hitT) / 8.0f, 0.0f, 1.0f));
        }
        return RET;
    }

    const bool frontFacing = dot(-D, vnW) > 0.0f; // :170-171
    if(!frontFacing) vnW = normalize(-vnW);
    F.frontFacing = frontFacing ? 1 : 0;
    F.vn = vnW;

    const float dp = fmaxf(dot(-c.L, vnW), 0.2f); // :174-175
    F.baseColor = dp * F.diffuse;

    F.shadowColor = f3(1.0f); // :187-204
    if(dot(-c.L, vnW) > 0.07f) {
        if(P.recDepth < c.maxRec) {
            P.hitValue = f3(1.0f);
            P.rayType = RT_SHADOW_TRACE;
            P.recDepth++;
            set_ray(out, origin, tmin * 10.0f, -c.L, tmax, 0xffu, 1u);
            F.stage = G_SHADOW;
            return EMIT_FRAME;
        }
    }
    else {
        const float shadowModulation = powf(F.transparency, 2.0f);
        F.shadowColor = F.transparency < 1.0f ? mix(f3(1.0f), F.diffuse * shadowModulation, F.transparency) : f3(0.4f);
    }
    return generic_reflection(c, F, P, out);
}

// Continuation of a suspended frame once its nested trace returned.
__device__ int resume(const Ctx& c, Frame& F, Payload& P, OutRay& out)
{
    switch(F.stage) {
    case G_SHADOW:
        P.recDepth--;
        F.shadowColor = P.hitValue;
        P.rayType = RT_GENERIC;
        return generic_reflection(c, F, P, out);
    case G_REFL:
        P.recDepth -= F.rayConsumption;
        F.reflectColor = P.hitValue * F.specular;
        F.reflectDepth = P.depth;
        return generic_refraction(c, F, P, out);
    case G_REFR: {
        P.recDepth--;
        F3 refractColor;
        if(F.frontFacing) {
            refractColor = P.hitValue;
        }
        else {
            const F3 transmittanceModulation = mix(f3(1.0f), F.diffuse, logf(1.0f + F.t));
            refractColor = transmittanceModulation * P.hitValue;
        }
        generic_finish(F, P, refractColor);
        return RET;
    }
    case SI_TRACE:
        P.recDepth--;
        if(P.depth < 1000.0f) {
            P.hitValue = P.hitValue * F.shadowColor;
            return RET;
        }
        else {
            const float eta = F.ior / 1.0f;
            const F3 dir = refract(F.dir, F.vn, eta);
            const float dp = powf(dot(c.L, dir), 5.0f) + 0.75f;
            F.shadowColor = F.shadowColor * dp;
            set_ray(out, F.origin, 0.01f, -dir, 1000.0f, 0x0u, 0u);
            F.stage = SI_SKY;
            return EMIT_FRAME;
        }
    case SI_SKY:
        P.hitValue = F.shadowColor + 0.1f * P.hitValue;
        return RET;
    case ST_INTERNAL:
    default:
        P.recDepth--;
        return RET;
    }
}

__device__ __forceinline__ uint16_t to_half_bits(float x)
{
    const _Float16 h = (_Float16)x;
    return __builtin_bit_cast(uint16_t, h);
}

__device__ __forceinline__ void write_rgba16(uint16_t* img, uint32_t pix, float a, float b, float c, float d)
{
    ushort4 v;
    v.x = to_half_bits(a);
    v.y = to_half_bits(b);
    v.z = to_half_bits(c);
    v.w = to_half_bits(d);
    reinterpret_cast<ushort4*>(img)[pix] = v;
}

// raygen.h:105-114 accumulation and raygen.rgen:35-38 stores.
// Accumulator record: hitValue, normal, rv[4], reflectContribution, depth.
// With stochastic sampling the colour accumulated is the sample's composite
// mix(hitValue, roughValue.rgb, reflectContribution) (what the post pass
// would form from this one sample), and its squared luminance is summed too.
__device__ void accumulate(const ShadeParams& p, uint32_t pix, const Payload& P)
{
    float4* a = reinterpret_cast<float4*>(p.st.accum) + (size_t)pix * 3;
    float4 a0 = a[0], a1 = a[1], a2 = a[2];
    F3 col = P.hitValue;
    if(p.sampling == RGRT_SAMPLING_STOCHASTIC) {
        col = mix(P.hitValue, f3(P.rv[0], P.rv[1], P.rv[2]), P.reflectContribution);
        const float l = (0.299f * col.x + 0.587f * col.y) + 0.114f * col.z;
        p.st.lum2[pix] = p.st.lum2[pix] + l * l;
    }
    a0.x = a0.x + col.x;
    a0.y = a0.y + col.y;
    a0.z = a0.z + col.z;
    a0.w = a0.w + P.normal.x;
    a1.x = a1.x + P.normal.y;
    a1.y = a1.y + P.normal.z;
    a1.z = a1.z + P.rv[0];
    a1.w = a1.w + P.rv[1];
    a2.x = a2.x + P.rv[2];
    a2.y = a2.y + P.rv[3];
    a2.z = a2.z + P.reflectContribution;
    a2.w = a2.w + P.depth;
    a[0] = a0;
    a[1] = a1;
    a[2] = a2;
}

// Pixel of wavefront slot `slot`. Slots run over 8x8 tiles (row-major
// within a tile, tiles row-major within bands of 8 rows; the last band and
// the last tile of a band may be smaller), so the 64 rays of a wave come from
// one compact tile and take similar paths through the trees.
__device__ __forceinline__ void slot_pixel(uint32_t slot, uint32_t W, uint32_t H, uint32_t& px, uint32_t& py)
{
    const uint32_t band = slot / (8u * W);
    const uint32_t y0 = band * 8u;
    const uint32_t bh = min(8u, H - y0);
    const uint32_t o = slot - y0 * W;
    const uint32_t tc = o / (8u * bh);
    const uint32_t i = o - tc * 8u * bh;
    const uint32_t x0 = tc * 8u;
    const uint32_t tw = min(8u, W - x0);
    px = x0 + i % tw;
    py = y0 + i / tw;
}

// Tile-order pixel index of a wavefront slot (slots of sample group g
// follow those of group g - 1).
__device__ __forceinline__ uint32_t slot_image(const ShadeParams& p, uint32_t slot)
{
    return p.st.groups == 1u ? slot : slot % p.st.nimg;
}

// Record of one finished sample of tile-order pixel i (sample groups > 1):
// the values accumulate() would add, summed later in sample order.
__device__ void store_sample(const ShadeParams& p, uint32_t i, int sample, const Payload& P)
{
    F3 col = P.hitValue;
    const size_t at = (size_t)sample * p.st.nimg + i;
    if(p.sampling == RGRT_SAMPLING_STOCHASTIC) {
        col = mix(P.hitValue, f3(P.rv[0], P.rv[1], P.rv[2]), P.reflectContribution);
        const float l = (0.299f * col.x + 0.587f * col.y) + 0.114f * col.z;
        p.st.sampleLum2[at] = l * l;
    }
    float4* r = reinterpret_cast<float4*>(p.st.sampleAcc) + at * 3;
    r[0] = make_float4(col.x, col.y, col.z, P.normal.x);
    r[1] = make_float4(P.normal.y, P.normal.z, P.rv[0], P.rv[1]);
    r[2] = make_float4(P.rv[2], P.rv[3], P.reflectContribution, P.depth);
}

__device__ void finalize(const ShadeParams& p, uint32_t slot)
{
    const float4* av = reinterpret_cast<const float4*>(p.st.accum) + (size_t)slot * 3;
    const float4 a0 = av[0], a1 = av[1], a2 = av[2];
    const float a[12] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w};
    const float ns = (float)p.ubo.numSamples;
    const bool stochastic = p.sampling == RGRT_SAMPLING_STOCHASTIC;
    const float r0[4] = {a[0] / ns, a[1] / ns, a[2] / ns, stochastic ? 0.0f : a[10] / ns};
    const float r1[4] = {a[3] / ns, a[4] / ns, a[5] / ns, (logf(a[11]) * 0.25f) / ns};
    const float r2[4] = {a[6] / ns, a[7] / ns, a[8] / ns, a[9] / ns};
    uint32_t px, py;
    slot_pixel(slot, p.width, p.height, px, py);
    const uint32_t pix = py * p.width + px;
    write_rgba16(p.img.base, pix, r0[0], r0[1], r0[2], r0[3]);
    write_rgba16(p.img.normal, pix, r1[0], r1[1], r1[2], r1[3]);
    write_rgba16(p.img.rough, pix, r2[0], r2[1], r2[2], r2[3]);
    if(stochastic) {
        const float lm = (0.299f * r0[0] + 0.587f * r0[1]) + 0.114f * r0[2];
        const float var = fmaxf(p.st.lum2[slot] / ns - lm * lm, 0.0f) / ns;
        write_rgba16(p.img.features, 2 * pix + 1, a[11] / ns, var, 0.0f, 0.0f);
    }
    if(p.keepF32) {
        reinterpret_cast<float4*>(p.img.base32)[pix] = make_float4(r0[0], r0[1], r0[2], r0[3]);
        reinterpret_cast<float4*>(p.img.normal32)[pix] = make_float4(r1[0], r1[1], r1[2], r1[3]);
        reinterpret_cast<float4*>(p.img.rough32)[pix] = make_float4(r2[0], r2[1], r2[2], r2[3]);
    }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ void write_ray(float* rays, size_t n, uint32_t slot, const OutRay& r, uint32_t pix, uint32_t flags)
{
    rays[0 * n + slot] = r.o.x;
    rays[1 * n + slot] = r.o.y;
    rays[2 * n + slot] = r.o.z;
    rays[3 * n + slot] = r.tmin;
    rays[4 * n + slot] = r.d.x;
    rays[5 * n + slot] = r.d.y;
    rays[6 * n + slot] = r.d.z;
    rays[7 * n + slot] = r.tmax;
    rays[8 * n + slot] = __uint_as_float(pix);
    rays[9 * n + slot] = __uint_as_float(flags);
}

// Primary ray of the band's slot `q`, at position q of the band's queue 0.
__device__ __forceinline__ void raygen_slot(const ShadeParams& p, uint32_t q)
{
    if(q == 0) {
        uint32_t* c = p.st.counters;
        c[0] = p.st.npix; // queue 0 holds every slot's primary ray
        c[1] = 0;
        c[2] = 0;
        c[3] = 0;
        for(int k = 4; k < 12; ++k) c[k] = 0; // ray / sky / zero-dir counters (u64)
    }
    if(q >= p.st.npix) return;
    const uint32_t pix = p.st.chunkFirst * 64u + q;
    // slot `pix` starts sample g (its group) of tile-order pixel i
    const uint32_t g = p.st.groups == 1u ? 0u : pix / p.st.nimg;
    const uint32_t i = pix - g * p.st.nimg;
    uint32_t px, py;
    slot_pixel(i, p.width, p.height, px, py);
    OutRay r;
    primary_ray(p.ubo, px, py, p.width, p.height, (int)g, p.sampling, p.seed, r);
    Payload P;
    init_payload(P);
    store_state(p, pix, P, 0, (int)g);
    if(g == 0u) {
        float4* a = reinterpret_cast<float4*>(p.st.accum) + (size_t)i * 3;
        for(int k = 0; k < 3; ++k) a[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if(p.sampling == RGRT_SAMPLING_STOCHASTIC) p.st.lum2[i] = 0.0f;
    }
    // only sample 0's primary ray carries the primary-hit flag (AOV, features)
    write_ray(p.st.rays[0], p.st.npix, q, r, pix, (g == 0u ? kFlagPrimary : 0u) | (0xffu << kMaskShift));
}

__global__ __launch_bounds__(kShadeBlock) void k_raygen(ShadeParams p)
{
    raygen_slot(p, blockIdx.x * blockDim.x + threadIdx.x);
}

// Ray generation of every scene of a batch: block b covers slots
// [256 (b - start[s]), +256) of scene s.
__global__ __launch_bounds__(kShadeBlock) void k_raygen_multi(const ShadeParams* __restrict__ ps, const uint32_t* __restrict__ start,
                                                              int numScenes)
{
    const int s = block_scene(start, numScenes, blockIdx.x);
    raygen_slot(ps[s], (blockIdx.x - start[s]) * kShadeBlock + threadIdx.x);
}

// Sample groups > 1: every pixel's sample records summed in sample order
// (the additions accumulate() makes one sample at a time), then the outputs.
__global__ __launch_bounds__(kShadeBlock) void k_finalize_samples(ShadeParams p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= p.st.nimg) return;
    float4 a0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), a1 = a0, a2 = a0;
    float l2 = 0.0f;
    const bool stochastic = p.sampling == RGRT_SAMPLING_STOCHASTIC;
    for(uint32_t s = 0; s < p.st.numSamples; ++s) {
        const size_t at = (size_t)s * p.st.nimg + i;
        const float4* r = reinterpret_cast<const float4*>(p.st.sampleAcc) + at * 3;
        const float4 r0 = r[0], r1 = r[1], r2 = r[2];
        a0.x = a0.x + r0.x;
        a0.y = a0.y + r0.y;
        a0.z = a0.z + r0.z;
        a0.w = a0.w + r0.w;
        a1.x = a1.x + r1.x;
        a1.y = a1.y + r1.y;
        a1.z = a1.z + r1.z;
        a1.w = a1.w + r1.w;
        a2.x = a2.x + r2.x;
        a2.y = a2.y + r2.y;
        a2.z = a2.z + r2.z;
        a2.w = a2.w + r2.w;
        if(stochastic) l2 = l2 + p.st.sampleLum2[at];
    }
    float4* a = reinterpret_cast<float4*>(p.st.accum) + (size_t)i * 3;
    a[0] = a0;
    a[1] = a1;
    a[2] = a2;
    if(stochastic) p.st.lum2[i] = l2;
    finalize(p, i);
}

// One closest-hit or miss stage for a traced ray of pixel `pix`, then the
// continuations it returns into. Returns true with `out` set when the pixel
// issues its next ray; false once its last sample is accumulated and the
// outputs are written. kNextSample = false compiles out the start of a
// slot's next sample (its camera ray needs the frame's two matrices, ~22
// VGPRs at the kernels' peak): valid when every slot runs one sample, i.e.
// sample groups >= samples (the launchers pick; see run_one_sample).
template<bool kNextSample>
__device__ bool advance_pixel(const Ctx& c0, uint32_t pix, uint32_t flags, F3 O, F3 D, const HitRec& h, Payload& P, int& sp,
                              int& sample, OutRay& out)
{
    const ShadeParams& p = c0.p;
    Ctx c{c0.p, c0.L, c0.maxRec, 0u};
    const uint32_t tile = slot_image(p, pix); // the pixel in tile order
    uint32_t imgPix = 0;
    if(p.sampling == RGRT_SAMPLING_STOCHASTIC) {
        uint32_t px, py;
        slot_pixel(tile, p.width, p.height, px, py);
        imgPix = py * p.width + px;
        c.key = sample_key(p.seed, imgPix, (uint32_t)sample);
        if(flags & kFlagPrimary) { // denoiser features of sample 0's primary hit
            float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if(h.inst != kNoHit) {
                const rgrt_material& m = hit_material(p, h.inst, h.prim);
                f[0] = m.diffuse[0];
                f[1] = m.diffuse[1];
                f[2] = m.diffuse[2];
                f[3] = m.roughness;
            }
            write_rgba16(p.img.features, 2 * imgPix, f[0], f[1], f[2], f[3]);
        }
    }
    unsigned long long* skyCount = reinterpret_cast<unsigned long long*>(p.st.counters + 6);
    unsigned long long* zeroCount = reinterpret_cast<unsigned long long*>(p.st.counters + 8);
    if((flags & kFlagPrimary) && p.keepAov) {
        rgrt_hit a;
        const bool hit = h.inst != kNoHit;
        a.t = hit ? h.t : 0.0f;
        a.u = hit ? h.u : 0.0f;
        a.v = hit ? h.v : 0.0f;
        a.instance = hit ? p.inst[min(h.inst, p.numInstances - 1)].customIndex : kNoHit;
        a.primitive = hit ? h.prim : 0u;
        uint32_t px, py;
        slot_pixel(tile, p.width, p.height, px, py);
        p.img.primary[py * p.width + px] = a;
    }
    Frame F;
    int step;
    bool fresh = true; // F comes from the closest-hit stage, not from a resume
    if(h.inst != kNoHit) {
        step = closest_hit(c, P, h.inst, h.prim, h.t, h.u, h.v, O, D, F, out);
    }
    else {
        run_miss(c, P, D, flags & kFlagMiss1);
        step = RET;
    }
    for(;;) {
        if(step == EMIT_FRAME) {
            if(sp < kMaxFrames) store_frame(p, pix, sp, F, fresh);
            sp++;
            step = EMIT;
        }
        if(step == EMIT) {
            if(out.mask == 0u) {
                atomicAdd(skyCount, 1ull);
                run_miss(c, P, out.d, out.miss1);
                step = RET;
            }
            else if(out.d.x == 0.0f && out.d.y == 0.0f && out.d.z == 0.0f) {
                atomicAdd(zeroCount, 1ull);
                run_miss(c, P, out.d, out.miss1);
                step = RET;
            }
            else {
                return true;
            }
        }
        // step == RET: the innermost call returned
        if(sp == 0) {
            if(p.st.groups == 1u)
                accumulate(p, pix, P);
            else
                store_sample(p, tile, sample, P);
            sample += (int)p.st.groups;
            if(kNextSample && sample < p.ubo.numSamples) {
                init_payload(P);
                uint32_t px, py;
                slot_pixel(tile, p.width, p.height, px, py);
                primary_ray(p.ubo, px, py, p.width, p.height, sample, p.sampling, p.seed, out);
                if(p.sampling == RGRT_SAMPLING_STOCHASTIC) c.key = sample_key(p.seed, imgPix, (uint32_t)sample);
                step = EMIT;
                continue;
            }
            if(p.st.groups == 1u) finalize(p, pix); // otherwise k_finalize_samples does
            return false;
        }
        sp--;
        if(sp < kMaxFrames) load_frame(c, pix, sp, F);
        step = resume(c, F, P, out);
        fresh = false;
    }
}

__device__ __forceinline__ HitRec load_hit(const float* hits, size_t n, uint32_t i)
{
    HitRec h;
    h.t = hits[0 * n + i];
    h.u = hits[1 * n + i];
    h.v = hits[2 * n + i];
    h.inst = __float_as_uint(hits[3 * n + i]);
    h.prim = __float_as_uint(hits[4 * n + i]);
    return h;
}

// Shades queue entries [base, base + 64) of queue `cur` (one wave, one lane
// per entry), appending the emitted rays to the other queue with one atomic
// per wave.
template<bool kNextSample>
__device__ __forceinline__ void shade_wave(const ShadeParams& p, const Ctx& c, int cur, uint32_t count, uint32_t base, int lane,
                                           uint64_t ltMask)
{
    const size_t n = p.st.npix;
    const float* __restrict__ rays = p.st.rays[cur];
    float* __restrict__ outRays = p.st.rays[1 - cur];
    uint32_t* outCount = p.st.counters + (1 - cur);
    {
        const uint32_t i = base + lane;
        bool emit = false;
        OutRay out;
        uint32_t pix = 0;
        if(i < count) {
            pix = __float_as_uint(rays[8 * n + i]);
            const uint32_t flags = __float_as_uint(rays[9 * n + i]);
            const F3 O = f3(rays[0 * n + i], rays[1 * n + i], rays[2 * n + i]);
            const F3 D = f3(rays[4 * n + i], rays[5 * n + i], rays[6 * n + i]);
            const HitRec h = load_hit(p.st.hits, n, i);
            Payload P;
            int sp, sample;
            load_state(p, pix, P, sp, sample);
            emit = advance_pixel<kNextSample>(c, pix, flags, O, D, h, P, sp, sample, out);
            if(emit) store_state(p, pix, P, sp, sample);
        }
        // append the emitted rays to the next queue, one atomic per wave
        const uint64_t m = __ballot(emit);
        if(m) {
            uint32_t wbase = 0;
            if(lane == 0) wbase = atomicAdd(outCount, (uint32_t)__popcll(m));
            wbase = __shfl(wbase, 0);
            if(emit) write_ray(outRays, n, wbase + (uint32_t)__popcll(m & ltMask), out, pix, (out.mask << kMaskShift) | out.miss1);
        }
    }
}

// Wavefront shading pass over the rays traced by the preceding traversal pass.
template<bool kNextSample>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(kNextSample ? RGRT_SHADE_WAVES : RGRT_SHADE1_WAVES))) void k_shade(ShadeParams p, int cur)
{
    const uint32_t count = p.st.counters[cur];
    const int lane = threadIdx.x & 63;
    const uint64_t ltMask = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    const Ctx c{p, f3(p.ubo.lightDir[0], p.ubo.lightDir[1], p.ubo.lightDir[2]), p.ubo.maxRecursions, 0u};
    const uint32_t stride = gridDim.x * blockDim.x;
    for(uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < count; base += stride)
        shade_wave<kNextSample>(p, c, cur, count, base, lane, ltMask);
}

// Scene of a batch that chunk k (64 queue entries) belongs to, from the
// inclusive prefix of the scenes' chunk counts in LDS; wave-uniform.
__device__ __forceinline__ int chunk_scene(const uint32_t* chunkEnd, int numScenes, uint32_t k, int lane)
{
    return __popcll(__ballot(lane < numScenes && chunkEnd[lane < numScenes ? lane : 0] <= k));
}

// Inclusive prefix over the batch's scenes (lane s < numScenes <= 64) of the
// 64-entry chunks in queue `cur`, written to chunkEnd by the block's first
// wave; the other waves wait at the barrier.
__device__ __forceinline__ void batch_chunk_prefix(const ShadeParams* ps, int numScenes, int cur, uint32_t* chunkEnd)
{
    if(threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t v = lane < numScenes ? (ps[lane].st.counters[cur] + 63u) / 64u : 0u;
        for(int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(v, off);
            if(lane >= off) v += u;
        }
        if(lane < numScenes) chunkEnd[lane] = v;
    }
    __syncthreads();
}

// Shading pass of a whole batch: the queues `cur` of every scene in one
// launch, 64-entry chunks handed to waves in order (a chunk belongs to one
// scene, so a wave shades one scene at a time).
template<bool kNextSample>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(kNextSample ? RGRT_SHADE_WAVES : RGRT_SHADE1_WAVES))) void k_shade_multi(
    const ShadeParams* __restrict__ ps, int numScenes, int cur)
{
    __shared__ uint32_t chunkEnd[kMaxBatchScenes];
    batch_chunk_prefix(ps, numScenes, cur, chunkEnd);
    const uint32_t chunks = chunkEnd[numScenes - 1];
    const int lane = threadIdx.x & 63;
    const uint64_t ltMask = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t waves = gridDim.x * (kShadeBlock / 64);
    for(uint32_t k = blockIdx.x * (kShadeBlock / 64) + (threadIdx.x >> 6); k < chunks; k += waves) {
        const int s = __builtin_amdgcn_readfirstlane(chunk_scene(chunkEnd, numScenes, k, lane));
        const ShadeParams& p = ps[s];
        const uint32_t first = s ? chunkEnd[s - 1] : 0u;
        const Ctx c{p, f3(p.ubo.lightDir[0], p.ubo.lightDir[1], p.ubo.lightDir[2]), p.ubo.maxRecursions, 0u};
        shade_wave<kNextSample>(p, c, cur, p.st.counters[cur], (k - first) * 64u, lane, ltMask);
    }
}

constexpr int kDrainBlock = 256;
constexpr int kDrainLdsStack = 16;
constexpr int kDrainSpill = 368; // see traverse.h: 16 + 368 >= 381

// Drain: finishes every pixel still in flight after the wavefront passes.
// Each lane takes a queued ray, traces and shades it, and keeps tracing that
// pixel's next rays itself until the pixel is done; idle lanes refill from
// the queue (one atomic per wave, a ballot to hand out the entries). The
// recursion tail thus costs the longest remaining chain of one pixel rather
// than one launch per recursion step.
template<bool kNextSample, typename Stack>
__device__ __forceinline__ void drain_scene(const ShadeParams& p, const TraceScene& sc, int cur, Stack& stack)
{
    const uint32_t count = p.st.counters[cur];
    const size_t n = p.st.npix;
    const float* __restrict__ rays = p.st.rays[cur];
    uint32_t* fetch = p.st.counters + 2 + cur;
    const int lane = threadIdx.x & 63;
    const uint64_t ltMask = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    const Ctx c{p, f3(p.ubo.lightDir[0], p.ubo.lightDir[1], p.ubo.lightDir[2]), p.ubo.maxRecursions, 0u};
    bool active = false;
    bool exhausted = false;
    uint32_t pix = 0, flags = 0;
    RayIn r;
    Payload P;
    int sp = 0, sample = 0;
    uint64_t traced = 0;
    for(;;) {
        if(!exhausted) {
            const uint64_t idle = __ballot(!active);
            if(idle) {
                uint32_t b = 0;
                if(lane == 0) b = atomicAdd(fetch, (uint32_t)__popcll(idle));
                b = __shfl(b, 0);
                if(b + (uint32_t)__popcll(idle) >= count) exhausted = true;
                if(!active) {
                    const uint32_t i = b + (uint32_t)__popcll(idle & ltMask);
                    if(i < count) {
                        pix = __float_as_uint(rays[8 * n + i]);
                        flags = __float_as_uint(rays[9 * n + i]);
                        r.ox = rays[0 * n + i];
                        r.oy = rays[1 * n + i];
                        r.oz = rays[2 * n + i];
                        r.tmin = rays[3 * n + i];
                        r.dx = rays[4 * n + i];
                        r.dy = rays[5 * n + i];
                        r.dz = rays[6 * n + i];
                        r.tmax = rays[7 * n + i];
                        load_state(p, pix, P, sp, sample);
                        active = true;
                    }
                }
            }
        }
        if(!__ballot(active)) break;
        if(active) {
            // the active lanes start this traversal together, so the
            // wave-wide step schedule applies
            const HitRec h = traverse<RGRT_INST_EVERY, RGRT_TRI_EVERY, RGRT_SENT_EVERY>(sc, r, (flags >> kMaskShift) & 0xffu, stack);
            ++traced;
            OutRay out;
            if(advance_pixel<kNextSample>(c, pix, flags, f3(r.ox, r.oy, r.oz), f3(r.dx, r.dy, r.dz), h, P, sp, sample, out)) {
                r.ox = out.o.x;
                r.oy = out.o.y;
                r.oz = out.o.z;
                r.tmin = out.tmin;
                r.dx = out.d.x;
                r.dy = out.d.y;
                r.dz = out.d.z;
                r.tmax = out.tmax;
                flags = (out.mask << kMaskShift) | out.miss1;
            }
            else {
                active = false;
            }
        }
    }
    // rays traced here count like the wavefront passes' rays (counters[4..5])
    for(int off = 32; off > 0; off >>= 1) traced += __shfl_xor(traced, off);
    if(lane == 0 && traced) atomicAdd(reinterpret_cast<unsigned long long*>(p.st.counters + 4), (unsigned long long)traced);
}

template<bool kNextSample>
__global__ __launch_bounds__(kDrainBlock) __attribute__((amdgpu_waves_per_eu(kNextSample ? RGRT_DRAIN_WAVES : RGRT_DRAIN1_WAVES))) void k_drain(ShadeParams p, TraceScene sc, int cur, int32_t* __restrict__ spillBase)
{
    __shared__ int32_t ldsStack[kDrainLdsStack * kDrainBlock];
    const uint32_t gtid = blockIdx.x * kDrainBlock + threadIdx.x;
    __shared__ float boxRay[kBoxRayWords * kDrainBlock];
    TraversalStack<kDrainLdsStack, kDrainBlock, kDrainSpill> stack{ldsStack, spillBase + (size_t)gtid * kDrainSpill,
                                                                   (int)threadIdx.x, boxRay};
    drain_scene<kNextSample>(p, sc, cur, stack);
}

// Drain of a whole batch: every wave drains the scenes in order, moving to
// the next scene once the current one has no queued ray left for it and its
// lanes are idle, so the waves that finish a scene early work on the next
// one while the last chains of the earlier scenes are still traced.
template<bool kNextSample>
__global__ __launch_bounds__(kDrainBlock) __attribute__((amdgpu_waves_per_eu(kNextSample ? RGRT_DRAIN_WAVES : RGRT_DRAIN1_WAVES))) void k_drain_multi(
    const ShadeParams* __restrict__ ps, const TraceScene* __restrict__ scs, int numScenes, int cur,
    int32_t* __restrict__ spillBase)
{
    __shared__ int32_t ldsStack[kDrainLdsStack * kDrainBlock];
    const uint32_t gtid = blockIdx.x * kDrainBlock + threadIdx.x;
    __shared__ float boxRay[kBoxRayWords * kDrainBlock];
    TraversalStack<kDrainLdsStack, kDrainBlock, kDrainSpill> stack{ldsStack, spillBase + (size_t)gtid * kDrainSpill,
                                                                   (int)threadIdx.x, boxRay};
    for(int s = 0; s < numScenes; ++s) {
        const TraceScene sc = scs[s];
        drain_scene<kNextSample>(ps[s], sc, cur, stack);
    }
}

} // namespace

void launch_raygen(hipStream_t s, const ShadeParams& p)
{
    // at least one block: slot 0's thread resets the counters of an empty band too
    const uint32_t blocks = std::max(1u, (p.st.npix + kShadeBlock - 1) / kShadeBlock);
    k_raygen<<<blocks, kShadeBlock, 0, s>>>(p);
}

void launch_raygen_multi(hipStream_t s, const ShadeParams* ps, const uint32_t* start, int numScenes, uint32_t blocks)
{
    if(blocks) k_raygen_multi<<<blocks, kShadeBlock, 0, s>>>(ps, start, numScenes);
}

void launch_finalize_samples(hipStream_t s, const ShadeParams& p)
{
    if(p.st.groups > 1u) k_finalize_samples<<<(p.st.nimg + kShadeBlock - 1) / kShadeBlock, kShadeBlock, 0, s>>>(p);
}

// True when every wavefront slot runs exactly one sample of its pixel
// (sample groups >= samples), so no slot starts a next sample.
bool run_one_sample(const ShadeParams& p)
{
    return p.st.groups >= (uint32_t)std::max(1, p.ubo.numSamples);
}

void launch_shade(hipStream_t s, const ShadeParams& p, int cur, uint32_t gridBlocks)
{
    if(run_one_sample(p))
        k_shade<false><<<gridBlocks, kShadeBlock, 0, s>>>(p, cur);
    else
        k_shade<true><<<gridBlocks, kShadeBlock, 0, s>>>(p, cur);
}

uint32_t drain_spill_words_per_thread() { return kDrainSpill; }
uint32_t drain_block_size() { return kDrainBlock; }

void launch_drain(hipStream_t s, const ShadeParams& p, const TraceScene& sc, int cur, uint32_t gridBlocks, int32_t* spill)
{
    if(run_one_sample(p))
        k_drain<false><<<gridBlocks, kDrainBlock, 0, s>>>(p, sc, cur, spill);
    else
        k_drain<true><<<gridBlocks, kDrainBlock, 0, s>>>(p, sc, cur, spill);
}

void launch_shade_multi(hipStream_t s, const ShadeParams* ps, int numScenes, int cur, uint32_t gridBlocks, bool oneSample)
{
    if(oneSample)
        k_shade_multi<false><<<gridBlocks, kShadeBlock, 0, s>>>(ps, numScenes, cur);
    else
        k_shade_multi<true><<<gridBlocks, kShadeBlock, 0, s>>>(ps, numScenes, cur);
}

void launch_drain_multi(hipStream_t s, const ShadeParams* ps, const TraceScene* scs, int numScenes, int cur, uint32_t gridBlocks,
                        int32_t* spill, bool oneSample)
{
    if(oneSample)
        k_drain_multi<false><<<gridBlocks, kDrainBlock, 0, s>>>(ps, scs, numScenes, cur, spill);
    else
        k_drain_multi<true><<<gridBlocks, kDrainBlock, 0, s>>>(ps, scs, numScenes, cur, spill);
}

} // namespace rgrt
