/*
 * rt_oracle.cpp — TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A CPU restatement of Raygun's ray-tracing path, written from the
 * reference's GLSL and host C++ read as text:
 *   resources/shaders/raygen.h:69-115, raygen.rgen:30-39      -> raygen()
 *   resources/shaders/closesthit.rchit:74-268                   -> closest_hit()
 *   resources/shaders/miss.rmiss:38-83, shadowMiss.rmiss:30-34   -> miss(), shadow_miss()
 *   resources/shaders/payload.h:29-43                           -> Payload
 *   resources/shaders/rough_prepare.comp:28-54                  -> rough_prepare()
 *   resources/shaders/rough_blur.h:23-40 (+ _h/_v.comp)          -> rough_blur()
 *   resources/shaders/postprocess.comp:28-50                    -> postprocess()
 *   resources/shaders/fxaa.comp:29-36 (FXAA 3.11 quality, preset 39, from
 *     the published algorithm; knobs fxaa.h:178-216, 487-501)   -> fxaa()
 *   raygun/render/raytracer.cpp:87-147 (pass order)            -> orc_postprocess()
 *   raygun/render/acceleration_structure.cpp:34-194            -> BLAS/TLAS
 * The Vulkan driver's BVH and intersection are replaced by the LBVH the HIP
 * core also builds (Morton-30, stable sort, Karras 2012 hierarchy) and the
 * watertight test of Woop, Benthin and Wald (JCGT 2013); an FP64 brute-force
 * intersector states the FP32 error.
 *
 * Floating point: every expression is written out in the order the GLSL
 * evaluates it, built with -ffp-contract=off (no fused multiply-add), so
 * that the result is defined bit for bit up to the transcendental functions
 * (powf, logf), which come from the C library.
 *
 * Behaviour the reference leaves undefined, fixed here (parity unpinned
 * against the reference, pinned against this restatement):
 *  - refl/(transp+refl) at closesthit.rchit:257 is 0/0 for plain materials:
 *    weight 0 (the intended result, see docs/screenshot.png);
 *  - refract() returning 0 (total internal reflection) gives a ray with zero
 *    direction: it never hits, and the miss shader normalizes it to 0;
 *  - GLSL pow(x, y) with x < 0: C99 powf semantics;
 *  - out-of-bounds imageLoad returns 0;
 *  - fp32 -> fp16 stores round to nearest even, R8_SNORM stores round half
 *    to even;
 *  - the FXAA sampler filters in fp32 with 8-bit sub-texel weights.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// --------------------------------------------------------------------------
// Layouts (vertex.def, gpu_material.def, uniform_buffer_object.def,
// instance_offset_table.def), restated.
// --------------------------------------------------------------------------
struct Vertex {
    float px, py, pz;
    uint32_t matIndex;
    float nx, ny, nz;
    float pad1;
};
static_assert(sizeof(Vertex) == 32, "vertex.def is 32 bytes");

struct Material {
    float diffuse[3];
    float transparency;
    float specular[3];
    float reflectivity;
    float roughness;
    float ior;
    uint32_t effectId;
    uint32_t rayConsumption;
    float emission;
    float pad[3];
};
static_assert(sizeof(Material) == 64, "gpu_material.def is 64 bytes");

struct UBO {
    float viewInverse[16];
    float projInverse[16];
    float clearColor[3];
    int32_t numSamples;
    float lightDir[3];
    int32_t maxRecursions;
    float time;
    uint32_t showAlpha;
    float pad0, pad1;
    float fadeColor[4];
};
static_assert(sizeof(UBO) == 192, "uniform_buffer_object.def is 192 bytes (std140)");

struct Instance {
    float transform[12];
    uint32_t customIndex;
    uint32_t mask;
    uint32_t flags;
    uint32_t blas;
    uint32_t vertexBufferOffset;
    uint32_t indexBufferOffset;
    uint32_t materialBufferOffset;
    uint32_t pad;
};
static_assert(sizeof(Instance) == 80, "instance record");

struct CanonNode {
    float lo[2][3];
    float hi[2][3];
    int32_t child[2];
    uint32_t parent;
    uint32_t pad;
};
static_assert(sizeof(CanonNode) == 64, "canonical node");

struct HitOut {
    float t, u, v;
    uint32_t inst, prim;
};
static_assert(sizeof(HitOut) == 20, "hit record");

constexpr int32_t EMPTY_CHILD = (int32_t)0x80000000;
constexpr uint32_t NO_HIT = 0xffffffffu;

// --------------------------------------------------------------------------
// Small vector type, GLSL built-ins in GLSL evaluation order
// --------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 v3(float s) { return {s, s, s}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(V3 a, V3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// GLSL normalize(x) = x / length(x).
inline V3 normalize(V3 a)
{
    float l = length(a);
    return {a.x / l, a.y / l, a.z / l};
}
inline float distance(V3 a, V3 b) { return length(a - b); }
inline float mixf(float x, float y, float a) { return (x * (1.0f - a)) + (y * a); }
inline V3 mix(V3 x, V3 y, float a) { return {mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a)}; }
inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
inline float modf_glsl(float x, float y) { return x - y * std::floor(x / y); }
// reflect(I, N) = I - 2 * dot(N, I) * N
inline V3 reflect(V3 I, V3 N)
{
    float d2 = 2.0f * dot(N, I);
    return {I.x - d2 * N.x, I.y - d2 * N.y, I.z - d2 * N.z};
}
// refract(I, N, eta): k = 1 - eta^2 (1 - dot(N,I)^2); 0 if k < 0,
// else eta * I - (eta * dot(N, I) + sqrt(k)) * N
inline V3 refract(V3 I, V3 N, float eta)
{
    float d = dot(N, I);
    float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if(k < 0.0f) return v3(0.0f);
    float s = eta * d + std::sqrt(k);
    return {eta * I.x - s * N.x, eta * I.y - s * N.y, eta * I.z - s * N.z};
}
inline float powg(float x, float y) { return std::pow(x, y); }

// --------------------------------------------------------------------------
// Storage conversions
// --------------------------------------------------------------------------
uint16_t f32_to_f16(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if(ax >= 0x7f800000u) { // inf or nan
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
    }
    if(ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); // rounds to >= 65520 -> inf
    if(ax < 0x38800000u) { // below the smallest normal half: subnormal or zero
        if(ax < 0x33000000u) return (uint16_t)sign; // below 2^-25 rounds to 0
        // |f| = m * 2^(e-150); in units of 2^-24 (the half subnormal step)
        // that is m * 2^(e-126) = m >> (126 - e), rounded to nearest even.
        uint32_t e = ax >> 23;
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t drop = 126u - e;
        uint32_t q = m >> drop;
        uint32_t rem = m & ((1u << drop) - 1u);
        uint32_t half = 1u << (drop - 1u);
        if(rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t e = (ax >> 23) - 127 + 15;
    uint32_t m = ax & 0x7fffffu;
    uint32_t q = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fffu;
    if(rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}

float f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    uint32_t x;
    if(e == 0) {
        if(m == 0) {
            x = sign;
        }
        else { // subnormal: m * 2^-24
            float f = (float)m * 5.9604644775390625e-8f;
            std::memcpy(&x, &f, 4);
            x |= sign;
        }
    }
    else if(e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    }
    else {
        x = sign | ((e - 15 + 127) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

inline float round_f16(float f) { return f16_to_f32(f32_to_f16(f)); }

int8_t f32_to_snorm8(float f)
{
    float c = clampf(f, -1.0f, 1.0f);
    return (int8_t)std::nearbyint(c * 127.0f);
}
inline float snorm8_to_f32(int8_t v) { return std::fmax((float)v / 127.0f, -1.0f); }

// --------------------------------------------------------------------------
// LBVH (shared by BLAS and TLAS)
// --------------------------------------------------------------------------
struct Box {
    float lo[3], hi[3];
};
inline Box empty_box()
{
    const float inf = std::numeric_limits<float>::infinity();
    return {{inf, inf, inf}, {-inf, -inf, -inf}};
}
inline Box box_union(const Box& a, const Box& b)
{
    Box r;
    for(int k = 0; k < 3; ++k) {
        r.lo[k] = std::fmin(a.lo[k], b.lo[k]);
        r.hi[k] = std::fmax(a.hi[k], b.hi[k]);
    }
    return r;
}

inline uint32_t expand_bits10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Quantize a coordinate in [0,1] to 10 bits: min(max(x * 1024, 0), 1023).
inline uint32_t quantize10(float x)
{
    float q = std::fmin(std::fmax(x * 1024.0f, 0.0f), 1023.0f);
    return (uint32_t)q;
}

uint32_t morton3(float x, float y, float z)
{
    return (expand_bits10(quantize10(x)) << 2) | (expand_bits10(quantize10(y)) << 1) | expand_bits10(quantize10(z));
}

// 4-wide node, 128 bytes: per-axis box bounds of the four children, then
// the child references (>= 0 wide node, < 0 leaf ~k, EMPTY_CHILD unused).
struct Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "wide node");

// Compressed 4-wide node, 64 bytes: the node box's min corner, a
// power-of-two scale per axis (exponent bytes) and each child box as 8-bit
// multiples of the scale, rounded outwards (lo down, hi up).
struct Node4q {
    float ox, oy, oz;
    int8_t ex, ey, ez, pad0;
    uint8_t qlox[4], qhix[4], qloy[4], qhiy[4];
    uint8_t qloz[4], qhiz[4];
    uint32_t pad1[2];
    int32_t child[4];
};
static_assert(sizeof(Node4q) == 64, "compressed wide node");

struct BVH {
    std::vector<CanonNode> nodes;
    std::vector<Node4> wide;        // the binary tree collapsed to 4 children per node
    std::vector<Node4q> wideq;      // the same nodes compressed (what the traversal reads)
    std::vector<uint32_t> leafPrim; // Morton order -> original primitive index
    std::vector<uint32_t> morton;   // original order
    std::vector<Box> primBox;       // original order
    Box root;
};

// delta(i, j) of Karras 2012 with the index tie-break for equal keys.
inline int delta(const std::vector<uint32_t>& k, int n, int i, int j)
{
    if(j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if(a == b) return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
    return __builtin_clz(a ^ b);
}

void compute_node_boxes(BVH& b)
{
    const int n = (int)b.leafPrim.size();
    if(n <= 1) {
        CanonNode& r = b.nodes[0];
        Box e = empty_box();
        Box c0 = n == 1 ? b.primBox[b.leafPrim[0]] : e;
        for(int k = 0; k < 3; ++k) {
            r.lo[0][k] = c0.lo[k];
            r.hi[0][k] = c0.hi[k];
            r.lo[1][k] = e.lo[k];
            r.hi[1][k] = e.hi[k];
        }
        b.root = c0;
        return;
    }
    // Post-order over internal nodes (min/max is exact, so any order gives
    // the bit-identical boxes a bottom-up GPU refit produces).
    std::vector<Box> nodeBox(n - 1);
    std::vector<int> stack;
    std::vector<uint8_t> state(n - 1, 0);
    stack.push_back(0);
    while(!stack.empty()) {
        int i = stack.back();
        CanonNode& nd = b.nodes[i];
        if(state[i] == 0) {
            state[i] = 1;
            for(int c = 1; c >= 0; --c)
                if(nd.child[c] >= 0) stack.push_back(nd.child[c]);
            continue;
        }
        stack.pop_back();
        Box cb[2];
        for(int c = 0; c < 2; ++c) {
            int ch = nd.child[c];
            cb[c] = ch >= 0 ? nodeBox[ch] : b.primBox[b.leafPrim[~ch]];
            for(int k = 0; k < 3; ++k) {
                nd.lo[c][k] = cb[c].lo[k];
                nd.hi[c][k] = cb[c].hi[k];
            }
        }
        nodeBox[i] = box_union(cb[0], cb[1]);
    }
    b.root = nodeBox[0];
}

inline float half_area(const float lo[3], const float hi[3])
{
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return ((dx * dy) + (dy * dz)) + (dz * dx);
}

// Collapses the binary tree into 4-wide nodes. Each wide node starts from a
// binary node's two children and repeatedly opens the internal child of
// largest surface area (the first one on equal areas) until it has four
// children or only leaves. Wide nodes are numbered in breadth-first order.
void collapse4(BVH& b)
{
    const int n = (int)b.leafPrim.size();
    const float inf = std::numeric_limits<float>::infinity();
    auto empty_node = [&]() {
        Node4 w;
        for(int c = 0; c < 4; ++c) {
            w.lox[c] = w.loy[c] = w.loz[c] = inf;
            w.hix[c] = w.hiy[c] = w.hiz[c] = -inf;
            w.child[c] = EMPTY_CHILD;
            w.pad[c] = 0;
        }
        return w;
    };
    struct E {
        int32_t ref;
        float lo[3], hi[3];
    };
    auto slot = [&](int node, int c) {
        const CanonNode& nd = b.nodes[node];
        E e;
        e.ref = nd.child[c];
        for(int k = 0; k < 3; ++k) {
            e.lo[k] = nd.lo[c][k];
            e.hi[k] = nd.hi[c][k];
        }
        return e;
    };
    b.wide.clear();
    std::vector<int> queue{0};
    for(size_t qi = 0; qi < queue.size(); ++qi) {
        const int r = queue[qi];
        E S[4];
        int ns = 0;
        for(int c = 0; c < 2; ++c)
            if(b.nodes[r].child[c] != EMPTY_CHILD) S[ns++] = slot(r, c);
        while(ns < 4) {
            int j = -1;
            float best = -1.0f;
            for(int k = 0; k < ns; ++k) {
                if(S[k].ref < 0) continue;
                const float a = half_area(S[k].lo, S[k].hi);
                if(a > best) {
                    best = a;
                    j = k;
                }
            }
            if(j < 0) break;
            const int e = S[j].ref;
            for(int k = ns; k > j + 1; --k) S[k] = S[k - 1];
            S[j] = slot(e, 0);
            S[j + 1] = slot(e, 1);
            ++ns;
        }
        Node4 w = empty_node();
        for(int k = 0; k < ns; ++k) {
            w.lox[k] = S[k].lo[0];
            w.loy[k] = S[k].lo[1];
            w.loz[k] = S[k].lo[2];
            w.hix[k] = S[k].hi[0];
            w.hiy[k] = S[k].hi[1];
            w.hiz[k] = S[k].hi[2];
            if(S[k].ref >= 0) {
                w.child[k] = (int32_t)queue.size();
                queue.push_back(S[k].ref);
            }
            else {
                w.child[k] = S[k].ref;
            }
        }
        b.wide.push_back(w);
    }
    (void)n;
}

// 2^e as a float, e in [-126, 127].
inline float pow2f(int e)
{
    uint32_t b = (uint32_t)(e + 127) << 23;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// Smallest e in [-126, 127] with 254 * 2^e >= ext: the 8-bit grid spans the
// node box with one step to spare for outward rounding.
inline int quant_exponent(float ext)
{
    if(!(ext > 0.0f)) return -126;
    uint32_t bits;
    std::memcpy(&bits, &ext, 4);
    int e = (int)((bits >> 23) & 0xffu) - 127 - 7;
    e = std::min(127, std::max(-126, e));
    while(e < 127 && 254.0f * pow2f(e) < ext) ++e;
    while(e > -126 && 254.0f * pow2f(e - 1) >= ext) --e;
    return e;
}

// Sign of the exact (o + p) - x for floats o, p, x: the rounded sum and its
// error from TwoSum (Knuth) decide it without rounding.
inline int sum_cmp(float o, float p, float x)
{
    const float sum = o + p;
    if(sum < x) return -1;
    if(sum > x) return 1;
    const float bv = sum - o;
    const float err = (o - (sum - bv)) + (p - bv);
    return err < 0.0f ? -1 : (err > 0.0f ? 1 : 0);
}

// Outward-rounded 8-bit offsets of [lo, hi] from origin o on the grid 2^e:
// the largest a with o + a*2^e <= lo and the smallest b with o + b*2^e >= hi,
// exactly (a*2^e is exact for a < 256).
inline void quantize_range(float o, int e, float lo, float hi, uint8_t& qlo, uint8_t& qhi)
{
    const float s = pow2f(e), inv = pow2f(-std::min(e, 126));
    int a = (int)std::floor((lo - o) * inv);
    a = std::min(255, std::max(0, a));
    while(a > 0 && sum_cmp(o, (float)a * s, lo) > 0) --a;
    int b = (int)std::ceil((hi - o) * inv);
    b = std::min(255, std::max(0, b));
    while(b < 255 && sum_cmp(o, (float)b * s, hi) < 0) ++b;
    qlo = (uint8_t)a;
    qhi = (uint8_t)b;
}

void quantize4(BVH& b)
{
    b.wideq.resize(b.wide.size());
    for(size_t i = 0; i < b.wide.size(); ++i) {
        const Node4& w = b.wide[i];
        Node4q q;
        std::memset(&q, 0, sizeof(q));
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const float* clo[3] = {w.lox, w.loy, w.loz};
        const float* chi[3] = {w.hix, w.hiy, w.hiz};
        for(int c = 0; c < 4; ++c) {
            if(w.child[c] == EMPTY_CHILD) continue;
            for(int k = 0; k < 3; ++k) {
                lo[k] = std::fmin(lo[k], clo[k][c]);
                hi[k] = std::fmax(hi[k], chi[k][c]);
            }
        }
        int e[3];
        float org[3];
        for(int k = 0; k < 3; ++k) {
            const bool any = lo[k] <= hi[k];
            org[k] = any ? lo[k] : 0.0f;
            e[k] = any ? quant_exponent(hi[k] - lo[k]) : -126;
        }
        q.ox = org[0];
        q.oy = org[1];
        q.oz = org[2];
        q.ex = (int8_t)e[0];
        q.ey = (int8_t)e[1];
        q.ez = (int8_t)e[2];
        uint8_t* qlo[3] = {q.qlox, q.qloy, q.qloz};
        uint8_t* qhi[3] = {q.qhix, q.qhiy, q.qhiz};
        for(int c = 0; c < 4; ++c) {
            q.child[c] = w.child[c];
            if(w.child[c] == EMPTY_CHILD) continue;
            for(int k = 0; k < 3; ++k) quantize_range(org[k], e[k], clo[k][c], chi[k][c], qlo[k][c], qhi[k][c]);
        }
        b.wideq[i] = q;
    }
}

// Parallel locally-ordered clustering (Meister & Bittner 2018) over the
// Morton-ordered primitives: each cluster's nearest neighbour is the
// cluster within +-kPlocRadius positions whose union box has the smallest
// surface area (the lowest position on equal areas); mutual nearest
// neighbours merge (the left one keeps the merged cluster), and the
// clusters are compacted in order. Internal nodes are numbered from n-2
// down in merge order, so the last merge is the root, node 0.
constexpr int kPlocRadius = 16;

void build_ploc(BVH& b)
{
    const int n = (int)b.leafPrim.size();
    struct C {
        int32_t ref;
        float lo[3], hi[3];
    };
    std::vector<C> cl(n), nxt;
    for(int i = 0; i < n; ++i) {
        const Box& pb = b.primBox[b.leafPrim[i]];
        cl[i].ref = ~i;
        for(int k = 0; k < 3; ++k) {
            cl[i].lo[k] = pb.lo[k];
            cl[i].hi[k] = pb.hi[k];
        }
    }
    auto merged_area = [&](const C& a, const C& c) {
        float lo[3], hi[3];
        for(int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(a.lo[k], c.lo[k]);
            hi[k] = std::fmax(a.hi[k], c.hi[k]);
        }
        return half_area(lo, hi);
    };
    int nextNode = n - 2;
    std::vector<int> nn;
    for(int iter = 0; (int)cl.size() > 1 && iter < 100000; ++iter) {
        const int m = (int)cl.size();
        nn.assign(m, -1);
        for(int i = 0; i < m; ++i) {
            float best = std::numeric_limits<float>::infinity();
            int bj = -1;
            for(int j = std::max(0, i - kPlocRadius); j <= std::min(m - 1, i + kPlocRadius); ++j) {
                if(j == i) continue;
                const float a = merged_area(cl[i], cl[j]);
                if(a < best) {
                    best = a;
                    bj = j;
                }
            }
            nn[i] = bj;
        }
        nxt.clear();
        int merges = 0;
        for(int i = 0; i < m; ++i) {
            const int j = nn[i];
            const bool mutual = j >= 0 && nn[j] == i;
            if(mutual && j < i) continue; // merged into the left cluster
            if(mutual) {
                const int node = nextNode--;
                CanonNode& nd = b.nodes[node];
                nd.child[0] = cl[i].ref;
                nd.child[1] = cl[j].ref;
                for(int k = 0; k < 3; ++k) {
                    nd.lo[0][k] = cl[i].lo[k];
                    nd.hi[0][k] = cl[i].hi[k];
                    nd.lo[1][k] = cl[j].lo[k];
                    nd.hi[1][k] = cl[j].hi[k];
                }
                for(int c = 0; c < 2; ++c)
                    if(nd.child[c] >= 0) b.nodes[nd.child[c]].parent = (uint32_t)node;
                C u;
                u.ref = node;
                for(int k = 0; k < 3; ++k) {
                    u.lo[k] = std::fmin(cl[i].lo[k], cl[j].lo[k]);
                    u.hi[k] = std::fmax(cl[i].hi[k], cl[j].hi[k]);
                }
                nxt.push_back(u);
                ++merges;
            }
            else {
                nxt.push_back(cl[i]);
            }
        }
        cl.swap(nxt);
        (void)merges;
    }
    b.nodes[0].parent = 0xffffffffu;
    b.root = Box{{cl[0].lo[0], cl[0].lo[1], cl[0].lo[2]}, {cl[0].hi[0], cl[0].hi[1], cl[0].hi[2]}};
}

// Hierarchy of the next builds: 0 = Karras 2012 (fast build), 1 = PLOC
// (fast trace).
int g_hierarchy = 0;

// Builds the canonical LBVH over primitive boxes.
void build_lbvh(BVH& b, int hierarchy)
{
    const int n = (int)b.primBox.size();
    // Scene box: union of all primitive boxes.
    Box sb = empty_box();
    for(const Box& pb: b.primBox) sb = box_union(sb, pb);
    float ext[3];
    for(int k = 0; k < 3; ++k) ext[k] = sb.hi[k] - sb.lo[k];

    b.morton.resize(n);
    for(int i = 0; i < n; ++i) {
        const Box& pb = b.primBox[i];
        float c[3];
        for(int k = 0; k < 3; ++k) {
            float centroid = (pb.lo[k] + pb.hi[k]) * 0.5f;
            c[k] = ext[k] > 0.0f ? (centroid - sb.lo[k]) / ext[k] : 0.0f;
        }
        b.morton[i] = morton3(c[0], c[1], c[2]);
    }
    b.leafPrim.resize(n);
    for(int i = 0; i < n; ++i) b.leafPrim[i] = (uint32_t)i;
    std::stable_sort(b.leafPrim.begin(), b.leafPrim.end(), [&](uint32_t a, uint32_t c) { return b.morton[a] < b.morton[c]; });

    std::vector<uint32_t> keys(n);
    for(int i = 0; i < n; ++i) keys[i] = b.morton[b.leafPrim[i]];

    const int numNodes = n <= 1 ? 1 : n - 1;
    b.nodes.assign(numNodes, CanonNode{});
    for(auto& nd: b.nodes) {
        nd.parent = 0xffffffffu;
        nd.child[0] = nd.child[1] = EMPTY_CHILD;
    }
    if(n == 1) b.nodes[0].child[0] = ~0;
    if(n >= 2 && hierarchy == 1) {
        build_ploc(b);
    }
    else if(n >= 2) {
        for(int i = 0; i < n - 1; ++i) {
            int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
            int dmin = delta(keys, n, i, i - d);
            int lmax = 2;
            while(delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
            int l = 0;
            for(int t = lmax / 2; t >= 1; t /= 2)
                if(delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
            int j = i + l * d;
            int dnode = delta(keys, n, i, j);
            int s = 0;
            for(int div = 2;; div *= 2) {
                int t = (l + div - 1) / div;
                if(delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
                if(t <= 1) break;
            }
            int gamma = i + s * d + std::min(d, 0);
            int lo = std::min(i, j), hi = std::max(i, j);
            int left = (lo == gamma) ? ~gamma : gamma;
            int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
            b.nodes[i].child[0] = left;
            b.nodes[i].child[1] = right;
            if(left >= 0) b.nodes[left].parent = (uint32_t)i;
            if(right >= 0) b.nodes[right].parent = (uint32_t)i;
        }
    }
    compute_node_boxes(b);
    collapse4(b);
    quantize4(b);
}

// --------------------------------------------------------------------------
// Scene
// --------------------------------------------------------------------------
struct Tri {
    float v[3][3];
};

struct Blas {
    uint32_t firstVertex, indexOffset, numFaces;
    BVH bvh;
    std::vector<Tri> tris; // original primitive order
};

struct InstanceX {
    Instance rec;
    float w2o[12]; // world -> object, row-major 3x4
};

} // namespace

struct orc_scene {
    std::vector<Vertex> vb;
    std::vector<uint32_t> ib;
    std::vector<Material> mats;
    std::vector<Blas> blas;
    std::vector<InstanceX> inst;
    BVH tlas;
    bool hasTlas = false;
};

namespace {

void load_tris(const orc_scene& s, Blas& b)
{
    b.tris.resize(b.numFaces);
    b.bvh.primBox.resize(b.numFaces);
    for(uint32_t f = 0; f < b.numFaces; ++f) {
        Box bx = empty_box();
        for(int k = 0; k < 3; ++k) {
            uint32_t vi = b.firstVertex + s.ib[b.indexOffset + 3 * f + k];
            const Vertex& v = s.vb[vi];
            b.tris[f].v[k][0] = v.px;
            b.tris[f].v[k][1] = v.py;
            b.tris[f].v[k][2] = v.pz;
            bx.lo[0] = std::fmin(bx.lo[0], v.px);
            bx.lo[1] = std::fmin(bx.lo[1], v.py);
            bx.lo[2] = std::fmin(bx.lo[2], v.pz);
            bx.hi[0] = std::fmax(bx.hi[0], v.px);
            bx.hi[1] = std::fmax(bx.hi[1], v.py);
            bx.hi[2] = std::fmax(bx.hi[2], v.pz);
        }
        b.bvh.primBox[f] = bx;
    }
}

// Inverse of a 3x4 affine map in double, rounded to float. Linear part by
// the adjugate, translation by -R^-1 t.
void invert_affine(const float m[12], float out[12])
{
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double r = 1.0 / det;
    double inv[9] = {A * r, -(b * i - c * h) * r, (b * f - c * e) * r, B * r, (a * i - c * g) * r, -(a * f - c * d) * r,
                     C * r, -(a * h - b * g) * r, (a * e - b * d) * r};
    double tx = m[3], ty = m[7], tz = m[11];
    for(int row = 0; row < 3; ++row) {
        out[row * 4 + 0] = (float)inv[row * 3 + 0];
        out[row * 4 + 1] = (float)inv[row * 3 + 1];
        out[row * 4 + 2] = (float)inv[row * 3 + 2];
        out[row * 4 + 3] = (float)(-(inv[row * 3 + 0] * tx + inv[row * 3 + 1] * ty + inv[row * 3 + 2] * tz));
    }
}

// World box of a transformed object box (Arvo): per output axis,
// t + sum_j min/max(m_ij lo_j, m_ij hi_j), summed left to right.
Box transform_box(const float m[12], const Box& b)
{
    if(!(b.lo[0] <= b.hi[0])) return b; // empty stays empty
    Box r;
    for(int row = 0; row < 3; ++row) {
        float lo = m[row * 4 + 3], hi = m[row * 4 + 3];
        for(int j = 0; j < 3; ++j) {
            float p = m[row * 4 + j] * b.lo[j];
            float q = m[row * 4 + j] * b.hi[j];
            lo = lo + std::fmin(p, q);
            hi = hi + std::fmax(p, q);
        }
        r.lo[row] = lo;
        r.hi[row] = hi;
    }
    return r;
}

// --------------------------------------------------------------------------
// Ray / box and ray / triangle
// --------------------------------------------------------------------------
struct RayPre {
    float o[3], d[3], inv[3], no[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

// Reciprocal for the slab test: a zero component becomes +-1e-20 so the
// fused form lo*inv - o*inv stays finite.
inline float safe_rcp(float d)
{
    float a = std::fabs(d) > 1e-20f ? d : std::copysign(1e-20f, d);
    return 1.0f / a;
}

void ray_pre(RayPre& r, const float o[3], const float d[3])
{
    for(int k = 0; k < 3; ++k) {
        r.o[k] = o[k];
        r.d[k] = d[k];
        r.inv[k] = safe_rcp(d[k]);
        r.no[k] = -o[k] * r.inv[k];
    }
    int kz = 0;
    if(std::fabs(d[1]) > std::fabs(d[kz])) kz = 1;
    if(std::fabs(d[2]) > std::fabs(d[kz])) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if(d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sz = 1.0f / d[kz]; // one division; the shears reuse it (as the HIP traversal does)
    r.Sx = d[kx] * r.Sz;
    r.Sy = d[ky] * r.Sz;
}

// Watertight ray/triangle (Woop, Benthin, Wald 2013). Accepts t in
// [tmin, tlimit]. u, v are the weights of vertex 1 and vertex 2.
inline bool ray_tri(const RayPre& r, const Tri& tri, float tmin, float tlimit, float& t, float& u, float& v)
{
    const int kx = r.kx, ky = r.ky, kz = r.kz;
    float A[3], B[3], C[3];
    for(int k = 0; k < 3; ++k) {
        A[k] = tri.v[0][k] - r.o[k];
        B[k] = tri.v[1][k] - r.o[k];
        C[k] = tri.v[2][k] - r.o[k];
    }
    float Ax = A[kx] - r.Sx * A[kz];
    float Ay = A[ky] - r.Sy * A[kz];
    float Bx = B[kx] - r.Sx * B[kz];
    float By = B[ky] - r.Sy * B[kz];
    float Cx = C[kx] - r.Sx * C[kz];
    float Cy = C[ky] - r.Sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if(U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if(det == 0.0f) return false;
    float Az = r.Sz * A[kz];
    float Bz = r.Sz * B[kz];
    float Cz = r.Sz * C[kz];
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float tt = T * rcp;
    if(!(tt >= tmin && tt <= tlimit)) return false;
    t = tt;
    u = V * rcp;
    v = W * rcp;
    return true;
}

struct Hit {
    float t = 0, u = 0, v = 0;
    int32_t inst = -1; // index into the instance array
    uint32_t prim = 0;
    bool tie = false;
    uint32_t nodes = 0, tris = 0;
    uint32_t instLeaves = 0, blasEntries = 0; // instance leaves reached, BLASes entered
};

constexpr float TIE_REL = 1e-6f;

// Closest hit over the TLAS, on the 4-wide trees. Depth-first over child
// references (>= 0 a wide node, < 0 a leaf ~k): the hit children of a node
// are ordered by entry distance with the five-comparator sorting network
// (0,1) (2,3) (0,2) (1,3) (1,2), swapping only on a strictly smaller
// distance, and pushed farthest first so the nearest is visited next.
// Candidates are accepted if t < best, so the first found wins an exact tie;
// near-ties are recorded for the parity tests.
bool trace_closest(const orc_scene& s, const float o[3], const float d[3], float tmin, float tmax, uint32_t cullMask, Hit& h)
{
    h = Hit{};
    if(!s.hasTlas || s.inst.empty()) return false;
    RayPre wr;
    ray_pre(wr, o, d);
    float best = tmax;
    bool found = false;
    struct Cand {
        float t;
        int32_t inst;
        uint32_t prim;
    };
    static thread_local std::vector<Cand> cands;
    cands.clear();
    // The box and triangle tests run against a slightly widened limit so
    // that near-tie candidates are seen; acceptance still needs t < best.
    auto widen = [](float b) { return b * (1.0f + 8.0f * TIE_REL); };

    // Child boxes of a compressed node: plane distance t = q * (2^e / d) +
    // (origin - o) / d, one fused multiply-add per plane. b = (origin - o) / d
    // carries two roundings, an absolute error up to |b| * 2^-23 that does not
    // shrink with t (a ray starting far from the node's corner), so besides
    // the relative widening each axis's near distance is lowered and its far
    // distance raised by |b| * 2^-21 (per axis: an axis with a tiny direction
    // component has a huge b, and its planes only bind when that far away).
    auto visit = [&](const std::vector<Node4q>& nodes, const RayPre& r, int32_t ref, int32_t* stack, int& sp) {
        const Node4q& q = nodes[ref];
        h.nodes++;
        const float ax = pow2f(q.ex) * r.inv[0], ay = pow2f(q.ey) * r.inv[1], az = pow2f(q.ez) * r.inv[2];
        const float bx = (q.ox - r.o[0]) * r.inv[0], by = (q.oy - r.o[1]) * r.inv[1], bz = (q.oz - r.o[2]) * r.inv[2];
        const float mx = std::fabs(bx) * 0x1p-21f, my = std::fabs(by) * 0x1p-21f, mz = std::fabs(bz) * 0x1p-21f;
        const float lim = widen(best);
        float tn[4];
        int32_t rf[4];
        for(int c = 0; c < 4; ++c) {
            const float x0 = std::fma((float)q.qlox[c], ax, bx), x1 = std::fma((float)q.qhix[c], ax, bx);
            const float y0 = std::fma((float)q.qloy[c], ay, by), y1 = std::fma((float)q.qhiy[c], ay, by);
            const float z0 = std::fma((float)q.qloz[c], az, bz), z1 = std::fma((float)q.qhiz[c], az, bz);
            const float n = std::fmax(std::fmax(std::fmin(x0, x1) - mx, std::fmin(y0, y1) - my), std::fmax(std::fmin(z0, z1) - mz, tmin));
            const float f = std::fmin(std::fmin(std::fmax(x0, x1) + mx, std::fmax(y0, y1) + my), std::fmin(std::fmax(z0, z1) + mz, lim));
            const bool hit = q.child[c] != EMPTY_CHILD && n <= f * 1.0000004f;
            tn[c] = hit ? n : std::numeric_limits<float>::infinity();
            rf[c] = hit ? q.child[c] : EMPTY_CHILD;
        }
        auto cx = [&](int i, int j) {
            if(tn[j] < tn[i]) {
                std::swap(tn[i], tn[j]);
                std::swap(rf[i], rf[j]);
            }
        };
        cx(0, 1);
        cx(2, 3);
        cx(0, 2);
        cx(1, 3);
        cx(1, 2);
        for(int c = 3; c >= 0; --c)
            if(rf[c] != EMPTY_CHILD) stack[sp++] = rf[c];
    };

    // A wide node pushes at most 3 entries and a tree is at most 63 levels
    // deep, so 3 * 63 + 1 entries per level always suffice.
    int32_t tstack[192];
    int tsp = 0;
    tstack[tsp++] = 0;
    while(tsp) {
        int32_t ref = tstack[--tsp];
        if(ref >= 0) {
            visit(s.tlas.wideq, wr, ref, tstack, tsp);
            continue;
        }
        uint32_t ii = s.tlas.leafPrim[~ref];
        const InstanceX& ix = s.inst[ii];
        h.nodes++; // the 64-byte instance leaf record is fetched like a node
        h.instLeaves++;
        if((ix.rec.mask & cullMask & 0xffu) == 0) continue;
        const Blas& b = s.blas[ix.rec.blas];
        if(b.numFaces == 0) continue;
        h.blasEntries++;
        const float* w = ix.w2o;
        float oo[3], od[3];
        for(int row = 0; row < 3; ++row) {
            oo[row] = ((w[row * 4 + 0] * o[0] + w[row * 4 + 1] * o[1]) + w[row * 4 + 2] * o[2]) + w[row * 4 + 3];
            od[row] = (w[row * 4 + 0] * d[0] + w[row * 4 + 1] * d[1]) + w[row * 4 + 2] * d[2];
        }
        RayPre r;
        ray_pre(r, oo, od);
        int32_t bstack[192];
        int bsp = 0;
        bstack[bsp++] = 0;
        while(bsp) {
            int32_t bref = bstack[--bsp];
            if(bref >= 0) {
                visit(b.bvh.wideq, r, bref, bstack, bsp);
                continue;
            }
            uint32_t prim = b.bvh.leafPrim[~bref];
            h.tris++;
            float t, u, v;
            if(!ray_tri(r, b.tris[prim], tmin, widen(best), t, u, v)) continue;
            cands.push_back({t, (int32_t)ii, prim});
            if(t < best) {
                best = t;
                found = true;
                h.t = t;
                h.u = u;
                h.v = v;
                h.inst = (int32_t)ii;
                h.prim = prim;
            }
        }
    }
    if(found) {
        for(const Cand& c: cands) {
            if(c.inst == h.inst && c.prim == h.prim) continue;
            if(std::fabs(c.t - h.t) <= TIE_REL * h.t) h.tie = true;
        }
    }
    return found;
}

// FP32 brute force: the same instance mapping, ray setup and watertight test
// as trace_closest, over every triangle of every instance whose mask passes
// (no tree). A box test that culls a child holding the closest hit makes
// trace_closest differ from this.
bool brute_fp32(const orc_scene& s, const float o[3], const float d[3], float tmin, float tmax, uint32_t cullMask, float& bt,
                uint32_t& binst, uint32_t& bprim)
{
    bool found = false;
    bt = tmax;
    for(size_t ii = 0; ii < s.inst.size(); ++ii) {
        const InstanceX& ix = s.inst[ii];
        if((ix.rec.mask & cullMask & 0xffu) == 0) continue;
        const Blas& b = s.blas[ix.rec.blas];
        const float* w = ix.w2o;
        float oo[3], od[3];
        for(int row = 0; row < 3; ++row) {
            oo[row] = ((w[row * 4 + 0] * o[0] + w[row * 4 + 1] * o[1]) + w[row * 4 + 2] * o[2]) + w[row * 4 + 3];
            od[row] = (w[row * 4 + 0] * d[0] + w[row * 4 + 1] * d[1]) + w[row * 4 + 2] * d[2];
        }
        RayPre r;
        ray_pre(r, oo, od);
        for(uint32_t p = 0; p < b.numFaces; ++p) {
            float t, u, v;
            if(!ray_tri(r, b.tris[p], tmin, bt, t, u, v) || !(t < bt)) continue;
            bt = t;
            binst = s.inst[ii].rec.customIndex;
            bprim = p;
            found = true;
        }
    }
    return found;
}

// FP64 brute force: Moller-Trumbore in double over every triangle of every
// instance whose mask passes, with the ray mapped by the double inverse.
bool brute_fp64(const orc_scene& s, const float o[3], const float d[3], float tmin, float tmax, uint32_t cullMask, double& bt,
                uint32_t& binst, uint32_t& bprim, double& bu, double& bv)
{
    bool found = false;
    bt = tmax;
    for(size_t ii = 0; ii < s.inst.size(); ++ii) {
        const InstanceX& ix = s.inst[ii];
        if((ix.rec.mask & cullMask & 0xffu) == 0) continue;
        const float* m = ix.rec.transform;
        // exact double inverse of the float matrix
        double a = m[0], b = m[1], c = m[2], dd = m[4], e = m[5], f = m[6], g = m[8], hh = m[9], i = m[10];
        double A = e * i - f * hh, B = -(dd * i - f * g), C = dd * hh - e * g;
        double det = a * A + b * B + c * C;
        double inv[9] = {A / det, -(b * i - c * hh) / det, (b * f - c * e) / det, B / det, (a * i - c * g) / det,
                         -(a * f - c * dd) / det, C / det, -(a * hh - b * g) / det, (a * e - b * dd) / det};
        double po[3] = {o[0] - (double)m[3], o[1] - (double)m[7], o[2] - (double)m[11]};
        double oo[3], od[3];
        for(int r = 0; r < 3; ++r) {
            oo[r] = inv[r * 3] * po[0] + inv[r * 3 + 1] * po[1] + inv[r * 3 + 2] * po[2];
            od[r] = inv[r * 3] * d[0] + inv[r * 3 + 1] * d[1] + inv[r * 3 + 2] * d[2];
        }
        const Blas& bl = s.blas[ix.rec.blas];
        for(uint32_t p = 0; p < bl.numFaces; ++p) {
            const Tri& t = bl.tris[p];
            double e1[3], e2[3], pv[3], tv[3], qv[3];
            for(int k = 0; k < 3; ++k) {
                e1[k] = (double)t.v[1][k] - t.v[0][k];
                e2[k] = (double)t.v[2][k] - t.v[0][k];
            }
            pv[0] = od[1] * e2[2] - od[2] * e2[1];
            pv[1] = od[2] * e2[0] - od[0] * e2[2];
            pv[2] = od[0] * e2[1] - od[1] * e2[0];
            double dt = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
            if(dt == 0.0) continue;
            double idt = 1.0 / dt;
            for(int k = 0; k < 3; ++k) tv[k] = oo[k] - t.v[0][k];
            double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * idt;
            if(u < 0.0 || u > 1.0) continue;
            qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
            qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
            qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
            double v = (od[0] * qv[0] + od[1] * qv[1] + od[2] * qv[2]) * idt;
            if(v < 0.0 || u + v > 1.0) continue;
            double tt = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * idt;
            if(tt < tmin || tt >= bt) continue;
            bt = tt;
            binst = ix.rec.customIndex;
            bprim = p;
            bu = u;
            bv = v;
            found = true;
        }
    }
    return found;
}

// --------------------------------------------------------------------------
// Shading restatement
// --------------------------------------------------------------------------
constexpr int RT_GENERIC = 0, RT_SHADOW_TRACE = 1, RT_SHADOW_INTERNAL = 2;

struct Payload { // payload.h:29-39
    V3 hitValue;
    float reflectContribution;
    V3 normal;
    float roughValue[4];
    float depth;
    float curIOR;
    float refDepth;
    int rayType;
    int recDepth;
};

// Random numbers of the stochastic sampling mode (no counterpart in the
// reference): the "lowbias32" finaliser of C. Wellons' hash-prospector
// (published constants 0x7feb352d, 0x846ca68b, shifts 16/15/16), keyed by
// (seed, image pixel, sample) and, for glossy reflections, the bits of the
// hit position; [0, 1) floats from the top 24 bits.
inline uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
inline uint32_t sample_key(uint32_t seed, uint32_t pixel, uint32_t sample)
{
    return hash32(seed ^ hash32(pixel ^ hash32(sample ^ 0x9e3779b9u)));
}
inline float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }
inline uint32_t fbits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

struct Ctx {
    const orc_scene* s;
    const UBO* ubo;
    V3 lightDir;
    uint32_t sampling = 0, seed = 0, key = 0; // stochastic sampling mode
    float primAlbedo[3] = {0, 0, 0}, primRough = 0; // sample 0's primary hit material
    uint64_t rays = 0, sky = 0, zeroDir = 0;
    uint64_t nodes = 0, tris = 0, hits = 0; // traversal work of the traced rays
    uint64_t instLeaves = 0, blasEntries = 0;
    bool tie = false;
    bool recordPrimary = false;
    HitOut primary{0, 0, 0, NO_HIT, 0};
};

inline V3 mat4_mul_dir(const float* m, V3 v) // column-major mat4 * vec4(v, 0)
{
    return {((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * 0.0f, ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * 0.0f,
            ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * 0.0f};
}
inline void mat4_mul(const float* m, const float v[4], float out[4])
{
    for(int r = 0; r < 4; ++r) out[r] = ((m[r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

// miss.rmiss:38-74 (skyMix). rayDir of a zero direction is 0 (defined here).
V3 sky_mix(V3 worldDir, V3 lightDir, V3 sunTone, V3 skyTone, V3 scatterTone, float scatterFactor, float powFactor)
{
    V3 rayDir = (worldDir.x == 0.0f && worldDir.y == 0.0f && worldDir.z == 0.0f) ? v3(0.0f) : normalize(worldDir);
    float y = std::fabs(worldDir.y + 1.5f) / 3.0f;

    float sun = 1.0f - distance(rayDir, normalize(-lightDir));
    sun = clampf(sun, 0.0f, 2.0f);

    float glow = sun;
    glow = clampf(glow, 0.0f, 1.0f);

    sun = powg(sun, powFactor);
    sun *= 1000.0f;
    sun = clampf(sun, 0.0f, 16.0f);

    glow = powg(glow, 6.0f) * 1.0f;
    glow = powg(glow, y);
    glow = clampf(glow, 0.0f, 1.0f);

    sun *= powg(y * y, (float)(1.0 / 1.65));
    glow *= powg(y * y, (float)(1.0 / 2.0));
    sun += glow;

    V3 sunColor = sunTone * sun;
    float atmosphere = std::sqrt(1.0f - y);
    float scatter = powg(4.0f - lightDir.y, (float)(1.0 / 15.0));
    scatter = 1.0f - clampf(scatter, 0.8f, 1.0f);
    V3 scatterColor = mix(v3(1.0f), scatterTone * 1.5f, scatter);
    V3 skyScatter = mix(skyTone, scatterColor, atmosphere / scatterFactor);
    return sunColor + skyScatter;
}

V3 sky(V3 dir, V3 lightDir)
{
    return sky_mix(dir, lightDir, v3(1.0f, 0.6f, 0.05f), v3(0.2f, 0.4f, 0.8f), v3(1.0f, 0.3f, 0.0f), 1.3f, 80.0f);
}

void miss(Ctx& c, Payload& P, V3 dir) // miss.rmiss:76-83
{
    V3 res = sky(dir, c.lightDir);
    P.hitValue = res;
    P.roughValue[0] = res.x;
    P.roughValue[1] = res.y;
    P.roughValue[2] = res.z;
    P.roughValue[3] = 0.0f;
    P.depth = 10000.0f;
}

void shadow_miss(Payload& P) { P.hitValue = v3(1.0f); } // shadowMiss.rmiss:30-34

void closest_hit(Ctx& c, Payload& P, const Hit& h, V3 O, V3 D);

// traceRayEXT(topLevelAS, Opaque, cullMask, 0, 0, missIndex, O, tmin, D, tmax, 0)
void trace_ray(Ctx& c, Payload& P, uint32_t cullMask, int missIndex, V3 O, float tmin, V3 D, float tmax, bool primary = false)
{
    Hit h;
    bool hit = false;
    if(cullMask == 0) {
        c.sky++;
    }
    else if(D.x == 0.0f && D.y == 0.0f && D.z == 0.0f) {
        c.zeroDir++;
    }
    else {
        c.rays++;
        float o[3] = {O.x, O.y, O.z}, d[3] = {D.x, D.y, D.z};
        hit = trace_closest(*c.s, o, d, tmin, tmax, cullMask, h);
        if(hit && h.tie) c.tie = true;
        c.nodes += h.nodes;
        c.tris += h.tris;
        c.instLeaves += h.instLeaves;
        c.blasEntries += h.blasEntries;
        c.hits += hit ? 1 : 0;
    }
    if(primary && c.recordPrimary) {
        c.primary = hit ? HitOut{h.t, h.u, h.v, c.s->inst[h.inst].rec.customIndex, h.prim} : HitOut{0, 0, 0, NO_HIT, 0};
        // denoiser features of the stochastic mode: the hit's material (vertex 0's matIndex)
        c.primAlbedo[0] = c.primAlbedo[1] = c.primAlbedo[2] = c.primRough = 0.0f;
        if(hit) {
            const orc_scene& s = *c.s;
            const Instance& table = s.inst[s.inst[h.inst].rec.customIndex].rec;
            const Vertex& v0 = s.vb[table.vertexBufferOffset + s.ib[table.indexBufferOffset + 3 * h.prim]];
            const Material& gm = s.mats[table.materialBufferOffset + v0.matIndex];
            for(int k = 0; k < 3; ++k) c.primAlbedo[k] = gm.diffuse[k];
            c.primRough = gm.roughness;
        }
    }
    if(hit) {
        closest_hit(c, P, h, O, D);
        return;
    }
    if(missIndex == 0) miss(c, P, D);
    else shadow_miss(P);
}

struct Mat {
    V3 diffuse;
    float transparency;
    V3 specular;
    float reflectivity;
    float roughness;
    float ior;
    uint32_t effectId;
    uint32_t rayConsumption;
    float emission;
};

void grid_effect(const Payload& P, float hitT, Mat& mat, V3 pos) // closesthit.rchit:74-91
{
    float aa = ((P.refDepth + hitT) + 8.0f) / 30.0f;
    float aa2 = aa / 2.0f;
    float minmod = std::fmin(std::fabs(modf_glsl((pos.x + 1000.0f) * 10.0f + aa2, 20.0f) - aa2),
                             std::fabs(modf_glsl((pos.z + 1000.0f) * 10.0f + aa2, 20.0f) - aa2));
    if(minmod < aa2) {
        minmod -= aa2 - powg(aa, 2.0f) / 3.0f;
        minmod *= 3.0f / powg(aa, 2.0f);
        float m = mixf(aa / 10.0f, 1.0f, minmod);
        mat.diffuse = mat.diffuse * m;
        mat.specular = mat.specular * m;
        mat.reflectivity *= m;
    }
    if(modf_glsl((pos.x + 1000.0f) * 5.0f, 20.0f) < 10.0f && modf_glsl((pos.z + 1000.0f) * 5.0f, 20.0f) < 10.0f) {
        mat.reflectivity *= 1.5f;
    }
}

void closest_hit(Ctx& c, Payload& P, const Hit& h, V3 O, V3 D) // closesthit.rchit:93-268
{
    const orc_scene& s = *c.s;
    const UBO& ubo = *c.ubo;
    const InstanceX& ix = s.inst[h.inst];
    const uint32_t custom = ix.rec.customIndex;
    const Instance& table = s.inst[custom].rec; // instanceOffsetTable.e[gl_InstanceCustomIndexEXT]

    const float b0 = (1.0f - h.u) - h.v, b1 = h.u, b2 = h.v;
    uint32_t ibo = table.indexBufferOffset;
    uint32_t i0 = s.ib[ibo + 3 * h.prim + 0];
    uint32_t i1 = s.ib[ibo + 3 * h.prim + 1];
    uint32_t i2 = s.ib[ibo + 3 * h.prim + 2];
    uint32_t vbo = table.vertexBufferOffset;
    const Vertex& v0 = s.vb[vbo + i0];
    const Vertex& v1 = s.vb[vbo + i1];
    const Vertex& v2 = s.vb[vbo + i2];
    const Material& gm = s.mats[table.materialBufferOffset + v0.matIndex];
    Mat mat{v3(gm.diffuse[0], gm.diffuse[1], gm.diffuse[2]),
            gm.transparency,
            v3(gm.specular[0], gm.specular[1], gm.specular[2]),
            gm.reflectivity,
            gm.roughness,
            gm.ior,
            gm.effectId,
            gm.rayConsumption,
            gm.emission};

    const float tmin = 0.01f, tmax = 1000.0f;
    const float hitT = h.t;
    V3 origin = O + D * hitT;

    V3 vn = (v3(v0.nx, v0.ny, v0.nz) * b0 + v3(v1.nx, v1.ny, v1.nz) * b1) + v3(v2.nx, v2.ny, v2.nz) * b2;
    const float* m = ix.rec.transform; // object-to-world, row-major 3x4
    V3 wn = {(m[0] * vn.x + m[1] * vn.y) + m[2] * vn.z, (m[4] * vn.x + m[5] * vn.y) + m[6] * vn.z,
             (m[8] * vn.x + m[9] * vn.y) + m[10] * vn.z};
    V3 vnW = normalize(wn);

    if(mat.effectId == 1) grid_effect(P, hitT, mat, origin);

    if(P.rayType == RT_SHADOW_INTERNAL) {
        float thicknessModulation = clampf((hitT * (1.0f - mat.transparency)) * 10.0f, 0.0f, 1.0f);
        V3 shadowCol = P.hitValue - mix(v3(0.0f), normalize(v3(1.1f) - mat.diffuse) + v3(0.1f), thicknessModulation);
        if(P.recDepth < ubo.maxRecursions) {
            P.rayType = RT_SHADOW_TRACE;
            P.recDepth++;
            trace_ray(c, P, 0xFF, 0, origin, tmin, D, tmax);
            P.recDepth--;
            if(P.depth < 1000.0f) {
                P.hitValue = P.hitValue * shadowCol;
            }
            else {
                float eta = mat.ior / 1.0f;
                V3 dir = refract(D, vnW, eta);
                float dp = powg(dot(c.lightDir, dir), 5.0f) + 0.75f;
                shadowCol = shadowCol * dp;
                trace_ray(c, P, 0x0, 0, origin, tmin, -dir, tmax);
                P.hitValue = shadowCol + 0.1f * P.hitValue;
            }
        }
        else {
            P.hitValue = shadowCol * v3(0.4f);
        }
        return;
    }
    if(P.rayType == RT_SHADOW_TRACE) {
        if(mat.transparency > 0.0f) {
            if(P.recDepth < ubo.maxRecursions) {
                P.rayType = RT_SHADOW_INTERNAL;
                P.recDepth++;
                trace_ray(c, P, 0xFF, 1, origin, tmin, D, tmax);
                P.recDepth--;
            }
        }
        else {
            P.hitValue = P.hitValue * mix(v3(0.4f), v3(0.8f), clampf(std::log(hitT) / 8.0f, 0.0f, 1.0f));
        }
        return;
    }

    bool frontFacing = dot(-D, vnW) > 0.0f;
    if(!frontFacing) vnW = normalize(-vnW);

    float dp = std::fmax(dot(-c.lightDir, vnW), 0.2f);
    V3 baseColor = dp * mat.diffuse;

    V3 shadowColor = v3(1.0f);
    if(dot(-c.lightDir, vnW) > 0.07f) {
        if(P.recDepth < ubo.maxRecursions) {
            P.hitValue = v3(1.0f);
            P.rayType = RT_SHADOW_TRACE;
            P.recDepth++;
            trace_ray(c, P, 0xFF, 1, origin, tmin * 10.0f, -c.lightDir, tmax);
            P.recDepth--;
            shadowColor = P.hitValue;
            P.rayType = RT_GENERIC;
        }
    }
    else {
        float shadowModulation = powg(mat.transparency, 2.0f);
        shadowColor = mat.transparency < 1.0f ? mix(v3(1.0f), mat.diffuse * shadowModulation, mat.transparency) : v3(0.4f);
    }

    V3 reflectColor = v3(1.0f);
    float reflectDepth = 0.0f;
    if(P.recDepth < ubo.maxRecursions && mat.reflectivity > 0.0f) {
        V3 dir = reflect(D, vnW);
        if(c.sampling == 1 && mat.roughness > 0.0f) {
            // glossy spread: normalize(r + roughness * s), s uniform in [-1, 1]^3, kept above the surface
            const uint32_t hh = hash32(c.key ^ hash32(fbits(origin.x) ^ hash32(fbits(origin.y) ^ hash32(fbits(origin.z)))));
            const V3 sp = v3(u01(hash32(hh ^ 1u)) * 2.0f - 1.0f, u01(hash32(hh ^ 2u)) * 2.0f - 1.0f,
                             u01(hash32(hh ^ 3u)) * 2.0f - 1.0f);
            const V3 g = normalize(dir + sp * mat.roughness);
            if(dot(g, vnW) > 0.0f) dir = g;
        }
        P.recDepth += (int)mat.rayConsumption;
        P.refDepth += hitT;
        trace_ray(c, P, 0xff, 0, origin, tmin, dir, tmax);
        P.recDepth -= (int)mat.rayConsumption;
        reflectColor = P.hitValue * mat.specular;
        reflectDepth = P.depth;
    }

    V3 refractColor = v3(1.0f);
    if(P.recDepth < ubo.maxRecursions && mat.transparency > 0.0f) {
        if(frontFacing) {
            float eta = P.curIOR / mat.ior;
            V3 dir = refract(D, vnW, eta);
            P.recDepth++;
            P.curIOR = mat.ior;
            trace_ray(c, P, 0xff, 0, origin, tmin, dir, tmax);
            P.recDepth--;
            refractColor = P.hitValue;
        }
        else {
            float eta = mat.ior / 1.0f;
            V3 dir = refract(D, vnW, eta);
            P.recDepth++;
            P.curIOR = 1.0f;
            trace_ray(c, P, 0xff, 0, origin, tmin, dir, tmax);
            P.recDepth--;
            V3 transmittanceModulation = mix(v3(1.0f), mat.diffuse, std::log(1.0f + hitT));
            refractColor = transmittanceModulation * P.hitValue;
        }
    }

    baseColor = baseColor * shadowColor;
    baseColor = baseColor + mat.emission * mat.diffuse;
    float totalContrib = std::fmax(mat.transparency, mat.reflectivity);
    float wsum = mat.transparency + mat.reflectivity;
    float w = wsum == 0.0f ? 0.0f : mat.reflectivity / wsum;
    V3 roughCol = mix(refractColor, reflectColor, w);
    P.hitValue = mix(baseColor, roughCol, totalContrib);

    if(P.recDepth == 0) {
        P.hitValue = baseColor;
        P.normal = vnW;
        P.roughValue[0] = roughCol.x;
        P.roughValue[1] = roughCol.y;
        P.roughValue[2] = roughCol.z;
        P.roughValue[3] = std::fmin((reflectDepth / 50.0f) * mat.roughness, mat.roughness / 2.1f);
        P.reflectContribution = totalContrib;
    }
    P.depth = hitT;
}

const float kAAOffsets[9][8][2] = { // raygen.h:37-67
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

// raygen.h:69-115: returns the three rows of the mat3x4 result. With the
// stochastic sampling mode the pixel offsets are hashed instead of the SSAA
// table, the colour summed is each sample's composite mix(hitValue,
// roughValue.rgb, reflectContribution) with alpha 0, and feat receives the
// denoiser features (sample 0's primary material, mean primary t, variance
// of the mean luminance).
void raygen(Ctx& c, uint32_t px, uint32_t py, uint32_t W, uint32_t H, float out[3][4], float feat[2][4] = nullptr)
{
    const UBO& ubo = *c.ubo;
    const int numSamples = ubo.numSamples;
    V3 color = v3(0.0f), normal = v3(0.0f);
    float rough[4] = {0, 0, 0, 0};
    float reflectContrib = 0.0f, depth = 0.0f;

    const float zero_one[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float org4[4];
    mat4_mul(ubo.viewInverse, zero_one, org4);
    V3 origin = v3(org4[0], org4[1], org4[2]);
    const float tmin = 0.001f, tmax = 10000.0f;

    float lum2 = 0.0f;
    for(int i = 0; i < numSamples; ++i) {
        const float* off = kAAOffsets[std::min(numSamples, 8)][i % 8];
        float offx = off[0], offy = off[1];
        if(c.sampling == 1) {
            c.key = sample_key(c.seed, py * W + px, (uint32_t)i);
            offx = u01(hash32(c.key ^ 0x68bc21ebu)) - 0.5f;
            offy = u01(hash32(c.key ^ 0x02e5be93u)) - 0.5f;
        }
        float pcx = ((float)px + 0.5f) + offx;
        float pcy = ((float)py + 0.5f) + offy;
        float ux = pcx / (float)W, uy = pcy / (float)H;
        float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
        const float t4[4] = {dx, dy, 1.0f, 1.0f};
        float target[4];
        mat4_mul(ubo.projInverse, t4, target);
        V3 dn = normalize(v3(target[0], target[1], target[2]));
        V3 direction = mat4_mul_dir(ubo.viewInverse, dn);

        Payload P;
        P.hitValue = v3(0.0f);
        P.normal = v3(0.0f);
        P.roughValue[0] = P.roughValue[1] = P.roughValue[2] = P.roughValue[3] = 0.0f;
        P.depth = 0.0f;
        P.refDepth = 0.0f;
        P.curIOR = 1.0f;
        P.rayType = RT_GENERIC;
        P.recDepth = 0;
        P.reflectContribution = 0.0f;

        c.recordPrimary = (i == 0);
        trace_ray(c, P, 0xff, 0, origin, tmin, direction, tmax, true);
        c.recordPrimary = false;

        if(c.sampling == 1) {
            const V3 col = mix(P.hitValue, v3(P.roughValue[0], P.roughValue[1], P.roughValue[2]), P.reflectContribution);
            const float l = (0.299f * col.x + 0.587f * col.y) + 0.114f * col.z;
            lum2 = lum2 + l * l;
            color = color + col;
        }
        else {
            color = color + P.hitValue;
        }
        normal = normal + P.normal;
        for(int k = 0; k < 4; ++k) rough[k] += P.roughValue[k];
        reflectContrib += P.reflectContribution;
        depth += P.depth;
    }
    const float n = (float)numSamples;
    out[0][0] = color.x / n;
    out[0][1] = color.y / n;
    out[0][2] = color.z / n;
    out[0][3] = c.sampling == 1 ? 0.0f : reflectContrib / n;
    out[1][0] = normal.x / n;
    out[1][1] = normal.y / n;
    out[1][2] = normal.z / n;
    out[1][3] = (std::log(depth) * 0.25f) / n;
    for(int k = 0; k < 4; ++k) out[2][k] = rough[k] / n;
    if(feat) {
        for(int k = 0; k < 3; ++k) feat[0][k] = c.primAlbedo[k];
        feat[0][3] = c.primRough;
        const float lm = (0.299f * out[0][0] + 0.587f * out[0][1]) + 0.114f * out[0][2];
        feat[1][0] = depth / n;
        feat[1][1] = std::fmax(lum2 / n - lm * lm, 0.0f) / n;
        feat[1][2] = feat[1][3] = 0.0f;
    }
}

// --------------------------------------------------------------------------
// Compute post passes
// --------------------------------------------------------------------------
struct Img16 {
    uint16_t* p;
    int W, H;
    void load(int x, int y, float out[4]) const
    {
        if(x < 0 || y < 0 || x >= W || y >= H) {
            out[0] = out[1] = out[2] = out[3] = 0.0f;
            return;
        }
        const uint16_t* q = p + 4 * ((size_t)y * W + x);
        for(int k = 0; k < 4; ++k) out[k] = f16_to_f32(q[k]);
    }
    void store(int x, int y, const float v[4])
    {
        uint16_t* q = p + 4 * ((size_t)y * W + x);
        for(int k = 0; k < 4; ++k) q[k] = f32_to_f16(v[k]);
    }
};

inline float dist4(const float a[4], const float b[4])
{
    float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2], d3 = a[3] - b[3];
    return std::sqrt(((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3);
}

void rough_prepare(Img16 rough, Img16 normal, int8_t* trans, Img16 A, Img16 B) // rough_prepare.comp:28-54
{
    for(int y = 0; y < rough.H; ++y)
        for(int x = 0; x < rough.W; ++x) {
            float r[4], ru[4], rd[4], rl[4], rr[4], n[4], nu[4], nd[4], nl[4], nr[4];
            rough.load(x, y, r);
            rough.load(x, y + 1, ru);
            rough.load(x, y - 1, rd);
            rough.load(x + 1, y, rl);
            rough.load(x - 1, y, rr);
            normal.load(x, y, n);
            normal.load(x, y + 1, nu);
            normal.load(x, y - 1, nd);
            normal.load(x + 1, y, nl);
            normal.load(x - 1, y, nr);
            float uf = std::fmin(ru[3], r[3]) * clampf(1.0f - dist4(nu, n) * 10.0f, 0.0f, 1.0f);
            float df = std::fmin(rd[3], r[3]) * clampf(1.0f - dist4(nd, n) * 10.0f, 0.0f, 1.0f);
            float lf = std::fmin(rl[3], r[3]) * clampf(1.0f - dist4(nl, n) * 10.0f, 0.0f, 1.0f);
            float rf = std::fmin(rr[3], r[3]) * clampf(1.0f - dist4(nr, n) * 10.0f, 0.0f, 1.0f);
            float t = std::fmin(std::fmin(std::fmin(uf, df), lf), rf);
            trans[(size_t)y * rough.W + x] = f32_to_snorm8(t);
            A.store(x, y, r);
            B.store(x, y, r);
        }
}

void rough_blur(const int8_t* trans, Img16 in, Img16 out, int ox, int oy) // rough_blur.h:23-40
{
    for(int y = 0; y < in.H; ++y)
        for(int x = 0; x < in.W; ++x) {
            float t = snorm8_to_f32(trans[(size_t)y * in.W + x]);
            if(t < 0.001f) continue;
            float r[4], r1[4], r2[4];
            in.load(x, y, r);
            in.load(x + ox, y + oy, r1);
            in.load(x - ox, y - oy, r2);
            float w0 = (1.0f - t) - t;
            float o[4];
            for(int k = 0; k < 3; ++k) o[k] = ((r[k] * w0) + (r1[k] * t)) + (r2[k] * t);
            o[3] = r[3];
            out.store(x, y, o);
        }
}

void postprocess(const UBO& ubo, Img16 base, Img16 A, Img16 fin) // postprocess.comp:28-50 (without showAlpha)
{
    for(int y = 0; y < base.H; ++y)
        for(int x = 0; x < base.W; ++x) {
            float b[4], r[4];
            base.load(x, y, b);
            A.load(x, y, r);
            float col[3];
            for(int k = 0; k < 3; ++k) col[k] = clampf(mixf(b[k], r[k], b[3]), 0.0f, 1.0f);
            float luma = ((col[0] * 0.299f) + (col[1] * 0.587f)) + (col[2] * 0.114f);
            for(int k = 0; k < 3; ++k) col[k] = mixf(col[k], ubo.fadeColor[k], ubo.fadeColor[3]);
            float o[4] = {col[0], col[1], col[2], luma};
            fin.store(x, y, o);
        }
}

// Bilinear clamp-to-edge sample at normalized coordinates (the reference's
// compute sampler: linear filter, clamp to edge, compute_system.cpp:76-85),
// with 8 bits of sub-texel precision. (ox, oy) is an integer texel offset.
void tex_sample(const Img16& img, float px, float py, int ox, int oy, float out[4])
{
    float u = (px * (float)img.W - 0.5f) + (float)ox;
    float v = (py * (float)img.H - 0.5f) + (float)oy;
    float fu = std::floor(u), fv = std::floor(v);
    float au = std::nearbyint((u - fu) * 256.0f) / 256.0f;
    float av = std::nearbyint((v - fv) * 256.0f) / 256.0f;
    int x0 = (int)fu, y0 = (int)fv;
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = std::min(std::max(x0, 0), img.W - 1);
    x1 = std::min(std::max(x1, 0), img.W - 1);
    y0 = std::min(std::max(y0, 0), img.H - 1);
    y1 = std::min(std::max(y1, 0), img.H - 1);
    float t00[4], t10[4], t01[4], t11[4];
    img.load(x0, y0, t00);
    img.load(x1, y0, t10);
    img.load(x0, y1, t01);
    img.load(x1, y1, t11);
    for(int k = 0; k < 4; ++k) {
        float a = mixf(t00[k], t10[k], au);
        float b = mixf(t01[k], t11[k], au);
        out[k] = mixf(a, b, av);
    }
}

// FXAA, quality variant, with Raygun's settings (fxaa.comp:29-36: subpixel
// quality 1.0, edge threshold 0.063, edge threshold minimum 0.0312; preset
// 39's strides 1,1,1,1,1,1.5,2,2,2,2,4,8 from fxaa.h:487-501; green as luma,
// fxaa.h:201), restated from the published algorithm (Lottes, "FXAA", NVIDIA
// white paper 2009; FXAA 3.11 quality): contrast gate on the 4-neighbour
// luma range; edge orientation from 1-2-1 weighted second differences of
// the 3x3 luma; the neighbour across the edge with the larger luma step
// picks the side and the reference luma (mean of the two); a walk along the
// half-pixel line in both directions until the luma leaves the reference by
// a quarter of that step; a shift across the edge of 0.5 - near / span when
// the nearer end turns away from the pixel, bounded below by the squared
// smoothstep of the 3x3 mean's contrast; resample there, alpha = luma.
// Walk distances are exact texel counts (the strides' running sums).
struct FxaaNeighbourhood {
    // luma[r][c] for rows y-1..y+1 and columns x-1..x+1
    float luma[3][3];
    float centre[4];
};

float sample_green(const Img16& img, float u, float v, int ox, int oy)
{
    float c[4];
    tex_sample(img, u, v, ox, oy, c);
    return c[1];
}

float walk_reach(int k)
{
    static const float reach[12] = {1.0f, 2.0f, 3.0f, 4.0f, 5.0f, 6.5f, 8.5f, 10.5f, 12.5f, 14.5f, 18.5f, 26.5f};
    return reach[k];
}

void fxaa_pixel(const Img16& tex, int x, int y, float out[4])
{
    const float u0 = ((float)x + 0.5f) / (float)tex.W, v0 = ((float)y + 0.5f) / (float)tex.H;
    const float texel[2] = {1.0f / (float)tex.W, 1.0f / (float)tex.H};
    FxaaNeighbourhood nb;
    tex_sample(tex, u0, v0, 0, 0, nb.centre);
    nb.luma[1][1] = nb.centre[1];
    // the four edge neighbours first: the gate needs only them
    nb.luma[0][1] = sample_green(tex, u0, v0, 0, -1);
    nb.luma[2][1] = sample_green(tex, u0, v0, 0, 1);
    nb.luma[1][0] = sample_green(tex, u0, v0, -1, 0);
    nb.luma[1][2] = sample_green(tex, u0, v0, 1, 0);
    const float c = nb.luma[1][1];
    const float n = nb.luma[0][1], s = nb.luma[2][1], w = nb.luma[1][0], e = nb.luma[1][2];
    const float hi = std::fmax(std::fmax(std::fmax(n, s), std::fmax(w, e)), c);
    const float lo = std::fmin(std::fmin(std::fmin(n, s), std::fmin(w, e)), c);
    const float range = hi - lo;
    if(range < std::fmax(0.0312f, hi * 0.063f)) {
        for(int k = 0; k < 4; ++k) out[k] = nb.centre[k];
        return;
    }
    for(int r = 0; r < 3; r += 2)
        for(int q = 0; q < 3; q += 2) nb.luma[r][q] = sample_green(tex, u0, v0, q - 1, r - 1);
    const float nw = nb.luma[0][0], ne = nb.luma[0][2], sw = nb.luma[2][0], se = nb.luma[2][2];

    // |second difference| of a line of three lumas
    auto curv = [](float a, float b, float d) { return std::fabs((a + d) - 2.0f * b); };
    const float rowsChange = 2.0f * curv(n, c, s) + (curv(nw, w, sw) + curv(ne, e, se));
    const float colsChange = 2.0f * curv(w, c, e) + (curv(nw, n, ne) + curv(sw, s, se));
    const int alongAxis = rowsChange >= colsChange ? 0 : 1; // 0: the edge runs along x
    const int acrossAxis = 1 - alongAxis;

    const float mean3x3 = (2.0f * ((n + s) + (w + e)) + ((nw + ne) + (sw + se))) / 12.0f;
    const float t = clampf(std::fabs(mean3x3 - c) / range, 0.0f, 1.0f);
    const float smooth = (t * t) * (3.0f - 2.0f * t);
    const float subpixelShift = (smooth * smooth) * 1.0f;

    // the two neighbours across the edge: (before, after) along the across axis
    const float before = alongAxis == 0 ? n : w, after = alongAxis == 0 ? s : e;
    const float dBefore = before - c, dAfter = after - c;
    const bool sideBefore = std::fabs(dBefore) >= std::fabs(dAfter);
    const float acrossStep = sideBefore ? -texel[acrossAxis] : texel[acrossAxis];
    const float reference = (c + (sideBefore ? before : after)) * 0.5f;
    const float stopAt = std::fmax(std::fabs(dBefore), std::fabs(dAfter)) * 0.25f;
    const bool pixelDarker = (c - reference) < 0.0f;

    float start[2] = {u0, v0};
    start[acrossAxis] = start[acrossAxis] + acrossStep * 0.5f;
    float reach[2], endDelta[2];
    for(int side = 0; side < 2; ++side) { // side 0 walks toward -along, side 1 toward +along
        const float dir = side == 0 ? -texel[alongAxis] : texel[alongAxis];
        reach[side] = walk_reach(11);
        endDelta[side] = 0.0f;
        for(int k = 0; k < 11; ++k) {
            float p[2] = {start[0], start[1]};
            p[alongAxis] = p[alongAxis] + dir * walk_reach(k);
            endDelta[side] = sample_green(tex, p[0], p[1], 0, 0) - reference;
            if(std::fabs(endDelta[side]) >= stopAt) {
                reach[side] = walk_reach(k);
                break;
            }
        }
    }
    const int nearer = reach[0] < reach[1] ? 0 : 1;
    const bool opposite = (endDelta[nearer] < 0.0f) != pixelDarker;
    const float edgeShift = opposite ? 0.5f - reach[nearer] / (reach[0] + reach[1]) : 0.0f;
    const float shift = std::fmax(edgeShift, subpixelShift);
    float fin[2] = {u0, v0};
    fin[acrossAxis] = fin[acrossAxis] + shift * acrossStep;
    float col[4];
    tex_sample(tex, fin[0], fin[1], 0, 0, col);
    out[0] = col[0];
    out[1] = col[1];
    out[2] = col[2];
    out[3] = c;
}

template<typename F>
void parallel_rows(int rowBegin, int rowEnd, int threads, F f)
{
    if(threads <= 1) {
        for(int y = rowBegin; y < rowEnd; ++y) f(y, 0);
        return;
    }
    std::atomic<int> next{rowBegin};
    std::vector<std::thread> pool;
    for(int t = 0; t < threads; ++t)
        pool.emplace_back([&, t]() {
            for(;;) {
                int y = next.fetch_add(1);
                if(y >= rowEnd) break;
                f(y, t);
            }
        });
    for(auto& th: pool) th.join();
}

} // namespace

// ==========================================================================
// C API
// ==========================================================================
extern "C" {

orc_scene* orc_scene_create(const void* vertices, uint64_t numVertices, const uint32_t* indices, uint64_t numIndices,
                            const void* materials, uint32_t numMaterials)
{
    auto* s = new orc_scene();
    s->vb.resize(numVertices);
    if(numVertices) std::memcpy(s->vb.data(), vertices, numVertices * sizeof(Vertex));
    s->ib.assign(indices, indices + numIndices);
    s->mats.resize(numMaterials);
    if(numMaterials) std::memcpy(s->mats.data(), materials, numMaterials * sizeof(Material));
    return s;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

int32_t orc_build_blas(orc_scene* s, uint32_t firstVertex, uint32_t indexOffsetBytes, uint32_t numFaces)
{
    if(indexOffsetBytes % 4) return -1;
    Blas b;
    b.firstVertex = firstVertex;
    b.indexOffset = indexOffsetBytes / 4;
    b.numFaces = numFaces;
    if((uint64_t)b.indexOffset + 3ull * numFaces > s->ib.size()) return -1;
    load_tris(*s, b);
    build_lbvh(b.bvh, g_hierarchy);
    s->blas.push_back(std::move(b));
    return (int32_t)s->blas.size() - 1;
}

int32_t orc_update_vertices(orc_scene* s, uint64_t firstVertex, const void* vertices, uint64_t count)
{
    if(firstVertex + count > s->vb.size()) return -1;
    std::memcpy(s->vb.data() + firstVertex, vertices, count * sizeof(Vertex));
    return 0;
}

int32_t orc_refit_blas(orc_scene* s, int32_t id)
{
    if(id < 0 || id >= (int32_t)s->blas.size()) return -1;
    Blas& b = s->blas[id];
    load_tris(*s, b);
    compute_node_boxes(b.bvh);
    collapse4(b.bvh); // the wide topology follows the refitted areas
    quantize4(b.bvh);
    return 0;
}

// Full rebuild of an existing BLAS from the current vertices, keeping its id
// (what a deforming mesh gets each frame: new Morton order, new hierarchy).
int32_t orc_rebuild_blas(orc_scene* s, int32_t id)
{
    if(id < 0 || id >= (int32_t)s->blas.size()) return -1;
    Blas& b = s->blas[id];
    b.bvh = BVH{};
    load_tris(*s, b);
    build_lbvh(b.bvh, g_hierarchy);
    return 0;
}

uint32_t orc_blas_num_nodes(orc_scene* s, int32_t id) { return (uint32_t)s->blas[id].bvh.nodes.size(); }

uint32_t orc_blas_num_wide(orc_scene* s, int32_t id) { return (uint32_t)s->blas[id].bvh.wide.size(); }

int32_t orc_export_blas_wide(orc_scene* s, int32_t id, int32_t compressed, void* nodes)
{
    const BVH& b = s->blas[id].bvh;
    if(compressed) std::memcpy(nodes, b.wideq.data(), b.wideq.size() * sizeof(Node4q));
    else std::memcpy(nodes, b.wide.data(), b.wide.size() * sizeof(Node4));
    return 0;
}

uint32_t orc_tlas_num_wide(orc_scene* s) { return (uint32_t)s->tlas.wide.size(); }

int32_t orc_export_tlas_wide(orc_scene* s, int32_t compressed, void* nodes)
{
    if(compressed) std::memcpy(nodes, s->tlas.wideq.data(), s->tlas.wideq.size() * sizeof(Node4q));
    else std::memcpy(nodes, s->tlas.wide.data(), s->tlas.wide.size() * sizeof(Node4));
    return 0;
}

int32_t orc_export_blas(orc_scene* s, int32_t id, void* nodes, uint32_t* leafPrim)
{
    const BVH& b = s->blas[id].bvh;
    if(nodes) std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(CanonNode));
    if(leafPrim) std::memcpy(leafPrim, b.leafPrim.data(), b.leafPrim.size() * 4);
    return 0;
}

int32_t orc_blas_morton(orc_scene* s, int32_t id, uint32_t* codes)
{
    const BVH& b = s->blas[id].bvh;
    std::memcpy(codes, b.morton.data(), b.morton.size() * 4);
    return 0;
}

int32_t orc_build_tlas(orc_scene* s, const void* instances, uint32_t n)
{
    s->inst.resize(n);
    const Instance* in = (const Instance*)instances;
    s->tlas = BVH{};
    s->tlas.primBox.resize(n);
    for(uint32_t i = 0; i < n; ++i) {
        if(in[i].blas >= s->blas.size()) return -1;
        s->inst[i].rec = in[i];
        invert_affine(in[i].transform, s->inst[i].w2o);
        s->tlas.primBox[i] = transform_box(in[i].transform, s->blas[in[i].blas].bvh.root);
    }
    build_lbvh(s->tlas, 0);
    s->hasTlas = true;
    return 0;
}

// Refit of the TLAS for moved instances: same count and BLAS per position
// (the entity DFS keeps positions when only transforms change), new boxes,
// the existing hierarchy and leaf order, then collapse and compression as in
// a build (the wide topology follows the refitted areas).
int32_t orc_refit_tlas(orc_scene* s, const void* instances, uint32_t n)
{
    if(!s->hasTlas || n != s->inst.size()) return -1;
    const Instance* in = (const Instance*)instances;
    for(uint32_t i = 0; i < n; ++i)
        if(in[i].blas != s->inst[i].rec.blas) return -1;
    for(uint32_t i = 0; i < n; ++i) {
        s->inst[i].rec = in[i];
        invert_affine(in[i].transform, s->inst[i].w2o);
        s->tlas.primBox[i] = transform_box(in[i].transform, s->blas[in[i].blas].bvh.root);
    }
    compute_node_boxes(s->tlas);
    collapse4(s->tlas);
    quantize4(s->tlas);
    return 0;
}

uint32_t orc_tlas_num_nodes(orc_scene* s) { return (uint32_t)s->tlas.nodes.size(); }

int32_t orc_export_tlas(orc_scene* s, void* nodes, uint32_t* leafInstance)
{
    if(nodes) std::memcpy(nodes, s->tlas.nodes.data(), s->tlas.nodes.size() * sizeof(CanonNode));
    if(leafInstance) std::memcpy(leafInstance, s->tlas.leafPrim.data(), s->tlas.leafPrim.size() * 4);
    return 0;
}

int32_t orc_trace_rays(orc_scene* s, const float* rays, uint32_t n, uint32_t cullMask, void* hits, uint8_t* tie,
                       uint32_t* nodes, uint32_t* tris, int32_t threads)
{
    HitOut* out = (HitOut*)hits;
    const int chunk = 1024;
    const int nchunks = (int)((n + chunk - 1) / chunk);
    parallel_rows(0, nchunks, threads, [&](int c, int) {
        uint32_t end = std::min<uint32_t>(n, (uint32_t)(c + 1) * chunk);
        for(uint32_t i = (uint32_t)c * chunk; i < end; ++i) {
            const float* r = rays + 8 * (size_t)i;
            Hit h;
            bool hit = trace_closest(*s, r, r + 4, r[3], r[7], cullMask, h);
            out[i] = hit ? HitOut{h.t, h.u, h.v, s->inst[h.inst].rec.customIndex, h.prim} : HitOut{0, 0, 0, NO_HIT, 0};
            if(tie) tie[i] = hit && h.tie;
            if(nodes) nodes[i] = h.nodes;
            if(tris) tris[i] = h.tris;
        }
    });
    return 0;
}

int32_t orc_trace_rays_fp64(orc_scene* s, const float* rays, uint32_t n, uint32_t cullMask, double* t_out, uint32_t* inst_out,
                            uint32_t* prim_out, double* uv_out, int32_t threads)
{
    parallel_rows(0, (int)n, threads, [&](int i, int) {
        const float* r = rays + 8 * (size_t)i;
        double t, u = 0, v = 0;
        uint32_t inst = NO_HIT, prim = 0;
        bool hit = brute_fp64(*s, r, r + 4, r[3], r[7], cullMask, t, inst, prim, u, v);
        t_out[i] = hit ? t : 0.0;
        inst_out[i] = hit ? inst : NO_HIT;
        prim_out[i] = prim;
        if(uv_out) {
            uv_out[2 * i] = u;
            uv_out[2 * i + 1] = v;
        }
    });
    return 0;
}

int32_t orc_trace_rays_brute32(orc_scene* s, const float* rays, uint32_t n, uint32_t cullMask, float* t_out, uint32_t* inst_out,
                               uint32_t* prim_out, int32_t threads)
{
    parallel_rows(0, (int)n, threads, [&](int i, int) {
        const float* r = rays + 8 * (size_t)i;
        float t;
        uint32_t inst = NO_HIT, prim = 0;
        const bool hit = brute_fp32(*s, r, r + 4, r[3], r[7], cullMask, t, inst, prim);
        t_out[i] = hit ? t : 0.0f;
        inst_out[i] = hit ? inst : NO_HIT;
        prim_out[i] = prim;
    });
    return 0;
}

int32_t orc_render(orc_scene* s, const void* uboPtr, uint32_t W, uint32_t H, uint16_t* base16, uint16_t* normal16,
                   uint16_t* rough16, float* base32, float* normal32, float* rough32, void* primary, uint8_t* tie,
                   uint64_t* stats, int32_t threads, int32_t rowBegin, int32_t rowEnd)
{
    if(!s->hasTlas) return -1;
    UBO ubo;
    std::memcpy(&ubo, uboPtr, sizeof(UBO));
    if(rowEnd <= 0 || rowEnd > (int32_t)H) rowEnd = (int32_t)H;
    if(rowBegin < 0) rowBegin = 0;
    const int nt = std::max(1, (int)threads);
    std::vector<uint64_t> rays(nt, 0), sky(nt, 0), zd(nt, 0);
    parallel_rows(rowBegin, rowEnd, nt, [&](int y, int t) {
        Ctx c;
        c.s = s;
        c.ubo = &ubo;
        c.lightDir = v3(ubo.lightDir[0], ubo.lightDir[1], ubo.lightDir[2]);
        for(uint32_t x = 0; x < W; ++x) {
            c.tie = false;
            float out[3][4];
            raygen(c, x, (uint32_t)y, W, H, out);
            size_t p = (size_t)y * W + x;
            for(int k = 0; k < 4; ++k) {
                if(base16) base16[4 * p + k] = f32_to_f16(out[0][k]);
                if(normal16) normal16[4 * p + k] = f32_to_f16(out[1][k]);
                if(rough16) rough16[4 * p + k] = f32_to_f16(out[2][k]);
                if(base32) base32[4 * p + k] = out[0][k];
                if(normal32) normal32[4 * p + k] = out[1][k];
                if(rough32) rough32[4 * p + k] = out[2][k];
            }
            if(primary) ((HitOut*)primary)[p] = c.primary;
            if(tie) tie[p] = c.tie;
        }
        rays[t] += c.rays;
        sky[t] += c.sky;
        zd[t] += c.zeroDir;
    });
    if(stats) {
        stats[0] = stats[1] = stats[2] = 0;
        for(int t = 0; t < nt; ++t) {
            stats[0] += rays[t];
            stats[1] += sky[t];
            stats[2] += zd[t];
        }
    }
    return 0;
}

// orc_render with the sampling mode (0 Whitted, 1 stochastic) and the
// denoiser feature image (2 x rgba16f per pixel) of the stochastic mode.
int32_t orc_render_sampled(orc_scene* s, const void* uboPtr, uint32_t W, uint32_t H, uint32_t sampling, uint32_t seed,
                           uint16_t* base16, uint16_t* normal16, uint16_t* rough16, float* base32, uint16_t* features16,
                           uint8_t* tie, uint64_t* stats, int32_t threads)
{
    if(!s->hasTlas) return -1;
    UBO ubo;
    std::memcpy(&ubo, uboPtr, sizeof(UBO));
    const int nt = std::max(1, (int)threads);
    std::vector<uint64_t> rays(nt, 0);
    parallel_rows(0, (int)H, nt, [&](int y, int t) {
        Ctx c;
        c.s = s;
        c.ubo = &ubo;
        c.lightDir = v3(ubo.lightDir[0], ubo.lightDir[1], ubo.lightDir[2]);
        c.sampling = sampling;
        c.seed = seed;
        for(uint32_t x = 0; x < W; ++x) {
            c.tie = false;
            float out[3][4], feat[2][4];
            raygen(c, x, (uint32_t)y, W, H, out, feat);
            size_t p = (size_t)y * W + x;
            for(int k = 0; k < 4; ++k) {
                if(base16) base16[4 * p + k] = f32_to_f16(out[0][k]);
                if(normal16) normal16[4 * p + k] = f32_to_f16(out[1][k]);
                if(rough16) rough16[4 * p + k] = f32_to_f16(out[2][k]);
                if(base32) base32[4 * p + k] = out[0][k];
                if(features16) {
                    features16[8 * p + k] = f32_to_f16(feat[0][k]);
                    features16[8 * p + 4 + k] = f32_to_f16(feat[1][k]);
                }
            }
            if(tie) tie[p] = c.tie;
        }
        rays[t] += c.rays;
    });
    if(stats) {
        stats[0] = 0;
        for(int t = 0; t < nt; ++t) stats[0] += rays[t];
    }
    return 0;
}

int32_t orc_render_counts(orc_scene* s, const void* uboPtr, uint32_t W, uint32_t H, const int32_t* rows, int32_t nrows,
                          int32_t threads, uint64_t* out)
{
    if(!s->hasTlas) return -1;
    UBO ubo;
    std::memcpy(&ubo, uboPtr, sizeof(UBO));
    const int nt = std::max(1, (int)threads);
    constexpr int K = 7;
    std::vector<uint64_t> acc(K * nt, 0);
    parallel_rows(0, nrows, nt, [&](int r, int t) {
        Ctx c;
        c.s = s;
        c.ubo = &ubo;
        c.lightDir = v3(ubo.lightDir[0], ubo.lightDir[1], ubo.lightDir[2]);
        const uint32_t y = (uint32_t)rows[r];
        for(uint32_t x = 0; x < W; ++x) {
            float o[3][4];
            raygen(c, x, y, W, H, o);
        }
        acc[K * t + 0] += c.rays;
        acc[K * t + 1] += c.nodes;
        acc[K * t + 2] += c.tris;
        acc[K * t + 3] += c.hits;
        acc[K * t + 4] += c.sky;
        acc[K * t + 5] += c.instLeaves;
        acc[K * t + 6] += c.blasEntries;
    });
    for(int k = 0; k < K; ++k) {
        out[k] = 0;
        for(int t = 0; t < nt; ++t) out[k] += acc[K * t + k];
    }
    return 0;
}

int32_t orc_postprocess(const void* uboPtr, uint32_t W, uint32_t H, uint16_t* base, uint16_t* normal, uint16_t* rough,
                        uint16_t* final_, int8_t* transitions, uint16_t* colorsA, uint16_t* colorsB, int32_t useFxaa)
{
    UBO ubo;
    std::memcpy(&ubo, uboPtr, sizeof(UBO));
    Img16 iBase{base, (int)W, (int)H}, iNormal{normal, (int)W, (int)H}, iRough{rough, (int)W, (int)H};
    Img16 iFinal{final_, (int)W, (int)H}, iA{colorsA, (int)W, (int)H}, iB{colorsB, (int)W, (int)H};
    rough_prepare(iRough, iNormal, transitions, iA, iB);
    for(int i = 0; i < 10; ++i) {
        rough_blur(transitions, iA, iB, 1, 0);
        rough_blur(transitions, iB, iA, 0, 1);
    }
    postprocess(ubo, iBase, iA, iFinal);
    if(ubo.showAlpha) {
        // postprocess.comp:41-49: every image shows its alpha channel
        Img16* imgs[] = {&iFinal, &iNormal, &iRough, &iBase, &iA, &iB};
        for(Img16* im: imgs)
            for(int y = 0; y < (int)H; ++y)
                for(int x = 0; x < (int)W; ++x) {
                    float v[4];
                    im->load(x, y, v);
                    float a[4] = {v[3], v[3], v[3], v[3]};
                    im->store(x, y, a);
                }
        // roughTransitions is single-channel: .aaaa of an R8 image reads a = 1
        for(size_t p = 0; p < (size_t)W * H; ++p) transitions[p] = f32_to_snorm8(1.0f);
    }
    if(useFxaa) {
        for(int y = 0; y < (int)H; ++y)
            for(int x = 0; x < (int)W; ++x) {
                float o[4];
                fxaa_pixel(iFinal, x, y, o);
                iBase.store(x, y, o);
            }
    }
    return 0;
}

uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

void orc_sky(const float* dir, const float* lightDir, float* rgb)
{
    V3 r = sky(v3(dir[0], dir[1], dir[2]), v3(lightDir[0], lightDir[1], lightDir[2]));
    rgb[0] = r.x;
    rgb[1] = r.y;
    rgb[2] = r.z;
}

uint32_t orc_morton3(float x, float y, float z) { return morton3(x, y, z); }

void orc_set_hierarchy(int32_t h) { g_hierarchy = h; }

void orc_primary_ray(const void* uboPtr, uint32_t x, uint32_t y, uint32_t W, uint32_t H, int32_t sample, float* origin, float* dir)
{
    UBO ubo;
    std::memcpy(&ubo, uboPtr, sizeof(UBO));
    const int numSamples = ubo.numSamples;
    const float* off = kAAOffsets[std::min(numSamples, 8)][sample % 8];
    float pcx = ((float)x + 0.5f) + off[0];
    float pcy = ((float)y + 0.5f) + off[1];
    float ux = pcx / (float)W, uy = pcy / (float)H;
    float dx = ux * 2.0f - 1.0f, dy = uy * 2.0f - 1.0f;
    const float t4[4] = {dx, dy, 1.0f, 1.0f};
    float target[4];
    mat4_mul(ubo.projInverse, t4, target);
    V3 dn = normalize(v3(target[0], target[1], target[2]));
    V3 d = mat4_mul_dir(ubo.viewInverse, dn);
    const float zero_one[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float org4[4];
    mat4_mul(ubo.viewInverse, zero_one, org4);
    origin[0] = org4[0];
    origin[1] = org4[1];
    origin[2] = org4[2];
    dir[0] = d.x;
    dir[1] = d.y;
    dir[2] = d.z;
}

} // extern "C"
