// rgrt_device.h — device-side records and GLSL built-ins for the HIP core.
//
// Layouts are the C ABI's (include/rgrt.h). The built-ins follow the GLSL
// 4.60 definitions, evaluated in the order written; files that include this
// header compile with floating-point contraction off (see the pragma at the
// top of each .hip file) so results do not depend on fused multiply-add
// choices of the compiler.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rgrt.h"

namespace rgrt {

constexpr int32_t kEmptyChild = (int32_t)0x80000000;
constexpr int32_t kSentinel = (int32_t)0x80000000; // stack marker: return to the TLAS
constexpr uint32_t kNoHit = 0xffffffffu;

// Loads through pointers that the compiler cannot prove global (they are
// read from other records) would otherwise be flat instructions, which also
// wait on LDS traffic; these go through the global address space.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u32x4 ld_global16(const void* p)
{
    return *(const __attribute__((address_space(1))) u32x4*)p;
}
__device__ __forceinline__ u32x2 ld_global8(const void* p)
{
    return *(const __attribute__((address_space(1))) u32x2*)p;
}

// ---------------------------------------------------------------------------
// BVH records in HBM
// ---------------------------------------------------------------------------
// Node: 64 bytes, the canonical rgrt_bvh_node layout; the traversal reads it
// as four 16-byte vectors.
struct __align__(16) Node {
    float lo0x, lo0y, lo0z, lo1x; // lo[0][0..2], lo[1][0]
    float lo1y, lo1z, hi0x, hi0y; // lo[1][1..2], hi[0][0..1]
    float hi0z, hi1x, hi1y, hi1z; // hi[0][2], hi[1][0..2]
    int32_t child0, child1;
    uint32_t parent, pad;
};
static_assert(sizeof(Node) == 64, "node is 64 bytes");
static_assert(sizeof(Node) == sizeof(rgrt_bvh_node), "canonical node layout");

// 4-wide node the traversal walks, 128 bytes: the per-axis bounds of the
// four child boxes (one 16-byte vector each), then the four child refs
// (>= 0 wide node, < 0 leaf ~k, kEmptyChild for an unused slot).
struct __align__(16) Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "wide node is 128 bytes");

// Compressed 4-wide node, 64 bytes — what the traversal reads. The node's
// box min corner, a power-of-two scale 2^e per axis, and every child box as
// 8-bit multiples of the scale from that corner, rounded outwards. Read as
// four 16-byte vectors: (origin, exponent bytes), (qlo/qhi x, qlo/qhi y),
// (qlo/qhi z, pad), child refs.
struct __align__(16) Node4q {
    float ox, oy, oz;
    int8_t ex, ey, ez, pad0;
    uint8_t qlox[4], qhix[4], qloy[4], qhiy[4];
    uint8_t qloz[4], qhiz[4];
    uint32_t pad1[2];
    int32_t child[4];
};
static_assert(sizeof(Node4q) == 64, "compressed wide node is 64 bytes");

// Leaf triangle in Morton order: the three vertices and the original
// primitive index (gl_PrimitiveID), 48 bytes. A BLAS stores its triangles
// three times, [3][numFaces]: copy r holds every vertex's coordinates in the
// order (r + 1, r + 2, r) mod 3, the axis order of the watertight test for
// a ray whose largest direction component is r, so the test reads its
// permuted coordinates directly (no per-lane selects).
struct __align__(16) Tri {
    float v0x, v0y, v0z;
    uint32_t prim;
    float v1x, v1y, v1z;
    uint32_t pad0;
    float v2x, v2y, v2z;
    uint32_t pad1;
};
static_assert(sizeof(Tri) == 48, "triangle is 48 bytes");

// Shading record of one triangle of the index buffer (triangle g = indices
// 3g .. 3g + 2), 48 bytes: the three vertex normals and vertex 0's material
// index as the closest-hit stage reads them from vertex offset
// `vertexOffset` (closesthit.rchit:95-110), so a hit reads one record
// instead of three indices and three vertices. Written by every BLAS build
// over its index range with its first vertex; kNoShadeRecord marks a
// triangle no BLAS has written since the geometry was uploaded.
struct __align__(16) ShadeTri {
    float n0[3];
    uint32_t matIndex;
    float n1[3];
    uint32_t vertexOffset;
    float n2[3];
    uint32_t pad;
};
static_assert(sizeof(ShadeTri) == 48, "shading record is 48 bytes");
constexpr uint32_t kNoShadeRecord = 0xffffffffu;

// TLAS leaf record in Morton order, 64 bytes.
struct __align__(16) DevInstance {
    float w2o[12];     // world -> object, row-major 3x4
    uint32_t blas;     // BLAS handle
    uint32_t mask;     // instance mask (low 8 bits)
    uint32_t position; // index in the caller's instance array
    uint32_t customIndex;
};
static_assert(sizeof(DevInstance) == 64, "instance leaf is 64 bytes");

struct __align__(16) DevBlas {
    const Node4q* nodes;
    const Tri* tris;
    uint32_t numFaces;
    uint32_t pad[3];
};
static_assert(sizeof(DevBlas) == 32, "BLAS table entry is 32 bytes");

// Per-instance data the shading reads, indexed by instance position.
struct __align__(16) InstanceShade {
    float o2w[12];
    rgrt_instance_offsets offsets; // the offset table entry at this position
    uint32_t customIndex;
};
static_assert(sizeof(InstanceShade) == 64, "shade record is 64 bytes");

// Scene seen by the traversal.
struct TraceScene {
    const Node4q* tlasNodes;
    const DevInstance* instances; // Morton order
    const DevBlas* blas;
    uint32_t numInstances;
    uint32_t tlasNodeCap; // TLAS wide nodes allocated
    uint32_t numBlas;     // entries of the BLAS table
    uint32_t* errors;     // sticky flags of indices found out of range (kTravErr*)
};

// Flags a traversal sets when a tree or table index is out of range (the
// ray then ends as a miss): a node, an instance leaf, a BLAS handle, a
// triangle leaf.
constexpr uint32_t kTravErrNode = 1u, kTravErrInstance = 2u, kTravErrBlas = 4u, kTravErrTriangle = 8u;

// ---------------------------------------------------------------------------
// GLSL built-ins (fp32, evaluated as written)
// ---------------------------------------------------------------------------
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3 f3(float s) { return {s, s, s}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
__device__ __forceinline__ float length(F3 a) { return __builtin_sqrtf(dot(a, a)); }
__device__ __forceinline__ F3 normalize(F3 a)
{
    const float l = length(a);
    return {a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float distance(F3 a, F3 b) { return length(a - b); }
__device__ __forceinline__ float mixf(float x, float y, float a) { return (x * (1.0f - a)) + (y * a); }
__device__ __forceinline__ F3 mix(F3 x, F3 y, float a) { return {mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a)}; }

// ---------------------------------------------------------------------------
// Random numbers of the stochastic sampling mode: the "lowbias32" integer
// finaliser of C. Wellons' hash-prospector, keyed by (seed, image pixel,
// sample) and, for glossy reflections, the hit position's bits. Integer only,
// so the oracle reproduces every number bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t sample_key(uint32_t seed, uint32_t pixel, uint32_t sample)
{
    return hash32(seed ^ hash32(pixel ^ hash32(sample ^ 0x9e3779b9u)));
}
// [0, 1) in steps of 2^-24
__device__ __forceinline__ float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

// Glossy reflection: the mirror direction r spread by roughness * s, s
// uniform in [-1, 1]^3 from the hit position's hash; r itself if the spread
// direction falls below the surface (normal n).
__device__ __forceinline__ F3 glossy_dir(uint32_t key, F3 origin, F3 r, F3 n, float roughness)
{
    const uint32_t h =
        hash32(key ^ hash32(__float_as_uint(origin.x) ^ hash32(__float_as_uint(origin.y) ^ hash32(__float_as_uint(origin.z)))));
    const F3 sp = f3(u01(hash32(h ^ 1u)) * 2.0f - 1.0f, u01(hash32(h ^ 2u)) * 2.0f - 1.0f, u01(hash32(h ^ 3u)) * 2.0f - 1.0f);
    const F3 g = normalize(r + sp * roughness);
    return dot(g, n) > 0.0f ? g : r;
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float glsl_mod(float x, float y) { return x - y * floorf(x / y); }
__device__ __forceinline__ F3 reflect(F3 I, F3 N)
{
    const float d2 = 2.0f * dot(N, I);
    return {I.x - d2 * N.x, I.y - d2 * N.y, I.z - d2 * N.z};
}
__device__ __forceinline__ F3 refract(F3 I, F3 N, float eta)
{
    const float d = dot(N, I);
    const float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if(k < 0.0f) return f3(0.0f);
    const float s = eta * d + __builtin_sqrtf(k);
    return {eta * I.x - s * N.x, eta * I.y - s * N.y, eta * I.z - s * N.z};
}

} // namespace rgrt
