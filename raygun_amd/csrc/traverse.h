// traverse.h — closest-hit traversal of the two-level LBVH.
//
// Stands in for the driver's traversal behind traceRayEXT
// (resources/shaders/raygen.h:102, closesthit.rchit:133,144,158,193,215,232,244):
// opaque geometry, no face culling (the instance flag set by
// raygun/render/acceleration_structure.cpp:41), instance mask test
// (mask & cullMask), t measured along the unnormalised world ray.
//
// Depth-first over the 4-wide trees (child refs >= 0 wide node, < 0 leaf ~k):
// the hit children of a node are ordered by entry distance, the nearest is
// visited next and the others wait on the stack. An instance leaf maps the
// ray to object space, pushes a sentinel and continues at the BLAS root;
// popping the sentinel restores the world ray. A candidate is accepted when
// tmin <= t < tbest, so the first triangle found wins an exact tie.
//
// The stack lives in LDS (kLdsStack entries per thread, lane-interleaved so
// a wave's pushes hit distinct banks) and spills to a per-thread global
// region beyond that depth. A wide node pushes at most 3 entries and an
// LBVH is at most 63 levels deep, so two levels need at most
// 2 x (3 x 63 + 1) + 1 = 381 entries: kLdsStack + spill = 384.
#pragma once

#include "rgrt_device.h"

namespace rgrt {

struct RayIn {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
};

struct HitRec {
    float t, u, v;
    uint32_t inst; // instance position, kNoHit on a miss
    uint32_t prim;
};

// Reciprocal for the slab test: a zero component becomes +-1e-20 so the
// fused form lo*inv - o*inv stays finite. The box test only culls, so the
// hardware reciprocal (v_rcp_f32, within 1 ulp) is enough; the far distance
// is widened for its error (qbox_enter).
__device__ __forceinline__ float safe_rcp(float d)
{
    const float a = fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d);
    return __builtin_amdgcn_rcpf(a);
}

constexpr int kBoxRayWords = 6; // origin, 1/d

struct BoxRay {
    float ox, oy, oz; // origin
    float ix, iy, iz; // 1/d (zero components made finite)
};

__device__ __forceinline__ float pow2i(int e) { return __uint_as_float((uint32_t)(e + 127) << 23); }

__device__ __forceinline__ void cswap(float& ta, int32_t& ra, float& tb, int32_t& rb)
{
    const bool sw = tb < ta;
    const float t = sw ? tb : ta;
    const int32_t r = sw ? rb : ra;
    tb = sw ? ta : tb;
    rb = sw ? ra : rb;
    ta = t;
    ra = r;
}

// Entry distance of child c of a compressed node, +inf on a miss. A plane at
// origin + q * 2^e lies at t = q * (2^e / d) + (origin - o) / d: one fused
// multiply-add per plane (the box test only culls). The words nx..fz hold
// the byte offsets of the near and far plane of each axis: for a direction
// component >= 0 the low bound is nearer, otherwise the high bound, and
// fma(q, a, b) is monotonic in q, so these are the min and max of the two
// planes' distances. Error bounds: b = (origin - o) / d carries two
// roundings, an absolute error up to |b| * 2^-23 that does not shrink with t
// (a ray starting far from the node's corner), so near planes use
// bn = b - |b| * 2^-21 and far planes bf = b + |b| * 2^-21 (per axis: an
// axis with a tiny direction component has a huge b, and its planes only
// bind when they are as far away); the relative part is 1 + 2*gamma(3)
// (Ize 2013) plus twice the reciprocal's 1-ulp error (near and far may come
// from different axes): 1 + 14 * 2^-24.
__device__ __forceinline__ float qbox_enter(uint32_t nx, uint32_t fx, uint32_t ny, uint32_t fy, uint32_t nz, uint32_t fz, int c,
                                            float ax, float bxn, float bxf, float ay, float byn, float byf, float az, float bzn,
                                            float bzf, float tmin, float tmax)
{
    const int sh = 8 * c;
    const float xn = __builtin_fmaf((float)((nx >> sh) & 0xffu), ax, bxn), xf = __builtin_fmaf((float)((fx >> sh) & 0xffu), ax, bxf);
    const float yn = __builtin_fmaf((float)((ny >> sh) & 0xffu), ay, byn), yf = __builtin_fmaf((float)((fy >> sh) & 0xffu), ay, byf);
    const float zn = __builtin_fmaf((float)((nz >> sh) & 0xffu), az, bzn), zf = __builtin_fmaf((float)((fz >> sh) & 0xffu), az, bzf);
    const float n = fmaxf(fmaxf(xn, yn), fmaxf(zn, tmin));
    const float f = fminf(fminf(xf, yf), fminf(zf, tmax));
    return n <= f * 1.0000008f ? n : INFINITY;
}

// Visits a compressed wide node (64 bytes, four 16-byte loads): tests the
// four child boxes, orders them by entry distance with the comparator
// network (0,1) (2,3) (0,2) (1,3) (1,2) (a swap only on a strictly smaller
// distance; missed and empty children sort last at +inf), pushes the hit
// children but the nearest farthest first and returns the nearest
// (kEmptyChild if none was hit).
template<typename Stack>
__device__ __forceinline__ int32_t visit_wide(const Node4q* nodes, int32_t ref, const BoxRay& br, float tmin, float tbest, Stack& stack)
{
    const u32x4* np = reinterpret_cast<const u32x4*>(nodes + ref);
    const u32x4 v0 = ld_global16(np), v1 = ld_global16(np + 1), v2 = ld_global16(np + 2), v3 = ld_global16(np + 3);
    const int4 ch = make_int4((int)v3.x, (int)v3.y, (int)v3.z, (int)v3.w);
    const int ex = (int)(int8_t)(v0.w & 0xffu), ey = (int)(int8_t)((v0.w >> 8) & 0xffu), ez = (int)(int8_t)((v0.w >> 16) & 0xffu);
    const float ax = pow2i(ex) * br.ix, ay = pow2i(ey) * br.iy, az = pow2i(ez) * br.iz;
    const float bx = (__uint_as_float(v0.x) - br.ox) * br.ix;
    const float by = (__uint_as_float(v0.y) - br.oy) * br.iy;
    const float bz = (__uint_as_float(v0.z) - br.oz) * br.iz;
    const float bxn = __builtin_fmaf(-fabsf(bx), 0x1p-21f, bx), bxf = __builtin_fmaf(fabsf(bx), 0x1p-21f, bx);
    const float byn = __builtin_fmaf(-fabsf(by), 0x1p-21f, by), byf = __builtin_fmaf(fabsf(by), 0x1p-21f, by);
    const float bzn = __builtin_fmaf(-fabsf(bz), 0x1p-21f, bz), bzf = __builtin_fmaf(fabsf(bz), 0x1p-21f, bz);
    const bool sx = br.ix < 0.0f, sy = br.iy < 0.0f, sz = br.iz < 0.0f;
    const uint32_t nx = sx ? v1.y : v1.x, fx = sx ? v1.x : v1.y;
    const uint32_t ny = sy ? v1.w : v1.z, fy = sy ? v1.z : v1.w;
    const uint32_t nz = sz ? v2.y : v2.x, fz = sz ? v2.x : v2.y;
    // empty slots hold zero offsets, so their (discarded) tests are harmless
    // and all four run without branches
    const float e0 = qbox_enter(nx, fx, ny, fy, nz, fz, 0, ax, bxn, bxf, ay, byn, byf, az, bzn, bzf, tmin, tbest);
    const float e1 = qbox_enter(nx, fx, ny, fy, nz, fz, 1, ax, bxn, bxf, ay, byn, byf, az, bzn, bzf, tmin, tbest);
    const float e2 = qbox_enter(nx, fx, ny, fy, nz, fz, 2, ax, bxn, bxf, ay, byn, byf, az, bzn, bzf, tmin, tbest);
    const float e3 = qbox_enter(nx, fx, ny, fy, nz, fz, 3, ax, bxn, bxf, ay, byn, byf, az, bzn, bzf, tmin, tbest);
    float t0 = ch.x != kEmptyChild ? e0 : INFINITY;
    float t1 = ch.y != kEmptyChild ? e1 : INFINITY;
    float t2 = ch.z != kEmptyChild ? e2 : INFINITY;
    float t3 = ch.w != kEmptyChild ? e3 : INFINITY;
    int32_t r0 = ch.x, r1 = ch.y, r2 = ch.z, r3 = ch.w;
    cswap(t0, r0, t1, r1);
    cswap(t2, r2, t3, r3);
    cswap(t0, r0, t2, r2);
    cswap(t1, r1, t3, r3);
    cswap(t1, r1, t2, r2);
    if(t3 != INFINITY) stack.push(r3);
    if(t2 != INFINITY) stack.push(r2);
    if(t1 != INFINITY) stack.push(r1);
    return t0 != INFINITY ? r0 : kEmptyChild;
}

// Per-ray setup of the watertight test (Woop, Benthin, Wald 2013). kz is
// the axis of the largest direction component, kx = kz + 1 and ky = kz + 2
// (mod 3); the paper swaps kx and ky when d[kz] < 0 to keep the triangles'
// winding, which only negates U, V, W, det and T exactly (the roundings are
// symmetric), so t, u, v and the accept decision are the same without it.
// The triangle copy `rot` = kz holds the vertices in (kx, ky, kz) order.
struct WtRay {
    int rot;
    float Sx, Sy, Sz;
    float okx, oky, okz; // origin in (kx, ky, kz) order
};

__device__ __forceinline__ float sel3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ void wt_setup(WtRay& w, float ox, float oy, float oz, float dx, float dy, float dz)
{
    int kz = 0;
    float m = fabsf(dx);
    if(fabsf(dy) > m) {
        kz = 1;
        m = fabsf(dy);
    }
    if(fabsf(dz) > m) kz = 2;
    const int kx = kz + 1 == 3 ? 0 : kz + 1;
    const int ky = kx + 1 == 3 ? 0 : kx + 1;
    w.rot = kz;
    const float dkz = sel3(dx, dy, dz, kz);
    const float dkx = sel3(dx, dy, dz, kx), dky = sel3(dx, dy, dz, ky);
    w.Sz = 1.0f / dkz; // one division; the shears reuse it
    w.Sx = dkx * w.Sz;
    w.Sy = dky * w.Sz;
    w.okx = sel3(ox, oy, oz, kx);
    w.oky = sel3(ox, oy, oz, ky);
    w.okz = sel3(ox, oy, oz, kz);
}

// Watertight ray/triangle test on a triangle copy in the ray's axis order;
// on a hit with tmin <= t < tbest returns true with t and the weights u
// (vertex 1), v (vertex 2).
__device__ __forceinline__ bool wt_intersect(const WtRay& w, const Tri& tri, float tmin, float tbest, float& t, float& u,
                                             float& v)
{
    const float Akx = tri.v0x - w.okx;
    const float Aky = tri.v0y - w.oky;
    const float Akz = tri.v0z - w.okz;
    const float Bkx = tri.v1x - w.okx;
    const float Bky = tri.v1y - w.oky;
    const float Bkz = tri.v1z - w.okz;
    const float Ckx = tri.v2x - w.okx;
    const float Cky = tri.v2y - w.oky;
    const float Ckz = tri.v2z - w.okz;
    const float Ax = Akx - w.Sx * Akz;
    const float Ay = Aky - w.Sy * Akz;
    const float Bx = Bkx - w.Sx * Bkz;
    const float By = Bky - w.Sy * Bkz;
    const float Cx = Ckx - w.Sx * Ckz;
    const float Cy = Cky - w.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if(U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if(det == 0.0f) return false;
    const float Az = w.Sz * Akz;
    const float Bz = w.Sz * Bkz;
    const float Cz = w.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float tt = T * rcp;
    if(!(tt >= tmin && tt < tbest)) return false;
    t = tt;
    u = V * rcp;
    v = W * rcp;
    return true;
}

template<int kLdsStack, int kBlock, int kSpillWords>
struct TraversalStack {
    int32_t* lds;   // base of this block's [kLdsStack][kBlock] array
    int32_t* spill; // this thread's global overflow region (kSpillWords)
    int tid;
    // base of this block's [kBoxRayWords][kBlock] array: the box test's ray
    // (origin, 1/d) of the tree being walked. It changes only at instance
    // entries and exits; held here it is no loop-carried register state,
    // which a wave's divergent step kinds would otherwise copy between
    // registers in every iteration.
    float* brl;
    int sp = 0;
    __device__ __forceinline__ void set_box_ray(float ox, float oy, float oz, float ix, float iy, float iz)
    {
        brl[0 * kBlock + tid] = ox;
        brl[1 * kBlock + tid] = oy;
        brl[2 * kBlock + tid] = oz;
        brl[3 * kBlock + tid] = ix;
        brl[4 * kBlock + tid] = iy;
        brl[5 * kBlock + tid] = iz;
    }
    __device__ __forceinline__ BoxRay box_ray() const
    {
        BoxRay b;
        b.ox = brl[0 * kBlock + tid];
        b.oy = brl[1 * kBlock + tid];
        b.oz = brl[2 * kBlock + tid];
        b.ix = brl[3 * kBlock + tid];
        b.iy = brl[4 * kBlock + tid];
        b.iz = brl[5 * kBlock + tid];
        return b;
    }
    // kLdsStack + kSpillWords >= 384 covers the deepest two-level tree, so
    // no entry is ever dropped; the bound check only keeps every access
    // inside the allocation.
    __device__ __forceinline__ void push(int32_t v)
    {
        if(sp < kLdsStack) lds[sp * kBlock + tid] = v;
        else if(sp < kLdsStack + kSpillWords) spill[sp - kLdsStack] = v;
        ++sp;
    }
    __device__ __forceinline__ int32_t pop()
    {
        --sp;
        if(sp < kLdsStack) return lds[sp * kBlock + tid];
        if(sp < kLdsStack + kSpillWords) return spill[sp - kLdsStack];
        return kSentinel;
    }
};

// Object-space mapping of a world ray by a 3x4 row-major matrix, summed
// left to right without fused operations.
__device__ __forceinline__ void xform_ray(const float* w, float ox, float oy, float oz, float dx, float dy, float dz, float& rox,
                                          float& roy, float& roz, float& rdx, float& rdy, float& rdz)
{
    rox = ((w[0] * ox + w[1] * oy) + w[2] * oz) + w[3];
    roy = ((w[4] * ox + w[5] * oy) + w[6] * oz) + w[7];
    roz = ((w[8] * ox + w[9] * oy) + w[10] * oz) + w[11];
    rdx = (w[0] * dx + w[1] * dy) + w[2] * dz;
    rdy = (w[4] * dx + w[5] * dy) + w[6] * dz;
    rdz = (w[8] * dx + w[9] * dy) + w[10] * dz;
}

// Traversal state of one ray between steps (the box test's ray is in the
// stack's LDS, see TraversalStack::brl; h.t is tbest once h.inst is set).
struct Traversal {
    RayIn ray;
    float wix, wiy, wiz; // 1/d of the world ray (safe_rcp)
    WtRay wt;
    const Node4q* nodes;
    const Tri* tris;
    float tbest;
    int32_t ref;
    uint32_t cullMask;
    uint32_t curInst;
    uint32_t nodeCap;  // nodes of the tree being walked
    uint32_t numFaces; // triangles of the BLAS being walked
    bool inBlas;
    HitRec h;
};

// Every index a traversal step uses is checked against the table it
// indexes; a bad one (a corrupt tree or table) ends the ray as a miss and
// raises a sticky flag that the host turns into an error.
__device__ __noinline__ void trav_error(const TraceScene& sc, uint32_t flag) { atomicOr(sc.errors, flag); }

// Starts a ray at the TLAS root; false if there is nothing to traverse.
template<typename Stack>
__device__ __forceinline__ bool trav_begin(Traversal& T, const TraceScene& sc, const RayIn& ray, uint32_t cullMask, Stack& stack)
{
    T.ray = ray;
    T.h.u = 0.0f;
    T.h.v = 0.0f;
    T.h.inst = kNoHit;
    T.h.prim = 0;
    T.nodes = sc.tlasNodes;
    T.tris = nullptr;
    T.inBlas = false;
    T.curInst = 0;
    T.tbest = ray.tmax;
    T.ref = 0;
    T.cullMask = cullMask;
    T.nodeCap = sc.tlasNodeCap;
    T.numFaces = 0;
    stack.sp = 0;
    T.wix = safe_rcp(ray.dx);
    T.wiy = safe_rcp(ray.dy);
    T.wiz = safe_rcp(ray.dz);
    stack.set_box_ray(ray.ox, ray.oy, ray.oz, T.wix, T.wiy, T.wiz);
    return sc.numInstances != 0;
}

// One step: a wide node, a sentinel, an instance leaf or a triangle.
// Returns false once the ray is finished (T.h holds the closest hit).
//
// A wave runs every kind of step its lanes need in each iteration, so the
// rare kinds (instance entry, triangle test) cost their instructions in
// nearly every iteration. With kInstEvery / kTriEvery > 1 those steps are
// taken only in iterations `iter` that are multiples of it; a lane that
// reaches one in another iteration waits (its state is unchanged), which
// delays that ray by a few iterations but leaves the wave's other
// iterations without the path. Results do not change: every ray takes the
// same steps in the same order. `iter` must count the wave's iterations
// (all lanes start the ray together).
template<int kInstEvery = 1, int kTriEvery = 1, int kSentEvery = 1, typename Stack>
__device__ __forceinline__ bool trav_step(Traversal& T, const TraceScene& sc, Stack& stack, uint32_t iter = 0)
{
    if(T.ref >= 0) {
        if((uint32_t)T.ref >= T.nodeCap) {
            trav_error(sc, kTravErrNode);
            T.h.inst = kNoHit;
            return false;
        }
        const int32_t nearest = visit_wide(T.nodes, T.ref, stack.box_ray(), T.ray.tmin, T.tbest, stack);
        if(nearest != kEmptyChild) {
            T.ref = nearest;
            return true;
        }
    }
    else if(T.ref == kSentinel) {
        if(kSentEvery > 1 && iter % kSentEvery != 0u) return true; // taken in a later iteration
        T.inBlas = false;
        T.nodes = sc.tlasNodes;
        T.nodeCap = sc.tlasNodeCap;
        stack.set_box_ray(T.ray.ox, T.ray.oy, T.ray.oz, T.wix, T.wiy, T.wiz);
    }
    else if(!T.inBlas) {
        if(kInstEvery > 1 && iter % kInstEvery != 0u) return true; // taken in a later iteration
        const int32_t leaf = ~T.ref;
        if((uint32_t)leaf >= sc.numInstances) {
            trav_error(sc, kTravErrInstance);
            T.h.inst = kNoHit;
            return false;
        }
        const u32x4* ip = reinterpret_cast<const u32x4*>(sc.instances + leaf);
        const u32x4 i3 = ld_global16(ip + 3); // blas, mask, position, customIndex
        if(i3.x >= sc.numBlas) {
            trav_error(sc, kTravErrBlas);
            T.h.inst = kNoHit;
            return false;
        }
        const u32x4* bp = reinterpret_cast<const u32x4*>(sc.blas + i3.x);
        const u32x4 b0 = ld_global16(bp);     // nodes, tris
        const uint32_t numFaces = ld_global8(bp + 1).x;
        if((i3.y & T.cullMask) != 0u && numFaces > 0u) {
            float w[12];
            for(int q = 0; q < 3; ++q) {
                const u32x4 r = ld_global16(ip + q);
                w[4 * q + 0] = __uint_as_float(r.x);
                w[4 * q + 1] = __uint_as_float(r.y);
                w[4 * q + 2] = __uint_as_float(r.z);
                w[4 * q + 3] = __uint_as_float(r.w);
            }
            float rox, roy, roz, rdx, rdy, rdz;
            xform_ray(w, T.ray.ox, T.ray.oy, T.ray.oz, T.ray.dx, T.ray.dy, T.ray.dz, rox, roy, roz, rdx, rdy, rdz);
            stack.set_box_ray(rox, roy, roz, safe_rcp(rdx), safe_rcp(rdy), safe_rcp(rdz));
            wt_setup(T.wt, rox, roy, roz, rdx, rdy, rdz);
            stack.push(kSentinel);
            T.nodes = reinterpret_cast<const Node4q*>((uint64_t)b0.x | ((uint64_t)b0.y << 32));
            // the copy of the triangles in this ray's axis order
            T.tris = reinterpret_cast<const Tri*>((uint64_t)b0.z | ((uint64_t)b0.w << 32)) + (size_t)T.wt.rot * numFaces;
            T.inBlas = true;
            T.curInst = i3.z;
            T.numFaces = numFaces;
            T.nodeCap = numFaces > 2u ? numFaces - 1u : 1u;
            T.ref = 0;
            return true;
        }
    }
    else {
        if(kTriEvery > 1 && iter % kTriEvery != 0u) return true; // taken in a later iteration
        const int32_t leaf = ~T.ref;
        if((uint32_t)leaf >= T.numFaces) {
            trav_error(sc, kTravErrTriangle);
            T.h.inst = kNoHit;
            return false;
        }
        const u32x4* tp = reinterpret_cast<const u32x4*>(T.tris + leaf);
        const u32x4 a = ld_global16(tp), b = ld_global16(tp + 1), cc = ld_global16(tp + 2);
        Tri tri;
        tri.v0x = __uint_as_float(a.x);
        tri.v0y = __uint_as_float(a.y);
        tri.v0z = __uint_as_float(a.z);
        tri.prim = a.w;
        tri.v1x = __uint_as_float(b.x);
        tri.v1y = __uint_as_float(b.y);
        tri.v1z = __uint_as_float(b.z);
        tri.v2x = __uint_as_float(cc.x);
        tri.v2y = __uint_as_float(cc.y);
        tri.v2z = __uint_as_float(cc.z);
        float t, u, v;
        if(wt_intersect(T.wt, tri, T.ray.tmin, T.tbest, t, u, v)) {
            T.tbest = t;
            T.h.u = u;
            T.h.v = v;
            T.h.inst = T.curInst;
            T.h.prim = tri.prim;
        }
    }
    if(stack.sp == 0) return false;
    T.ref = stack.pop();
    return true;
}

// Iterations between a wave's instance entries / triangle tests / returns
// to the TLAS (trav_step): the schedule of the traversal kernels, whose
// lanes start their rays together (the queue pass, the ray queries and
// each round of the drain).
#ifndef RGRT_INST_EVERY
#define RGRT_INST_EVERY 4
#endif
#ifndef RGRT_TRI_EVERY
#define RGRT_TRI_EVERY 1
#endif
#ifndef RGRT_SENT_EVERY
#define RGRT_SENT_EVERY 1
#endif

// Whole traversal of one ray. Each node and leaf is visited at most once
// per level, so a well-formed two-level tree ends far below the step cap;
// the cap only bounds the loop.
template<int kInstEvery = 1, int kTriEvery = 1, int kSentEvery = 1, typename Stack>
__device__ HitRec traverse(const TraceScene& sc, const RayIn& ray, uint32_t cullMask, Stack& stack)
{
    Traversal T;
    if(trav_begin(T, sc, ray, cullMask, stack)) {
        for(uint32_t guard = 0; guard < (1u << 26); ++guard)
            if(!trav_step<kInstEvery, kTriEvery, kSentEvery>(T, sc, stack, guard)) break;
    }
    T.h.t = T.h.inst != kNoHit ? T.tbest : 0.0f;
    return T.h;
}

} // namespace rgrt
