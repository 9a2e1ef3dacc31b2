// trace.hip — traversal kernels: the wavefront queue pass and plain
// closest-hit queries (rgrt_trace_rays).
#pragma clang fp contract(off)

#include "rgrt_internal.h"
#include "traverse.h"

namespace rgrt {

namespace {

constexpr int kTraceBlock = 256;
constexpr int kLdsStack = 16;
constexpr int kSpill = 368; // with the 16 LDS entries: the 381-entry worst case of two wide LBVH levels

// Wavefront trace pass: rays [0, count) of queue `cur` -> hit records.
// Waves fetch 64 rays at a time from a global counter (persistent grid).
// Waves per SIMD the traversal pass is compiled for: the register budget is
// 512 / waves VGPRs (7: 73, which its 71 fit without spilling; more waves
// hide more of the node-fetch latency; 8 spills).
#ifndef RGRT_TRACE_WAVES
#define RGRT_TRACE_WAVES 7
#endif

__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(RGRT_TRACE_WAVES))) void k_trace_queue(TraceScene sc, const float* __restrict__ rays,
                                                              float* __restrict__ hits, uint32_t npix,
                                                              uint32_t* __restrict__ counters, int cur,
                                                              int32_t* __restrict__ spillBase, uint32_t* hostSlot)
{
    __shared__ int32_t ldsStack[kLdsStack * kTraceBlock];
    const uint32_t count = counters[cur];
    if(blockIdx.x == 0 && threadIdx.x == 0) {
        *hostSlot = count;         // host-visible: lets the host stop once the queues drain
        counters[1 - cur] = 0;     // the output queue of the shading pass that follows
        counters[2 + (1 - cur)] = 0; // the fetch counter of the next traversal pass
        // ray statistics (64-bit) at counters[4..5]
        atomicAdd(reinterpret_cast<unsigned long long*>(counters + 4), (unsigned long long)count);
    }
    const int lane = threadIdx.x & 63;
    const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
    __shared__ float boxRay[kBoxRayWords * kTraceBlock];
    TraversalStack<kLdsStack, kTraceBlock, kSpill> stack{ldsStack, spillBase + (size_t)gtid * kSpill, (int)threadIdx.x, boxRay};
    uint32_t* fetch = counters + 2 + cur;
    // Wave w first takes rays [64w, 64w + 64); rays beyond the grid's static
    // share are fetched 64 at a time from a counter, so a pass with fewer
    // rays than the grid has lanes does no atomics at all.
    const uint32_t wavesInGrid = gridDim.x * (kTraceBlock / 64);
    const uint32_t dynStart = wavesInGrid * 64u;
    uint32_t base = (blockIdx.x * (kTraceBlock / 64) + (threadIdx.x >> 6)) * 64u;
    for(;;) {
        if(base >= count) break;
        const uint32_t i = base + lane;
        if(i < count) {
            RayIn r;
            r.ox = rays[0 * (size_t)npix + i];
            r.oy = rays[1 * (size_t)npix + i];
            r.oz = rays[2 * (size_t)npix + i];
            r.tmin = rays[3 * (size_t)npix + i];
            r.dx = rays[4 * (size_t)npix + i];
            r.dy = rays[5 * (size_t)npix + i];
            r.dz = rays[6 * (size_t)npix + i];
            r.tmax = rays[7 * (size_t)npix + i];
            const uint32_t flags = __float_as_uint(rays[9 * (size_t)npix + i]);
            const HitRec h = traverse<RGRT_INST_EVERY, RGRT_TRI_EVERY, RGRT_SENT_EVERY>(sc, r, (flags >> kMaskShift) & 0xffu, stack);
            hits[0 * (size_t)npix + i] = h.t;
            hits[1 * (size_t)npix + i] = h.u;
            hits[2 * (size_t)npix + i] = h.v;
            hits[3 * (size_t)npix + i] = __uint_as_float(h.inst);
            hits[4 * (size_t)npix + i] = __uint_as_float(h.prim);
        }
        if(count <= dynStart) break; // no rays beyond the static share
        uint32_t b = 0;
        if(lane == 0) b = dynStart + atomicAdd(fetch, 64u);
        base = __shfl(b, 0);
    }
}

// Traversal pass of a whole batch: the queues `cur` of every scene in one
// launch. The work is cut into 64-ray chunks per scene (a chunk never spans
// two scenes, so a wave traces one scene at a time); wave w first takes
// chunk w, then chunks from the batch's counter. Block 0 does each scene's
// bookkeeping of a pass: clears the output queue and the drain's fetch
// counter, adds the queue to the ray statistics, and clears the batch
// counter of the next pass.
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(RGRT_TRACE_WAVES))) void k_trace_multi(
    const ShadeParams* __restrict__ ps, const TraceScene* __restrict__ scs, int numScenes, int cur,
    int32_t* __restrict__ spillBase, uint32_t* __restrict__ batchCounters)
{
    __shared__ int32_t ldsStack[kLdsStack * kTraceBlock];
    __shared__ uint32_t chunkEnd[kMaxBatchScenes];
    const int lane = threadIdx.x & 63;
    if(threadIdx.x < 64) {
        uint32_t count = 0, v = 0;
        if(lane < numScenes) {
            count = ps[lane].st.counters[cur];
            v = (count + 63u) / 64u;
        }
        for(int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(v, off);
            if(lane >= off) v += u;
        }
        if(lane < numScenes) chunkEnd[lane] = v;
        if(blockIdx.x == 0) {
            if(lane < numScenes) {
                uint32_t* counters = ps[lane].st.counters;
                counters[1 - cur] = 0;       // the output queue of the shading pass that follows
                counters[2 + (1 - cur)] = 0; // the fetch counter of the drain
                atomicAdd(reinterpret_cast<unsigned long long*>(counters + 4), (unsigned long long)count);
            }
            if(lane == 0) batchCounters[1 - cur] = 0;
        }
    }
    __syncthreads();
    const uint32_t chunks = chunkEnd[numScenes - 1];
    const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
    __shared__ float boxRay[kBoxRayWords * kTraceBlock];
    TraversalStack<kLdsStack, kTraceBlock, kSpill> stack{ldsStack, spillBase + (size_t)gtid * kSpill, (int)threadIdx.x, boxRay};
    const uint32_t wavesInGrid = gridDim.x * (kTraceBlock / 64);
    uint32_t k = blockIdx.x * (kTraceBlock / 64) + (threadIdx.x >> 6);
    for(;;) {
        if(k >= chunks) break;
        const int s = __builtin_amdgcn_readfirstlane(__popcll(__ballot(lane < numScenes && chunkEnd[lane < numScenes ? lane : 0] <= k)));
        const uint32_t first = s ? chunkEnd[s - 1] : 0u;
        const WavefrontState& st = ps[s].st;
        const uint32_t count = st.counters[cur];
        const uint32_t i = (k - first) * 64u + (uint32_t)lane;
        if(i < count) {
            const size_t npix = st.npix;
            const float* __restrict__ rays = st.rays[cur];
            float* __restrict__ hits = st.hits;
            const TraceScene sc = scs[s];
            RayIn r;
            r.ox = rays[0 * npix + i];
            r.oy = rays[1 * npix + i];
            r.oz = rays[2 * npix + i];
            r.tmin = rays[3 * npix + i];
            r.dx = rays[4 * npix + i];
            r.dy = rays[5 * npix + i];
            r.dz = rays[6 * npix + i];
            r.tmax = rays[7 * npix + i];
            const uint32_t flags = __float_as_uint(rays[9 * npix + i]);
            const HitRec h = traverse<RGRT_INST_EVERY, RGRT_TRI_EVERY, RGRT_SENT_EVERY>(sc, r, (flags >> kMaskShift) & 0xffu, stack);
            hits[0 * npix + i] = h.t;
            hits[1 * npix + i] = h.u;
            hits[2 * npix + i] = h.v;
            hits[3 * npix + i] = __uint_as_float(h.inst);
            hits[4 * npix + i] = __uint_as_float(h.prim);
        }
        if(chunks <= wavesInGrid) break; // no chunks beyond the static share
        uint32_t b = 0;
        if(lane == 0) b = wavesInGrid + atomicAdd(batchCounters + cur, 1u);
        k = __shfl(b, 0);
    }
}

// Closest-hit queries, grid-stride over a persistent grid (the spill
// region is sized per launched thread).
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(TraceScene sc, const rgrt_ray* __restrict__ rays, uint32_t n,
                                                             uint32_t cullMask, rgrt_hit* __restrict__ hits,
                                                             int32_t* __restrict__ spillBase)
{
    __shared__ int32_t ldsStack[kLdsStack * kTraceBlock];
    const uint32_t gtid = blockIdx.x * kTraceBlock + threadIdx.x;
    __shared__ float boxRay[kBoxRayWords * kTraceBlock];
    TraversalStack<kLdsStack, kTraceBlock, kSpill> stack{ldsStack, spillBase + (size_t)gtid * kSpill, (int)threadIdx.x, boxRay};
    for(uint32_t i = gtid; i < n; i += gridDim.x * kTraceBlock) {
        const rgrt_ray rr = rays[i];
        RayIn r{rr.origin[0], rr.origin[1], rr.origin[2], rr.tmin, rr.dir[0], rr.dir[1], rr.dir[2], rr.tmax};
        const HitRec h = traverse<RGRT_INST_EVERY, RGRT_TRI_EVERY, RGRT_SENT_EVERY>(sc, r, cullMask & 0xffu, stack);
        rgrt_hit o;
        o.t = h.t;
        o.u = h.u;
        o.v = h.v;
        o.instance = h.inst; // position; mapped to customIndex below
        o.primitive = h.prim;
        hits[i] = o;
    }
}

__global__ void k_hit_custom_index(rgrt_hit* hits, uint32_t n, const InstanceShade* __restrict__ inst, uint32_t numInstances)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const uint32_t p = hits[i].instance;
    if(p != kNoHit) hits[i].instance = p < numInstances ? inst[p].customIndex : kNoHit;
}

} // namespace

uint32_t trace_spill_words_per_thread() { return kSpill; }
uint32_t trace_block_size() { return kTraceBlock; }

void launch_trace_queue_impl(hipStream_t s, const TraceScene& sc, const WavefrontState& st, int cur, uint32_t gridBlocks,
                             int32_t* spill, uint32_t* hostSlot)
{
    k_trace_queue<<<gridBlocks, kTraceBlock, 0, s>>>(sc, st.rays[cur], st.hits, st.npix, st.counters, cur, spill, hostSlot);
}

void launch_trace_multi(hipStream_t s, const ShadeParams* ps, const TraceScene* scs, int numScenes, int cur, uint32_t gridBlocks,
                        int32_t* spill, uint32_t* batchCounters)
{
    k_trace_multi<<<gridBlocks, kTraceBlock, 0, s>>>(ps, scs, numScenes, cur, spill, batchCounters);
}

namespace {
__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) p[i] = 0u;
}
} // namespace

// Clears n words with a kernel (capturable into a graph, unlike a memset node).
void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n)
{
    if(n) k_zero_u32<<<(n + 255) / 256, 256, 0, s>>>(p, n);
}

void launch_trace_rays_impl(hipStream_t s, const TraceScene& sc, const rgrt_ray* rays, uint32_t n, uint32_t cullMask,
                            rgrt_hit* hits, const InstanceShade* inst, int32_t* spill, uint32_t gridBlocks)
{
    if(!n) return;
    const uint32_t blocks = std::min<uint32_t>(gridBlocks, (n + kTraceBlock - 1) / kTraceBlock);
    k_trace_rays<<<blocks, kTraceBlock, 0, s>>>(sc, rays, n, cullMask, hits, spill);
    k_hit_custom_index<<<(n + 255) / 256, 256, 0, s>>>(hits, n, inst, sc.numInstances);
}

} // namespace rgrt
