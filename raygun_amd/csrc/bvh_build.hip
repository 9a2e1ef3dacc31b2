// bvh_build.hip — LBVH construction on the device (BLAS and TLAS).
//
// Replaces the driver builds behind vkCmdBuildAccelerationStructuresKHR
// (raygun/render/acceleration_structure.cpp:134 TLAS, :193 BLAS). Steps:
//   1. primitive boxes (triangles, or transformed instance boxes)
//   2. scene box (atomic min/max on order-preserving integer encodings)
//   3. 30-bit Morton codes of box centroids normalised to the scene box
//   4. stable LSD radix sort (4 x 8-bit digits), ranks from wave64 ballots
//   5. Karras 2012 hierarchy, ties between equal codes broken by index
//   6. bottom-up box refit, second arrival at a node continues upwards
// Every step is deterministic, so the tree equals the CPU oracle's node for
// node.
#pragma clang fp contract(off)

#include "rgrt_device.h"
#include "rgrt_internal.h"

namespace rgrt {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
static_assert(kSortTile == (int)kLbvhSortTile, "the scratch layout assumes this sort tile");
constexpr int kWaves = kSortThreads / 64;

__device__ __forceinline__ uint32_t float_to_ordered(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(b);
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ __forceinline__ uint32_t quantize10(float x) { return (uint32_t)fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f); }

// ---------------------------------------------------------------------------
// 1. primitive boxes
// ---------------------------------------------------------------------------
__global__ void k_tri_boxes(const rgrt_vertex* __restrict__ vb, const uint32_t* __restrict__ ib, uint32_t firstVertex,
                            uint32_t indexOffset, uint32_t n, float4* __restrict__ boxLo, float4* __restrict__ boxHi)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= n) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(int k = 0; k < 3; ++k) {
        const rgrt_vertex& v = vb[firstVertex + ib[indexOffset + 3 * f + k]];
        for(int c = 0; c < 3; ++c) {
            lo[c] = fminf(lo[c], v.position[c]);
            hi[c] = fmaxf(hi[c], v.position[c]);
        }
    }
    boxLo[f] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    boxHi[f] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

// World box of each instance (Arvo's method on the BLAS root box), the
// world-to-object inverse in double, and the per-position shading record.
__device__ __forceinline__ void inst_prep_one(const rgrt_instance& r, const float4* __restrict__ blasRoot, float wlo[3],
                                              float whi[3], float w[12], InstanceShade& s)
{
    const float* m = r.transform;
    const float4 rlo = blasRoot[2 * r.blas], rhi = blasRoot[2 * r.blas + 1];
    const float blo[3] = {rlo.x, rlo.y, rlo.z}, bhi[3] = {rhi.x, rhi.y, rhi.z};
    if(!(blo[0] <= bhi[0])) {
        for(int c = 0; c < 3; ++c) {
            wlo[c] = blo[c];
            whi[c] = bhi[c];
        }
    }
    else {
        for(int row = 0; row < 3; ++row) {
            float lo = m[row * 4 + 3], hi = m[row * 4 + 3];
            for(int j = 0; j < 3; ++j) {
                const float p = m[row * 4 + j] * blo[j];
                const float q = m[row * 4 + j] * bhi[j];
                lo = lo + fminf(p, q);
                hi = hi + fmaxf(p, q);
            }
            wlo[row] = lo;
            whi[row] = hi;
        }
    }

    // inverse of the affine map: adjugate of the 3x3 part in double
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    const double rr = 1.0 / det;
    const double inv[9] = {A * rr, -(b * k - c * h) * rr, (b * f - c * e) * rr, B * rr, (a * k - c * g) * rr,
                           -(a * f - c * d) * rr, C * rr, -(a * h - b * g) * rr, (a * e - b * d) * rr};
    const double tx = m[3], ty = m[7], tz = m[11];
    for(int row = 0; row < 3; ++row) {
        w[row * 4 + 0] = (float)inv[row * 3 + 0];
        w[row * 4 + 1] = (float)inv[row * 3 + 1];
        w[row * 4 + 2] = (float)inv[row * 3 + 2];
        w[row * 4 + 3] = (float)(-(inv[row * 3 + 0] * tx + inv[row * 3 + 1] * ty + inv[row * 3 + 2] * tz));
    }
    for(int q = 0; q < 12; ++q) s.o2w[q] = m[q];
    s.offsets = r.offsets;
    s.customIndex = r.customIndex;
}

__global__ void k_inst_prep(const rgrt_instance* __restrict__ in, uint32_t n, const float4* __restrict__ blasRoot,
                            float4* __restrict__ boxLo, float4* __restrict__ boxHi, float* __restrict__ w2oOut,
                            InstanceShade* __restrict__ shade)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const rgrt_instance r = in[i];
    float wlo[3], whi[3], w[12];
    InstanceShade s;
    inst_prep_one(r, blasRoot, wlo, whi, w, s);
    boxLo[i] = make_float4(wlo[0], wlo[1], wlo[2], 0.0f);
    boxHi[i] = make_float4(whi[0], whi[1], whi[2], 0.0f);
    for(int q = 0; q < 12; ++q) w2oOut[12 * (size_t)i + q] = w[q];
    shade[i] = s;
}

// ---------------------------------------------------------------------------
// 2. scene box  3. Morton codes
// ---------------------------------------------------------------------------
__global__ void k_scene_box_init(uint32_t* sb)
{
    if(threadIdx.x < 3) sb[threadIdx.x] = float_to_ordered(INFINITY);
    else if(threadIdx.x < 6) sb[threadIdx.x] = float_to_ordered(-INFINITY);
}

__global__ void k_scene_box(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi, uint32_t n, uint32_t* sb)
{
    __shared__ uint32_t red[6];
    if(threadIdx.x < 3) red[threadIdx.x] = float_to_ordered(INFINITY);
    else if(threadIdx.x < 6) red[threadIdx.x] = float_to_ordered(-INFINITY);
    __syncthreads();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = boxLo[i], b = boxHi[i];
        lo[0] = fminf(lo[0], a.x);
        lo[1] = fminf(lo[1], a.y);
        lo[2] = fminf(lo[2], a.z);
        hi[0] = fmaxf(hi[0], b.x);
        hi[1] = fmaxf(hi[1], b.y);
        hi[2] = fmaxf(hi[2], b.z);
    }
    for(int c = 0; c < 3; ++c) {
        atomicMin(&red[c], float_to_ordered(lo[c]));
        atomicMax(&red[3 + c], float_to_ordered(hi[c]));
    }
    __syncthreads();
    if(threadIdx.x < 3) atomicMin(&sb[threadIdx.x], red[threadIdx.x]);
    else if(threadIdx.x < 6) atomicMax(&sb[threadIdx.x], red[threadIdx.x]);
}

// Morton code of a box centroid in the scene box sb (order-preserving integer encodings).
__device__ __forceinline__ uint32_t morton_key(const float lo[3], const float hi[3], const uint32_t* sb)
{
    const float slo[3] = {ordered_to_float(sb[0]), ordered_to_float(sb[1]), ordered_to_float(sb[2])};
    const float shi[3] = {ordered_to_float(sb[3]), ordered_to_float(sb[4]), ordered_to_float(sb[5])};
    uint32_t q[3];
    for(int c = 0; c < 3; ++c) {
        const float ext = shi[c] - slo[c];
        const float centroid = (lo[c] + hi[c]) * 0.5f;
        const float x = ext > 0.0f ? (centroid - slo[c]) / ext : 0.0f;
        q[c] = quantize10(x);
    }
    return (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
}

__global__ void k_morton(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi, uint32_t n,
                         const uint32_t* __restrict__ sb, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    const float4 a = boxLo[i], b = boxHi[i];
    const float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
    keys[i] = morton_key(lo, hi, sb);
    vals[i] = i;
}

// ---------------------------------------------------------------------------
// 4. stable LSD radix sort
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                             uint32_t* __restrict__ hist, uint32_t numBlocks)
{
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for(int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if(i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * numBlocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive prefix sum of the sort histogram over many workgroups (it has
// ~n/8 entries for n keys): kScanBlocks blocks each total one contiguous
// segment, one block scans the totals, then every block rescans its segment
// from its offset, 256 consecutive entries at a time (coalesced).
constexpr int kScanBlocks = 256;
constexpr int kScanThreads = 256;

__device__ __forceinline__ uint32_t block_sum256(uint32_t v, uint32_t* red)
{
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if(lane == 0) red[wave] = v;
    __syncthreads();
    const uint32_t t = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ hist, uint32_t total,
                                                              uint32_t* __restrict__ partial)
{
    __shared__ uint32_t red[kScanThreads / 64];
    const uint32_t seg = (total + kScanBlocks - 1) / kScanBlocks;
    const uint32_t b = blockIdx.x * seg, e = min(total, b + seg);
    uint32_t s = 0;
    for(uint32_t i = b + threadIdx.x; i < e; i += kScanThreads) s += hist[i];
    s = block_sum256(s, red);
    if(threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_partials(uint32_t* __restrict__ partial)
{
    __shared__ uint32_t part[kScanBlocks];
    const uint32_t v = partial[threadIdx.x];
    part[threadIdx.x] = v;
    __syncthreads();
    for(uint32_t off = 1; off < kScanBlocks; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    partial[threadIdx.x] = part[threadIdx.x] - v; // exclusive
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(uint32_t* __restrict__ hist, uint32_t total,
                                                             const uint32_t* __restrict__ partial)
{
    __shared__ uint32_t waveTot[kScanThreads / 64];
    const uint32_t seg = (total + kScanBlocks - 1) / kScanBlocks;
    const uint32_t b = blockIdx.x * seg, e = min(total, b + seg);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t carry = partial[blockIdx.x];
    for(uint32_t base = b; base < e; base += kScanThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t c = i < e ? hist[i] : 0u;
        uint32_t incl = c;
        for(int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off);
            if(lane >= off) incl += t;
        }
        if(lane == 63) waveTot[wave] = incl;
        __syncthreads();
        uint32_t before = 0, all = 0;
        for(int w = 0; w < kScanThreads / 64; ++w) {
            before += w < wave ? waveTot[w] : 0u;
            all += waveTot[w];
        }
        if(i < e) hist[i] = carry + before + incl - c;
        carry += all;
        __syncthreads();
    }
}

// Scatter of one tile, in tile order: round r holds keys [r*256, r*256+256),
// wave w its 64 consecutive keys. A key's rank is the count of equal digits
// before it: earlier rounds (running), earlier waves of the round (LDS
// counts) and lower lanes of its wave (ballot match mask).
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint32_t* __restrict__ keysIn,
                                                                const uint32_t* __restrict__ valsIn,
                                                                uint32_t* __restrict__ keysOut,
                                                                uint32_t* __restrict__ valsOut, uint32_t n, int shift,
                                                                const uint32_t* __restrict__ hist, uint32_t numBlocks)
{
    __shared__ uint32_t gofs[256];
    __shared__ uint32_t waveCnt[kWaves][256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    gofs[tid] = hist[tid * numBlocks + blockIdx.x];
    for(int w = 0; w < kWaves; ++w) waveCnt[w][tid] = 0;
    __syncthreads();
    const uint64_t ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t base = blockIdx.x * kSortTile;
    for(int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kSortThreads + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? keysIn[i] : 0u;
        const uint32_t val = valid ? valsIn[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
        for(int bit = 0; bit < 8; ++bit) {
            const bool set = (d >> bit) & 1u;
            const uint64_t bb = __ballot(valid && set);
            m &= set ? bb : ~bb;
        }
        const uint32_t rankInWave = (uint32_t)__popcll(m & ltMask);
        const bool leader = valid && ((m & ltMask) == 0ull);
        if(leader) waveCnt[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if(valid) {
            uint32_t pre = 0;
            for(int w = 0; w < wave; ++w) pre += waveCnt[w][d];
            const uint32_t pos = gofs[d] + pre + rankInWave;
            if(pos < n) { // always true: the histogram counted the same keys
                keysOut[pos] = key;
                valsOut[pos] = val;
            }
        }
        __syncthreads();
        {
            uint32_t add = 0;
            for(int w = 0; w < kWaves; ++w) {
                add += waveCnt[w][tid];
                waveCnt[w][tid] = 0;
            }
            gofs[tid] += add;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 5. Karras hierarchy
// ---------------------------------------------------------------------------
__device__ __forceinline__ int karras_delta(const uint32_t* k, int n, int i, int j)
{
    if(j < 0 || j >= n) return -1;
    const uint32_t a = k[i], b = k[j];
    if(a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

// Children of internal node i (Karras 2012): refs >= 0 are internal nodes,
// ~k the leaf at sorted position k; gamma is the split position.
__device__ __forceinline__ void karras_node(const uint32_t* keys, int n, int i, int& left, int& right, int& gamma, int& rlo,
                                            int& rhi)
{
    const int d = (karras_delta(keys, n, i, i + 1) - karras_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = karras_delta(keys, n, i, i - d);
    int lmax = 2;
    while(karras_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for(int t = lmax / 2; t >= 1; t /= 2)
        if(karras_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = karras_delta(keys, n, i, j);
    int s = 0;
    for(int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if(karras_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if(t <= 1) break;
    }
    gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    left = (lo == gamma) ? ~gamma : gamma;
    right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    rlo = lo;
    rhi = hi;
}

__global__ void k_karras(const uint32_t* __restrict__ keys, int n, Node* __restrict__ nodes, uint32_t* __restrict__ leafParent,
                         uint2* __restrict__ range)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n - 1) return;
    int left, right, gamma, lo, hi;
    karras_node(keys, n, i, left, right, gamma, lo, hi);
    range[i] = make_uint2((uint32_t)lo, (uint32_t)hi);
    Node& nd = nodes[i];
    nd.child0 = left;
    nd.child1 = right;
    nd.pad = 0;
    if(i == 0) nd.parent = 0xffffffffu;
    if(left >= 0) nodes[left].parent = (uint32_t)i;
    else leafParent[gamma] = (uint32_t)i;
    if(right >= 0) nodes[right].parent = (uint32_t)i;
    else leafParent[gamma + 1] = (uint32_t)i;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// 6. bottom-up refit
// ---------------------------------------------------------------------------
// The walk from every leaf to the root combines two child boxes per node.
// A single walk across the whole grid would need a device-wide release and
// acquire per node (each XCD has its own L2), which costs more than the work
// itself. The leaves are therefore cut into chunks of kRefitChunk
// consecutive leaves: a workgroup finishes every node whose leaf range lies
// inside its chunk with LDS counters and LDS copies of the child boxes, and
// records each arrival at a node reaching outside the chunk. One workgroup
// then finishes those upper nodes from the recorded arrivals (workgroup-scope
// ordering only). A node's box is always fminf/fmaxf(slot 0, slot 1), so the
// result does not depend on the order of arrival.
constexpr int kRefitChunk = 1024;
constexpr int kRefitTopThreads = 1024;

__device__ __forceinline__ void write_slot(Node* nd, int slot, const float lo[3], const float hi[3])
{
    float* f = reinterpret_cast<float*>(nd);
    // lo[slot] at float offset 3*slot, hi[slot] at 6 + 3*slot
    for(int c = 0; c < 3; ++c) {
        f[3 * slot + c] = lo[c];
        f[6 + 3 * slot + c] = hi[c];
    }
}

// counters[n - 1] counts the recorded arrivals, exits[] holds their nodes.
__global__ __launch_bounds__(kRefitChunk) void k_refit_chunk(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi,
                                                             const uint32_t* __restrict__ leafPrim,
                                                             const uint32_t* __restrict__ leafParent,
                                                             const uint2* __restrict__ range, Node* nodes, uint32_t* counters,
                                                             uint32_t* exits, int n, float4* rootBox)
{
    __shared__ uint32_t cnt[kRefitChunk];
    __shared__ float box[kRefitChunk][12]; // node - c0: slot 0 lo, slot 1 lo, slot 0 hi, slot 1 hi
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * kRefitChunk;
    const int c1 = min(c0 + kRefitChunk, n) - 1; // last leaf of the chunk
    cnt[t] = 0u;
    __syncthreads();
    const int k = c0 + t;
    if(k >= n) return;
    const uint32_t prim = leafPrim[k];
    const float4 a = boxLo[prim], b = boxHi[prim];
    float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
    int child = ~k;
    uint32_t p = leafParent[k];
    // A Karras tree is at most 63 levels deep; the cap only guarantees that
    // the walk ends even on a malformed tree.
    for(int level = 0; level < 128; ++level) {
        if(p >= (uint32_t)(n - 1)) return;
        Node* nd = nodes + p;
        const int slot = nd->child0 == child ? 0 : 1;
        const uint2 r = range[p];
        write_slot(nd, slot, lo, hi);
        const int q = (int)p - c0; // a node lies inside its own leaf range
        if((int)r.x < c0 || (int)r.y > c1 || q < 0 || q >= kRefitChunk) {
            // p reaches outside the chunk: the top pass takes it from here
            const uint32_t e = atomicAdd(&counters[n - 1], 1u);
            if(e < (uint32_t)n) exits[e] = p;
            return;
        }
        float* f = box[q];
        for(int c = 0; c < 3; ++c) {
            f[3 * slot + c] = lo[c];
            f[6 + 3 * slot + c] = hi[c];
        }
        const uint32_t prev = __hip_atomic_fetch_add(&cnt[q], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if(prev == 0u) return; // the sibling's thread finishes this node
        for(int c = 0; c < 3; ++c) {
            lo[c] = fminf(f[c], f[3 + c]);
            hi[c] = fmaxf(f[6 + c], f[9 + c]);
        }
        if(p == 0) {
            rootBox[0] = make_float4(lo[0], lo[1], lo[2], 0.0f);
            rootBox[1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
            return;
        }
        child = (int)p;
        p = nd->parent;
    }
}

// The nodes above the chunks, from the recorded arrivals, in one workgroup.
__global__ __launch_bounds__(kRefitTopThreads) void k_refit_top(Node* nodes, uint32_t* counters, const uint32_t* __restrict__ exits,
                                                                int n, float4* rootBox)
{
    const uint32_t numExits = min(counters[n - 1], (uint32_t)n);
    for(uint32_t e = threadIdx.x; e < numExits; e += kRefitTopThreads) {
        uint32_t p = exits[e];
        for(int level = 0; level < 128; ++level) {
            if(p >= (uint32_t)(n - 1)) break;
            Node* nd = nodes + p;
            const uint32_t prev = __hip_atomic_fetch_add(&counters[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if(prev == 0u) break; // the other child's arrival finishes this node
            const float* f = reinterpret_cast<const float*>(nd);
            float lo[3], hi[3];
            for(int c = 0; c < 3; ++c) {
                lo[c] = fminf(f[c], f[3 + c]);
                hi[c] = fmaxf(f[6 + c], f[9 + c]);
            }
            if(p == 0) {
                rootBox[0] = make_float4(lo[0], lo[1], lo[2], 0.0f);
                rootBox[1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
                break;
            }
            const uint32_t q = nd->parent;
            if(q >= (uint32_t)(n - 1)) break;
            Node* pn = nodes + q;
            write_slot(pn, pn->child0 == (int32_t)p ? 0 : 1, lo, hi);
            p = q;
        }
    }
}

// Leaf range of every internal node of an existing tree: the leftmost leaf
// below child 0 and the rightmost below child 1.
__global__ void k_node_ranges(const Node* __restrict__ nodes, int n, uint2* __restrict__ range)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n - 1) return;
    int32_t a = nodes[i].child0, b = nodes[i].child1;
    for(int level = 0; level < 128 && a >= 0 && a < n - 1; ++level) a = nodes[a].child0;
    for(int level = 0; level < 128 && b >= 0 && b < n - 1; ++level) b = nodes[b].child1;
    range[i] = make_uint2(a < 0 ? (uint32_t)~a : 0u, b < 0 ? (uint32_t)~b : (uint32_t)(n - 1));
}

// Trees of 0 or 1 primitives: one node with an empty second child.
__global__ void k_small_tree(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi, int n, Node* nodes,
                             float4* rootBox)
{
    if(threadIdx.x != 0) return;
    Node nd;
    const float inf = INFINITY;
    float lo0[3] = {inf, inf, inf}, hi0[3] = {-inf, -inf, -inf};
    if(n == 1) {
        const float4 a = boxLo[0], b = boxHi[0];
        lo0[0] = a.x;
        lo0[1] = a.y;
        lo0[2] = a.z;
        hi0[0] = b.x;
        hi0[1] = b.y;
        hi0[2] = b.z;
    }
    nd.lo0x = lo0[0];
    nd.lo0y = lo0[1];
    nd.lo0z = lo0[2];
    nd.hi0x = hi0[0];
    nd.hi0y = hi0[1];
    nd.hi0z = hi0[2];
    nd.lo1x = nd.lo1y = nd.lo1z = inf;
    nd.hi1x = nd.hi1y = nd.hi1z = -inf;
    nd.child0 = n == 1 ? ~0 : kEmptyChild;
    nd.child1 = kEmptyChild;
    nd.parent = 0xffffffffu;
    nd.pad = 0;
    nodes[0] = nd;
    rootBox[0] = make_float4(lo0[0], lo0[1], lo0[2], 0.0f);
    rootBox[1] = make_float4(hi0[0], hi0[1], hi0[2], 0.0f);
}

// ---------------------------------------------------------------------------
// 7. collapse to 4-wide nodes
// ---------------------------------------------------------------------------
struct WideEntry {
    int32_t ref;
    float lo[3], hi[3];
};

// Child slot c of binary node `node`, read from the node array in HBM.
struct GlobalNodes {
    const Node* nodes;
    __device__ __forceinline__ WideEntry slot(int node, int c) const
    {
        const float* f = reinterpret_cast<const float*>(nodes + node);
        WideEntry e;
        e.ref = reinterpret_cast<const int32_t*>(nodes + node)[12 + c];
        for(int k = 0; k < 3; ++k) {
            e.lo[k] = f[3 * c + k];
            e.hi[k] = f[6 + 3 * c + k];
        }
        return e;
    }
};

__device__ __forceinline__ float half_area(const float lo[3], const float hi[3])
{
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return ((dx * dy) + (dy * dz)) + (dz * dx);
}

constexpr int kCollapseThreads = 1024;

__device__ Node4q quantize_node(const Node4& w);

// Field-wise c ? a : b (keeps the entries in registers).
__device__ __forceinline__ WideEntry pick(bool c, const WideEntry& a, const WideEntry& b)
{
    WideEntry r;
    r.ref = c ? a.ref : b.ref;
#pragma unroll
    for(int k = 0; k < 3; ++k) {
        r.lo[k] = c ? a.lo[k] : b.lo[k];
        r.hi[k] = c ? a.hi[k] : b.hi[k];
    }
    return r;
}

// The children of a wide node rooted at binary node r: its binary node's two
// children, then repeatedly the internal child of largest surface area
// (the first on equal areas) opened in place, until there are four or only
// leaves. Returns how many of S[0..4) are used.
template <class Src>
__device__ __forceinline__ int wide_children(const Src& src, int r, WideEntry S[4])
{
    const WideEntry e0 = src.slot(r, 0), e1 = src.slot(r, 1);
    const bool h0 = e0.ref != kEmptyChild, h1 = e1.ref != kEmptyChild;
    S[0] = pick(h0, e0, e1);
    S[1] = e1;
    int ns = (int)h0 + (int)h1;
    while(ns < 4) {
        int j = -1, e = 0;
        float best = -1.0f;
#pragma unroll
        for(int q = 0; q < 4; ++q) {
            if(q < ns && S[q].ref >= 0) {
                const float a = half_area(S[q].lo, S[q].hi);
                if(a > best) {
                    best = a;
                    j = q;
                    e = S[q].ref;
                }
            }
        }
        if(j < 0) break;
        const WideEntry a0 = src.slot(e, 0), a1 = src.slot(e, 1);
#pragma unroll
        for(int q = 3; q >= 1; --q) S[q] = pick(q > j + 1 && q <= ns, S[q - 1], S[q]);
#pragma unroll
        for(int q = 0; q < 4; ++q) S[q] = pick(q == j, a0, pick(q == j + 1, a1, S[q]));
        ++ns;
    }
    return ns;
}

// One workgroup of kThreads walks the binary tree breadth-first. A
// wide node starts from its binary node's two children and opens the
// internal child of largest surface area (the first on equal areas) until it
// has four children or only leaves; its internal children become the next
// level's wide nodes, numbered in order (a block-wide prefix sum over the
// per-node counts: a shuffle scan in each wave, then the wave totals). The
// frontiers hold at most n - 1 entries each; `waveTot` is
// kThreads / 64 ints of LDS. With outq set, every wide node is also
// written compressed. Returns the wide node count (in every thread).
template <int kThreads, class Src>
__device__ int collapse_bfs(const Src src, Node4* __restrict__ out, Node4q* __restrict__ outq, int32_t* frontA,
                            int32_t* frontB, int32_t* waveTot)
{
    constexpr int kWavesPerBlock = kThreads / 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    int32_t* front = frontA;
    int32_t* next = frontB;
    if(tid == 0) front[0] = 0;
    __syncthreads();
    int fcount = 1, base = 0;
    for(int level = 0; level < 128 && fcount > 0; ++level) {
        int nextCount = 0;
        for(int chunk = 0; chunk < fcount; chunk += kThreads) {
            const int k = chunk + tid;
            const bool valid = k < fcount;
            WideEntry S[4];
            int ns = 0;
            if(valid) ns = wide_children(src, front[k], S);
            int cnt = 0;
#pragma unroll
            for(int q = 0; q < 4; ++q) cnt += (q < ns && S[q].ref >= 0) ? 1 : 0;
            int incl = cnt;
#pragma unroll
            for(int off = 1; off < 64; off <<= 1) {
                const int t = __shfl_up(incl, off);
                if(lane >= off) incl += t;
            }
            if(lane == 63) waveTot[wave] = incl;
            __syncthreads();
            int excl = incl - cnt, total = 0;
#pragma unroll
            for(int w = 0; w < kWavesPerBlock; ++w) {
                const int t = waveTot[w];
                excl += w < wave ? t : 0;
                total += t;
            }
            if(valid) {
                Node4 w;
#pragma unroll
                for(int c = 0; c < 4; ++c) {
                    w.lox[c] = w.loy[c] = w.loz[c] = INFINITY;
                    w.hix[c] = w.hiy[c] = w.hiz[c] = -INFINITY;
                    w.child[c] = kEmptyChild;
                    w.pad[c] = 0;
                }
                int rank = 0;
#pragma unroll
                for(int q = 0; q < 4; ++q) {
                    if(q < ns) {
                        w.lox[q] = S[q].lo[0];
                        w.loy[q] = S[q].lo[1];
                        w.loz[q] = S[q].lo[2];
                        w.hix[q] = S[q].hi[0];
                        w.hiy[q] = S[q].hi[1];
                        w.hiz[q] = S[q].hi[2];
                        if(S[q].ref >= 0) {
                            const int idx = nextCount + excl + rank;
                            w.child[q] = base + fcount + idx;
                            next[idx] = S[q].ref;
                            ++rank;
                        }
                        else {
                            w.child[q] = S[q].ref;
                        }
                    }
                }
                out[base + k] = w;
                if(outq) outq[base + k] = quantize_node(w);
            }
            nextCount += total;
            __syncthreads(); // waveTot[] and next[] settle before the next chunk / level
        }
        base += fcount;
        fcount = nextCount;
        int32_t* t = front;
        front = next;
        next = t;
        __threadfence_block();
        __syncthreads();
    }
    return base;
}

// The collapse of a tree in HBM in two steps. Which children a wide node
// gets depends only on the subtree below its binary node, so the first step
// works that out for every binary node at once (k_wide_candidates, one
// thread per node: candidate wide nodes whose internal child refs are still
// binary node indices). The breadth-first numbering then needs one 128-byte
// load per wide node instead of a chain of dependent node reads
// (k_collapse_numbered, one workgroup); the result is the wide tree
// collapse_bfs builds, node for node.
__global__ void k_wide_candidates(const Node* __restrict__ nodes, int numInternal, Node4* __restrict__ cand)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if(r >= numInternal) return;
    WideEntry S[4];
    const int ns = wide_children(GlobalNodes{nodes}, r, S);
    Node4 w;
#pragma unroll
    for(int q = 0; q < 4; ++q) {
        const bool used = q < ns;
        w.lox[q] = used ? S[q].lo[0] : INFINITY;
        w.loy[q] = used ? S[q].lo[1] : INFINITY;
        w.loz[q] = used ? S[q].lo[2] : INFINITY;
        w.hix[q] = used ? S[q].hi[0] : -INFINITY;
        w.hiy[q] = used ? S[q].hi[1] : -INFINITY;
        w.hiz[q] = used ? S[q].hi[2] : -INFINITY;
        w.child[q] = used ? S[q].ref : kEmptyChild;
        w.pad[q] = 0;
    }
    cand[r] = w;
}

constexpr int kCollapseItems = 4; // frontier entries per thread and chunk (consecutive; 8 and 16 measured slower)

// Breadth-first numbering of the wide tree (one workgroup): `order` receives
// the binary node of every wide node in wide-node order (the frontiers of
// all levels one after another, level L's at [base, base + count)), and
// each wide node's child refs, renumbered, are written to its slot in `out`
// (16 bytes; k_wide_boxes copies the rest of the node in parallel). Internal
// children in entry and slot order get consecutive next-level numbers.
__global__ __launch_bounds__(kCollapseThreads) void k_collapse_numbered(const Node4* __restrict__ cand, int numInternal,
                                                                        Node4* __restrict__ out, int32_t* order,
                                                                        uint32_t* numWide)
{
    __shared__ int32_t waveTot[kCollapseThreads / 64];
    constexpr int kWavesPerBlock = kCollapseThreads / 64;
    constexpr int kChunk = kCollapseThreads * kCollapseItems;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if(tid == 0) order[0] = 0;
    __syncthreads();
    int fcount = 1, base = 0;
    for(int level = 0; level < 128 && fcount > 0; ++level) {
        const int32_t* front = order + base;
        int32_t* next = order + base + fcount;
        int nextCount = 0;
        for(int chunk = 0; chunk < fcount; chunk += kChunk) {
            // thread tid takes frontier entries k0 .. k0 + kCollapseItems - 1 (in order)
            const int k0 = chunk + tid * kCollapseItems;
            int r[kCollapseItems];
            u32x4 ch[kCollapseItems];
            int cnt = 0;
            // the entries' binary nodes first, then their candidates' child refs
            // (each group of loads issued back to back)
#pragma unroll
            for(int j = 0; j < kCollapseItems; ++j) {
                const int k = k0 + j;
                const int f = k < fcount ? front[k] : -1;
                r[j] = k < fcount ? (f < numInternal && f >= 0 ? f : 0) : -1;
            }
#pragma unroll
            for(int j = 0; j < kCollapseItems; ++j) {
                ch[j] = u32x4{0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u}; // kEmptyChild
                if(r[j] >= 0) ch[j] = ld_global16(reinterpret_cast<const u32x4*>(cand + r[j]) + 6);
            }
#pragma unroll
            for(int j = 0; j < kCollapseItems; ++j)
                cnt += ((int)ch[j].x >= 0) + ((int)ch[j].y >= 0) + ((int)ch[j].z >= 0) + ((int)ch[j].w >= 0);
            int incl = cnt;
#pragma unroll
            for(int off = 1; off < 64; off <<= 1) {
                const int t = __shfl_up(incl, off);
                if(lane >= off) incl += t;
            }
            if(lane == 63) waveTot[wave] = incl;
            __syncthreads();
            int excl = incl - cnt, total = 0;
#pragma unroll
            for(int wv = 0; wv < kWavesPerBlock; ++wv) {
                const int t = waveTot[wv];
                excl += wv < wave ? t : 0;
                total += t;
            }
            int at = nextCount + excl;
#pragma unroll
            for(int j = 0; j < kCollapseItems; ++j) {
                if(r[j] < 0) continue;
                const uint32_t c0 = ch[j].x, c1 = ch[j].y, c2 = ch[j].z, c3 = ch[j].w;
                const int i0 = at, i1 = i0 + ((int)c0 >= 0), i2 = i1 + ((int)c1 >= 0), i3 = i2 + ((int)c2 >= 0);
                at = i3 + ((int)c3 >= 0);
                if((int)c0 >= 0) next[i0] = (int32_t)c0;
                if((int)c1 >= 0) next[i1] = (int32_t)c1;
                if((int)c2 >= 0) next[i2] = (int32_t)c2;
                if((int)c3 >= 0) next[i3] = (int32_t)c3;
                reinterpret_cast<u32x4*>(out + base + k0 + j)[6] =
                    u32x4{(int)c0 >= 0 ? (uint32_t)(base + fcount + i0) : c0, (int)c1 >= 0 ? (uint32_t)(base + fcount + i1) : c1,
                          (int)c2 >= 0 ? (uint32_t)(base + fcount + i2) : c2, (int)c3 >= 0 ? (uint32_t)(base + fcount + i3) : c3};
            }
            nextCount += total;
            __syncthreads(); // waveTot[] and next[] settle before the next chunk / level
        }
        base += fcount;
        fcount = nextCount;
        __threadfence_block();
        __syncthreads();
    }
    if(tid == 0) *numWide = (uint32_t)base;
}

// The rest of every wide node (the child boxes and the padding, seven of its
// eight 16-byte pieces) from its candidate, in parallel: piece q of wide node
// i comes from candidate order[i].
__global__ void k_wide_boxes(const Node4* __restrict__ cand, const int32_t* __restrict__ order,
                             const uint32_t* __restrict__ numWide, int capacity, Node4* __restrict__ out)
{
    const long piece = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long i = piece >> 3;
    const int q = (int)(piece & 7);
    const long nw = min((long)*numWide, (long)capacity);
    if(i >= nw || q == 6) return;
    const int32_t r = order[i];
    if(r < 0 || r >= capacity) return;
    reinterpret_cast<u32x4*>(out + i)[q] = ld_global16(reinterpret_cast<const u32x4*>(cand + r) + q);
}

// ---------------------------------------------------------------------------
// 8. compress the wide nodes
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pow2f(int e) { return __uint_as_float((uint32_t)(e + 127) << 23); }

// Smallest e in [-126, 127] with 254 * 2^e >= ext (one grid step to spare
// for the outward rounding).
__device__ __forceinline__ int quant_exponent(float ext)
{
    if(!(ext > 0.0f)) return -126;
    int e = (int)((__float_as_uint(ext) >> 23) & 0xffu) - 127 - 7;
    e = min(127, max(-126, e));
    while(e < 127 && 254.0f * pow2f(e) < ext) ++e;
    while(e > -126 && 254.0f * pow2f(e - 1) >= ext) --e;
    return e;
}

// Sign of the exact (o + p) - x: the rounded sum and its TwoSum error
// (Knuth) decide it without rounding.
__device__ __forceinline__ int sum_cmp(float o, float p, float x)
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
__device__ __forceinline__ void quantize_range(float o, int e, float lo, float hi, uint8_t& qlo, uint8_t& qhi)
{
    const float s = pow2f(e), inv = pow2f(-min(e, 126));
    int a = (int)floorf((lo - o) * inv);
    a = min(255, max(0, a));
    while(a > 0 && sum_cmp(o, (float)a * s, lo) > 0) --a;
    int b = (int)ceilf((hi - o) * inv);
    b = min(255, max(0, b));
    while(b < 255 && sum_cmp(o, (float)b * s, hi) < 0) ++b;
    qlo = (uint8_t)a;
    qhi = (uint8_t)b;
}

__device__ Node4q quantize_node(const Node4& w)
{
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for(int c = 0; c < 4; ++c) {
        if(w.child[c] == kEmptyChild) continue;
        lo[0] = fminf(lo[0], w.lox[c]);
        lo[1] = fminf(lo[1], w.loy[c]);
        lo[2] = fminf(lo[2], w.loz[c]);
        hi[0] = fmaxf(hi[0], w.hix[c]);
        hi[1] = fmaxf(hi[1], w.hiy[c]);
        hi[2] = fmaxf(hi[2], w.hiz[c]);
    }
    int e[3];
    float org[3];
#pragma unroll
    for(int k = 0; k < 3; ++k) {
        const bool any = lo[k] <= hi[k];
        org[k] = any ? lo[k] : 0.0f;
        e[k] = any ? quant_exponent(hi[k] - lo[k]) : -126;
    }
    // child offsets packed four bytes to a word (byte c = child c)
    uint32_t plo[3] = {0u, 0u, 0u}, phi[3] = {0u, 0u, 0u};
#pragma unroll
    for(int c = 0; c < 4; ++c) {
        if(w.child[c] == kEmptyChild) continue;
        const float clo[3] = {w.lox[c], w.loy[c], w.loz[c]};
        const float chi[3] = {w.hix[c], w.hiy[c], w.hiz[c]};
#pragma unroll
        for(int k = 0; k < 3; ++k) {
            uint8_t qa, qb;
            quantize_range(org[k], e[k], clo[k], chi[k], qa, qb);
            plo[k] |= (uint32_t)qa << (8 * c);
            phi[k] |= (uint32_t)qb << (8 * c);
        }
    }
    Node4q q;
    q.ox = org[0];
    q.oy = org[1];
    q.oz = org[2];
    q.ex = (int8_t)e[0];
    q.ey = (int8_t)e[1];
    q.ez = (int8_t)e[2];
    q.pad0 = 0;
    __builtin_memcpy(q.qlox, &plo[0], 4);
    __builtin_memcpy(q.qhix, &phi[0], 4);
    __builtin_memcpy(q.qloy, &plo[1], 4);
    __builtin_memcpy(q.qhiy, &phi[1], 4);
    __builtin_memcpy(q.qloz, &plo[2], 4);
    __builtin_memcpy(q.qhiz, &phi[2], 4);
    q.pad1[0] = q.pad1[1] = 0;
#pragma unroll
    for(int c = 0; c < 4; ++c) q.child[c] = w.child[c];
    return q;
}

__global__ void k_quantize(const Node4* __restrict__ wide, const uint32_t* __restrict__ numWide, uint32_t capacity,
                           Node4q* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= min(*numWide, capacity)) return;
    out[i] = quantize_node(wide[i]);
}

// ---------------------------------------------------------------------------
// leaf payloads
// ---------------------------------------------------------------------------
// Shading records (ShadeTri) of triangles indexOffset / 3 + k, with the
// gathers clamped to the buffers exactly as the closest-hit stage clamps
// them, so a record equals what that stage would gather.
__global__ void k_shade_records(const rgrt_vertex* __restrict__ vb, const uint32_t* __restrict__ ib, uint64_t nv, uint64_t ni,
                                uint32_t firstVertex, uint32_t indexOffset, uint32_t n, ShadeTri* __restrict__ srec)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n) return;
    const uint64_t lastIndex = ni ? ni - 1 : 0;
    const uint64_t lastVertex = nv ? nv - 1 : 0;
    const uint64_t ib0 = (uint64_t)indexOffset + 3ull * k;
    const uint32_t i0 = ib[ib0 + 0 < lastIndex ? ib0 + 0 : lastIndex];
    const uint32_t i1 = ib[ib0 + 1 < lastIndex ? ib0 + 1 : lastIndex];
    const uint32_t i2 = ib[ib0 + 2 < lastIndex ? ib0 + 2 : lastIndex];
    const uint64_t a0 = (uint64_t)firstVertex + i0, a1 = (uint64_t)firstVertex + i1, a2 = (uint64_t)firstVertex + i2;
    const rgrt_vertex& v0 = vb[a0 < lastVertex ? a0 : lastVertex];
    const rgrt_vertex& v1 = vb[a1 < lastVertex ? a1 : lastVertex];
    const rgrt_vertex& v2 = vb[a2 < lastVertex ? a2 : lastVertex];
    ShadeTri r;
    for(int j = 0; j < 3; ++j) {
        r.n0[j] = v0.normal[j];
        r.n1[j] = v1.normal[j];
        r.n2[j] = v2.normal[j];
    }
    r.matIndex = v0.matIndex;
    r.vertexOffset = firstVertex;
    r.pad = 0;
    srec[indexOffset / 3 + k] = r;
}

__global__ void k_tri_gather(const rgrt_vertex* __restrict__ vb, const uint32_t* __restrict__ ib, uint32_t firstVertex,
                             uint32_t indexOffset, const uint32_t* __restrict__ leafPrim, uint32_t n, Tri* __restrict__ tris)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n) return;
    const uint32_t prim = leafPrim[k];
    const uint32_t* ix = ib + indexOffset + 3 * prim;
    const rgrt_vertex& a = vb[firstVertex + ix[0]];
    const rgrt_vertex& b = vb[firstVertex + ix[1]];
    const rgrt_vertex& c = vb[firstVertex + ix[2]];
#pragma unroll
    for(int r = 0; r < 3; ++r) {
        const int x = r == 2 ? 0 : r + 1, y = r == 0 ? 2 : r - 1; // (r + 1, r + 2) mod 3
        Tri t;
        t.v0x = a.position[x];
        t.v0y = a.position[y];
        t.v0z = a.position[r];
        t.prim = prim;
        t.v1x = b.position[x];
        t.v1y = b.position[y];
        t.v1z = b.position[r];
        t.pad0 = 0;
        t.v2x = c.position[x];
        t.v2y = c.position[y];
        t.v2z = c.position[r];
        t.pad1 = 0;
        tris[(size_t)r * n + k] = t;
    }
}

__global__ void k_inst_gather(const float* __restrict__ w2o, const rgrt_instance* __restrict__ in,
                              const uint32_t* __restrict__ leafInst, uint32_t n, DevInstance* __restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n) return;
    const uint32_t i = leafInst[k];
    DevInstance d;
    for(int q = 0; q < 12; ++q) d.w2o[q] = w2o[12 * (size_t)i + q];
    d.blas = in[i].blas;
    d.mask = in[i].mask & 0xffu;
    d.position = i;
    d.customIndex = in[i].customIndex;
    out[k] = d;
}

// ---------------------------------------------------------------------------
// TLAS of 2..kFusedTlasMax instances in one workgroup
// ---------------------------------------------------------------------------
// The same steps as the multi-kernel path (instance boxes, scene box,
// Morton codes, stable sort, Karras, refit, collapse, compression, leaf
// records), with the intermediate arrays in LDS and one launch instead of
// about twenty. The sort ranks each key by counting the keys that precede
// it in (key, index) order, which is the order the stable radix sort
// produces; the refit combines child boxes in the same order as
// k_refit_chunk / k_refit_top (fminf/fmaxf of slot 0 and slot 1).
// The trees are therefore identical to the multi-kernel path's.
// Instantiated for kMax = 256 and 1024 instances (one thread per instance):
// the smaller form holds 28 KB of LDS instead of 116 KB, so with a frame's
// other kernels resident it finds a CU sooner.
constexpr int kFusedTlasMax = kCollapseThreads;
constexpr int kFusedTlasSmall = 256;

// Binary nodes held in LDS: box[node][12] in Node's float order, child refs.
struct LdsNodes {
    const float* box;
    const int32_t* child0;
    const int32_t* child1;
    __device__ __forceinline__ WideEntry slot(int node, int c) const
    {
        WideEntry e;
        e.ref = c ? child1[node] : child0[node];
        const float* f = box + 12 * node;
        for(int k = 0; k < 3; ++k) {
            e.lo[k] = f[3 * c + k];
            e.hi[k] = f[6 + 3 * c + k];
        }
        return e;
    }
};

template <int kMax>
__global__ __launch_bounds__(kMax) void k_tlas_fused(
    const rgrt_instance* __restrict__ in, int n, const float4* __restrict__ blasRoot, float* __restrict__ w2oOut,
    InstanceShade* __restrict__ shade, Node* __restrict__ nodes, uint32_t* __restrict__ leafOut, float4* __restrict__ rootBox,
    Node4* __restrict__ wide, Node4q* __restrict__ wideq, uint32_t* __restrict__ numWide, DevInstance* __restrict__ instOut,
    int refit)
{
    __shared__ float ilo[3][kMax], ihi[3][kMax]; // instance world boxes
    __shared__ uint32_t key0[kMax];              // by instance
    __shared__ uint32_t keys[kMax];              // sorted
    __shared__ uint32_t order[kMax];             // sorted position -> instance
    __shared__ int32_t child0[kMax], child1[kMax];
    __shared__ uint32_t parent[kMax], leafParent[kMax], arrive[kMax];
    __shared__ float box[12 * (kMax - 1)];
    __shared__ int32_t frontA[kMax], frontB[kMax], waveTot[kMax / 64];
    __shared__ uint32_t sb[6];

    const int tid = threadIdx.x;
    const bool mine = tid < n;
    if(tid < 3) sb[tid] = float_to_ordered(INFINITY);
    else if(tid < 6) sb[tid] = float_to_ordered(-INFINITY);
    __syncthreads();

    // instance boxes, inverses, shading records; scene box
    float w[12];
    uint32_t blas = 0, mask = 0, custom = 0;
    float lo[3], hi[3];
    if(mine) {
        const rgrt_instance r = in[tid];
        InstanceShade sh;
        inst_prep_one(r, blasRoot, lo, hi, w, sh);
        for(int q = 0; q < 12; ++q) w2oOut[12 * (size_t)tid + q] = w[q];
        shade[tid] = sh;
        blas = r.blas;
        mask = r.mask & 0xffu;
        custom = r.customIndex;
        for(int c = 0; c < 3; ++c) {
            ilo[c][tid] = lo[c];
            ihi[c][tid] = hi[c];
            atomicMin(&sb[c], float_to_ordered(lo[c]));
            atomicMax(&sb[3 + c], float_to_ordered(hi[c]));
        }
    }
    __syncthreads();

    uint32_t rank = 0;
    if(refit) {
        // the existing leaf order: key0 holds each instance's position in it
        if(mine) {
            const uint32_t i = leafOut[tid];
            order[tid] = i < (uint32_t)n ? i : 0u;
            if(i < (uint32_t)n) key0[i] = (uint32_t)tid;
        }
        __syncthreads();
        if(mine) rank = min(key0[tid], (uint32_t)n - 1u);
    }
    else {
        // Morton codes, then the rank of each in (key, index) order
        uint32_t key = 0;
        if(mine) {
            key = morton_key(lo, hi, sb);
            key0[tid] = key;
        }
        __syncthreads();
        if(mine) {
            for(int j = 0; j < n; ++j) {
                const uint32_t kj = key0[j];
                rank += (kj < key || (kj == key && j < tid)) ? 1u : 0u;
            }
            keys[rank] = key;
            order[rank] = (uint32_t)tid;
        }
    }
    if(mine) {
        // instance leaf at its Morton position
        DevInstance d;
        for(int q = 0; q < 12; ++q) d.w2o[q] = w[q];
        d.blas = blas;
        d.mask = mask;
        d.position = (uint32_t)tid;
        d.customIndex = custom;
        instOut[rank] = d;
    }
    __syncthreads();

    // hierarchy: Karras, or the existing one for a refit
    if(tid < n - 1) {
        int left, right;
        if(refit) {
            const Node nd = nodes[tid];
            left = nd.child0;
            right = nd.child1;
        }
        else {
            int gamma;
            int rlo, rhi;
            karras_node(keys, n, tid, left, right, gamma, rlo, rhi);
        }
        child0[tid] = left;
        child1[tid] = right;
        arrive[tid] = 0;
        if(tid == 0) parent[0] = 0xffffffffu;
        // every ref stays inside the tables (a refit reads nodes this
        // library wrote; the check only keeps the accesses in bounds)
        if(left >= 0) {
            if(left < n - 1) parent[left] = (uint32_t)tid;
        }
        else if(~left < n) leafParent[~left] = (uint32_t)tid;
        if(right >= 0) {
            if(right < n - 1) parent[right] = (uint32_t)tid;
        }
        else if(~right < n) leafParent[~right] = (uint32_t)tid;
    }
    __syncthreads();

    // bottom-up refit: the second arrival at a node combines its two slots
    if(mine) {
        const uint32_t prim = order[tid];
        float blo[3], bhi[3];
        for(int c = 0; c < 3; ++c) {
            blo[c] = ilo[c][prim];
            bhi[c] = ihi[c][prim];
        }
        int child = ~tid;
        uint32_t p = leafParent[tid];
        for(int level = 0; level < 128; ++level) {
            if(p >= (uint32_t)(n - 1)) break;
            float* f = box + 12 * p;
            const int slot = child0[p] == child ? 0 : 1;
            for(int c = 0; c < 3; ++c) {
                f[3 * slot + c] = blo[c];
                f[6 + 3 * slot + c] = bhi[c];
            }
            __threadfence_block();
            if(atomicAdd(&arrive[p], 1u) == 0u) break; // the sibling's thread finishes this node
            __threadfence_block();
            for(int c = 0; c < 3; ++c) {
                blo[c] = fminf(f[c], f[3 + c]);
                bhi[c] = fmaxf(f[6 + c], f[9 + c]);
            }
            if(p == 0) {
                rootBox[0] = make_float4(blo[0], blo[1], blo[2], 0.0f);
                rootBox[1] = make_float4(bhi[0], bhi[1], bhi[2], 0.0f);
                break;
            }
            child = (int)p;
            p = parent[p];
        }
    }
    __syncthreads();

    // binary nodes and leaf order for export
    if(tid < n - 1) {
        Node nd;
        float* f = reinterpret_cast<float*>(&nd);
        for(int q = 0; q < 12; ++q) f[q] = box[12 * tid + q];
        nd.child0 = child0[tid];
        nd.child1 = child1[tid];
        nd.parent = parent[tid];
        nd.pad = 0;
        nodes[tid] = nd;
    }
    if(mine) leafOut[tid] = order[tid];

    // wide + compressed nodes
    const int nw = collapse_bfs<kMax>(LdsNodes{box, child0, child1}, wide, wideq, frontA, frontB, waveTot);
    if(tid == 0) *numWide = (uint32_t)nw;
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

} // namespace

// ---------------------------------------------------------------------------
// Host-side drivers
// ---------------------------------------------------------------------------
void LbvhBuilder::reserve(uint32_t n)
{
    if(n <= capacity) return;
    release();
    const uint32_t cap = n < 1024 ? 1024 : n;
    const uint32_t nb = blocks_for(cap, kSortTile);
    check(hipMalloc(&keysA, cap * 4ull));
    check(hipMalloc(&keysB, cap * 4ull));
    check(hipMalloc(&valsA, cap * 4ull));
    check(hipMalloc(&valsB, cap * 4ull));
    check(hipMalloc(&hist, 256ull * nb * 4ull));
    check(hipMalloc(&boxLo, cap * 16ull));
    check(hipMalloc(&boxHi, cap * 16ull));
    check(hipMalloc(&leafParent, cap * 4ull));
    check(hipMalloc(&counters, cap * 4ull));
    check(hipMalloc(&sceneBox, 64));
    check(hipMalloc(&nodeRange, cap * 8ull));
    check(hipMalloc(&wideTmp, cap * sizeof(Node4)));
    capacity = cap;
}

void LbvhBuilder::bind(void* base, uint32_t cap)
{
    char* p = static_cast<char*>(base);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += lbvh_scratch_align(bytes);
        return q;
    };
    const uint32_t nb = blocks_for(cap, kSortTile);
    keysA = reinterpret_cast<uint32_t*>(take(cap * 4ull));
    keysB = reinterpret_cast<uint32_t*>(take(cap * 4ull));
    valsA = reinterpret_cast<uint32_t*>(take(cap * 4ull));
    valsB = reinterpret_cast<uint32_t*>(take(cap * 4ull));
    hist = reinterpret_cast<uint32_t*>(take(256ull * nb * 4ull));
    boxLo = reinterpret_cast<float4*>(take(cap * 16ull));
    boxHi = reinterpret_cast<float4*>(take(cap * 16ull));
    leafParent = reinterpret_cast<uint32_t*>(take(cap * 4ull));
    counters = reinterpret_cast<uint32_t*>(take(cap * 4ull));
    sceneBox = reinterpret_cast<uint32_t*>(take(64));
    nodeRange = reinterpret_cast<uint2*>(take(cap * 8ull));
    wideTmp = reinterpret_cast<Node4*>(take(cap * sizeof(Node4)));
    static_assert(sizeof(Node4) == 128, "the scratch layout assumes 128-byte wide nodes");
    if((size_t)(p - static_cast<char*>(base)) != lbvh_scratch_bytes(cap))
        throw std::logic_error("LBVH scratch layout differs from lbvh_scratch_bytes");
    capacity = cap;
}

void LbvhBuilder::release()
{
    for(void* p: {(void*)keysA, (void*)keysB, (void*)valsA, (void*)valsB, (void*)hist, (void*)boxLo, (void*)boxHi,
                  (void*)leafParent, (void*)counters, (void*)sceneBox, (void*)nodeRange, (void*)wideTmp})
        if(p) (void)hipFree(p);
    keysA = keysB = valsA = valsB = hist = leafParent = counters = sceneBox = nullptr;
    boxLo = boxHi = nullptr;
    nodeRange = nullptr;
    wideTmp = nullptr;
    capacity = 0;
}

// Builds the hierarchy over boxLo/boxHi[0..n) (already filled). Writes
// max(n-1,1) nodes, the Morton-ordered primitive ids and the root box.
void LbvhBuilder::build(hipStream_t s, uint32_t n, Node* nodes, uint32_t* leafPrimOut, float4* rootBox)
{
    if(n <= 1) {
        k_small_tree<<<1, 64, 0, s>>>(boxLo, boxHi, (int)n, nodes, rootBox);
        if(n == 1) k_fill_u32<<<1, 64, 0, s>>>(leafPrimOut, 1, 0u);
        return;
    }
    k_scene_box_init<<<1, 64, 0, s>>>(sceneBox);
    const uint32_t sbBlocks = std::min<uint32_t>(blocks_for(n, 256), 1024);
    k_scene_box<<<sbBlocks, 256, 0, s>>>(boxLo, boxHi, n, sceneBox);
    k_morton<<<blocks_for(n, 256), 256, 0, s>>>(boxLo, boxHi, n, sceneBox, keysA, valsA);
    const uint32_t nb = blocks_for(n, kSortTile);
    uint32_t *kin = keysA, *vin = valsA, *kout = keysB, *vout = valsB;
    for(int shift = 0; shift < 32; shift += 8) {
        k_sort_hist<<<nb, kSortThreads, 0, s>>>(kin, n, shift, hist, nb);
        // counters[] is free until the refit: it holds the segment totals
        k_scan_reduce<<<kScanBlocks, kScanThreads, 0, s>>>(hist, 256u * nb, counters);
        k_scan_partials<<<1, kScanThreads, 0, s>>>(counters);
        k_scan_apply<<<kScanBlocks, kScanThreads, 0, s>>>(hist, 256u * nb, counters);
        k_sort_scatter<<<nb, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, shift, hist, nb);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // after 4 passes the sorted data is back in keysA / valsA
    k_karras<<<blocks_for(n - 1, 256), 256, 0, s>>>(kin, (int)n, nodes, leafParent, nodeRange);
    refit_boxes(s, n, vin, nodes, rootBox);
    check(hipMemcpyAsync(leafPrimOut, vin, n * 4ull, hipMemcpyDeviceToDevice, s));
}

void LbvhBuilder::refit_boxes(hipStream_t s, uint32_t n, const uint32_t* leafPrim, Node* nodes, float4* rootBox)
{
    // counters[0, n-1) per node, counters[n-1] the arrivals the top pass
    // takes over; keysB (free after the sort) holds those arrivals
    k_fill_u32<<<blocks_for(n, 256), 256, 0, s>>>(counters, n, 0u);
    k_refit_chunk<<<blocks_for(n, kRefitChunk), kRefitChunk, 0, s>>>(boxLo, boxHi, leafPrim, leafParent, nodeRange, nodes,
                                                                     counters, keysB, (int)n, rootBox);
    if(n > (uint32_t)kRefitChunk) k_refit_top<<<1, kRefitTopThreads, 0, s>>>(nodes, counters, keysB, (int)n, rootBox);
}

// Refit with the existing topology: new primitive boxes, same leaf order.
void LbvhBuilder::refit(hipStream_t s, uint32_t n, Node* nodes, const uint32_t* leafPrim, float4* rootBox)
{
    if(n <= 1) {
        k_small_tree<<<1, 64, 0, s>>>(boxLo, boxHi, (int)n, nodes, rootBox);
        return;
    }
    // leafParent and the leaf ranges from the existing nodes
    recompute_leaf_parents(s, n, nodes);
    k_node_ranges<<<blocks_for(n - 1, 256), 256, 0, s>>>(nodes, (int)n, nodeRange);
    refit_boxes(s, n, leafPrim, nodes, rootBox);
}

namespace {
__global__ void k_leaf_parents(const Node* __restrict__ nodes, uint32_t numInternal, uint32_t* __restrict__ leafParent)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= numInternal) return;
    const Node nd = nodes[i];
    if(nd.child0 < 0 && nd.child0 != kEmptyChild) leafParent[~nd.child0] = i;
    if(nd.child1 < 0 && nd.child1 != kEmptyChild) leafParent[~nd.child1] = i;
}
} // namespace

void LbvhBuilder::collapse(hipStream_t s, uint32_t n, const Node* nodes, Node4* wide, uint32_t* numWide)
{
    // the sort's second key buffer is free after the build: it holds the
    // breadth-first order of the wide nodes (at most n - 1 entries)
    const int numInternal = (int)std::max<uint32_t>(n, 2) - 1;
    k_wide_candidates<<<blocks_for(numInternal, 256), 256, 0, s>>>(nodes, numInternal, wideTmp);
    k_collapse_numbered<<<1, kCollapseThreads, 0, s>>>(wideTmp, numInternal, wide, reinterpret_cast<int32_t*>(keysB), numWide);
    k_wide_boxes<<<blocks_for(8u * (uint32_t)numInternal, 256), 256, 0, s>>>(wideTmp, reinterpret_cast<const int32_t*>(keysB),
                                                                             numWide, numInternal, wide);
}

void LbvhBuilder::quantize(hipStream_t s, uint32_t capacity, const Node4* wide, const uint32_t* numWide, Node4q* out)
{
    k_quantize<<<blocks_for(capacity, 256), 256, 0, s>>>(wide, numWide, capacity, out);
}

void LbvhBuilder::recompute_leaf_parents(hipStream_t s, uint32_t n, const Node* nodes)
{
    k_leaf_parents<<<blocks_for(n - 1, 256), 256, 0, s>>>(nodes, n - 1, leafParent);
}

void launch_tri_boxes(hipStream_t s, const rgrt_vertex* vb, const uint32_t* ib, uint32_t firstVertex, uint32_t indexOffset,
                      uint32_t n, float4* boxLo, float4* boxHi)
{
    if(n) k_tri_boxes<<<blocks_for(n, 256), 256, 0, s>>>(vb, ib, firstVertex, indexOffset, n, boxLo, boxHi);
}

void launch_shade_records(hipStream_t s, const rgrt_vertex* vb, const uint32_t* ib, uint64_t nv, uint64_t ni,
                          uint32_t firstVertex, uint32_t indexOffset, uint32_t n, ShadeTri* srec)
{
    if(n && srec && indexOffset % 3 == 0)
        k_shade_records<<<blocks_for(n, 256), 256, 0, s>>>(vb, ib, nv, ni, firstVertex, indexOffset, n, srec);
}

void launch_tri_gather(hipStream_t s, const rgrt_vertex* vb, const uint32_t* ib, uint32_t firstVertex, uint32_t indexOffset,
                       const uint32_t* leafPrim, uint32_t n, Tri* tris)
{
    if(n) k_tri_gather<<<blocks_for(n, 256), 256, 0, s>>>(vb, ib, firstVertex, indexOffset, leafPrim, n, tris);
}

void launch_inst_prep(hipStream_t s, const rgrt_instance* in, uint32_t n, const float4* blasRoot, float4* boxLo,
                      float4* boxHi, float* w2o, InstanceShade* shade)
{
    if(n) k_inst_prep<<<blocks_for(n, 128), 128, 0, s>>>(in, n, blasRoot, boxLo, boxHi, w2o, shade);
}

void launch_inst_gather(hipStream_t s, const float* w2o, const rgrt_instance* in, const uint32_t* leafInst, uint32_t n,
                        DevInstance* out)
{
    if(n) k_inst_gather<<<blocks_for(n, 128), 128, 0, s>>>(w2o, in, leafInst, n, out);
}

bool launch_tlas_fused(hipStream_t s, const rgrt_instance* in, uint32_t n, const float4* blasRoot, float* w2o,
                       InstanceShade* shade, Node* nodes, uint32_t* leaf, float4* rootBox, Node4* wide, Node4q* wideq,
                       uint32_t* numWide, DevInstance* inst, bool refit)
{
    if(n < 2 || n > (uint32_t)kFusedTlasMax) return false;
    if(n <= (uint32_t)kFusedTlasSmall) {
        k_tlas_fused<kFusedTlasSmall><<<1, kFusedTlasSmall, 0, s>>>(in, (int)n, blasRoot, w2o, shade, nodes, leaf, rootBox, wide,
                                                                    wideq, numWide, inst, refit ? 1 : 0);
        return true;
    }
    k_tlas_fused<kFusedTlasMax><<<1, kFusedTlasMax, 0, s>>>(in, (int)n, blasRoot, w2o, shade, nodes, leaf, rootBox, wide, wideq, numWide,
                                                inst, refit ? 1 : 0);
    return true;
}

} // namespace rgrt
