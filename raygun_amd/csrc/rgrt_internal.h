// rgrt_internal.h — host-side structures shared by the HIP core's sources.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "rgrt_device.h"

namespace rgrt {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
};

inline void check(hipError_t e)
{
    if(e != hipSuccess) throw HipError(e, hipGetErrorString(e));
}

// Reusable scratch of the LBVH builder (device memory).
// Primitives per radix-sort workgroup of the LBVH builder (256 threads x 8).
constexpr uint32_t kLbvhSortTile = 2048;

// Bytes of the LBVH builder's working arrays for `cap` primitives (cap >=
// 1024) when they sit in one block (LbvhBuilder::bind): four key/value
// arrays, the sort histogram, two box arrays, leaf parents, counters, the
// scene box, leaf ranges and the candidate wide nodes, each 256-B aligned.
constexpr size_t lbvh_scratch_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t lbvh_scratch_bytes(uint32_t cap)
{
    return 4 * lbvh_scratch_align(cap * 4ull) + lbvh_scratch_align(256ull * ((cap + kLbvhSortTile - 1) / kLbvhSortTile) * 4) +
           2 * lbvh_scratch_align(cap * 16ull) + 2 * lbvh_scratch_align(cap * 4ull) + lbvh_scratch_align(64) +
           lbvh_scratch_align(cap * 8ull) + lbvh_scratch_align(cap * 128ull);
}

struct LbvhBuilder {
    uint32_t capacity = 0;
    uint32_t *keysA = nullptr, *keysB = nullptr, *valsA = nullptr, *valsB = nullptr;
    uint32_t* hist = nullptr;
    float4 *boxLo = nullptr, *boxHi = nullptr;
    uint32_t* leafParent = nullptr;
    uint32_t* counters = nullptr;
    uint32_t* sceneBox = nullptr;
    uint2* nodeRange = nullptr; // first and last leaf of every internal node
    Node4* wideTmp = nullptr;   // candidate wide node of every binary node (collapse)

    void reserve(uint32_t n);
    void release();
    // Points the working arrays into `base` (device memory of
    // lbvh_scratch_bytes(cap) bytes, 256-B aligned, owned by the caller);
    // release() must not be called on such a builder.
    void bind(void* base, uint32_t cap);
    void build(hipStream_t s, uint32_t n, Node* nodes, uint32_t* leafPrimOut, float4* rootBox);
    void refit(hipStream_t s, uint32_t n, Node* nodes, const uint32_t* leafPrim, float4* rootBox);
    void recompute_leaf_parents(hipStream_t s, uint32_t n, const Node* nodes);
    // Bottom-up boxes of a tree whose leafParent and nodeRange are set.
    void refit_boxes(hipStream_t s, uint32_t n, const uint32_t* leafPrim, Node* nodes, float4* rootBox);
    // Collapses the binary tree of n leaves into 4-wide nodes (breadth-first);
    // writes the wide node count to *numWide (device).
    void collapse(hipStream_t s, uint32_t n, const Node* nodes, Node4* wide, uint32_t* numWide);
    // Compresses the wide nodes (count read from *numWide, at most capacity).
    void quantize(hipStream_t s, uint32_t capacity, const Node4* wide, const uint32_t* numWide, Node4q* out);
};

void launch_tri_boxes(hipStream_t s, const rgrt_vertex* vb, const uint32_t* ib, uint32_t firstVertex, uint32_t indexOffset,
                      uint32_t n, float4* boxLo, float4* boxHi);
// Shading records of the n triangles from index offset `indexOffset` (a
// multiple of 3) read with vertex offset firstVertex, into srec[indexOffset / 3 ..].
void launch_shade_records(hipStream_t s, const rgrt_vertex* vb, const uint32_t* ib, uint64_t nv, uint64_t ni,
                          uint32_t firstVertex, uint32_t indexOffset, uint32_t n, ShadeTri* srec);
void launch_tri_gather(hipStream_t s, const rgrt_vertex* vb, const uint32_t* ib, uint32_t firstVertex, uint32_t indexOffset,
                       const uint32_t* leafPrim, uint32_t n, Tri* tris);
void launch_inst_prep(hipStream_t s, const rgrt_instance* in, uint32_t n, const float4* blasRoot, float4* boxLo,
                      float4* boxHi, float* w2o, InstanceShade* shade);
void launch_inst_gather(hipStream_t s, const float* w2o, const rgrt_instance* in, const uint32_t* leafInst, uint32_t n,
                        DevInstance* out);
// Whole TLAS build (instance prep through leaf records) in one workgroup for
// 2 <= n <= 1024 instances; returns false, launching nothing, otherwise.
// With refit set, the hierarchy and leaf order already in nodes / leaf are
// kept and only the boxes, wide nodes and instance records are redone.
bool launch_tlas_fused(hipStream_t s, const rgrt_instance* in, uint32_t n, const float4* blasRoot, float* w2o,
                       InstanceShade* shade, Node* nodes, uint32_t* leaf, float4* rootBox, Node4* wide, Node4q* wideq,
                       uint32_t* numWide, DevInstance* inst, bool refit);

// ---------------------------------------------------------------------------
// Wavefront state (device memory, structure of arrays, one slot per pixel)
// ---------------------------------------------------------------------------
// A pixel's state between passes: the payload fields a later pass reads
// (hit value, depth, IOR, reference depth, recursion depth) and the frame
// stack depth, ray type and sample number, 8 words.
constexpr int kPayloadWords = 8;
constexpr int kAccumWords = 12;
// A suspended closest-hit frame: stage, material index, hit distance,
// origin, direction, normal, the reference depth at the hit, shadow and
// reflection colours and reflection depth (the material's values and the
// grid effect are recomputed from these on resume), 19 words + 1 of padding.
constexpr int kFrameStride = 20;
constexpr uint32_t kMaxMaterials = 1u << 28; // the material index shares a frame word with the stage
constexpr int kMaxFrames = 8; // maxRecursions <= 7 keeps at most 7 frames waiting
constexpr int kRayWords = 10; // origin, tmin, dir, tmax, pixel, flags
constexpr int kHitWords = 5;
constexpr int kTraversalErrorWord = 12; // in WavefrontState::counters, sticky (cleared at allocation)
constexpr int kCounterBlock = 16;       // counter words per band (the error word is band 0's)
constexpr uint32_t kMaxBands = 4;

// Ray flags in a queue entry
constexpr uint32_t kFlagMiss1 = 1u;     // miss shader index 1 (shadow miss)
constexpr uint32_t kFlagPrimary = 2u;   // primary ray of sample 0 (AOV record)
constexpr uint32_t kMaskShift = 8u;     // cull mask in bits 8..15

struct WavefrontState {
    uint32_t npix = 0; // wavefront slots: image pixels x sample groups (of the band, see chunkFirst)
    uint32_t nimg = 0; // image pixels
    // A frame traced in bands: a band takes the npix slots from the frame's
    // 64-slot chunk chunkFirst on (its slot q is frame slot chunkFirst * 64
    // + q), with its own queues, hit records and counter block; the per-slot
    // arrays (payload, accum, frames, lum2) are the whole frame's, frames
    // with frameSlots slots per level. The frame in one piece: chunkFirst 0,
    // frameSlots = npix.
    uint32_t chunkFirst = 0;
    uint32_t frameSlots = 0;
    // Samples of a pixel traced side by side: slot g * nimg + i runs samples
    // g, g + groups, ... of the pixel in tile order i and stores each
    // sample's record; k_finalize_samples sums them in sample order.
    uint32_t groups = 1;
    uint32_t numSamples = 1;     // of the frame being traced
    float* sampleAcc = nullptr;  // [numSamples][nimg][kAccumWords] (groups > 1)
    float* sampleLum2 = nullptr; // [numSamples][nimg] (groups > 1, stochastic sampling)
    // Per-slot records are contiguous (a lane reads its slot's record with
    // 16-byte loads, and consecutive lanes read consecutive records).
    float* payload = nullptr; // [npix][kPayloadWords]
    float* accum = nullptr;   // [npix][kAccumWords]
    float* frames = nullptr;  // [kMaxFrames][npix][kFrameStride]
    float* rays[2] = {nullptr, nullptr}; // [kRayWords][npix]
    float* hits = nullptr;    // [kHitWords][npix]
    uint32_t* counters = nullptr; // [0..1] queue counts, [2..3] fetch, [4..11] stats (u64 pairs), [12] traversal errors
    float* lum2 = nullptr;        // [npix]: sum of squared sample luminances (stochastic sampling)
};

struct Images {
    uint32_t w = 0, h = 0;
    uint16_t* base = nullptr;
    uint16_t* normal = nullptr;
    uint16_t* rough = nullptr;
    uint16_t* final_ = nullptr;
    int8_t* transitions = nullptr;
    uint16_t* colorsA = nullptr;
    uint16_t* colorsB = nullptr;
    float* base32 = nullptr;
    float* normal32 = nullptr;
    float* rough32 = nullptr;
    rgrt_hit* primary = nullptr;
    uint16_t* features = nullptr; // [npix][2][4] fp16 (stochastic sampling)
};

// Arguments of the shading kernels
struct ShadeParams {
    const rgrt_vertex* vb;
    const uint32_t* ib;
    const rgrt_material* mats;
    const InstanceShade* inst; // by instance position
    const ShadeTri* srec;      // by triangle of the index buffer (numIndices / 3)
    uint32_t numInstances;
    uint64_t numVertices, numIndices;
    uint32_t numMaterials;
    rgrt_ubo ubo;
    uint32_t width, height;
    WavefrontState st;
    Images img;
    int32_t keepF32;
    int32_t keepAov;
    uint32_t sampling; // RGRT_SAMPLING_*
    uint32_t seed;
};

// ---------------------------------------------------------------------------
// Learned denoiser (denoise.hip)
// ---------------------------------------------------------------------------
class Denoiser {
  public:
    static constexpr int kNumLayers = 11;
    // activation and gradient tensors (bf16 NHWC), in workspace order
    enum Buf {
        X0, A1, S1, P1, A2, S2, P2, A3, B3, C2, A4, U2, C1, A5, U1, OUT, // forward
        DZO, GA, GC1, GU2, GS1X, GC2, GB3, GS2X, GL2A, GP2, GL1A, GP1, GS1, GB, GS2, // backward
        kNumBufs
    };
    static constexpr int kNumFwdBufs = OUT + 1;

    explicit Denoiser(uint32_t seed);
    ~Denoiser();
    Denoiser(const Denoiser&) = delete;
    Denoiser& operator=(const Denoiser&) = delete;

    size_t num_params() const { return numParams; }
    // true once parameters were loaded (set_params) or trained (train);
    // the He-initialised network of a fresh denoiser is not a denoiser
    bool has_trained_params() const { return paramsLoaded; }
    void set_params(hipStream_t s, const float* host);
    void get_params(hipStream_t s, float* host);
    // forward + loss + gradients on host inputs (x0: N*H*W*16 bf16 bits,
    // target: N*H*W*3 fp32); returns the loss, gradients to gradHost
    float loss_and_grads(hipStream_t s, const uint16_t* x0Host, const float* targetHost, int N, int H, int W,
                         float* gradHost);
    void read_output(hipStream_t s, uint16_t* host); // o of the last forward pass
    // forward pass alone on host inputs (x0: N*H*W*16 bf16 bits); o to outHost
    void infer(hipStream_t s, const uint16_t* x0Host, int N, int H, int W, uint16_t* outHost);
    // training set in device memory: packed inputs and fp32 targets
    void dataset_reserve(uint32_t count, int H, int W);
    void dataset_add_input(hipStream_t s, const Images& img, uint32_t slot);
    void dataset_add_target(hipStream_t s, const Images& img, uint32_t slot, float weight, bool accumulate);
    void dataset_read(hipStream_t s, uint32_t first, uint32_t count, int H, int W, uint16_t* x0Host, float* targetHost);
    // steps of Adam on random batches of the first `used` samples; returns the mean loss of the last steps
    float train(hipStream_t s, uint32_t steps, uint32_t batch, float lr, uint32_t seed, uint32_t used);
    // denoises img.base in place (rgb; fallback tiles keep the traced frame); returns the tile count
    uint32_t denoise(hipStream_t s, Images& img, float fallbackA, float fallbackB);
    uint32_t count_fallback(hipStream_t s, uint32_t tiles);

  private:
    void reserve(int N, int H, int W, bool training);
    void prep_weights(hipStream_t s);
    // keepConcat: also store the decoder's concatenated inputs (the backward pass reads them)
    void forward(hipStream_t s, int N, int H, int W, bool keepConcat);
    void backward(hipStream_t s, const float* target, int N, int H, int W);
    void adam_step(hipStream_t s, float lr);

    size_t numParams = 0;
    size_t wOff[kNumLayers], bOff[kNumLayers], wbfOff[kNumLayers];
    float *params = nullptr, *grads = nullptr, *adamM = nullptr, *adamV = nullptr;
    uint16_t *wFwd = nullptr, *wBwd = nullptr;
    bool weightsReady = false;
    bool paramsLoaded = false;
    uint32_t stepCount = 0;
    void* ws = nullptr;
    size_t wsBytes = 0;
    uint16_t* buf[kNumBufs] = {};
    float* partial = nullptr;
    float* lossPart = nullptr;
    uint8_t* fallback = nullptr;
    int wsN = 0, wsH = 0, wsW = 0;
    bool wsTraining = false;
    bool ds = false;
    uint16_t* dsX = nullptr;
    float* dsT = nullptr;
    uint32_t dsCount = 0;
    int dsH = 0, dsW = 0;
};

// Scenes a batch traces in one launch per pass at most (RGRT_BATCH_PERSISTENT).
constexpr int kMaxBatchScenes = 64;

// Scene of block b of a batch launch from the scenes' block starts
// (start[s] <= b < start[s + 1]); uniform over the block.
__device__ __forceinline__ int block_scene(const uint32_t* start, int numScenes, uint32_t b)
{
    int s = 0;
    while(s + 1 < numScenes && start[s + 1] <= b) ++s;
    return s;
}

void launch_raygen(hipStream_t s, const ShadeParams& p);
// One pass / the drain of every scene of a batch in one launch; ps and scs
// are device arrays of numScenes entries (1..kMaxBatchScenes).
void launch_trace_multi(hipStream_t s, const ShadeParams* ps, const TraceScene* scs, int numScenes, int cur, uint32_t gridBlocks,
                        int32_t* spill, uint32_t* batchCounters);
// oneSample: every scene runs one sample per wavefront slot (run_one_sample)
void launch_shade_multi(hipStream_t s, const ShadeParams* ps, int numScenes, int cur, uint32_t gridBlocks, bool oneSample);
void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n);
// rough + composite and FXAA of every scene of a batch, one launch each
// (start: block starts, see post.hip); tile counts per scene for the starts
void launch_post_multi(hipStream_t s, const ShadeParams* ps, const uint32_t* start, int numScenes, uint32_t roughBlocks,
                       uint32_t fxaaBlocks);
uint32_t post_tiles(uint32_t w, uint32_t h);
uint32_t fxaa_tiles(uint32_t w, uint32_t h);
// Ray generation of every scene of a batch in one launch (start: numScenes + 1
// block starts of 256 slots each)
void launch_raygen_multi(hipStream_t s, const ShadeParams* ps, const uint32_t* start, int numScenes, uint32_t blocks);
void launch_drain_multi(hipStream_t s, const ShadeParams* ps, const TraceScene* scs, int numScenes, int cur, uint32_t gridBlocks,
                        int32_t* spill, bool oneSample);
// True when every wavefront slot of the frame runs exactly one sample of its
// pixel (sample groups >= samples): the shading kernels then leave out the
// start of a next sample.
bool run_one_sample(const ShadeParams& p);
void launch_trace_queue(hipStream_t s, const TraceScene& sc, const WavefrontState& st, int cur, uint32_t gridBlocks);
void launch_shade(hipStream_t s, const ShadeParams& p, int cur, uint32_t gridBlocks);
void launch_finalize_samples(hipStream_t s, const ShadeParams& p);
void launch_drain(hipStream_t s, const ShadeParams& p, const TraceScene& sc, int cur, uint32_t gridBlocks, int32_t* spill);
uint32_t drain_spill_words_per_thread();
uint32_t drain_block_size();
void launch_trace_rays(hipStream_t s, const TraceScene& sc, const rgrt_ray* rays, uint32_t n, uint32_t cullMask, rgrt_hit* hits);
void launch_post(hipStream_t s, const Images& img, const rgrt_ubo& ubo, int useFxaa, hipEvent_t roughEnd);
// rgba16f -> 8-bit RGBA / BGRA (optionally sRGB-encoded), 4 bytes per pixel
void launch_to_rgba8(hipStream_t s, const uint16_t* src, uint8_t* dst, size_t npix, int bgra, int srgb);

} // namespace rgrt
