/*
 * rgrt.h — C ABI of the MI355X-native ray-tracing core for Raygun.
 *
 * Plain C: no exceptions cross this boundary, every entry point returns an
 * rgrt_status (0 = OK), handles are opaque, host arrays are owned by the
 * caller, device memory is owned by the library. A context is not
 * thread-safe: one host thread drives it, as in the reference where the main
 * thread records and submits all GPU work.
 *
 * Record layouts are byte-identical to the reference's X-macro `.def` files
 * that are shared by its C++ and GLSL code, so a Raygun scene packs into the
 * same buffers it packs into today.
 */
#ifndef RGRT_H
#define RGRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
typedef int32_t rgrt_status;
#define RGRT_OK 0
#define RGRT_ERR_INVALID_ARG 1
#define RGRT_ERR_OUT_OF_MEMORY 2
#define RGRT_ERR_HIP 3
#define RGRT_ERR_NOT_READY 4   /* e.g. trace before TLAS / frame constants */
#define RGRT_ERR_NO_DEVICE 5

/* ------------------------------------------------------------------------ */
/* Records shared with the engine (layout contracts)                          */
/* ------------------------------------------------------------------------ */

/* Vertex, 32 bytes. Replaces resources/shaders/vertex.def:3-7 and
 * raygun/render/vertex.hpp:27-29 (struct Vertex). */
typedef struct rgrt_vertex {
    float position[3];
    uint32_t matIndex;
    float normal[3];
    float pad1;
} rgrt_vertex;

/* Material, 64 bytes. Replaces resources/shaders/gpu_material.def:11-26 and
 * raygun/gpu/gpu_material.hpp:27-31 (gpu::Material). Defaults of the
 * reference: diffuse (1,0,1), transparency 0, specular (1,1,1),
 * reflectivity 0, roughness 0, ior 1, effectId 0, rayConsumption 1,
 * emission 0. */
typedef struct rgrt_material {
    float diffuse[3];
    float transparency;
    float specular[3];
    float reflectivity;
    float roughness;
    float ior;
    uint32_t effectId;
    uint32_t rayConsumption;
    float emission;
    float pad0, pad1, pad2;
} rgrt_material;

/* Per-frame constants, 192 bytes, std140. Replaces
 * resources/shaders/uniform_buffer_object.def:3-17 and
 * raygun/gpu/uniform_buffer.hpp:29-31. Matrices are column-major 4x4
 * (glm). `showAlpha` is a 4-byte boolean slot here; the reference's C++
 * `bool` occupies the first byte of the same slot and the rest is zeroed by
 * memset (raygun/render/render_system.cpp:239), so the bytes agree. */
typedef struct rgrt_ubo {
    float viewInverse[16];
    float projInverse[16];
    float clearColor[3];
    int32_t numSamples;
    float lightDir[3];
    int32_t maxRecursions;
    float time;
    uint32_t showAlpha;
    float pad0;
    float pad1;
    float fadeColor[4];
} rgrt_ubo;

/* Per-instance offsets into the shared buffers, 12 bytes, in elements.
 * Replaces resources/shaders/instance_offset_table.def:1-3 and
 * raygun/render/acceleration_structure.hpp:75-78. */
typedef struct rgrt_instance_offsets {
    uint32_t vertexBufferOffset;
    uint32_t indexBufferOffset;
    uint32_t materialBufferOffset;
} rgrt_instance_offsets;

/* BLAS handle: index of a bottom-level structure inside a context. */
typedef uint32_t rgrt_blas;

/* Geometry range of one BLAS, in the packed vertex/index buffers.
 * Mirrors vk::AccelerationStructureBuildRangeInfoKHR as filled by
 * raygun/render/acceleration_structure.cpp:167-170: indices are mesh-local,
 * vertex k of a triangle is vertices[firstVertex + indices[...]]. */
typedef struct rgrt_blas_desc {
    uint32_t firstVertex;      /* vertexBufferRef.offsetInElements()   */
    uint32_t indexOffsetBytes; /* indexBufferRef.offsetInBytes          */
    uint32_t numFaces;         /* mesh.numFaces()                       */
    uint32_t maxVertex;        /* mesh.vertices.size()                  */
} rgrt_blas_desc;

/* One TLAS instance. Replaces vk::AccelerationStructureInstanceKHR as filled
 * by instanceFromEntity (raygun/render/acceleration_structure.cpp:34-51)
 * plus the matching InstanceOffsetTableEntry (:79-82). `transform` is the
 * object-to-world affine map as a 3x4 row-major matrix
 * (transpose(globalTransform().toMat4()) rows 0..2). */
typedef struct rgrt_instance {
    float transform[12];
    uint32_t customIndex; /* gl_InstanceCustomIndexEXT: DFS order         */
    uint32_t mask;        /* low 8 bits used, reference uses 0xff          */
    uint32_t flags;       /* RGRT_INSTANCE_* (cull-disable is the only mode) */
    rgrt_blas blas;
    rgrt_instance_offsets offsets;
    uint32_t pad;
} rgrt_instance; /* 80 bytes */

#define RGRT_INSTANCE_TRIANGLE_CULL_DISABLE 0x1u

/* ------------------------------------------------------------------------ */
/* Ray queries (closest hit) — the traversal on its own                       */
/* ------------------------------------------------------------------------ */
typedef struct rgrt_ray {
    float origin[3];
    float tmin;
    float dir[3];
    float tmax;
} rgrt_ray; /* 32 bytes */

#define RGRT_NO_HIT 0xffffffffu

typedef struct rgrt_hit {
    float t;
    float u; /* barycentric weight of vertex 1 (attribs.x) */
    float v; /* barycentric weight of vertex 2 (attribs.y) */
    uint32_t instance; /* customIndex, RGRT_NO_HIT on a miss */
    uint32_t primitive; /* gl_PrimitiveID                    */
} rgrt_hit; /* 20 bytes */

/* ------------------------------------------------------------------------ */
/* Canonical BVH2 export (used to compare builds node by node)               */
/* ------------------------------------------------------------------------ */
/* Child reference: >= 0 internal node index; < 0 leaf, ~child = position of
 * the primitive in Morton order. An empty child has child == RGRT_EMPTY_CHILD
 * and an inverted box. */
#define RGRT_EMPTY_CHILD ((int32_t)0x80000000)

typedef struct rgrt_bvh_node {
    float lo[2][3];
    float hi[2][3];
    int32_t child[2];
    uint32_t parent; /* 0xffffffff for the root */
    uint32_t pad;
} rgrt_bvh_node; /* 64 bytes */

/* ------------------------------------------------------------------------ */
/* Context                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct rgrt_ctx rgrt_ctx;

typedef struct rgrt_config {
    int32_t device;       /* HIP device ordinal                           */
    uint32_t width;       /* render target size (may be changed later)    */
    uint32_t height;
    uint32_t flags;       /* RGRT_CFG_*                                    */
} rgrt_config;

#define RGRT_CFG_FP32_RADIANCE 0x1u /* also keep fp32 copies of the 3 raygen outputs */
#define RGRT_CFG_PRIMARY_AOV 0x2u   /* keep primary-ray (t,u,v,instance,primitive)    */
#define RGRT_CFG_TIMING 0x4u        /* record hipEvents around every stage and kernel */

/* Replaces the ray-tracing parts of VulkanContext() and Raytracer()
 * (raygun/vulkan_context.cpp:48-69, raygun/render/raytracer.cpp:38-52). */
rgrt_status rgrt_create(rgrt_ctx** out, const rgrt_config* cfg);
/* No counterpart in the reference (it renders one frame at a time; its
 * frames in flight are swapchain images over one scene). A second context on
 * the same device that shares the parent's geometry, materials and BLASes
 * (handles are the same in both) with its own TLAS, images, wavefront state
 * and stream: frames rendered alternately on the two overlap on the GPU.
 * Every call that changes the shared scene (upload_geometry,
 * upload_materials, update_vertices, build_blas, refit_blas) first waits
 * on the host for the work queued on every sharing context;
 * update_vertices_device and rebuild_blas are ordered against the sharing
 * contexts' streams on the GPU instead. New geometry drops every sharing
 * context's TLAS. Calls on sharing contexts must not run
 * concurrently from several threads while one changes the scene. The scene
 * is freed with the last context that holds it; contexts may be destroyed
 * in any order. */
rgrt_status rgrt_create_shared(rgrt_ctx* parent, const rgrt_config* cfg, rgrt_ctx** out);
void rgrt_destroy(rgrt_ctx* ctx);
const char* rgrt_last_error(const rgrt_ctx* ctx);
/* Stream all work of the context is ordered on (hipStream_t as void*). */
void* rgrt_stream(rgrt_ctx* ctx);
rgrt_status rgrt_synchronize(rgrt_ctx* ctx);

/* Replaces RenderSystem::setupModelBuffers / updateModelBuffers
 * (raygun/render/render_system.cpp:192-230, 270-330): the caller passes the
 * already concatenated vertex / index / material buffers. Invalidates every
 * BLAS built on earlier geometry. Device memory: the buffers plus a 48-byte
 * shading record per triangle (numIndices / 3), which BLAS builds fill. */
rgrt_status rgrt_upload_geometry(rgrt_ctx* ctx, const rgrt_vertex* vertices, uint64_t numVertices, const uint32_t* indices,
                                 uint64_t numIndices);
/* At most 2^28 materials (a suspended closest-hit frame keeps the material
 * index in 28 bits); more returns RGRT_ERR_INVALID_ARG. */
rgrt_status rgrt_upload_materials(rgrt_ctx* ctx, const rgrt_material* materials, uint32_t numMaterials);
/* Rewrites vertex positions/normals of a range in place (deforming meshes);
 * follow with rgrt_refit_blas or rgrt_build_blas when positions change.
 * Normals and material indices take effect in the next frame without one
 * (the per-triangle shading records of every BLAS over the range are
 * gathered again on the context's stream). */
rgrt_status rgrt_update_vertices(rgrt_ctx* ctx, uint64_t firstVertex, const rgrt_vertex* vertices, uint64_t count);
/* No counterpart in the reference. The same from device memory (vertices
 * animated on the GPU): ordered on the context's stream, returns at once;
 * contexts sharing the scene order their later work after it. */
rgrt_status rgrt_update_vertices_device(rgrt_ctx* ctx, uint64_t firstVertex, const rgrt_vertex* deviceVertices,
                                        uint64_t count);

/* Replaces BottomLevelAS(cmd, mesh) (raygun/render/acceleration_structure.cpp:140-194).
 * LBVH build on the device; blocks until the build is complete, as the
 * reference waits on a fence (raygun/render/raytracer.cpp:71-73). */
rgrt_status rgrt_build_blas(rgrt_ctx* ctx, const rgrt_blas_desc* desc, rgrt_blas* out);
/* Refit of an existing BLAS after rgrt_update_vertices (no counterpart in
 * the reference, which never updates a BLAS): same topology, new boxes. */
rgrt_status rgrt_refit_blas(rgrt_ctx* ctx, rgrt_blas blas);
/* No counterpart in the reference. Full LBVH rebuild of an existing BLAS
 * from its current vertices (deforming meshes), into its own buffers: the
 * handle stays valid. Ordered on the context's stream and returns at once
 * (no host wait); contexts sharing the scene order their later work after
 * it. Rebuild the TLAS (or refit it) afterwards for the new root box. */
rgrt_status rgrt_rebuild_blas(rgrt_ctx* ctx, rgrt_blas blas);
/* No counterpart in the reference: for callers that cannot allocate (code
 * under hipGraph capture, tight serving loops). rgrt_rebuild_blas with the
 * builder's working arrays in a fixed per-GPU device array (48 MiB, held
 * until the process ends) instead of memory the context allocates: it calls
 * no allocation routine. At most 262144 triangles (RGRT_ERR_INVALID_ARG
 * beyond). Same tree as rgrt_rebuild_blas. Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other
 * contexts) race and must be ordered by the caller. */
rgrt_status rgrt_rebuild_blas_noalloc(rgrt_ctx* ctx, rgrt_blas blas);
/* Number of internal nodes / leaves and the root box of a BLAS. */
rgrt_status rgrt_blas_info(rgrt_ctx* ctx, rgrt_blas blas, uint32_t* numNodes, uint32_t* numLeaves, float rootLo[3],
                           float rootHi[3]);
/* Canonical BVH2 of a BLAS: numFaces-1 nodes (1 if numFaces <= 1) and the
 * original primitive index of every leaf in Morton order. */
rgrt_status rgrt_export_blas(rgrt_ctx* ctx, rgrt_blas blas, rgrt_bvh_node* nodes, uint32_t nodeCapacity,
                             uint32_t* leafPrimitive, uint32_t leafCapacity);
/* The 4-wide tree collapsed from the binary tree breadth-first. With
 * compressed == 0: 128-byte nodes (per-axis lo/hi of 4 children, 4 child
 * refs, 16 bytes padding). With compressed != 0: the 64-byte form the
 * traversal walks (node box min corner as 3 floats, a power-of-two exponent
 * byte per axis and a pad byte, the children's lo/hi as 8-bit grid offsets
 * per axis rounded outwards, 8 bytes padding, 4 child refs). Child refs are
 * >= 0 for a wide node, ~k for leaf k, INT32_MIN for an empty slot. With
 * nodes == NULL only *numNodes is written. */
rgrt_status rgrt_export_blas_wide(rgrt_ctx* ctx, rgrt_blas blas, int32_t compressed, void* nodes, uint32_t capacity,
                                  uint32_t* numNodes);
rgrt_status rgrt_export_tlas_wide(rgrt_ctx* ctx, int32_t compressed, void* nodes, uint32_t capacity, uint32_t* numNodes);

/* Replaces TopLevelAS(cmd, scene) + instanceOffsetTable()
 * (raygun/render/acceleration_structure.cpp:55-138). The caller keeps the
 * scene DFS (:63-85) so customIndex is the DFS order. Asynchronous on the
 * context stream. */
rgrt_status rgrt_build_tlas(rgrt_ctx* ctx, const rgrt_instance* instances, uint32_t numInstances);
/* Per-frame update of the TLAS when only instance transforms changed (the
 * scene DFS of raygun/render/acceleration_structure.cpp:63-85 then yields
 * the same instances in the same order, entity.cpp:180-199 moving them):
 * same count and BLAS per position as the last rgrt_build_tlas, new
 * transforms / masks / offsets. Keeps the hierarchy and the leaf order,
 * recomputes every box bottom-up, then the wide and compressed nodes. No
 * counterpart in the reference, which rebuilds the TLAS every frame
 * (raygun/render/raytracer.cpp:76-85). RGRT_ERR_INVALID_ARG if the count or
 * a BLAS differs. Asynchronous on the context stream. */
rgrt_status rgrt_refit_tlas(rgrt_ctx* ctx, const rgrt_instance* instances, uint32_t numInstances);
rgrt_status rgrt_export_tlas(rgrt_ctx* ctx, rgrt_bvh_node* nodes, uint32_t nodeCapacity, uint32_t* leafInstance,
                             uint32_t leafCapacity, uint32_t* numNodes);

/* Replaces RenderSystem::updateUniformBuffer (raygun/render/render_system.cpp:246-268). */
rgrt_status rgrt_set_frame(rgrt_ctx* ctx, const rgrt_ubo* ubo);
rgrt_status rgrt_resize(rgrt_ctx* ctx, uint32_t width, uint32_t height);

/* Closest-hit queries over the current TLAS, host arrays in and out
 * (synchronous). cullMask follows traceRayEXT's (instance.mask & cullMask). */
rgrt_status rgrt_trace_rays(rgrt_ctx* ctx, const rgrt_ray* rays, uint32_t numRays, uint32_t cullMask, rgrt_hit* hits);
/* Same on device pointers, asynchronous on the context stream. */
rgrt_status rgrt_trace_rays_device(rgrt_ctx* ctx, const rgrt_ray* rays, uint32_t numRays, uint32_t cullMask, rgrt_hit* hits);

/* Replaces Raytracer::doRaytracing (raygun/render/raytracer.cpp:87-147):
 * raygen + Whitted recursion (traceRaysKHR), rough prepare, 10x (H,V) blur,
 * postprocess, and FXAA followed by swap(base, final) when useFxaa != 0.
 * Asynchronous on the context stream. */
rgrt_status rgrt_render(rgrt_ctx* ctx, int32_t useFxaa);
/* Sampling of rgrt_trace. RGRT_SAMPLING_WHITTED is the reference's
 * deterministic tracer (the SSAA offsets of raygen.h:37-67). With
 * RGRT_SAMPLING_STOCHASTIC (no counterpart in the reference) every sample
 * jitters its pixel position uniformly and every reflection ray of a material
 * with roughness > 0 is spread around the mirror direction: dir =
 * normalize(reflect + roughness * s), s uniform in [-1, 1]^3, kept only above
 * the surface. The random numbers are integer hashes of (seed, pixel, sample,
 * hit position), so a frame is reproducible. BASE then holds the mean of the
 * per-sample composite mix(base, rough, reflectContribution) with alpha 0, so
 * the post passes reduce to clamp + luma (+ FXAA): the glossy spread replaces
 * the screen-space roughness blur. FEATURES is written for the denoiser. */
typedef struct rgrt_sampling {
    uint32_t mode; /* RGRT_SAMPLING_* */
    uint32_t seed;
    uint32_t pad[2];
} rgrt_sampling;
#define RGRT_SAMPLING_WHITTED 0u
#define RGRT_SAMPLING_STOCHASTIC 1u
rgrt_status rgrt_set_sampling(rgrt_ctx* ctx, const rgrt_sampling* sampling);

/* ------------------------------------------------------------------------ */
/* Learned denoiser (no counterpart in the reference: it stands in for the   */
/* screen-space roughness blur of raygun/compute/rough_blur.h and for high   */
/* sample counts). A 3-level U-Net (16/32/64 channels, 3x3 convolutions,     */
/* skip concatenation; 11 input channels: noisy RGB, albedo, normal,         */
/* log depth, roughness) in BF16 on the matrix cores with FP32 accumulation, */
/* trained on the device by its own forward and backward kernels (Adam).     */
/* It denoises a frame of RGRT_SAMPLING_STOCHASTIC in place (BASE.rgb), then */
/* rgrt_postprocess finishes it; 16x16 tiles where the network's change      */
/* exceeds fallbackA * (noise variance of the tile) + fallbackB keep the     */
/* traced frame. A denoiser belongs to one context and runs on its stream.   */
/* ------------------------------------------------------------------------ */
typedef struct rgrt_denoiser rgrt_denoiser;
rgrt_status rgrt_denoiser_create(rgrt_ctx* ctx, uint32_t seed, rgrt_denoiser** out);
void rgrt_denoiser_destroy(rgrt_denoiser* den);
/* fp32 parameters: per layer the weights [9][cin][cout] then the biases */
rgrt_status rgrt_denoiser_num_params(rgrt_denoiser* den, uint64_t* count);
rgrt_status rgrt_denoiser_get_params(rgrt_denoiser* den, float* params, uint64_t count);
rgrt_status rgrt_denoiser_set_params(rgrt_denoiser* den, const float* params, uint64_t count);
/* One forward + backward pass on host data: x0 = n*h*w*16 bf16 bit patterns
 * (the packed input), target = n*h*w*3 floats; writes the loss and, when the
 * pointers are given, every parameter's gradient and the network output
 * (n*h*w*16 bf16, the correction added to the noisy RGB). Synchronous. */
rgrt_status rgrt_denoiser_loss_and_grads(rgrt_denoiser* den, const uint16_t* x0, const float* target, uint32_t n,
                                         uint32_t height, uint32_t width, float* loss, float* grads, uint16_t* output);
/* The network's forward pass alone on host inputs: x0 = n*h*w*16 bf16 bit
 * patterns (the packed input), output = n*h*w*16 bf16 (the correction added
 * to the noisy RGB; channels 3..15 zero). h and w multiples of 4, at most
 * 2^22 pixels. Synchronous. No counterpart in the reference. */
rgrt_status rgrt_denoiser_infer(rgrt_denoiser* den, const uint16_t* x0, uint32_t n, uint32_t height, uint32_t width,
                                uint16_t* output);
/* Training set in device memory: `count` frames of height x width. add_input
 * packs the context's current stochastic frame into a slot; add_target
 * (accumulate != 0: adds) weight * BASE.rgb of the current frame to the slot's
 * target. */
rgrt_status rgrt_denoiser_dataset_reserve(rgrt_denoiser* den, uint32_t count, uint32_t height, uint32_t width);
rgrt_status rgrt_denoiser_dataset_add_input(rgrt_denoiser* den, uint32_t slot);
rgrt_status rgrt_denoiser_dataset_add_target(rgrt_denoiser* den, uint32_t slot, float weight, int32_t accumulate);
/* Copies slots [first, first + count) of the training set to host memory:
 * x0 = count*h*w*16 bf16 bit patterns (the packed inputs, as rgrt_denoise
 * packs a frame), target = count*h*w*3 floats. height and width must be the
 * set's (RGRT_ERR_INVALID_ARG otherwise). Either pointer may be null.
 * Synchronous. */
rgrt_status rgrt_denoiser_dataset_read(rgrt_denoiser* den, uint32_t first, uint32_t count, uint32_t height, uint32_t width,
                                       uint16_t* x0, float* target);
/* `steps` Adam steps (learning rate lr) on random batches of `batch` of the
 * first `used` slots; *loss = mean loss of the last steps. Synchronous. */
rgrt_status rgrt_denoiser_train(rgrt_denoiser* den, uint32_t steps, uint32_t batch, float lr, uint32_t seed, uint32_t used,
                                float* loss);
/* Denoises the context's current frame (asynchronous on the context stream).
 * With fallbackTiles != NULL waits and reports how many of the *tiles fell
 * back to the traced frame. RGRT_ERR_NOT_READY until the network has
 * parameters from rgrt_denoiser_set_params or rgrt_denoiser_train (a fresh
 * denoiser holds only its random initialisation). */
rgrt_status rgrt_denoise(rgrt_denoiser* den, float fallbackA, float fallbackB, uint32_t* fallbackTiles, uint32_t* tiles);

/* rgrt_frame flags */
#define RGRT_FRAME_FXAA 0x1u       /* FXAA + swap(base, final), as doRaytracing with m_useFXAA */
#define RGRT_FRAME_REFIT_TLAS 0x2u /* refit the TLAS (rgrt_refit_tlas) instead of building it */
#define RGRT_FRAME_GRAPH 0x4u      /* the whole frame as one hipGraph launch                   */

/* One frame of RenderSystem::render (raygun/render/render_system.cpp:88-100):
 * updateUniformBuffer (rgrt_set_frame), setupTopLevelAS (rgrt_build_tlas, or
 * rgrt_refit_tlas with RGRT_FRAME_REFIT_TLAS) and doRaytracing (rgrt_render).
 * With RGRT_FRAME_GRAPH the instance upload, the TLAS build, the wavefront
 * passes, the drain and the post passes are captured into one hipGraph and
 * launched at once. The graph is captured again every frame and the
 * executable graph updated in place (hipGraphExecUpdate), so frame
 * constants, instances and the render size may change between frames. In
 * graph form the pass count is fixed (min(wavefrontPasses, recursion bound))
 * and the drain always runs; the images equal the eager form's. With
 * rgrt_tuning.frameBands > 1 the graph form is several graphs instead: the
 * instance upload and TLAS, then each band's trace on a stream of its own,
 * then (after all bands) the post passes, all ordered after the context
 * stream's earlier work and before its later work; same images. The staging
 * of a frame's instances waits until the previous frame's upload has run
 * (one frame in flight, like the reference's per-frame fence).
 * Asynchronous on the context stream. */
rgrt_status rgrt_frame(rgrt_ctx* ctx, const rgrt_ubo* ubo, const rgrt_instance* instances, uint32_t numInstances,
                       uint32_t flags);
/* ------------------------------------------------------------------------ */
/* Batches of scenes (no counterpart in the reference, which renders one    */
/* scene): the frames of several contexts — scenes of any size and render   */
/* resolution on one device — submitted together so the GPU runs the        */
/* scenes' kernels side by side. Each frame is what                          */
/* rgrt_frame(..., RGRT_FRAME_GRAPH | flags) does for its context, captured  */
/* again every call and updated in place. Two submission modes:             */
/*  RGRT_BATCH_GRAPH_PER_SCENE (default): each scene's frame is one hipGraph */
/*   launched on its own context's stream after that stream's earlier work   */
/*   (vertex updates, BLAS rebuilds, the scene's previous frame) and nothing */
/*   else, so a scene's next frame may start while slower scenes still run. */
/*  RGRT_BATCH_ONE_GRAPH: one hipGraph with a branch per scene, launched on  */
/*   the first context's stream after the work queued on every scene's      */
/*   stream; work issued later on any of the contexts runs after the batch. */
/*  RGRT_BATCH_PERSISTENT: one hipGraph as in RGRT_BATCH_ONE_GRAPH, but the  */
/*   scenes are traced together: each wavefront pass of every scene is one  */
/*   traversal launch and one shading launch over all scenes' ray queues,   */
/*   and one drain launch finishes every scene's recursion tail (waves move */
/*   on to the next scene as they run out of work). The instance uploads,   */
/*   TLAS builds, ray generation and post passes stay one branch per scene.  */
/*   Launch shape and pass count: the first context's tuning. At most 64    */
/*   scenes. Same images as the other modes.                                 */
/* In both modes the batch's own stream (rgrt_batch_stream) is ordered after */
/* every scene of every batch frame submitted so far.                        */
/* How far the scenes overlap depends on the hardware queues the HIP runtime */
/* gives the process (GPU_MAX_HW_QUEUES, 4 by default): streams beyond that  */
/* share queues.                                                              */
/* ------------------------------------------------------------------------ */
typedef struct rgrt_batch rgrt_batch;
typedef struct rgrt_frame_desc {
    const rgrt_ubo* ubo;
    const rgrt_instance* instances;
    uint32_t numInstances;
    uint32_t flags; /* RGRT_FRAME_FXAA | RGRT_FRAME_REFIT_TLAS */
} rgrt_frame_desc;
rgrt_status rgrt_batch_create(rgrt_ctx* const* ctxs, uint32_t count, rgrt_batch** out);
void rgrt_batch_destroy(rgrt_batch* batch);
const char* rgrt_batch_last_error(const rgrt_batch* batch);
#define RGRT_BATCH_GRAPH_PER_SCENE 0u
#define RGRT_BATCH_ONE_GRAPH 1u
#define RGRT_BATCH_PERSISTENT 2u
rgrt_status rgrt_batch_set_mode(rgrt_batch* batch, uint32_t mode);
/* descs[i] is the frame of ctxs[i]. Asynchronous. */
rgrt_status rgrt_batch_frame(rgrt_batch* batch, const rgrt_frame_desc* descs);
/* The batch's stream (hipStream_t as void*): waits for each batch frame. */
void* rgrt_batch_stream(rgrt_batch* batch);

/* The traceRaysKHR part alone (writes base, normal, rough). */
rgrt_status rgrt_trace(rgrt_ctx* ctx);
/* The compute post part alone (reads base, normal, rough). */
rgrt_status rgrt_postprocess(rgrt_ctx* ctx, int32_t useFxaa);

/* Images, as in raygun/render/raytracer.cpp:173-195. FXAA writes the base
 * image and the reference then swaps base and final (:137-140); the same
 * swap happens here, so after a render with FXAA, FINAL holds the
 * anti-aliased result and BASE the composite it was computed from. */
typedef enum rgrt_image {
    RGRT_IMAGE_FINAL = 0,        /* rgba16f */
    RGRT_IMAGE_BASE = 1,         /* rgba16f */
    RGRT_IMAGE_NORMAL = 2,       /* rgba16f */
    RGRT_IMAGE_ROUGH = 3,        /* rgba16f */
    RGRT_IMAGE_ROUGH_TRANSITIONS = 4, /* r8 snorm, 1 byte per pixel */
    RGRT_IMAGE_ROUGH_COLORS_A = 5,   /* rgba16f */
    RGRT_IMAGE_ROUGH_COLORS_B = 6,   /* rgba16f */
    RGRT_IMAGE_BASE_F32 = 7,     /* rgba32f, RGRT_CFG_FP32_RADIANCE */
    RGRT_IMAGE_NORMAL_F32 = 8,   /* rgba32f, RGRT_CFG_FP32_RADIANCE */
    RGRT_IMAGE_ROUGH_F32 = 9,    /* rgba32f, RGRT_CFG_FP32_RADIANCE */
    RGRT_IMAGE_PRIMARY_HIT = 10, /* rgrt_hit per pixel, RGRT_CFG_PRIMARY_AOV */
    RGRT_IMAGE_FEATURES = 11     /* 2 x rgba16f per pixel, stochastic sampling:
                                    (albedo.rgb, roughness) of sample 0's primary
                                    hit, (mean primary t, variance of the mean
                                    luminance, 0, 0); zero on a miss */
} rgrt_image;

/* Copies an image to host memory (synchronous); `bytes` must equal its size. */
rgrt_status rgrt_read_image(rgrt_ctx* ctx, rgrt_image which, void* dst, uint64_t bytes);
/* Device pointer of an image (valid until the next resize / render swap). */
rgrt_status rgrt_image_device_ptr(rgrt_ctx* ctx, rgrt_image which, void** dptr, uint64_t* bytes);
/* 8-bit display formats for rgrt_read_image_rgba8 */
typedef enum rgrt_format {
    RGRT_FORMAT_RGBA8_UNORM = 0,
    RGRT_FORMAT_BGRA8_UNORM = 1,
    RGRT_FORMAT_RGBA8_SRGB = 2,
    RGRT_FORMAT_BGRA8_SRGB = 3
} rgrt_format;
/* Replaces the blit of the result image into the swapchain image
 * (raygun/render/render_system.cpp:102-147: nearest filter, same size, so a
 * per-pixel format conversion) followed by a readback: an rgba16f image
 * converted on the device to 4 bytes per pixel in the given format (clamp
 * to [0, 1], NaN to 0, sRGB encoding of RGB for the _SRGB formats, round to
 * nearest) and copied to host memory (synchronous). `bytes` must equal
 * width * height * 4. The reference's surface format is the driver's first
 * (raygun/vulkan_context.cpp:180), so which one it was is not fixed. */
rgrt_status rgrt_read_image_rgba8(rgrt_ctx* ctx, rgrt_image which, rgrt_format format, void* dst, uint64_t bytes);
/* Uploads base/normal/rough images (rgba16f) so the post passes can run on
 * given inputs. */
rgrt_status rgrt_write_image(rgrt_ctx* ctx, rgrt_image which, const void* src, uint64_t bytes);

/* ------------------------------------------------------------------------ */
/* Profiling: the reference profiler's stage names (raygun/profiler.def:6-10) */
/* ------------------------------------------------------------------------ */
typedef struct rgrt_stats {
    /* last frame, milliseconds, hipEvent pairs (RGRT_CFG_TIMING) */
    float asBuildMs;   /* TLAS build                     */
    float rtTotalMs;   /* trace + post                   */
    float rtOnlyMs;    /* trace (all wavefront passes)   */
    float postprocMs;  /* rough + composite + FXAA       */
    float roughMs;     /* rough prepare + blur           */
    /* counters of the last rgrt_trace / rgrt_render */
    uint64_t rays;        /* traversal queries with a non-zero cull mask   */
    uint64_t skyLookups;  /* cull-mask-0 rays (never traverse)             */
    uint64_t zeroDirRays; /* refraction rays at total internal reflection  */
    uint32_t waves;       /* wavefront iterations                          */
    uint32_t traceLaunches; /* launches of the traversal kernel            */
    float traceKernelMs;  /* summed duration of those launches (hipEvents) */
    float shadeKernelMs;  /* summed duration of the shading launches       */
    float drainKernelMs;  /* duration of the drain launch (0 if none)      */
    uint32_t drained;     /* 1 if the frame's tail ran in the drain kernel */
    float frameMs;        /* TLAS + trace + post of the last rgrt_frame (or of
                             build_tlas + render with RGRT_CFG_TIMING)       */
} rgrt_stats;

/* Launch-shape parameters (see rgrt_set_tuning). */
typedef struct rgrt_tuning {
    uint32_t wavefrontPasses;   /* wavefront passes before the persistent drain takes the tail
                                   (0 = drain everything, >= 64 = never drain)                */
    uint32_t traceBlocksPerCU;  /* persistent traversal grid, blocks of 256 per CU            */
    uint32_t shadeBlocksPerCU;  /* shading grid, blocks of 256 per CU                         */
    uint32_t drainBlocksPerCU;  /* drain grid, blocks of 256 per CU                           */
    uint32_t hostLag;           /* wavefront passes the GPU may run ahead of the host's check
                                   for an empty queue (1..16); the passes queued after the
                                   last ray are empty launches                               */
    uint32_t frameBands;        /* the frame split into this many strips of whole 8x8 tiles,
                                   each traced on a stream of its own, side by side (1..4;
                                   0 = 1), in graph frames and in eager traces (these then
                                   run every pass, with no host check for an empty queue;
                                   with RGRT_CFG_TIMING they stay in one piece); the images
                                   are those of the frame in one piece. The strips need
                                   hardware queues of their own (GPU_MAX_HW_QUEUES, 4 by
                                   default) to overlap                                        */
} rgrt_tuning;

/* Sets / reads the launch shape (defaults tuned for MI355X). Values outside
 * 1..32 blocks per CU, a hostLag outside 1..16 or more than 4 frameBands
 * are rejected. */
rgrt_status rgrt_set_tuning(rgrt_ctx* ctx, const rgrt_tuning* t);
rgrt_status rgrt_get_tuning(rgrt_ctx* ctx, rgrt_tuning* t);
/* No counterpart in the reference (its raygen loops over a pixel's samples,
 * raygen.h:84-112). With numSamples > 1 the samples of a pixel are
 * independent (each starts from a fresh payload), so up to maxGroups of
 * them are traced side by side in separate wavefront slots; each sample's
 * record is kept and the records are summed in sample order afterwards, so
 * the images equal those of one sample at a time bit for bit. 0 (default):
 * as many as fit 2^25 slots (~1.26 KB of device memory each); 1: one
 * sample at a time; at most 32. */
rgrt_status rgrt_set_sample_groups(rgrt_ctx* ctx, uint32_t maxGroups);

rgrt_status rgrt_get_stats(rgrt_ctx* ctx, rgrt_stats* out);
/* Rays traced in each wavefront pass of the last rgrt_trace / rgrt_render
 * (the recursion depth profile of the frame); *count = number of passes. */
rgrt_status rgrt_get_pass_counts(rgrt_ctx* ctx, uint32_t* counts, uint32_t capacity, uint32_t* count);

/* Library version string. */
const char* rgrt_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* RGRT_H */
