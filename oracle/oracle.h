/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of Raygun's ray-tracing path, used as the parity oracle for
 * the HIP core. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline, never as part of the product path.
 *
 * Record layouts are restated here on purpose (not included from the
 * product's header) so that a layout mistake in one does not hide in both.
 */
#ifndef RGRT_ORACLE_H
#define RGRT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Scene data: packed vertex (32 B), index (u32) and material (64 B) buffers. */
orc_scene* orc_scene_create(const void* vertices, uint64_t numVertices, const uint32_t* indices, uint64_t numIndices,
                            const void* materials, uint32_t numMaterials);
void orc_scene_destroy(orc_scene* s);

/* LBVH over one mesh range; returns the BLAS id (>= 0) or -1. */
int32_t orc_build_blas(orc_scene* s, uint32_t firstVertex, uint32_t indexOffsetBytes, uint32_t numFaces);
/* Re-reads vertex positions (after orc_update_vertices) and recomputes boxes. */
int32_t orc_refit_blas(orc_scene* s, int32_t blas);
/* Full LBVH rebuild of BLAS `blas` from the current vertices, keeping its id. */
int32_t orc_rebuild_blas(orc_scene* s, int32_t blas);
int32_t orc_update_vertices(orc_scene* s, uint64_t firstVertex, const void* vertices, uint64_t count);
uint32_t orc_blas_num_nodes(orc_scene* s, int32_t blas);
/* Canonical nodes (64 B each, rgrt_bvh_node layout) + leaf primitive ids. */
int32_t orc_export_blas(orc_scene* s, int32_t blas, void* nodes, uint32_t* leafPrim);
/* The 4-wide tree (breadth-first): 128-byte nodes, or with compressed != 0
 * the 64-byte quantised nodes the traversal reads. */
uint32_t orc_blas_num_wide(orc_scene* s, int32_t blas);
int32_t orc_export_blas_wide(orc_scene* s, int32_t blas, int32_t compressed, void* nodes);
uint32_t orc_tlas_num_wide(orc_scene* s);
int32_t orc_export_tlas_wide(orc_scene* s, int32_t compressed, void* nodes);
/* Hierarchy of later BLAS builds: 0 Karras (default), 1 PLOC. */
void orc_set_hierarchy(int32_t h);
/* Morton codes of the primitives in original order (for the sort check). */
int32_t orc_blas_morton(orc_scene* s, int32_t blas, uint32_t* codes);

/* Instances: 80-byte records (rgrt_instance layout). */
int32_t orc_build_tlas(orc_scene* s, const void* instances, uint32_t numInstances);
uint32_t orc_tlas_num_nodes(orc_scene* s);
int32_t orc_export_tlas(orc_scene* s, void* nodes, uint32_t* leafInstance);

/* Closest hit over the TLAS in FP32 (same LBVH, watertight test).
 * rays: 8 floats each (origin, tmin, dir, tmax); hits: 20 B records
 * (t, u, v, instance customIndex, primitive). tie[i] = 1 when another
 * primitive is hit within 1e-6 relative of the closest t. nodes/tris, if
 * not null, receive per-ray counts of visited BVH nodes / tested triangles. */
int32_t orc_trace_rays(orc_scene* s, const float* rays, uint32_t n, uint32_t cullMask, void* hits, uint8_t* tie,
                       uint32_t* nodes, uint32_t* tris, int32_t threads);
/* Same queries by brute force in FP64 over every triangle of every instance. */
int32_t orc_trace_rays_fp64(orc_scene* s, const float* rays, uint32_t n, uint32_t cullMask, double* t_out,
                            uint32_t* inst_out, uint32_t* prim_out, double* uv_out, int32_t threads);
/* Same queries by brute force in FP32: the instance mapping, ray setup and
 * watertight test of orc_trace_rays over every triangle (no tree), so the
 * closest t must equal orc_trace_rays' bit for bit. */
int32_t orc_trace_rays_brute32(orc_scene* s, const float* rays, uint32_t n, uint32_t cullMask, float* t_out,
                               uint32_t* inst_out, uint32_t* prim_out, int32_t threads);

/* Whole traceRaysKHR restatement: raygen + closest-hit/miss recursion.
 * ubo: 192-byte block. Outputs (any may be null):
 *   base/normal/rough16: W*H*4 half (uint16) — the stored rgba16f images
 *   base/normal/rough32: W*H*4 float — the values before the fp16 store
 *   primary: W*H hit records of the first sample's primary ray
 *   tie: W*H flags, 1 if any ray of the pixel had a near-tie closest hit
 * stats[0..2] = rays traversed, sky lookups (cull mask 0), zero-dir rays. */
int32_t orc_render(orc_scene* s, const void* ubo, uint32_t W, uint32_t H, uint16_t* base16, uint16_t* normal16,
                   uint16_t* rough16, float* base32, float* normal32, float* rough32, void* primary, uint8_t* tie,
                   uint64_t* stats, int32_t threads, int32_t rowBegin, int32_t rowEnd);

/* Traversal work of rendering the given rows: out[0..4] = rays traversed,
 * BVH nodes visited (instance leaves included), triangles tested, rays that
 * hit, sky lookups. */
int32_t orc_render_counts(orc_scene* s, const void* ubo, uint32_t W, uint32_t H, const int32_t* rows, int32_t nrows,
                          int32_t threads, uint64_t* out);

/* Compute post passes on given rgba16f images (as uint16 arrays):
 * rough_prepare, 10x (blur H, blur V), postprocess and, if useFxaa, FXAA
 * writing into `base` (the reference then swaps base and final).
 * transitions: W*H int8 (snorm8). All image arrays are in/out. */
int32_t orc_postprocess(const void* ubo, uint32_t W, uint32_t H, uint16_t* base, uint16_t* normal, uint16_t* rough,
                        uint16_t* final_, int8_t* transitions, uint16_t* colorsA, uint16_t* colorsB, int32_t useFxaa);

/* Helpers exposed for the closed-form tests. */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
void orc_sky(const float* dir, const float* lightDir, float* rgb);
uint32_t orc_morton3(float x, float y, float z); /* inputs in [0,1] */
/* The primary ray of raygen.h:78-90 for pixel (x, y) and SSAA sample. */
void orc_primary_ray(const void* ubo, uint32_t x, uint32_t y, uint32_t W, uint32_t H, int32_t sample, float* origin,
                     float* dir);

#ifdef __cplusplus
}
#endif

#endif
