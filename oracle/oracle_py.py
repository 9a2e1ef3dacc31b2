"""ctypes binding of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
only as the checker or as the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [P, u64, P, u64, P, u32]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_build_blas.restype = i32
        L.orc_build_blas.argtypes = [P, u32, u32, u32]
        L.orc_refit_blas.restype = i32
        L.orc_refit_blas.argtypes = [P, i32]
        L.orc_rebuild_blas.restype = i32
        L.orc_rebuild_blas.argtypes = [P, i32]
        L.orc_update_vertices.restype = i32
        L.orc_update_vertices.argtypes = [P, u64, P, u64]
        L.orc_blas_num_nodes.restype = u32
        L.orc_blas_num_nodes.argtypes = [P, i32]
        L.orc_export_blas.restype = i32
        L.orc_export_blas.argtypes = [P, i32, P, P]
        L.orc_blas_num_wide.restype = u32
        L.orc_blas_num_wide.argtypes = [P, i32]
        L.orc_export_blas_wide.restype = i32
        L.orc_export_blas_wide.argtypes = [P, i32, i32, P]
        L.orc_tlas_num_wide.restype = u32
        L.orc_tlas_num_wide.argtypes = [P]
        L.orc_export_tlas_wide.restype = i32
        L.orc_export_tlas_wide.argtypes = [P, i32, P]
        L.orc_set_hierarchy.restype = None
        L.orc_set_hierarchy.argtypes = [i32]
        L.orc_blas_morton.restype = i32
        L.orc_blas_morton.argtypes = [P, i32, P]
        L.orc_build_tlas.restype = i32
        L.orc_build_tlas.argtypes = [P, P, u32]
        L.orc_render_sampled.restype = i32
        L.orc_render_sampled.argtypes = [P, P, u32, u32, u32, u32, P, P, P, P, P, P, P, i32]
        L.orc_refit_tlas.restype = i32
        L.orc_refit_tlas.argtypes = [P, P, u32]
        L.orc_tlas_num_nodes.restype = u32
        L.orc_tlas_num_nodes.argtypes = [P]
        L.orc_export_tlas.restype = i32
        L.orc_export_tlas.argtypes = [P, P, P]
        L.orc_trace_rays.restype = i32
        L.orc_trace_rays.argtypes = [P, P, u32, u32, P, P, P, P, i32]
        L.orc_trace_rays_fp64.restype = i32
        L.orc_trace_rays_fp64.argtypes = [P, P, u32, u32, P, P, P, P, i32]
        L.orc_trace_rays_brute32.restype = i32
        L.orc_trace_rays_brute32.argtypes = [P, P, u32, u32, P, P, P, i32]
        L.orc_render.restype = i32
        L.orc_render.argtypes = [P, P, u32, u32, P, P, P, P, P, P, P, P, P, i32, i32, i32]
        L.orc_render_counts.restype = i32
        L.orc_render_counts.argtypes = [P, P, u32, u32, P, i32, i32, P]
        L.orc_postprocess.restype = i32
        L.orc_postprocess.argtypes = [P, u32, u32, P, P, P, P, P, P, P, i32]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_f16_to_f32.restype = ctypes.c_float
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_sky.argtypes = [P, P, P]
        L.orc_morton3.restype = u32
        L.orc_morton3.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_primary_ray.restype = None
        L.orc_primary_ray.argtypes = [P, u32, u32, u32, u32, i32, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


class OracleScene:
    """Scene buffers + LBVHs + instances on the CPU."""

    def __init__(self, vertices, indices, materials):
        from raygun_amd.records import MATERIAL, VERTEX
        self._v = np.ascontiguousarray(vertices, VERTEX)
        self._i = np.ascontiguousarray(indices, np.uint32)
        self._m = np.ascontiguousarray(materials, MATERIAL)
        self.h = lib().orc_scene_create(_p(self._v), len(self._v), _p(self._i), len(self._i), _p(self._m), len(self._m))
        self.blas_faces = []

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def build_blas(self, first_vertex, index_offset_bytes, num_faces):
        b = lib().orc_build_blas(self.h, first_vertex, index_offset_bytes, num_faces)
        if b < 0:
            raise ValueError("oracle BLAS build failed")
        self.blas_faces.append(num_faces)
        return b

    def refit_blas(self, b):
        assert lib().orc_refit_blas(self.h, b) == 0

    def rebuild_blas(self, b):
        """Full LBVH rebuild of BLAS b from the current vertices (same id)."""
        assert lib().orc_rebuild_blas(self.h, b) == 0

    def update_vertices(self, first, verts):
        from raygun_amd.records import VERTEX
        verts = np.ascontiguousarray(verts, VERTEX)
        assert lib().orc_update_vertices(self.h, first, _p(verts), len(verts)) == 0

    def export_blas(self, b):
        from raygun_amd.records import BVH_NODE
        n = lib().orc_blas_num_nodes(self.h, b)
        nodes = np.zeros(n, BVH_NODE)
        leaf = np.zeros(self.blas_faces[b], np.uint32)
        lib().orc_export_blas(self.h, b, _p(nodes), _p(leaf))
        return nodes, leaf

    def export_blas_wide(self, b, compressed=False):
        from raygun_amd.records import QWIDE_NODE, WIDE_NODE
        nodes = np.zeros(lib().orc_blas_num_wide(self.h, b), QWIDE_NODE if compressed else WIDE_NODE)
        lib().orc_export_blas_wide(self.h, b, int(compressed), _p(nodes))
        return nodes

    def export_tlas_wide(self, compressed=False):
        from raygun_amd.records import QWIDE_NODE, WIDE_NODE
        nodes = np.zeros(lib().orc_tlas_num_wide(self.h), QWIDE_NODE if compressed else WIDE_NODE)
        lib().orc_export_tlas_wide(self.h, int(compressed), _p(nodes))
        return nodes

    def blas_morton(self, b):
        codes = np.zeros(self.blas_faces[b], np.uint32)
        lib().orc_blas_morton(self.h, b, _p(codes))
        return codes

    def build_tlas(self, instances):
        from raygun_amd.records import INSTANCE
        self._inst = np.ascontiguousarray(instances, INSTANCE)
        assert lib().orc_build_tlas(self.h, _p(self._inst), len(self._inst)) == 0

    def refit_tlas(self, instances):
        from raygun_amd.records import INSTANCE
        inst = np.ascontiguousarray(instances, INSTANCE)
        assert lib().orc_refit_tlas(self.h, _p(inst), len(inst)) == 0
        self._inst = inst

    def export_tlas(self):
        from raygun_amd.records import BVH_NODE
        n = lib().orc_tlas_num_nodes(self.h)
        nodes = np.zeros(n, BVH_NODE)
        leaf = np.zeros(len(self._inst), np.uint32)
        lib().orc_export_tlas(self.h, _p(nodes), _p(leaf))
        return nodes, leaf

    def trace_rays(self, rays, cull_mask=0xFF, threads=None, counts=False):
        from raygun_amd.records import HIT, RAY
        rays = np.ascontiguousarray(rays, RAY)
        n = len(rays)
        hits = np.zeros(n, HIT)
        tie = np.zeros(n, np.uint8)
        nodes = np.zeros(n, np.uint32) if counts else None
        tris = np.zeros(n, np.uint32) if counts else None
        lib().orc_trace_rays(self.h, _p(rays), n, cull_mask, _p(hits), _p(tie), _p(nodes), _p(tris), threads or default_threads())
        if counts:
            return hits, tie.astype(bool), nodes, tris
        return hits, tie.astype(bool)

    def trace_rays_fp64(self, rays, cull_mask=0xFF, threads=None):
        from raygun_amd.records import RAY
        rays = np.ascontiguousarray(rays, RAY)
        n = len(rays)
        t = np.zeros(n)
        inst = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        uv = np.zeros((n, 2))
        lib().orc_trace_rays_fp64(self.h, _p(rays), n, cull_mask, _p(t), _p(inst), _p(prim), _p(uv), threads or default_threads())
        return t, inst, prim, uv

    def trace_rays_brute32(self, rays, cull_mask=0xFF, threads=None):
        """Closest hits by FP32 brute force (no tree): t, instance
        customIndex (NO_HIT on a miss), primitive."""
        from raygun_amd.records import RAY
        rays = np.ascontiguousarray(rays, RAY)
        n = len(rays)
        t = np.zeros(n, np.float32)
        inst = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        lib().orc_trace_rays_brute32(self.h, _p(rays), n, cull_mask, _p(t), _p(inst), _p(prim), threads or default_threads())
        return t, inst, prim

    def render(self, ubo, width, height, threads=None, rows=None):
        """Returns dict of images (fp16 as uint16 arrays HxWx4, fp32 HxWx4),
        primary hits, tie mask and ray counters."""
        from raygun_amd.records import HIT
        W, H = width, height
        out = {k: np.zeros((H, W, 4), np.uint16) for k in ("base16", "normal16", "rough16")}
        out.update({k: np.zeros((H, W, 4), np.float32) for k in ("base32", "normal32", "rough32")})
        out["primary"] = np.zeros((H, W), HIT)
        out["tie"] = np.zeros((H, W), np.uint8)
        stats = np.zeros(3, np.uint64)
        ubo = np.ascontiguousarray(ubo)
        r0, r1 = rows if rows else (0, H)
        rc = lib().orc_render(self.h, _p(ubo), W, H, _p(out["base16"]), _p(out["normal16"]), _p(out["rough16"]),
                              _p(out["base32"]), _p(out["normal32"]), _p(out["rough32"]), _p(out["primary"]),
                              _p(out["tie"]), _p(stats), threads or default_threads(), r0, r1)
        if rc != 0:
            raise RuntimeError("oracle render failed (no TLAS?)")
        out["rays"], out["sky"], out["zero_dir"] = (int(x) for x in stats)
        return out

    def render_sampled(self, ubo, width, height, sampling=1, seed=0, threads=None):
        """Render with the sampling mode (1: stochastic): base16 / normal16 /
        rough16 / features16 (H x W x 8) as uint16 fp16 bits, base32, tie."""
        W, H = width, height
        out = {k: np.zeros((H, W, 4), np.uint16) for k in ("base16", "normal16", "rough16")}
        out["base32"] = np.zeros((H, W, 4), np.float32)
        out["features16"] = np.zeros((H, W, 8), np.uint16)
        out["tie"] = np.zeros((H, W), np.uint8)
        stats = np.zeros(1, np.uint64)
        ubo = np.ascontiguousarray(ubo)
        rc = lib().orc_render_sampled(self.h, _p(ubo), W, H, sampling, seed, _p(out["base16"]), _p(out["normal16"]),
                                      _p(out["rough16"]), _p(out["base32"]), _p(out["features16"]), _p(out["tie"]),
                                      _p(stats), threads or default_threads())
        if rc != 0:
            raise RuntimeError("oracle render failed (no TLAS?)")
        out["rays"] = int(stats[0])
        return out

    def render_counts(self, ubo, width, height, rows, threads=None):
        """Traversal work of rendering `rows`: rays, nodes, tris, hits, sky."""
        rows = np.ascontiguousarray(rows, np.int32)
        out = np.zeros(7, np.uint64)
        ubo = np.ascontiguousarray(ubo)
        rc = lib().orc_render_counts(self.h, _p(ubo), width, height, _p(rows), len(rows), threads or default_threads(), _p(out))
        if rc != 0:
            raise RuntimeError("oracle render_counts failed")
        return dict(zip(("rays", "nodes", "tris", "hits", "sky", "inst_leaves", "blas_entries"), (int(x) for x in out)))


def postprocess(ubo, base16, normal16, rough16, use_fxaa=True):
    """rough_prepare, 10x (H, V) blur, postprocess, FXAA on copies of the inputs.

    Returns dict (names after the reference's swap): final (the FXAA result
    when use_fxaa, else the composite), base (the composite when use_fxaa,
    else the input base), normal, rough, transitions (int8), colorsA, colorsB."""
    H, W = base16.shape[:2]
    base = np.ascontiguousarray(base16, np.uint16).copy()
    normal = np.ascontiguousarray(normal16, np.uint16).copy()
    rough = np.ascontiguousarray(rough16, np.uint16).copy()
    final = np.zeros_like(base)
    trans = np.zeros((H, W), np.int8)
    A = np.zeros_like(base)
    B = np.zeros_like(base)
    ubo = np.ascontiguousarray(ubo)
    lib().orc_postprocess(_p(ubo), W, H, _p(base), _p(normal), _p(rough), _p(final), _p(trans), _p(A), _p(B), int(use_fxaa))
    if use_fxaa:
        # FXAA wrote the base image; the reference then swaps base and final
        # (raygun/render/raytracer.cpp:139), so "final" is the FXAA result.
        base, final = final, base
    return {"base": base, "final": final, "normal": normal, "rough": rough, "transitions": trans, "colorsA": A, "colorsB": B}


def f32_to_f16(x):
    return lib().orc_f32_to_f16(float(x))


def sky(direction, light_dir):
    d = np.asarray(direction, np.float32)
    l_ = np.asarray(light_dir, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_sky(_p(d), _p(l_), _p(out))
    return out


def morton3(x, y, z):
    return lib().orc_morton3(float(x), float(y), float(z))


def primary_ray(ubo, x, y, width, height, sample=0):
    ubo = np.ascontiguousarray(ubo)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().orc_primary_ray(_p(ubo), x, y, width, height, sample, _p(o), _p(d))
    return o, d
