"""ctypes binding of librgrt.so (the C ABI declared in include/rgrt.h).

The product path: every call goes to the HIP library; if the library is not
built or no GPU is present, the calls raise instead of falling back.
"""
import ctypes
import os

import numpy as np

from .records import BVH_NODE, HIT, INSTANCE, MATERIAL, RAY, STATS_FIELDS, TUNING, QWIDE_NODE, UBO, VERTEX, WIDE_NODE

HERE = os.path.dirname(os.path.abspath(__file__))
# RGRT_LIBRARY selects another build of the library (e.g. a tuning variant).
LIB_PATH = os.environ.get("RGRT_LIBRARY") or os.path.join(HERE, "librgrt.so")

CFG_FP32_RADIANCE = 0x1
CFG_PRIMARY_AOV = 0x2
CFG_TIMING = 0x4
FRAME_FXAA = 0x1
FRAME_REFIT_TLAS = 0x2
FRAME_GRAPH = 0x4
FORMAT_RGBA8_UNORM = 0
FORMAT_BGRA8_UNORM = 1
FORMAT_RGBA8_SRGB = 2
FORMAT_BGRA8_SRGB = 3

IMAGE_FINAL, IMAGE_BASE, IMAGE_NORMAL, IMAGE_ROUGH, IMAGE_ROUGH_TRANSITIONS, IMAGE_ROUGH_COLORS_A, IMAGE_ROUGH_COLORS_B, \
    IMAGE_BASE_F32, IMAGE_NORMAL_F32, IMAGE_ROUGH_F32, IMAGE_PRIMARY_HIT, IMAGE_FEATURES = range(12)
SAMPLING_WHITTED, SAMPLING_STOCHASTIC = 0, 1
BATCH_GRAPH_PER_SCENE, BATCH_ONE_GRAPH, BATCH_PERSISTENT = 0, 1, 2

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OUT_OF_MEMORY", 3: "HIP", 4: "NOT_READY", 5: "NO_DEVICE"}

# Every symbol include/rgrt.h declares.
EXPORTS = [
    "rgrt_create", "rgrt_create_shared", "rgrt_destroy", "rgrt_last_error", "rgrt_stream", "rgrt_synchronize",
    "rgrt_upload_geometry", "rgrt_upload_materials", "rgrt_update_vertices", "rgrt_update_vertices_device",
    "rgrt_build_blas", "rgrt_refit_blas", "rgrt_rebuild_blas", "rgrt_rebuild_blas_noalloc", "rgrt_blas_info", "rgrt_export_blas",
    "rgrt_build_tlas", "rgrt_refit_tlas", "rgrt_export_tlas", "rgrt_set_frame", "rgrt_resize",
    "rgrt_trace_rays", "rgrt_trace_rays_device", "rgrt_render", "rgrt_frame", "rgrt_trace", "rgrt_postprocess",
    "rgrt_read_image", "rgrt_read_image_rgba8", "rgrt_image_device_ptr", "rgrt_write_image", "rgrt_get_stats", "rgrt_get_pass_counts",
    "rgrt_set_tuning", "rgrt_get_tuning", "rgrt_set_sample_groups", "rgrt_set_sampling", "rgrt_denoiser_create", "rgrt_denoiser_destroy",
    "rgrt_denoiser_num_params", "rgrt_denoiser_get_params", "rgrt_denoiser_set_params", "rgrt_denoiser_loss_and_grads",
    "rgrt_denoiser_infer", "rgrt_denoiser_dataset_reserve", "rgrt_denoiser_dataset_add_input", "rgrt_denoiser_dataset_add_target", "rgrt_denoiser_dataset_read",
    "rgrt_denoiser_train", "rgrt_denoise", "rgrt_batch_create", "rgrt_batch_destroy", "rgrt_batch_last_error", "rgrt_batch_set_mode", "rgrt_batch_stream",
    "rgrt_batch_frame", "rgrt_export_blas_wide", "rgrt_export_tlas_wide", "rgrt_version",
]


class RgrtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"rgrt status {status} ({STATUS.get(status, '?')}): {msg}")
        self.status = status


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class _BlasDesc(ctypes.Structure):
    _fields_ = [("firstVertex", ctypes.c_uint32), ("indexOffsetBytes", ctypes.c_uint32), ("numFaces", ctypes.c_uint32),
                ("maxVertex", ctypes.c_uint32)]


STATS = np.dtype(STATS_FIELDS)
assert STATS.itemsize == 80

_lib = None


def lib():
    """Loads librgrt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RgrtError(5, f"{LIB_PATH} is not built; run raygun_amd/build.py or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
        sig = {
            "rgrt_create": (i32, [ctypes.POINTER(P), ctypes.POINTER(_Config)]),
            "rgrt_create_shared": (i32, [P, ctypes.POINTER(_Config), ctypes.POINTER(P)]),
            "rgrt_destroy": (None, [P]),
            "rgrt_last_error": (ctypes.c_char_p, [P]),
            "rgrt_stream": (P, [P]),
            "rgrt_synchronize": (i32, [P]),
            "rgrt_upload_geometry": (i32, [P, P, u64, P, u64]),
            "rgrt_upload_materials": (i32, [P, P, u32]),
            "rgrt_update_vertices": (i32, [P, u64, P, u64]),
            "rgrt_update_vertices_device": (i32, [P, u64, P, u64]),
            "rgrt_rebuild_blas": (i32, [P, u32]),
            "rgrt_rebuild_blas_noalloc": (i32, [P, u32]),
            "rgrt_build_blas": (i32, [P, ctypes.POINTER(_BlasDesc), ctypes.POINTER(u32)]),
            "rgrt_refit_blas": (i32, [P, u32]),
            "rgrt_blas_info": (i32, [P, u32, ctypes.POINTER(u32), ctypes.POINTER(u32), P, P]),
            "rgrt_export_blas": (i32, [P, u32, P, u32, P, u32]),
            "rgrt_build_tlas": (i32, [P, P, u32]),
            "rgrt_refit_tlas": (i32, [P, P, u32]),
            "rgrt_export_tlas": (i32, [P, P, u32, P, u32, ctypes.POINTER(u32)]),
            "rgrt_set_frame": (i32, [P, P]),
            "rgrt_resize": (i32, [P, u32, u32]),
            "rgrt_trace_rays": (i32, [P, P, u32, u32, P]),
            "rgrt_trace_rays_device": (i32, [P, P, u32, u32, P]),
            "rgrt_render": (i32, [P, i32]),
            "rgrt_frame": (i32, [P, P, P, u32, u32]),
            "rgrt_trace": (i32, [P]),
            "rgrt_postprocess": (i32, [P, i32]),
            "rgrt_read_image": (i32, [P, i32, P, u64]),
            "rgrt_read_image_rgba8": (i32, [P, i32, i32, P, u64]),
            "rgrt_image_device_ptr": (i32, [P, i32, ctypes.POINTER(P), ctypes.POINTER(u64)]),
            "rgrt_write_image": (i32, [P, i32, P, u64]),
            "rgrt_get_stats": (i32, [P, P]),
            "rgrt_get_pass_counts": (i32, [P, P, u32, ctypes.POINTER(u32)]),
            "rgrt_set_tuning": (i32, [P, P]),
            "rgrt_set_sample_groups": (i32, [P, u32]),
            "rgrt_set_sampling": (i32, [P, P]),
            "rgrt_export_blas_wide": (i32, [P, u32, i32, P, u32, ctypes.POINTER(u32)]),
            "rgrt_export_tlas_wide": (i32, [P, i32, P, u32, ctypes.POINTER(u32)]),
            "rgrt_get_tuning": (i32, [P, P]),
            "rgrt_version": (ctypes.c_char_p, []),
            "rgrt_denoiser_create": (i32, [P, u32, ctypes.POINTER(P)]),
            "rgrt_denoiser_destroy": (None, [P]),
            "rgrt_denoiser_num_params": (i32, [P, ctypes.POINTER(u64)]),
            "rgrt_denoiser_get_params": (i32, [P, P, u64]),
            "rgrt_denoiser_set_params": (i32, [P, P, u64]),
            "rgrt_denoiser_loss_and_grads": (i32, [P, P, P, u32, u32, u32, ctypes.POINTER(ctypes.c_float), P, P]),
            "rgrt_denoiser_infer": (i32, [P, P, u32, u32, u32, P]),
            "rgrt_denoiser_dataset_reserve": (i32, [P, u32, u32, u32]),
            "rgrt_denoiser_dataset_add_input": (i32, [P, u32]),
            "rgrt_denoiser_dataset_add_target": (i32, [P, u32, ctypes.c_float, i32]),
            "rgrt_denoiser_dataset_read": (i32, [P, u32, u32, u32, u32, P, P]),
            "rgrt_denoiser_train": (i32, [P, u32, u32, ctypes.c_float, u32, u32, ctypes.POINTER(ctypes.c_float)]),
            "rgrt_denoise": (i32, [P, ctypes.c_float, ctypes.c_float, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
            "rgrt_batch_create": (i32, [P, u32, ctypes.POINTER(P)]),
            "rgrt_batch_destroy": (None, [P]),
            "rgrt_batch_last_error": (ctypes.c_char_p, [P]),
            "rgrt_batch_frame": (i32, [P, P]),
            "rgrt_batch_set_mode": (i32, [P, u32]),
            "rgrt_batch_stream": (P, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One rgrt context: device buffers, BLASes, the TLAS and the images."""

    def __init__(self, width=0, height=0, flags=0, device=0, share=None):
        """share: another Context whose geometry, materials and BLASes this
        one uses too (rgrt_create_shared); its TLAS, images and stream are
        its own, so frames on the two overlap on the GPU."""
        self._L = lib()
        self.h = ctypes.c_void_p()
        cfg = _Config(device, width, height, flags)
        if share is None:
            st = self._L.rgrt_create(ctypes.byref(self.h), ctypes.byref(cfg))
            self._blas_faces = []
        else:
            st = self._L.rgrt_create_shared(share.h, ctypes.byref(cfg), ctypes.byref(self.h))
            self._blas_faces = share._blas_faces
        if st != 0:
            raise RgrtError(st, "rgrt_create_shared failed" if share is not None else "rgrt_create failed")
        self.width, self.height, self.flags, self.device = width, height, flags, device

    def shared(self, width=None, height=None, flags=None):
        """A new context sharing this one's scene."""
        return Context(self.width if width is None else width, self.height if height is None else height,
                       self.flags if flags is None else flags, self.device, share=self)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self._L.rgrt_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ok(self, st):
        if st != 0:
            raise RgrtError(st, self._L.rgrt_last_error(self.h).decode())

    # -- scene data ---------------------------------------------------------
    def upload_geometry(self, vertices, indices):
        self._v = np.ascontiguousarray(vertices, VERTEX)
        self._i = np.ascontiguousarray(indices, np.uint32)
        self._ok(self._L.rgrt_upload_geometry(self.h, _p(self._v), len(self._v), _p(self._i), len(self._i)))

    def upload_materials(self, materials):
        m = np.ascontiguousarray(materials, MATERIAL)
        self._ok(self._L.rgrt_upload_materials(self.h, _p(m), len(m)))

    def update_vertices(self, first, vertices):
        v = np.ascontiguousarray(vertices, VERTEX)
        self._ok(self._L.rgrt_update_vertices(self.h, first, _p(v), len(v)))

    def build_blas(self, first_vertex, index_offset_bytes, num_faces, max_vertex):
        d = _BlasDesc(first_vertex, index_offset_bytes, num_faces, max_vertex)
        out = ctypes.c_uint32()
        self._ok(self._L.rgrt_build_blas(self.h, ctypes.byref(d), ctypes.byref(out)))
        while len(self._blas_faces) <= out.value:
            self._blas_faces.append(0)
        self._blas_faces[out.value] = num_faces
        return out.value

    def refit_blas(self, b):
        self._ok(self._L.rgrt_refit_blas(self.h, b))

    def update_vertices_device(self, first, device_ptr, count):
        """count vertices (32 B each) from device memory at device_ptr (an
        int address, e.g. a torch tensor's data_ptr()); stream-ordered."""
        self._ok(self._L.rgrt_update_vertices_device(self.h, first, ctypes.c_void_p(device_ptr), count))

    def rebuild_blas(self, b):
        """Full LBVH rebuild of BLAS b from the current vertices; stream-ordered."""
        self._ok(self._L.rgrt_rebuild_blas(self.h, b))

    def rebuild_blas_noalloc(self, b):
        """rebuild_blas with the builder's arrays in the library's fixed
        per-GPU scratch (no allocation; at most 262144 triangles)."""
        self._ok(self._L.rgrt_rebuild_blas_noalloc(self.h, b))

    def blas_info(self, b):
        nn, nl = ctypes.c_uint32(), ctypes.c_uint32()
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        self._ok(self._L.rgrt_blas_info(self.h, b, ctypes.byref(nn), ctypes.byref(nl), _p(lo), _p(hi)))
        return nn.value, nl.value, lo, hi

    def export_blas(self, b):
        nn, nl, _, _ = self.blas_info(b)
        nodes = np.zeros(nn, BVH_NODE)
        leaf = np.zeros(nl, np.uint32)
        self._ok(self._L.rgrt_export_blas(self.h, b, _p(nodes), nn, _p(leaf), nl))
        return nodes, leaf

    def export_blas_wide(self, b, compressed=False):
        """4-wide tree of BLAS b: WIDE_NODE records, or QWIDE_NODE (the form
        the traversal walks) with compressed=True."""
        c = int(bool(compressed))
        nn = ctypes.c_uint32()
        self._ok(self._L.rgrt_export_blas_wide(self.h, b, c, None, 0, ctypes.byref(nn)))
        nodes = np.zeros(nn.value, QWIDE_NODE if c else WIDE_NODE)
        self._ok(self._L.rgrt_export_blas_wide(self.h, b, c, _p(nodes), nn.value, ctypes.byref(nn)))
        return nodes

    def export_tlas_wide(self, compressed=False):
        c = int(bool(compressed))
        nn = ctypes.c_uint32()
        self._ok(self._L.rgrt_export_tlas_wide(self.h, c, None, 0, ctypes.byref(nn)))
        nodes = np.zeros(nn.value, QWIDE_NODE if c else WIDE_NODE)
        self._ok(self._L.rgrt_export_tlas_wide(self.h, c, _p(nodes), nn.value, ctypes.byref(nn)))
        return nodes

    def build_tlas(self, instances):
        inst = np.ascontiguousarray(instances, INSTANCE)
        self._ok(self._L.rgrt_build_tlas(self.h, _p(inst), len(inst)))
        self._ninst = len(inst)

    def refit_tlas(self, instances):
        """Same instance count and BLAS per position as the last build: new
        transforms on the existing hierarchy."""
        inst = np.ascontiguousarray(instances, INSTANCE)
        self._ok(self._L.rgrt_refit_tlas(self.h, _p(inst), len(inst)))
        self._ninst = len(inst)

    def export_tlas(self):
        nn = ctypes.c_uint32()
        self._ok(self._L.rgrt_export_tlas(self.h, None, 0, None, 0, ctypes.byref(nn)))
        nodes = np.zeros(nn.value, BVH_NODE)
        leaf = np.zeros(self._ninst, np.uint32)
        self._ok(self._L.rgrt_export_tlas(self.h, _p(nodes), nn.value, _p(leaf), len(leaf), ctypes.byref(nn)))
        return nodes, leaf

    def set_frame(self, ubo):
        u = np.ascontiguousarray(np.asarray(ubo, UBO).reshape(()))
        self._ok(self._L.rgrt_set_frame(self.h, _p(u)))

    def resize(self, width, height):
        self._ok(self._L.rgrt_resize(self.h, width, height))
        self.width, self.height = width, height

    # -- work ---------------------------------------------------------------
    def trace_rays(self, rays, cull_mask=0xFF):
        r = np.ascontiguousarray(rays, RAY)
        hits = np.zeros(len(r), HIT)
        self._ok(self._L.rgrt_trace_rays(self.h, _p(r), len(r), cull_mask, _p(hits)))
        return hits

    def trace_rays_device(self, rays_ptr, n, cull_mask, hits_ptr):
        self._ok(self._L.rgrt_trace_rays_device(self.h, rays_ptr, n, cull_mask, hits_ptr))

    def set_sample_groups(self, max_groups):
        """Samples of a pixel traced side by side (0: as many as fit, 1: one at a time)."""
        self._ok(self._L.rgrt_set_sample_groups(self.h, max_groups))

    def set_sampling(self, mode, seed=0):
        """SAMPLING_WHITTED (the reference's tracer) or SAMPLING_STOCHASTIC
        (jittered pixels, glossy reflections, denoiser features)."""
        sm = np.array([mode, seed, 0, 0], np.uint32)
        self._ok(self._L.rgrt_set_sampling(self.h, _p(sm)))

    def read_image_rgba8(self, which=None, fmt=None):
        """An rgba16f image converted on the device to 8 bits per channel
        (the swapchain blit's conversion): (height, width, 4) uint8."""
        which = IMAGE_FINAL if which is None else which
        fmt = FORMAT_RGBA8_UNORM if fmt is None else fmt
        out = np.zeros((self.height, self.width, 4), np.uint8)
        self._ok(self._L.rgrt_read_image_rgba8(self.h, which, fmt, _p(out), out.nbytes))
        return out

    def frame(self, ubo, instances, fxaa=True, refit=False, graph=False):
        """One frame: frame constants, TLAS build (or refit) and render; with
        graph=True captured and launched as one hipGraph."""
        u = np.ascontiguousarray(ubo, UBO)
        inst = np.ascontiguousarray(instances, INSTANCE)
        flags = (FRAME_FXAA if fxaa else 0) | (FRAME_REFIT_TLAS if refit else 0) | (FRAME_GRAPH if graph else 0)
        self._ok(self._L.rgrt_frame(self.h, _p(u), _p(inst), len(inst), flags))
        self._ninst = len(inst)

    def render(self, use_fxaa=True):
        self._ok(self._L.rgrt_render(self.h, int(bool(use_fxaa))))

    def trace(self):
        self._ok(self._L.rgrt_trace(self.h))

    def postprocess(self, use_fxaa=True):
        self._ok(self._L.rgrt_postprocess(self.h, int(bool(use_fxaa))))

    def synchronize(self):
        self._ok(self._L.rgrt_synchronize(self.h))

    # -- images -------------------------------------------------------------
    def read_image(self, which):
        W, H = self.width, self.height
        if which == IMAGE_ROUGH_TRANSITIONS:
            out = np.zeros((H, W), np.int8)
        elif which in (IMAGE_BASE_F32, IMAGE_NORMAL_F32, IMAGE_ROUGH_F32):
            out = np.zeros((H, W, 4), np.float32)
        elif which == IMAGE_PRIMARY_HIT:
            out = np.zeros((H, W), HIT)
        elif which == IMAGE_FEATURES:
            out = np.zeros((H, W, 8), np.uint16)
        else:
            out = np.zeros((H, W, 4), np.uint16)
        self._ok(self._L.rgrt_read_image(self.h, which, _p(out), out.nbytes))
        return out

    def write_image(self, which, data):
        d = np.ascontiguousarray(data)
        self._ok(self._L.rgrt_write_image(self.h, which, _p(d), d.nbytes))

    def image_device_ptr(self, which):
        p, n = ctypes.c_void_p(), ctypes.c_uint64()
        self._ok(self._L.rgrt_image_device_ptr(self.h, which, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def stats(self):
        s = np.zeros((), STATS)
        self._ok(self._L.rgrt_get_stats(self.h, _p(s)))
        return {k: s[k].item() for k in STATS.names if not k.startswith("pad")}

    def tuning(self):
        t = np.zeros((), TUNING)
        self._ok(self._L.rgrt_get_tuning(self.h, _p(t)))
        return {k: int(t[k]) for k in TUNING.names}

    def set_tuning(self, **kw):
        t = np.zeros((), TUNING)
        self._ok(self._L.rgrt_get_tuning(self.h, _p(t)))
        for k, v in kw.items():
            t[k] = v
        self._ok(self._L.rgrt_set_tuning(self.h, _p(t)))

    def pass_counts(self):
        n = ctypes.c_uint32()
        self._ok(self._L.rgrt_get_pass_counts(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.uint32)
        self._ok(self._L.rgrt_get_pass_counts(self.h, _p(out), n.value, ctypes.byref(n)))
        return out


def version():
    return lib().rgrt_version().decode()


class Denoiser:
    """The learned denoiser of one context (rgrt_denoiser_*)."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        self._L = lib()
        self.h = ctypes.c_void_p()
        ctx._ok(self._L.rgrt_denoiser_create(ctx.h, seed, ctypes.byref(self.h)))
        n = ctypes.c_uint64()
        ctx._ok(self._L.rgrt_denoiser_num_params(self.h, ctypes.byref(n)))
        self.num_params = n.value

    def close(self):
        if self.h:
            self._L.rgrt_denoiser_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        self.ctx._ok(self._L.rgrt_denoiser_get_params(self.h, _p(p), p.size))
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self.ctx._ok(self._L.rgrt_denoiser_set_params(self.h, _p(p), p.size))

    def loss_and_grads(self, x0_bits, target, with_output=False):
        """x0_bits: (n, h, w, 16) uint16 bf16 patterns; target: (n, h, w, 3)."""
        x = np.ascontiguousarray(x0_bits, np.uint16)
        t = np.ascontiguousarray(target, np.float32)
        n, hh, ww, _ = x.shape
        g = np.zeros(self.num_params, np.float32)
        out = np.zeros((n, hh, ww, 16), np.uint16) if with_output else None
        loss = ctypes.c_float()
        self.ctx._ok(self._L.rgrt_denoiser_loss_and_grads(self.h, _p(x), _p(t), n, hh, ww, ctypes.byref(loss), _p(g),
                                                          _p(out)))
        return (loss.value, g, out) if with_output else (loss.value, g)

    def infer(self, x0_bits):
        """Forward pass alone: x0_bits (n, h, w, 16) uint16 bf16 patterns ->
        output (n, h, w, 16) uint16 bf16 patterns."""
        x = np.ascontiguousarray(x0_bits, np.uint16)
        n, hh, ww, _ = x.shape
        out = np.zeros((n, hh, ww, 16), np.uint16)
        self.ctx._ok(self._L.rgrt_denoiser_infer(self.h, _p(x), n, hh, ww, _p(out)))
        return out

    def dataset_reserve(self, count, height, width):
        self.ctx._ok(self._L.rgrt_denoiser_dataset_reserve(self.h, count, height, width))

    def dataset_add_input(self, slot):
        self.ctx._ok(self._L.rgrt_denoiser_dataset_add_input(self.h, slot))

    def dataset_add_target(self, slot, weight, accumulate):
        self.ctx._ok(self._L.rgrt_denoiser_dataset_add_target(self.h, slot, weight, int(bool(accumulate))))

    def dataset_read(self, first, count, height, width):
        """Slots [first, first + count) of the training set: packed inputs
        (count, h, w, 16) uint16 bf16 patterns and targets (count, h, w, 3)."""
        x = np.zeros((count, height, width, 16), np.uint16)
        t = np.zeros((count, height, width, 3), np.float32)
        self.ctx._ok(self._L.rgrt_denoiser_dataset_read(self.h, first, count, height, width, _p(x), _p(t)))
        return x, t

    def train(self, steps, batch, lr, seed, used):
        loss = ctypes.c_float()
        self.ctx._ok(self._L.rgrt_denoiser_train(self.h, steps, batch, lr, seed, used, ctypes.byref(loss)))
        return loss.value

    def denoise(self, fallback_a=8.0, fallback_b=2e-4, count=False):
        """Denoises the context's current frame; with count=True returns
        (fallback tiles, tiles)."""
        if not count:
            self.ctx._ok(self._L.rgrt_denoise(self.h, fallback_a, fallback_b, None, None))
            return None
        f, t = ctypes.c_uint32(), ctypes.c_uint32()
        self.ctx._ok(self._L.rgrt_denoise(self.h, fallback_a, fallback_b, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value


class _FrameDesc(ctypes.Structure):
    _fields_ = [("ubo", ctypes.c_void_p), ("instances", ctypes.c_void_p), ("numInstances", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class Batch:
    """Frames of several contexts submitted together (rgrt_batch_*): a
    hipGraph per scene on its own stream (BATCH_GRAPH_PER_SCENE), one
    hipGraph with a branch per scene (BATCH_ONE_GRAPH), or one hipGraph
    tracing all scenes in shared launches (BATCH_PERSISTENT)."""

    def __init__(self, contexts, mode=BATCH_GRAPH_PER_SCENE):
        self._L = lib()
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
        self.h = ctypes.c_void_p()
        st = self._L.rgrt_batch_create(arr, len(self.contexts), ctypes.byref(self.h))
        if st != 0:
            raise RgrtError(st, "rgrt_batch_create failed")
        self.set_mode(mode)

    def stream(self):
        """The batch's stream (int address): ordered after every batch frame."""
        return self._L.rgrt_batch_stream(self.h)

    def synchronize(self):
        for c in self.contexts:
            c.synchronize()

    def set_mode(self, mode):
        st = self._L.rgrt_batch_set_mode(self.h, mode)
        if st != 0:
            raise RgrtError(st, "rgrt_batch_set_mode failed")
        self.mode = mode

    def close(self):
        if self.h:
            self._L.rgrt_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def frame(self, frames):
        """frames: per context (ubo, instances, fxaa, refit)."""
        keep = []
        descs = (_FrameDesc * len(self.contexts))()
        for d, (ubo, inst, fxaa, refit) in zip(descs, frames):
            u = np.ascontiguousarray(ubo, UBO)
            i = np.ascontiguousarray(inst, INSTANCE)
            keep += [u, i]
            d.ubo = u.ctypes.data
            d.instances = i.ctypes.data if len(i) else None
            d.numInstances = len(i)
            d.flags = (FRAME_FXAA if fxaa else 0) | (FRAME_REFIT_TLAS if refit else 0)
        st = self._L.rgrt_batch_frame(self.h, descs)
        if st != 0:
            raise RgrtError(st, self._L.rgrt_batch_last_error(self.h).decode())
        for c, (_, inst, _, _) in zip(self.contexts, frames):
            c._ninst = len(inst)
