"""numpy dtypes of the records that cross the C ABI (include/rgrt.h).

Each layout mirrors one of the reference's X-macro `.def` files, which the
reference compiles into both its C++ host and its GLSL shaders:

- VERTEX      resources/shaders/vertex.def:3-7                  (32 B)
- MATERIAL    resources/shaders/gpu_material.def:11-26           (64 B)
- UBO         resources/shaders/uniform_buffer_object.def:3-17   (192 B, std140)
- OFFSETS     resources/shaders/instance_offset_table.def:1-3    (12 B)
- INSTANCE    vk::AccelerationStructureInstanceKHR fields used by
              raygun/render/acceleration_structure.cpp:34-51 + the offsets (80 B)
"""
import numpy as np

VERTEX = np.dtype([("position", "<f4", 3), ("matIndex", "<u4"), ("normal", "<f4", 3), ("pad1", "<f4")])

MATERIAL = np.dtype([
    ("diffuse", "<f4", 3), ("transparency", "<f4"),
    ("specular", "<f4", 3), ("reflectivity", "<f4"),
    ("roughness", "<f4"), ("ior", "<f4"),
    ("effectId", "<u4"), ("rayConsumption", "<u4"),
    ("emission", "<f4"), ("pad", "<f4", 3),
])

UBO = np.dtype([
    ("viewInverse", "<f4", 16), ("projInverse", "<f4", 16),
    ("clearColor", "<f4", 3), ("numSamples", "<i4"),
    ("lightDir", "<f4", 3), ("maxRecursions", "<i4"),
    ("time", "<f4"), ("showAlpha", "<u4"), ("pad0", "<f4"), ("pad1", "<f4"),
    ("fadeColor", "<f4", 4),
])

INSTANCE = np.dtype([
    ("transform", "<f4", 12), ("customIndex", "<u4"), ("mask", "<u4"), ("flags", "<u4"),
    ("blas", "<u4"), ("offsets", "<u4", 3), ("pad", "<u4"),
])

BLAS_DESC = np.dtype([("firstVertex", "<u4"), ("indexOffsetBytes", "<u4"), ("numFaces", "<u4"), ("maxVertex", "<u4")])

RAY = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])

HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("instance", "<u4"), ("primitive", "<u4")])

BVH_NODE = np.dtype([("lo", "<f4", (2, 3)), ("hi", "<f4", (2, 3)), ("child", "<i4", 2), ("parent", "<u4"), ("pad", "<u4")])

STATS_FIELDS = [
    ("asBuildMs", "<f4"), ("rtTotalMs", "<f4"), ("rtOnlyMs", "<f4"), ("postprocMs", "<f4"), ("roughMs", "<f4"),
    ("pad_", "<u4"),
    ("rays", "<u8"), ("skyLookups", "<u8"), ("zeroDirRays", "<u8"),
    ("waves", "<u4"), ("traceLaunches", "<u4"), ("traceKernelMs", "<f4"), ("shadeKernelMs", "<f4"),
    ("drainKernelMs", "<f4"), ("drained", "<u4"), ("frameMs", "<f4"), ("pad2_", "<u4"),
]

TUNING = np.dtype([("wavefrontPasses", "<u4"), ("traceBlocksPerCU", "<u4"), ("shadeBlocksPerCU", "<u4"),
                   ("drainBlocksPerCU", "<u4"), ("hostLag", "<u4"), ("frameBands", "<u4")])

WIDE_NODE = np.dtype([("lox", "<f4", 4), ("hix", "<f4", 4), ("loy", "<f4", 4), ("hiy", "<f4", 4), ("loz", "<f4", 4),
                      ("hiz", "<f4", 4), ("child", "<i4", 4), ("pad", "<i4", 4)])

QWIDE_NODE = np.dtype([("origin", "<f4", 3), ("exp", "i1", 3), ("pad0", "i1"),
                       ("qlox", "u1", 4), ("qhix", "u1", 4), ("qloy", "u1", 4), ("qhiy", "u1", 4),
                       ("qloz", "u1", 4), ("qhiz", "u1", 4), ("pad1", "<u4", 2), ("child", "<i4", 4)])

NO_HIT = 0xFFFFFFFF
EMPTY_CHILD = -(2 ** 31)

assert VERTEX.itemsize == 32
assert MATERIAL.itemsize == 64
assert UBO.itemsize == 192
assert INSTANCE.itemsize == 80
assert RAY.itemsize == 32
assert HIT.itemsize == 20
assert BVH_NODE.itemsize == 64
assert WIDE_NODE.itemsize == 128
assert QWIDE_NODE.itemsize == 64


def default_material(**kw):
    """gpu::Material with the reference defaults (gpu_material.def:11-23)."""
    m = np.zeros((), MATERIAL)
    m["diffuse"] = (1.0, 0.0, 1.0)
    m["transparency"] = 0.0
    m["specular"] = (1.0, 1.0, 1.0)
    m["reflectivity"] = 0.0
    m["roughness"] = 0.0
    m["ior"] = 1.0
    m["effectId"] = 0
    m["rayConsumption"] = 1
    m["emission"] = 0.0
    for k, v in kw.items():
        m[k] = v
    return m
