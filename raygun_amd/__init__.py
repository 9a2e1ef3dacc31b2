"""rgrt — an MI355X-native ray-tracing core for Raygun.

HIP kernels (LBVH build, two-level traversal, wavefront closest-hit/miss
shading, compute post passes) behind the C ABI in include/rgrt.h, plus a
host-side mirror of the reference's scene and renderer classes.
"""
from . import records, scene  # noqa: F401

__all__ = ["records", "scene", "procgen", "rgrt", "raytracer"]
