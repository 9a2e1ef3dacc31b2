"""Driver of the learned denoiser: seeded procedural training scenes rendered
by the tracer itself, the training set (4-spp stochastic frames and their
256-spp means), training, and the comparison of the learned path (4 spp +
denoiser) with the plain tracer (256 spp) at 1080p.

No counterpart in the reference; the scenes are the config-1 family (the
sphere + ground mesh of procgen.config1, raygun's material model) with seeded
cameras, lights, transforms and materials.
"""
import math

import numpy as np

from . import procgen, rgrt
from .records import INSTANCE, default_material
from .scene import Camera, camera_ubo, pack_models


def psnr(img, ref):
    """10 log10(1 / MSE) on RGB clamped to [0, 1]."""
    a = np.clip(np.asarray(img, np.float64)[..., :3], 0, 1)
    b = np.clip(np.asarray(ref, np.float64)[..., :3], 0, 1)
    mse = float(np.mean((a - b) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(1.0 / mse)


def half(img16):
    return np.asarray(img16, np.uint16).view(np.float16).astype(np.float32)


class GlossyScenes:
    """The config-1 mesh uploaded once into a context; scene(seed) draws a
    camera, light direction, sphere transform and two materials and returns
    the frame constants (numSamples set by the caller)."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, width, height
        setup = procgen.config1(width, height)
        self.model = setup.scene.models()[0]
        v, i, m = pack_models([self.model])
        ctx.upload_geometry(v, i)
        ctx.upload_materials(m)
        mesh = self.model.mesh
        self.blas = ctx.build_blas(mesh.first_vertex, mesh.index_offset_bytes, mesh.num_faces(), len(mesh.vertices))

    def scene(self, seed, spp, max_recursions=2):
        rng = np.random.default_rng(seed)
        sphere = default_material(diffuse=tuple(rng.uniform(0, 1, 3)), reflectivity=float(rng.uniform(0.2, 1.0)),
                                  roughness=float(rng.uniform(0.02, 0.6)))
        ground = default_material(diffuse=tuple(rng.uniform(0.2, 0.9, 3)), reflectivity=float(rng.uniform(0.0, 0.6)),
                                  roughness=float(rng.uniform(0.0, 0.6)), effectId=int(rng.uniform() < 0.5))
        self.ctx.upload_materials(np.array([sphere, ground]))
        inst = np.zeros(1, INSTANCE)
        s = float(rng.uniform(0.6, 1.4))
        inst["transform"] = [s, 0, 0, float(rng.uniform(-0.5, 0.5)), 0, s, 0, 0, 0, 0, s, float(rng.uniform(-0.5, 0.5))]
        inst["mask"] = 0xFF
        inst["blas"] = self.blas
        self.ctx.build_tlas(inst)
        cam = Camera(fov_deg=float(rng.uniform(40, 70)))
        az, el, dist = rng.uniform(0, 2 * math.pi), rng.uniform(0.05, 0.6), rng.uniform(2.8, 6.0)
        eye = (dist * math.cos(el) * math.sin(az), dist * math.sin(el) - 0.3, dist * math.cos(el) * math.cos(az))
        cam.look_at(eye, (0.0, float(rng.uniform(-0.5, 0.3)), 0.0))
        L = np.array([rng.uniform(-1, 1), -rng.uniform(0.3, 1.0), rng.uniform(-1, 1)])
        L /= np.linalg.norm(L)
        return camera_ubo(cam, self.width, self.height, light_dir=tuple(L), num_samples=spp, max_recursions=max_recursions)


def trace_mean(ctx, ubo, spp, frames, seed):
    """Traces `frames` stochastic frames of `spp` samples (seeds seed..);
    returns nothing (the caller reads or accumulates between frames)."""
    u = ubo.copy()
    u["numSamples"] = spp
    for k in range(frames):
        ctx.set_sampling(rgrt.SAMPLING_STOCHASTIC, seed + k)
        ctx.set_frame(u)
        ctx.trace()
        yield k


def build_dataset(ctx, den, scenes, count, first_seed=1000, spp_in=4, spp_ref=256, ref_spp_per_frame=32):
    """count training pairs: slot k = the 4-spp frame of scene first_seed + k
    (packed features) and the mean of spp_ref samples (fp32 target)."""
    den.dataset_reserve(count, scenes.height, scenes.width)
    frames = max(1, spp_ref // ref_spp_per_frame)
    for k in range(count):
        seed = first_seed + k
        ubo = scenes.scene(seed, spp_in)
        for f in trace_mean(ctx, ubo, ref_spp_per_frame, frames, seed * 1000 + 1):
            den.dataset_add_target(k, 1.0 / frames, f > 0)
        for _ in trace_mean(ctx, ubo, spp_in, 1, seed * 1000):
            den.dataset_add_input(k)
    ctx.synchronize()


def denoised_frame(ctx, den, ubo, spp, seed, fallback=(8.0, 2e-4)):
    """4-spp learned frame: trace, denoise, post (no FXAA); final image and
    (fallback tiles, tiles)."""
    for _ in trace_mean(ctx, ubo, spp, 1, seed):
        pass
    fb = den.denoise(*fallback, count=True)
    ctx.postprocess(False)
    return half(ctx.read_image(rgrt.IMAGE_FINAL)), fb


def noisy_frame(ctx, ubo, spp, seed):
    for _ in trace_mean(ctx, ubo, spp, 1, seed):
        pass
    ctx.postprocess(False)
    return half(ctx.read_image(rgrt.IMAGE_FINAL))


def reference_frame(ctx, ubo, spp_total, seed, spp_per_frame=32):
    """Mean of spp_total samples (fp32 BASE of each frame averaged on the host),
    clamped like the post pass."""
    frames = max(1, spp_total // spp_per_frame)
    acc = None
    for _ in trace_mean(ctx, ubo, spp_per_frame, frames, seed):
        b = half(ctx.read_image(rgrt.IMAGE_BASE))[..., :3].astype(np.float64)
        acc = b if acc is None else acc + b
    return acc / frames
