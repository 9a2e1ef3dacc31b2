"""Driver of the learned denoiser: seeded procedural training scenes rendered
by the tracer itself, the training set (4-spp stochastic frames and their
256-spp means), training (denoise_train: PyTorch on the GPU while the
library's own backward kernels are switched off), and the comparison of the
learned path (4 spp + denoiser, the library's MFMA inference) with the plain
tracer at 1080p.

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

    def scene(self, seed, spp, max_recursions=2, frame=None, tile=(0, 0)):
        """Frame constants of scene `seed`. With frame = (W, H) the camera is
        that of a W x H frame and the context's width x height launch renders
        the window of it whose top-left pixel is `tile` = (x0, y0), with the
        same pixel footprint (the projection is restricted to the window)."""
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
        if frame is None:
            return camera_ubo(cam, self.width, self.height, light_dir=tuple(L), num_samples=spp, max_recursions=max_recursions)
        W, H = frame
        ubo = camera_ubo(cam, W, H, light_dir=tuple(L), num_samples=spp, max_recursions=max_recursions)
        return window_ubo(ubo, W, H, tile[0], tile[1], self.width, self.height)


def window_ubo(ubo, W, H, x0, y0, w, h):
    """Frame constants whose w x h launch renders the window at (x0, y0) of
    the W x H frame `ubo` describes: raygen maps a launch pixel to
    d = 2 (p + 0.5) / size - 1 and the ray to projInverse (d, 1, 1), so the
    window is projInverse composed with d -> (w / W) d + (2 x0 + w) / W - 1
    per axis."""
    u = ubo.copy()
    p_inv = np.asarray(ubo["projInverse"], np.float64).reshape(4, 4).T  # column-major storage
    m = np.eye(4)
    m[0, 0], m[0, 2] = w / W, (2.0 * x0 + w) / W - 1.0
    m[1, 1], m[1, 2] = h / H, (2.0 * y0 + h) / H - 1.0
    u["projInverse"] = (p_inv @ m).T.reshape(16).astype(np.float32)
    return u


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


def build_dataset(ctx, den, scenes, count, first_seed=1000, spp_in=4, spp_ref=256, ref_spp_per_frame=32, frame=None):
    """count training pairs: slot k = the spp_in-sample frame of scene
    first_seed + k (packed features) and the mean of spp_ref samples (fp32
    target); with a sequence of sample counts, slot k takes
    spp_in[k % len(spp_in)] (the features carry each pixel's variance, so one
    network learns the noise levels together). With frame = (W, H) each slot
    is a seeded window of a W x H frame of its scene (the pixel footprint of
    that frame), else the whole scene at the context's size."""
    den.dataset_reserve(count, scenes.height, scenes.width)
    frames = max(1, spp_ref // ref_spp_per_frame)
    counts = tuple(spp_in) if isinstance(spp_in, (tuple, list)) else (spp_in,)
    for k in range(count):
        seed = first_seed + k
        spp = counts[k % len(counts)]
        if frame is None:
            ubo = scenes.scene(seed, spp)
        else:
            r = np.random.default_rng(seed + 7_000_000)
            tile = (int(r.integers(0, frame[0] - scenes.width + 1)), int(r.integers(0, frame[1] - scenes.height + 1)))
            ubo = scenes.scene(seed, spp, frame=frame, tile=tile)
        for f in trace_mean(ctx, ubo, ref_spp_per_frame, frames, seed * 1000 + 1):
            den.dataset_add_target(k, 1.0 / frames, f > 0)
        for _ in trace_mean(ctx, ubo, spp, 1, seed * 1000):
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


def render_dataset(ctx, den, scenes, count, first_seed=1000, spp_in=4, spp_ref=256, frame=None):
    """`count` training pairs rendered and packed by the library (the
    denoiser's dataset on `ctx`), read back: (x0 bf16 bits N x H x W x 16,
    targets N x H x W x 3)."""
    build_dataset(ctx, den, scenes, count, first_seed=first_seed, spp_in=spp_in, spp_ref=spp_ref, frame=frame)
    return den.dataset_read(0, count, scenes.height, scenes.width)


def train_params(x, t, init, steps=3000, batch=16, lr=1e-3, crop=None, seed=0, clamp_target=False, workdir=None,
                 timeout=900):
    """Trains the network on the GPU in a child process (denoise_train;
    PyTorch has a HIP runtime of its own) from `init`. Returns (parameters,
    first loss, last loss)."""
    import json
    import os
    import subprocess
    import sys
    import tempfile
    with tempfile.TemporaryDirectory(dir=workdir) as tmp:
        data, out = os.path.join(tmp, "train.npz"), os.path.join(tmp, "params.npy")
        np.savez(data, x0=x, target=t, init=init)
        cmd = [sys.executable, "-m", "raygun_amd.denoise_train", data, out, "--steps", str(steps), "--batch", str(batch),
               "--lr", str(lr), "--seed", str(seed), "--crop", str(crop or 0)] + (["--clamp-target"] if clamp_target else [])
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        r = subprocess.run(cmd, cwd=root, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"denoiser training failed ({r.returncode}):\n{r.stderr[-2000:]}")
        stats = json.loads(r.stdout.strip().splitlines()[-1])
        params = np.load(out, allow_pickle=False)
    return params, stats["loss_first"], stats["loss_last"]


def train_network(ctx, den, scenes, count, first_seed=1000, spp_in=4, spp_ref=256, steps=3000, batch=16, lr=1e-3, crop=None,
                  seed=0, workdir=None, timeout=900, frame=None, clamp_target=False):
    """Renders `count` training pairs with the library, trains the network on
    the GPU from the library's own initialisation and loads the result into
    `den`. Returns (first loss, last loss, seconds rendering, seconds
    training)."""
    import time
    t0 = time.perf_counter()
    x, t = render_dataset(ctx, den, scenes, count, first_seed=first_seed, spp_in=spp_in, spp_ref=spp_ref, frame=frame)
    t1 = time.perf_counter()
    params, first, last = train_params(x, t, den.get_params(), steps=steps, batch=batch, lr=lr, crop=crop, seed=seed,
                                       clamp_target=clamp_target, workdir=workdir, timeout=timeout)
    t2 = time.perf_counter()
    den.set_params(params)
    return first, last, t1 - t0, t2 - t1


# Multiply-accumulates per pixel of one inference as executed (input padded
# to 16 channels, output computed for 16 channels): full resolution layers
# 16-16, 16-16, 48-16, 16-16, 16-16; quarter resolution 16-32, 32-32, 96-32,
# 32-32; sixteenth resolution 32-64, 64-64; 9 taps each.
MACS_PER_PIXEL = 9 * ((16 * 16 * 4 + 48 * 16) + (16 * 32 + 32 * 32 + 96 * 32 + 32 * 32) / 4 + (32 * 64 + 64 * 64) / 16)


def _plain_image(ctx, ubo, spp, seed):
    """Mean of `spp` samples (frames of at most 32) through the post pass's
    clamp: the plain tracer's final image."""
    if spp <= 32:
        return noisy_frame(ctx, ubo, spp, seed)
    return np.clip(reference_frame(ctx, ubo, spp, seed), 0.0, 1.0)


def _equal_psnr_ms(points, target):
    """ms of the plain tracer at PSNR `target`: log(ms) interpolated linearly
    in PSNR between the sweep points around it (extrapolated from the two
    nearest when outside)."""
    pts = sorted(points)
    if len(pts) < 2:
        return None
    lo = max([k for k in range(len(pts) - 1) if pts[k][0] <= target] or [0])
    (p0, m0), (p1, m1) = pts[lo], pts[min(lo + 1, len(pts) - 1)]
    if p1 == p0:
        return m1
    f = (target - p0) / (p1 - p0)
    return math.exp(math.log(m0) + f * (math.log(m1) - math.log(m0)))


def learned_vs_plain(ctx, den, scenes, seeds, spp_learned=4, spp_ref=256, sweep=(4, 8, 16, 32, 64, 128),
                     fallback=(8.0, 2e-4), events=None, spp_more=()):
    """For each seeded scene at the context's size: the plain tracer's
    spp_ref frame (the reference), plain frames of each `sweep` sample count
    and the learned frame (spp_learned samples + the denoiser, fallback
    tiles traced), PSNR against the reference, and GPU milliseconds per frame
    (hipEvents on the context's stream: the traces, the denoiser and the
    post pass). The learned frame's cost is compared with the plain frame's
    at the same PSNR (interpolated over the sweep). Each count in spp_more
    adds the same figures for a learned frame of that many samples
    (`learned_more`)."""
    from .hiptime import HipEvents
    ev = events or HipEvents()
    stream = rgrt.lib().rgrt_stream(ctx.h)
    a, b = ev.create(), ev.create()

    def gpu_ms(fn):
        ctx.synchronize()
        ev.record(a, stream)
        fn()
        ev.record(b, stream)
        return ev.elapsed_ms(a, b)

    def plain_run(ubo, spp, seed):
        frames = max(1, (spp + 31) // 32)
        for _ in trace_mean(ctx, ubo, spp // frames, frames, seed):
            pass
        ctx.postprocess(False)

    def learned_run(ubo, seed, spp=spp_learned):
        for _ in trace_mean(ctx, ubo, spp, 1, seed):
            pass
        den.denoise(*fallback)
        ctx.postprocess(False)

    out = []
    for k, sd in enumerate(seeds):
        ubo = scenes.scene(sd, spp_learned)
        ref = np.clip(reference_frame(ctx, ubo, spp_ref, seed=sd * 7919 + 1), 0.0, 1.0)
        ref_ms = gpu_ms(lambda: plain_run(ubo, spp_ref, sd * 7919 + 1))
        points = []
        for spp in sweep:
            seed = sd * 7919 + 100 + spp
            img = _plain_image(ctx, ubo, spp, seed)
            ms = gpu_ms(lambda: plain_run(ubo, spp, seed))
            points.append((psnr(img, ref), ms))
        seed = sd * 7919 + 50
        learned_run(ubo, seed)  # warm
        ms_l = min(gpu_ms(lambda: learned_run(ubo, seed)) for _ in range(3))
        img_l, (fb, tiles) = denoised_frame(ctx, den, ubo, spp_learned, seed, fallback=fallback)
        img_all, _ = denoised_frame(ctx, den, ubo, spp_learned, seed, fallback=(1e30, 1e30))
        noisy = noisy_frame(ctx, ubo, spp_learned, seed)
        den_ms = min(gpu_ms(lambda: den.denoise(1e9, 1e9)) for _ in range(3))
        p_l = psnr(img_l, ref)
        eq = _equal_psnr_ms(points, p_l)
        more = []
        for spp in spp_more:
            seed_m = sd * 7919 + 60 + spp
            learned_run(ubo, seed_m, spp)  # warm
            ms_m = min(gpu_ms(lambda: learned_run(ubo, seed_m, spp)) for _ in range(3))
            img_m, (fb_m, _) = denoised_frame(ctx, den, ubo, spp, seed_m, fallback=fallback)
            p_m = psnr(img_m, ref)
            eq_m = _equal_psnr_ms(points, p_m)
            more.append({"spp": spp, "psnr_learned": round(p_m, 2), "psnr_noisy": round(psnr(noisy_frame(ctx, ubo, spp, seed_m), ref), 2),
                         "learned_ms": round(ms_m, 3), "fallback_tiles": fb_m,
                         "plain_ms_at_equal_psnr": None if eq_m is None else round(eq_m, 3),
                         "ratio_at_equal_psnr": None if eq_m is None else round(eq_m / ms_m, 2)})
        out.append({"seed": sd, "psnr_learned": round(p_l, 2), "psnr_noisy": round(psnr(noisy, ref), 2),
                    "psnr_learned_every_tile": round(psnr(img_all, ref), 2),
                    "learned_ms": round(ms_l, 3), "denoise_ms": round(den_ms, 3), "fallback_tiles": fb, "tiles": tiles,
                    "plain_sweep": [{"spp": s, "psnr": round(p, 2), "ms": round(m, 3)} for s, (p, m) in zip(sweep, points)],
                    "reference_spp": spp_ref, "reference_ms": round(ref_ms, 3),
                    "plain_ms_at_equal_psnr": None if eq is None else round(eq, 3),
                    "ratio_at_equal_psnr": None if eq is None else round(eq / ms_l, 2),
                    "ratio_vs_reference": round(ref_ms / ms_l, 2), "learned_more": more})
    ev.destroy(a)
    ev.destroy(b)
    return out


class SceneViews:
    """A procgen scene (SceneSetup) uploaded once into a context: geometry,
    materials, BLASes and TLAS. scene(seed, spp) returns frame constants of a
    seeded view of it (camera on a sphere of radius U[r0, r1] around the
    origin, looking at a point in the central box of half-size `spread`,
    vertical FOV U[40, 70]), or, with fixed_camera, the setup's own camera.
    The config-5 counterpart of GlossyScenes."""

    def __init__(self, ctx, setup, width, height, fixed_camera=False, r0=60.0, r1=170.0, spread=60.0, upload=True):
        self.ctx, self.setup, self.width, self.height = ctx, setup, width, height
        self.fixed_camera, self.r0, self.r1, self.spread = fixed_camera, r0, r1, spread
        if not upload:  # the context holds the scene and its TLAS already
            return
        models = setup.scene.models()
        v, i, m = pack_models(models)
        ctx.upload_geometry(v, i)
        ctx.upload_materials(m)
        blas_of = {}
        for mo in models:
            k = id(mo.mesh)
            if k not in blas_of:
                blas_of[k] = ctx.build_blas(mo.mesh.first_vertex, mo.mesh.index_offset_bytes, mo.mesh.num_faces(),
                                            len(mo.mesh.vertices))
            mo.blas = blas_of[k]
        ctx.build_tlas(setup.scene.collect_instances())

    def scene(self, seed, spp, max_recursions=None):
        rec = self.setup.max_recursions if max_recursions is None else max_recursions
        if self.fixed_camera:
            return camera_ubo(self.setup.scene.camera, self.width, self.height, num_samples=spp, max_recursions=rec)
        rng = np.random.default_rng(seed)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = d * rng.uniform(self.r0, self.r1)
        cam = Camera(fov_deg=float(rng.uniform(40, 70)))
        cam.look_at(tuple(eye), tuple(rng.uniform(-self.spread, self.spread, 3)))
        return camera_ubo(cam, self.width, self.height, num_samples=spp, max_recursions=rec)
