"""Display-format readback (rgrt_read_image_rgba8, the swapchain blit's
per-pixel conversion) against a numpy restatement of Vulkan's float ->
unorm8 rule (clamp to [0, 1], NaN to 0, x * 255 rounded to nearest even) and
the sRGB transfer function; and the C++ host's graph frames (rgrt_frame,
TLAS refit) equal to its eager frames, the last one written as PPM."""
import os
import subprocess

import numpy as np
import pytest

from helpers import gpu_context
from raygun_amd import procgen, rgrt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "raygun_amd", "host", "scene_demo")


def _expected(h16, fmt):
    f = np.asarray(h16, np.uint16).view(np.float16).astype(np.float32)
    f = np.where(np.isnan(f), np.float32(0), f)
    c = np.clip(f, np.float32(0), np.float32(1)).astype(np.float32)
    if fmt in (rgrt.FORMAT_RGBA8_SRGB, rgrt.FORMAT_BGRA8_SRGB):
        rgb = c[..., :3].astype(np.float64)
        enc = np.where(rgb <= 0.0031308, 12.92 * rgb, 1.055 * rgb ** (1 / 2.4) - 0.055)
        c = np.concatenate([enc.astype(np.float32), c[..., 3:]], axis=-1)
    q = np.rint(c * np.float32(255)).astype(np.uint8)
    if fmt in (rgrt.FORMAT_BGRA8_UNORM, rgrt.FORMAT_BGRA8_SRGB):
        q = q[..., [2, 1, 0, 3]]
    return q


def _special_image(h, w):
    rng = np.random.default_rng(5)
    f = rng.uniform(-0.2, 1.2, size=(h, w, 4)).astype(np.float16)
    flat = f.reshape(-1)
    k = np.arange(256)
    flat[:256] = (k / 255.0).astype(np.float16)           # every level
    flat[256:512] = ((k + 0.5) / 255.0).astype(np.float16)  # near the midpoints
    flat[512:520] = np.array([np.nan, np.inf, -np.inf, -0.0, 0.0, 1.0, 65504, -65504], np.float16)
    return f.view(np.uint16)


@pytest.mark.parametrize("fmt", [rgrt.FORMAT_RGBA8_UNORM, rgrt.FORMAT_BGRA8_UNORM, rgrt.FORMAT_RGBA8_SRGB,
                                 rgrt.FORMAT_BGRA8_SRGB])
def test_rgba8_conversion_of_written_image(fmt):
    print(f"[case] rgba8 fmt={fmt}", flush=True)
    ctx = rgrt.Context(40, 30, 0)
    try:
        img = _special_image(30, 40)
        ctx.write_image(rgrt.IMAGE_BASE, img)
        got = ctx.read_image_rgba8(rgrt.IMAGE_BASE, fmt)
        want = _expected(img, fmt)
        d = np.abs(got.astype(int) - want.astype(int))
        if fmt in (rgrt.FORMAT_RGBA8_UNORM, rgrt.FORMAT_BGRA8_UNORM):
            assert d.max() == 0
        else:  # powf against numpy's pow: one level at most, and rarely
            assert d.max() <= 1 and (d > 0).mean() < 0.01
    finally:
        ctx.close()


def test_rgba8_of_a_rendered_frame_and_bad_arguments():
    setup = procgen.config1(160, 90)
    print("[case] rgba8 rendered", flush=True)
    ctx, inst = gpu_context(setup)
    try:
        ctx.render(True)
        want = _expected(ctx.read_image(rgrt.IMAGE_FINAL), rgrt.FORMAT_RGBA8_UNORM)
        assert np.array_equal(ctx.read_image_rgba8(), want)
        with pytest.raises(rgrt.RgrtError):
            ctx.read_image_rgba8(rgrt.IMAGE_ROUGH_TRANSITIONS)
        with pytest.raises(rgrt.RgrtError):
            ctx.read_image_rgba8(rgrt.IMAGE_FINAL, 7)
    finally:
        ctx.close()


def test_cpp_host_graph_frames_equal_eager_frames(tmp_path):
    if not os.path.exists(DEMO):
        pytest.fail(f"{DEMO} is not built; run raygun_amd/build.py or __graft_entry__.build()")
    out = str(tmp_path / "last.ppm")
    print("[case] cpp-host frames 240x135 x6", flush=True)
    r = subprocess.run([DEMO, "frames", "240", "135", "6", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    data = open(out, "rb").read()
    assert data.startswith(b"P6\n240 135\n255\n") and len(data) == len(b"P6\n240 135\n255\n") + 240 * 135 * 3
