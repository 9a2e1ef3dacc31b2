"""Host logic of the learned-path driver (raygun_amd/denoise.py), CPU only:
a window's launch traces the same primary rays as the matching pixels of
the whole frame (the oracle's raygen on both), the equal-PSNR interpolation
on known points, and the executed multiply-accumulates per pixel against a
count taken layer by layer from the network's shapes."""
import numpy as np
import pytest

import oracle_py
import unet_oracle as U
from raygun_amd import denoise as D
from raygun_amd.scene import Camera, camera_ubo


@pytest.mark.parametrize("x0,y0,w,h", [(0, 0, 128, 128), (600, 333, 128, 128), (1792, 952, 128, 128), (37, 5, 64, 96)])
def test_window_rays_equal_the_frames_rays(x0, y0, w, h):
    W, H = 1920, 1080
    cam = Camera(fov_deg=55.0)
    cam.look_at((3.0, 1.0, 4.0), (0.0, -0.2, 0.0))
    full = camera_ubo(cam, W, H, num_samples=1, max_recursions=2)
    win = D.window_ubo(full, W, H, x0, y0, w, h)
    for (i, j) in [(0, 0), (w - 1, h - 1), (w // 2, 3), (5, h // 2)]:
        of, df = oracle_py.primary_ray(full, x0 + i, y0 + j, W, H)
        ow, dw = oracle_py.primary_ray(win, i, j, w, h)
        assert np.array_equal(of, ow)
        assert np.abs(df - dw).max() < 2e-6, (i, j, df, dw)


def test_equal_psnr_interpolation():
    pts = [(30.0, 10.0), (33.0, 20.0), (36.0, 40.0)]
    assert D._equal_psnr_ms(pts, 33.0) == pytest.approx(20.0)
    assert D._equal_psnr_ms(pts, 34.5) == pytest.approx(np.sqrt(800.0))   # geometric between 20 and 40
    assert D._equal_psnr_ms(pts, 39.0) == pytest.approx(80.0)              # one more doubling past the last
    assert D._equal_psnr_ms(pts, 27.0) == pytest.approx(5.0)               # one halving before the first


def test_macs_per_pixel_from_the_layer_shapes():
    scale = [1, 1, 4, 4, 16, 16, 4, 4, 1, 1, 1]  # pixels of each layer relative to the input
    macs = sum(9 * cin * cout / s for (cin, cout), s in zip(U.LAYERS, scale))
    assert D.MACS_PER_PIXEL == macs == 32256
