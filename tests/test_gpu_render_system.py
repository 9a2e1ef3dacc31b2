"""The Python mirror of the reference's RenderSystem / Raytracer
(raygun_amd/raytracer.py): load_scene + updateUniformBuffer + render gives
the image of the plain C-ABI calls (the mirror traces in three bands, the
plain context in one piece), and render_frame with three frames in flight
(contexts sharing the scene, TLAS refits, one band each) gives the frames of
render_frame one at a time."""
import numpy as np
import pytest

from helpers import gpu_context
from raygun_amd import procgen, rgrt
from raygun_amd.raytracer import RenderSystem

pytestmark = pytest.mark.gpu


def _camera_path(setup, k):
    cam = setup.scene.camera
    base = cam.transform.copy()
    for f in range(k):
        t = base.copy()
        t.move((0.1 * f, 0.03 * f, 0.0))
        cam.set_transform(t)
        yield f
    cam.set_transform(base)


def test_render_equals_plain_calls():
    setup = procgen.config1(192, 108)
    print("[case] python RenderSystem render", flush=True)
    rs = RenderSystem(192, 108)
    ref = None
    try:
        rs.load_scene(setup.scene)
        rs.updateUniformBuffer(setup.scene.camera, max_recursions=setup.max_recursions)
        rs.render(setup.scene)
        got = rs.read(rgrt.IMAGE_FINAL).copy()
        ref, _ = gpu_context(setup)
        ref.render(True)
        assert np.array_equal(got, ref.read_image(rgrt.IMAGE_FINAL))
    finally:
        if ref is not None:
            ref.close()
        rs.close()


def test_frames_in_flight_equal_frames_one_at_a_time():
    setup = procgen.config1(192, 108)
    print("[case] python RenderSystem frames in flight", flush=True)
    out = {}
    for n in (1, 3):
        rs = RenderSystem(192, 108)
        try:
            rs.set_frames_in_flight(n)
            # one frame in flight: traced in three bands; more: one band each
            assert rs.ctx.tuning()["frameBands"] == (3 if n == 1 else 1)
            rs.raytracer.refit_tlas = True
            rs.load_scene(setup.scene)
            rs.updateUniformBuffer(setup.scene.camera, max_recursions=setup.max_recursions)
            imgs = []
            for _ in _camera_path(setup, 6):
                rs.render_frame(setup.scene, graph=True)
                imgs.append(rs.read(rgrt.IMAGE_FINAL).copy())
            out[n] = imgs
        finally:
            rs.close()
    for f, (a, b) in enumerate(zip(out[1], out[3])):
        assert np.array_equal(a, b), f
    assert not np.array_equal(out[1][0], out[1][-1])  # the camera moved
