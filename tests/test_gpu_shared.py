"""rgrt_create_shared: two contexts over one scene (geometry, materials,
BLASes), each with its own TLAS, images and stream, so consecutive frames
overlap on the GPU. Frames rendered alternately on the pair, enqueued before
either is read back, equal the same frames rendered one at a time on two
independent contexts; scene changes made through either context reach both."""
import numpy as np
import pytest

from helpers import gpu_context, small_config2
from raygun_amd import procgen, rgrt
from raygun_amd.scene import pack_models

pytestmark = pytest.mark.gpu

IMAGES = (rgrt.IMAGE_FINAL, rgrt.IMAGE_BASE, rgrt.IMAGE_NORMAL, rgrt.IMAGE_ROUGH)


def _read_all(ctx):
    return [ctx.read_image(k).copy() for k in IMAGES]


def _frames(setup, k):
    cam = setup.scene.camera
    base_t = cam.transform.copy()
    out = []
    for f in range(k):
        t = base_t.copy()
        t.move((0.12 * f, 0.04 * f, 0.0))
        cam.set_transform(t)
        inst = setup.scene.collect_instances()
        tr = inst["transform"].reshape(-1, 3, 4)
        tr[:, 1, 3] += 0.08 * f
        inst["transform"] = tr.reshape(-1, 12)
        out.append((setup.ubo(), inst))
    cam.set_transform(base_t)
    return out


@pytest.mark.parametrize("graph,bands", [(False, 1), (True, 1), (True, 2), (False, 3)])
@pytest.mark.parametrize("which", ["config1", "config2-small"])
def test_alternating_shared_frames_equal_single_context(which, graph, bands):
    """bands > 1: the pair's frames (graph or eager) are traced in bands on
    streams of their own while the other context's frame runs (frameBands)."""
    setup = procgen.config1(320, 180) if which == "config1" else small_config2(320, 180, tris_per_mesh=5_000)
    print(f"[case] shared-frames {which} graph={graph} bands={bands}", flush=True)
    # reference: two independent contexts, each frame read back before the
    # next starts (each context builds its TLAS on its first frame, then refits)
    own = [gpu_context(setup)[0], gpu_context(setup)[0]]
    frames = _frames(setup, 6)  # after the BLAS handles are set on the models
    try:
        want = []
        for f, (ubo, inst) in enumerate(frames):
            own[f % 2].frame(ubo, inst, fxaa=True, refit=f > 1, graph=graph)
            want.append((_read_all(own[f % 2]), own[f % 2].stats()["rays"]))
    finally:
        for c in own:
            c.close()
    a, _ = gpu_context(setup)
    b = a.shared()
    try:
        pair = (a, b)
        for c in pair:
            c.set_tuning(frameBands=bands)
        for f in range(0, len(frames), 2):
            # frame f on a and frame f + 1 on b are both queued before either is read
            for k in range(2):
                ubo, inst = frames[f + k]
                pair[k].frame(ubo, inst, fxaa=True, refit=f > 0, graph=graph)
            for k in range(2):
                got = _read_all(pair[k])
                for img, (x, y) in zip(IMAGES, zip(got, want[f + k][0])):
                    assert np.array_equal(x, y), (f + k, img)
                assert pair[k].stats()["rays"] == want[f + k][1]
    finally:
        b.close()
        a.close()


def test_scene_changes_reach_every_sharing_context():
    setup = procgen.config1(256, 144)
    print("[case] shared-scene changes", flush=True)
    a, inst = gpu_context(setup)
    b = a.shared()
    ref = None
    try:
        ubo = setup.ubo()
        b.build_tlas(inst)
        b.set_frame(ubo)
        b.render(True)
        a.render(True)
        assert np.array_equal(b.read_image(rgrt.IMAGE_FINAL), a.read_image(rgrt.IMAGE_FINAL))
        # new materials through b (while a has a frame in flight) show in a's next frame
        mats = pack_models(setup.scene.models())[2].copy()
        mats["diffuse"] = mats["diffuse"][:, ::-1]
        mats["reflectivity"] = 0.5
        a.render(True)
        b.upload_materials(mats)
        a.render(True)
        got = a.read_image(rgrt.IMAGE_FINAL).copy()
        ref, _ = gpu_context(setup)
        ref.upload_materials(mats)
        ref.render(True)
        assert np.array_equal(got, ref.read_image(rgrt.IMAGE_FINAL))
        # a BLAS built through b is a handle a can use
        mesh = setup.scene.models()[0].mesh
        h = b.build_blas(mesh.first_vertex, mesh.index_offset_bytes, mesh.num_faces(), len(mesh.vertices))
        inst2 = inst.copy()
        inst2["blas"] = h
        a.build_tlas(inst2)
        a.render(True)
        assert np.array_equal(a.read_image(rgrt.IMAGE_FINAL), got)
        # new geometry through a drops b's TLAS too
        v, i, _ = pack_models(setup.scene.models())
        a.upload_geometry(v, i)
        with pytest.raises(rgrt.RgrtError):
            b.render(True)
        with pytest.raises(rgrt.RgrtError):
            a.render(True)
    finally:
        if ref is not None:
            ref.close()
        b.close()
        a.close()


def test_parent_destroyed_first():
    setup = procgen.config1(256, 144)
    print("[case] shared parent destroyed first", flush=True)
    a, inst = gpu_context(setup)
    a.render(True)
    want = a.read_image(rgrt.IMAGE_FINAL).copy()
    b = a.shared()
    a.close()
    try:
        b.build_tlas(inst)
        b.set_frame(setup.ubo())
        b.render(True)
        assert np.array_equal(b.read_image(rgrt.IMAGE_FINAL), want)
    finally:
        b.close()
