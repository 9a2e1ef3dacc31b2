"""rgrt_frame, eager and as one hipGraph: the frame (instance upload, TLAS
build or refit, wavefront passes, drain, post) launched as a graph gives the
same images as the eager calls, on every replay, with the program's own eager
work (ray queries, image reads, synchronisation) between replays, while the
camera, the instances and the render size change between frames."""
import numpy as np
import pytest

from helpers import gpu_context, random_rays, small_config2
from raygun_amd import procgen, rgrt

pytestmark = pytest.mark.gpu

IMAGES = (rgrt.IMAGE_FINAL, rgrt.IMAGE_BASE, rgrt.IMAGE_NORMAL, rgrt.IMAGE_ROUGH)


def _read_all(ctx):
    return [ctx.read_image(k).copy() for k in IMAGES]


def _frames(setup, k):
    """Camera positions and moved instances of k frames."""
    cam = setup.scene.camera
    base_t = cam.transform.copy()
    out = []
    for f in range(k):
        t = base_t.copy()
        t.move((0.15 * f, 0.05 * f, 0.0))
        cam.set_transform(t)
        ubo = setup.ubo()
        inst = setup.scene.collect_instances()
        tr = inst["transform"].reshape(-1, 3, 4)
        tr[:, 1, 3] += 0.1 * f
        inst["transform"] = tr.reshape(-1, 12)
        out.append((ubo, inst))
    cam.set_transform(base_t)
    return out


@pytest.mark.parametrize("which", ["config1", "config2-small"])
def test_graph_frames_equal_eager_frames(which):
    setup = procgen.config1(320, 180) if which == "config1" else small_config2(320, 180, tris_per_mesh=5_000)
    print(f"[case] frame-graph {which}", flush=True)
    eager, _ = gpu_context(setup)
    graph, _ = gpu_context(setup)
    try:
        frames = _frames(setup, 4)
        for f, (ubo, inst) in enumerate(frames):
            refit = f > 0
            eager.frame(ubo, inst, fxaa=True, refit=refit, graph=False)
            want = _read_all(eager)
            graph.frame(ubo, inst, fxaa=True, refit=refit, graph=True)
            got = _read_all(graph)
            for k, (a, b) in zip(IMAGES, zip(got, want)):
                assert np.array_equal(a, b), (f, k)
            assert graph.stats()["rays"] == eager.stats()["rays"]
            # eager work on the graph context between frames
            hits = graph.trace_rays(random_rays(256, seed=f))
            assert len(hits) == 256
            graph.synchronize()
    finally:
        eager.close()
        graph.close()


def test_graph_replays_of_one_frame_are_identical():
    setup = procgen.config1(320, 180)
    print("[case] frame-graph replays", flush=True)
    ctx, inst = gpu_context(setup)
    try:
        ubo = setup.ubo()
        ctx.frame(ubo, inst, fxaa=False, graph=False)
        want = _read_all(ctx)
        for r in range(5):
            ctx.frame(ubo, inst, fxaa=False, graph=True)
            got = _read_all(ctx)
            for k, (a, b) in zip(IMAGES, zip(got, want)):
                assert np.array_equal(a, b), (r, k)
            ctx.render(False)  # an eager frame in between
            got = _read_all(ctx)
            for k, (a, b) in zip(IMAGES, zip(got, want)):
                assert np.array_equal(a, b), ("eager", r, k)
    finally:
        ctx.close()


def test_graph_frame_after_resize_and_settings_change():
    setup = procgen.config1(320, 180)
    print("[case] frame-graph resize", flush=True)
    ctx, inst = gpu_context(setup)
    ref, _ = gpu_context(setup, 200, 120)
    try:
        ctx.frame(setup.ubo(), inst, graph=True)
        ctx.resize(200, 120)
        setup.max_recursions = 3
        ubo = setup.ubo(200, 120)
        ctx.frame(ubo, inst, graph=True)
        ref.frame(ubo, inst, graph=False)
        for k in IMAGES:
            assert np.array_equal(ctx.read_image(k), ref.read_image(k)), k
    finally:
        ctx.close()
        ref.close()


def test_frame_rejects_bad_flags_and_refit_mismatch():
    setup = procgen.config1(64, 64)
    ctx, inst = gpu_context(setup)
    try:
        with pytest.raises(rgrt.RgrtError):
            ctx._ok(ctx._L.rgrt_frame(ctx.h, None, None, 0, 0x100))
        two = np.concatenate([inst, inst])
        two["customIndex"] = [0, 1]
        with pytest.raises(rgrt.RgrtError):
            ctx.frame(setup.ubo(), two, refit=True, graph=True)
        ctx.frame(setup.ubo(), inst, refit=True, graph=True)  # still usable
    finally:
        ctx.close()


@pytest.mark.parametrize("which,bands,spp", [("config2-small", 2, 1), ("config2-small", 3, 1), ("config2-small", 4, 1),
                                             ("config5-small", 3, 4), ("config1-tiny", 4, 1)])
def test_frames_in_bands_equal_the_frame_in_one_piece(which, bands, spp):
    """frameBands: the graph frame traced as runs of tiles on separate
    streams gives the images, ray statistics and per-pass ray counts of the
    frame in one piece, over frames that move the camera and refit the TLAS;
    with several samples per slot group and with bands left empty (a 16x8
    image has two 64-slot chunks for four bands)."""
    if which == "config2-small":
        setup = small_config2(320, 180, tris_per_mesh=5_000)
    elif which == "config5-small":
        setup = procgen.config5(192, 108, tris_per_mesh=3_000)
    else:
        setup = procgen.config1(16, 8)
    setup.num_samples = spp
    print(f"[case] frame-bands {which} bands={bands} spp={spp}", flush=True)
    one, _ = gpu_context(setup)
    banded, _ = gpu_context(setup)
    try:
        banded.set_tuning(frameBands=bands, wavefrontPasses=3)
        one.set_tuning(wavefrontPasses=3)
        assert banded.tuning()["frameBands"] == bands
        for f, (ubo, inst) in enumerate(_frames(setup, 3)):
            refit = f > 0
            one.frame(ubo, inst, fxaa=True, refit=refit, graph=True)
            want = _read_all(one)
            banded.frame(ubo, inst, fxaa=True, refit=refit, graph=True)
            got = _read_all(banded)
            for k, (a, b) in zip(IMAGES, zip(got, want)):
                assert np.array_equal(a, b), (f, k)
            s1, sb = one.stats(), banded.stats()
            for key in ("rays", "skyLookups", "zeroDirRays"):
                assert sb[key] == s1[key], (f, key)
            assert np.array_equal(banded.pass_counts(), one.pass_counts()), f
        # an eager frame in between (in bands on the banded context too), then
        # graph frames in bands again
        ubo, inst = _frames(setup, 1)[0]
        for c in (one, banded):
            c.set_frame(ubo)
            c.render(True)
        assert np.array_equal(banded.read_image(rgrt.IMAGE_FINAL), one.read_image(rgrt.IMAGE_FINAL))
        assert banded.stats()["rays"] == one.stats()["rays"]
        pb, po = banded.pass_counts(), one.pass_counts()
        # the one-piece eager trace stops once its queue is empty; the bands run every pass
        assert len(po) <= len(pb) and np.array_equal(pb[:len(po)], po) and pb[len(po):].sum() == 0
        for c in (one, banded):
            c.frame(ubo, inst, fxaa=True, refit=True, graph=True)
        assert np.array_equal(banded.read_image(rgrt.IMAGE_FINAL), one.read_image(rgrt.IMAGE_FINAL))
    finally:
        one.close()
        banded.close()


def test_frame_bands_range_checked():
    setup = procgen.config1(64, 36)
    print("[case] frame-bands range", flush=True)
    ctx, _ = gpu_context(setup)
    try:
        with pytest.raises(RuntimeError):
            ctx.set_tuning(frameBands=5)
        ctx.set_tuning(frameBands=0)  # one piece
        assert ctx.tuning()["frameBands"] == 0
    finally:
        ctx.close()
