"""Batched scenes (rgrt_batch_frame: a hipGraph per scene on its own stream,
one hipGraph with a branch per scene, or one hipGraph tracing every scene in
shared launches) against the same scenes rendered one by one with the eager
calls: identical
final images per scene over animated frames (TLAS refits, a deforming mesh
refitted), a refused frame naming its scene, and the autotuner changing only
timing, never a frame."""
import numpy as np
import pytest

from raygun_amd import autotune, batched, rgrt

pytestmark = pytest.mark.gpu

RES = [(160, 90), (96, 64), (200, 100), (64, 64)]


def _workload(mode=rgrt.BATCH_GRAPH_PER_SCENE):
    return batched.BatchWorkload(num_scenes=4, seed=5, lo=2_000, hi=30_000, resolutions=RES, mode=mode)


@pytest.mark.parametrize("mode", [rgrt.BATCH_GRAPH_PER_SCENE, rgrt.BATCH_ONE_GRAPH, rgrt.BATCH_PERSISTENT])
def test_batched_frames_equal_sequential_frames(mode):
    print(f"[case] batch 4 scenes x 3 frames mode {mode}", flush=True)
    a, b = _workload(mode), _workload()
    try:
        assert a.info == b.info
        for f in range(3):
            t = 0.3 * f
            a.frame_batched(t)
            b.frame_sequential(t)
            for k, (sa, sb) in enumerate(zip(a.scenes, b.scenes)):
                ia, ib = sa.ctx.read_image(rgrt.IMAGE_FINAL), sb.ctx.read_image(rgrt.IMAGE_FINAL)
                assert np.array_equal(ia, ib), (f, k)
            assert a.rays() == b.rays() > 0
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("mode", [rgrt.BATCH_GRAPH_PER_SCENE, rgrt.BATCH_ONE_GRAPH, rgrt.BATCH_PERSISTENT])
def test_batch_refuses_a_bad_scene_and_recovers(mode):
    print(f"[case] batch error path mode {mode}", flush=True)
    w = _workload(mode)
    try:
        w.frame_batched(0.0)
        descs = w.descs(0.1)
        ubo, inst, fxaa, refit = descs[2]
        descs[2] = (ubo, inst[:-1], fxaa, True)  # refit with one instance fewer than the build
        with pytest.raises(rgrt.RgrtError, match="scene 2"):
            w.batch.frame(descs)
        w.frame_batched(0.2)
        w.synchronize()
    finally:
        w.close()


def test_autotune_changes_timing_not_frames():
    print("[case] autotune", flush=True)
    w = batched.BatchWorkload(num_scenes=1, seed=9, lo=20_000, hi=20_000, resolutions=[(160, 96)])
    try:
        s = w.scenes[0]
        inst = s.instances(0.0)
        s.ctx.frame(s.ubo, inst, fxaa=True)
        before = s.ctx.read_image(rgrt.IMAGE_FINAL).copy()
        cand = {"wavefrontPasses": [2, 6], "traceBlocksPerCU": [4, 8], "hostLag": [1, 2]}
        best, ms, table = autotune.tune(s.ctx, lambda: s.ctx.frame(s.ubo, inst, fxaa=True), candidates=cand, reps=1)
        assert len(table) >= 4 and ms > 0
        assert s.ctx.tuning() == best
        s.ctx.frame(s.ubo, inst, fxaa=True)
        assert np.array_equal(s.ctx.read_image(rgrt.IMAGE_FINAL), before)
    finally:
        w.close()
