"""The batched workload (log-normal scene sizes, mixed resolutions, turning
instances refitted into the TLAS every frame, one scene in four deforming
with a full LBVH rebuild per frame) at reduced sizes, rendered with
rgrt_batch_frame and compared scene by scene, frame by frame, with the
oracle following the same sequence (build, then refits; deformed vertices,
then an in-place rebuild): BLAS and TLAS byte-equal, fp32 radiance within
1e-3, primary instance/primitive/t identical off near-ties, ray counts, and
the post-processed images."""
import numpy as np
import pytest

from helpers import NO_HIT
import oracle_py
from raygun_amd import batched, rgrt

pytestmark = pytest.mark.gpu

RES = [(96, 54), (160, 90), (128, 72), (192, 108), (64, 64)]
FRAMES = 3


class _Twin:
    """The oracle's copy of one batch scene."""

    def __init__(self, s):
        self.osc = oracle_py.OracleScene(s.vertices, s.indices, s.materials)
        m, g = s.mesh, s.ground
        self.mesh_blas = self.osc.build_blas(m.first_vertex, m.index_offset_bytes, m.num_faces())
        self.ground_blas = self.osc.build_blas(g.first_vertex, g.index_offset_bytes, g.num_faces())

    def instances(self, s, t):
        inst = s.instances(t)
        inst["blas"] = np.where(inst["blas"] == s.mesh_blas, self.mesh_blas, self.ground_blas)
        return inst

    def step(self, s, frame, t):
        if s.deforms and frame:
            self.osc.update_vertices(s.mesh.first_vertex, s.pose_of_frame(frame))
            self.osc.rebuild_blas(self.mesh_blas)
        inst = self.instances(s, t)
        if frame == 0:
            self.osc.build_tlas(inst)
        else:
            self.osc.refit_tlas(inst)


@pytest.mark.parametrize("mode", [rgrt.BATCH_GRAPH_PER_SCENE, rgrt.BATCH_ONE_GRAPH, rgrt.BATCH_PERSISTENT])
def test_batched_frames_match_oracle(mode):
    flags = rgrt.CFG_FP32_RADIANCE | rgrt.CFG_PRIMARY_AOV
    w = batched.BatchWorkload(num_scenes=8, seed=21, lo=3_000, hi=80_000, resolutions=RES, mode=mode, flags=flags)
    print(f"[case] batch-vs-oracle mode {mode} scenes {w.info['triangles']} res {w.info['resolutions']}", flush=True)
    try:
        assert any(s.deforms for s in w.scenes)
        twins = [_Twin(s) for s in w.scenes]
        for f in range(FRAMES):
            t = 0.35 * f
            for s, tw in zip(w.scenes, twins):
                tw.step(s, f, t)
            w.frame_batched(t)
            w.batch.synchronize()
            w.synchronize()
            for k, (s, tw) in enumerate(zip(w.scenes, twins)):
                print(f"[case] frame {f} scene {k} {s.width}x{s.height} tris {s.mesh.num_faces()} deforms {s.deforms}",
                      flush=True)
                if s.deforms:
                    g_nodes, g_leaf = s.ctx.export_blas(s.mesh_blas)
                    o_nodes, o_leaf = tw.osc.export_blas(tw.mesh_blas)
                    assert np.array_equal(g_leaf, o_leaf), (f, k)
                    assert g_nodes.tobytes() == o_nodes.tobytes(), (f, k)
                g_tn, g_tl = s.ctx.export_tlas()
                o_tn, o_tl = tw.osc.export_tlas()
                assert np.array_equal(g_tl, o_tl) and g_tn.tobytes() == o_tn.tobytes(), (f, k)
                o = tw.osc.render(s.ubo, s.width, s.height)
                ok = ~o["tie"].astype(bool)
                base = s.ctx.read_image(rgrt.IMAGE_BASE_F32)
                fin = np.isfinite(o["base32"])
                assert np.array_equal(np.isfinite(base), fin), (f, k)
                d = np.where(fin, np.abs(base - o["base32"]), 0.0).max(axis=-1)
                assert d[ok].max() <= 1e-3, (f, k, float(d[ok].max()))
                pg, po = s.ctx.read_image(rgrt.IMAGE_PRIMARY_HIT), o["primary"]
                assert np.array_equal(pg["instance"][ok], po["instance"][ok]), (f, k)
                assert np.array_equal(pg["primitive"][ok], po["primitive"][ok]), (f, k)
                hit = ok & (po["instance"] != NO_HIT)
                assert np.array_equal(pg["t"][hit], po["t"][hit]), (f, k)
                st = s.ctx.stats()
                if o["tie"].sum() == 0:
                    assert st["rays"] == o["rays"], (f, k, st["rays"], o["rays"])
                ref = oracle_py.postprocess(s.ubo, o["base16"], o["normal16"], o["rough16"], use_fxaa=True)
                same = np.mean(np.all(s.ctx.read_image(rgrt.IMAGE_FINAL) == ref["final"], axis=-1))
                assert same >= 0.999, (f, k, same)
    finally:
        w.close()
