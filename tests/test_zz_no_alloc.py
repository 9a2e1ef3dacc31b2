"""rgrt_rebuild_blas_noalloc (the BLAS rebuild with its working arrays in the
library's fixed per-GPU scratch) against rgrt_rebuild_blas on the same
deformed meshes: byte-equal binary and compressed trees and equal closest
hits, on two calls in a row (the second on a smaller mesh, so stale scratch
contents of the larger one are present), and the oracle's rebuild of the
same vertices; a mesh beyond the scratch's limit is refused."""
import math

import numpy as np
import pytest

from helpers import compare_hits, random_rays
import oracle_py
from raygun_amd import procgen, rgrt
from raygun_amd.records import INSTANCE, default_material
from raygun_amd.scene import Model, pack_models

pytestmark = pytest.mark.gpu


def _wobble(verts, phase):
    v = verts.copy()
    p = v["position"].astype(np.float64)
    w = 1.0 + 0.08 * np.sin(4.0 * p[:, 1] + phase) * np.cos(3.0 * p[:, 0] - phase)
    v["position"] = (p * w[:, None]).astype(np.float32)
    return v


def _tlas(ctx, blas):
    inst = np.zeros(1, INSTANCE)
    inst["transform"] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    inst["mask"] = 0xFF
    inst["blas"] = blas
    ctx.build_tlas(inst)


def test_rebuild_noalloc_equals_rebuild():
    print("[case] rebuild_blas_noalloc vs rebuild_blas", flush=True)
    big = procgen.uv_sphere(360, 180, 1.0)      # ~129k triangles
    small = procgen.torus(64, 16, 1.0, 0.35)    # 2048 triangles
    models = [Model(big, np.array([default_material()])), Model(small, np.array([default_material()]))]
    v, i, m = pack_models(models)
    ctxs = [rgrt.Context(32, 32, 0), rgrt.Context(32, 32, 0)]
    osc = oracle_py.OracleScene(v, i, m)
    try:
        handles = []
        for ctx in ctxs:
            ctx.upload_geometry(v, i)
            ctx.upload_materials(m)
            handles.append([ctx.build_blas(mesh.first_vertex, mesh.index_offset_bytes, mesh.num_faces(), len(mesh.vertices))
                            for mesh in (big, small)])
        ob = [osc.build_blas(mesh.first_vertex, mesh.index_offset_bytes, mesh.num_faces()) for mesh in (big, small)]
        for call, k in enumerate((0, 1)):  # the larger mesh first, then the smaller one
            mesh = (big, small)[k]
            moved = _wobble(mesh.vertices, 0.7 + call)
            for ctx in ctxs:
                ctx.update_vertices(mesh.first_vertex, moved)
            osc.update_vertices(mesh.first_vertex, moved)
            ctxs[0].rebuild_blas(handles[0][k])
            ctxs[1].rebuild_blas_noalloc(handles[1][k])
            osc.rebuild_blas(ob[k])
            (na, la), (nb, lb) = ctxs[0].export_blas(handles[0][k]), ctxs[1].export_blas(handles[1][k])
            no, lo = osc.export_blas(ob[k])
            print(f"[case] call {call}: {mesh.num_faces()} triangles", flush=True)
            assert np.array_equal(la, lb) and np.array_equal(la, lo), call
            assert na.tobytes() == nb.tobytes() == no.tobytes(), call
            wa = ctxs[0].export_blas_wide(handles[0][k], compressed=True)
            wb = ctxs[1].export_blas_wide(handles[1][k], compressed=True)
            assert wa.tobytes() == wb.tobytes() == osc.export_blas_wide(ob[k], compressed=True).tobytes(), call
            for ctx, h in zip(ctxs, handles):
                _tlas(ctx, h[k])
            rays = random_rays(20_000, seed=40 + call, radius=4.0)
            ha, hb = ctxs[0].trace_rays(rays), ctxs[1].trace_rays(rays)
            assert ha.tobytes() == hb.tobytes(), call
            _tlas_o = np.zeros(1, INSTANCE)
            _tlas_o["transform"] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
            _tlas_o["mask"] = 0xFF
            _tlas_o["blas"] = ob[k]
            osc.build_tlas(_tlas_o)
            ho, tie = osc.trace_rays(rays)
            assert len(compare_hits(ha, ho, tie)) == 0, call
            assert (ha["instance"] != 0xFFFFFFFF).sum() > 2000
        # the larger BLAS rebuilt first is unaffected by the second call's use of the scratch
        assert ctxs[1].export_blas(handles[1][0])[0].tobytes() == ctxs[0].export_blas(handles[0][0])[0].tobytes()
    finally:
        for ctx in ctxs:
            ctx.close()


def test_rebuild_noalloc_refuses_beyond_its_scratch():
    mesh = procgen.uv_sphere(520, 260, 1.0)  # ~270k triangles > 262144
    assert mesh.num_faces() > 1 << 18
    models = [Model(mesh, np.array([default_material()]))]
    v, i, m = pack_models(models)
    ctx = rgrt.Context(16, 16, 0)
    try:
        ctx.upload_geometry(v, i)
        ctx.upload_materials(m)
        b = ctx.build_blas(mesh.first_vertex, mesh.index_offset_bytes, mesh.num_faces(), len(mesh.vertices))
        with pytest.raises(rgrt.RgrtError, match="at most 262144"):
            ctx.rebuild_blas_noalloc(b)
        ctx.rebuild_blas(b)  # the allocating form takes it
        ctx.synchronize()
    finally:
        ctx.close()
