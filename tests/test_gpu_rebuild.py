"""Deforming meshes on the GPU: rgrt_update_vertices_device writes vertices
from device memory and rgrt_rebuild_blas rebuilds a BLAS in place, both
ordered on the context's stream. The rebuilt tree is byte-identical to a
fresh rgrt_build_blas over the same vertices, frames traced after it equal a
context built from those vertices, and a context sharing the scene sees the
new geometry in its next frame."""
import numpy as np
import pytest

from helpers import gpu_context
from raygun_amd import procgen, rgrt
from raygun_amd.hiptime import DeviceBuffer
from raygun_amd.scene import pack_models

pytestmark = pytest.mark.gpu


def _wobble(v, phase):
    out = v.copy()
    p = out["position"].astype(np.float64)
    w = 1.0 + 0.08 * np.sin(4.0 * p[:, 1] + phase)
    out["position"] = (p * w[:, None]).astype(np.float32)
    return out


def _trees(ctx, h):
    nodes, leaves = ctx.export_blas(h)
    return nodes.tobytes(), leaves.tobytes()


def test_rebuild_equals_fresh_build():
    setup = procgen.config1(192, 108)
    print("[case] rebuild_blas vs build_blas", flush=True)
    a, inst = gpu_context(setup)
    models = setup.scene.models()
    mesh = models[0].mesh
    v, i, m = pack_models(models)
    first, count = mesh.first_vertex, len(mesh.vertices)
    buf = None
    fresh = None
    try:
        moved = v.copy()
        moved[first:first + count] = _wobble(v[first:first + count], 0.7)
        buf = DeviceBuffer(count * 32)
        buf.upload(moved[first:first + count])
        a.update_vertices_device(first, buf.address, count)
        a.rebuild_blas(models[0].blas)
        a.build_tlas(inst)
        a.render(True)
        got = a.read_image(rgrt.IMAGE_FINAL).copy()
        # a context built from the moved vertices
        fresh = rgrt.Context(setup.width, setup.height, 0)
        fresh.upload_geometry(moved, i)
        fresh.upload_materials(m)
        h = fresh.build_blas(first, mesh.index_offset_bytes, mesh.num_faces(), count)
        assert _trees(a, models[0].blas) == _trees(fresh, h)
        assert np.array_equal(a.export_blas_wide(models[0].blas, compressed=True), fresh.export_blas_wide(h, compressed=True))
        inst2 = inst.copy()
        inst2["blas"] = h
        fresh.build_tlas(inst2)
        fresh.set_frame(setup.ubo())
        fresh.render(True)
        assert np.array_equal(got, fresh.read_image(rgrt.IMAGE_FINAL))
    finally:
        if fresh is not None:
            fresh.close()
        a.close()
        if buf is not None:
            buf.close()


def test_rebuild_reaches_sharing_context():
    setup = procgen.config1(192, 108)
    print("[case] rebuild_blas with a sharing context", flush=True)
    a, inst = gpu_context(setup)
    b = a.shared()
    models = setup.scene.models()
    mesh = models[0].mesh
    v, i, m = pack_models(models)
    first, count = mesh.first_vertex, len(mesh.vertices)
    bufs = []
    try:
        b.build_tlas(inst)
        b.set_frame(setup.ubo())
        want = []
        for k in range(3):
            moved = v.copy()
            moved[first:first + count] = _wobble(v[first:first + count], 0.9 * k)
            ref, rinst = None, None
            try:
                ref = rgrt.Context(setup.width, setup.height, 0)
                ref.upload_geometry(moved, i)
                ref.upload_materials(m)
                h = ref.build_blas(first, mesh.index_offset_bytes, mesh.num_faces(), count)
                rinst = inst.copy()
                rinst["blas"] = h
                ref.build_tlas(rinst)
                ref.set_frame(setup.ubo())
                ref.render(True)
                want.append(ref.read_image(rgrt.IMAGE_FINAL).copy())
            finally:
                if ref is not None:
                    ref.close()
            buf = DeviceBuffer(count * 32)
            buf.upload(moved[first:first + count])
            bufs.append(buf)
        for k in range(3):
            b.render(True)  # a frame of b in flight while a rewrites the scene
            a.update_vertices_device(first, bufs[k].address, count)
            a.rebuild_blas(models[0].blas)
            b.build_tlas(inst)  # the new root box
            b.render(True)
            assert np.array_equal(b.read_image(rgrt.IMAGE_FINAL), want[k]), k
    finally:
        b.close()
        a.close()
        for buf in bufs:
            buf.close()


def test_device_update_range_checked():
    setup = procgen.config1(64, 36)
    print("[case] update_vertices_device range", flush=True)
    a, _ = gpu_context(setup)
    buf = DeviceBuffer(64)
    try:
        nv = sum(len(mm.mesh.vertices) for mm in setup.scene.models())
        with pytest.raises(rgrt.RgrtError):
            a.update_vertices_device(nv - 1, buf.address, 2)
        with pytest.raises(rgrt.RgrtError):
            a.rebuild_blas(999)
    finally:
        a.close()
        buf.close()
