"""Edge cases of the scene description on the GPU, each against the oracle:
no instances, a one-triangle BLAS, an empty (zero-face) BLAS next to a real
one, instance masks against the cull mask, and an instance whose transform
mirrors its mesh."""
import numpy as np
import pytest

import oracle_py
from helpers import NO_HIT, compare_hits, random_rays
from raygun_amd import rgrt
from raygun_amd.records import INSTANCE, VERTEX, default_material
from raygun_amd.scene import Camera, camera_ubo

pytestmark = pytest.mark.gpu


def _tris_mesh(tris):
    tris = np.asarray(tris, np.float64).reshape(-1, 3, 3)
    v = np.zeros(len(tris) * 3, VERTEX)
    v["position"] = tris.reshape(-1, 3)
    v["normal"] = (0, 0, 1)
    return v, np.arange(len(v), dtype=np.uint32)


def _instance(blas, transform=None, mask=0xFF, custom=0, offsets=(0, 0, 0)):
    r = np.zeros((), INSTANCE)
    r["transform"] = np.eye(4)[:3].reshape(12) if transform is None else np.asarray(transform, np.float32).reshape(12)
    r["mask"] = mask
    r["flags"] = 1
    r["blas"] = blas
    r["customIndex"] = custom
    r["offsets"] = offsets
    return r


def _both(v, i, mats, blas_ranges, instances, W=32, H=16):
    """Same scene in a GPU context and in the oracle."""
    ctx = rgrt.Context(W, H, rgrt.CFG_FP32_RADIANCE | rgrt.CFG_PRIMARY_AOV)
    ctx.upload_geometry(v, i)
    ctx.upload_materials(mats)
    osc = oracle_py.OracleScene(v, i, mats)
    for (first, io, nf) in blas_ranges:
        assert ctx.build_blas(first, io, nf, len(v)) == osc.build_blas(first, io, nf)
    inst = np.array(instances, INSTANCE)
    ctx.build_tlas(inst)
    osc.build_tlas(inst)
    return ctx, osc


def _render_both(ctx, osc, W=32, H=16):
    cam = Camera(fov_deg=60.0)
    cam.look_at((0.2, 0.3, 4.0), (0.0, 0.0, 0.0))
    ubo = camera_ubo(cam, W, H, max_recursions=3)
    ctx.set_frame(ubo)
    ctx.trace()
    g = ctx.read_image(rgrt.IMAGE_BASE_F32)
    gp = ctx.read_image(rgrt.IMAGE_PRIMARY_HIT)
    o = osc.render(ubo, W, H)
    ok = ~o["tie"].astype(bool)
    assert np.array_equal(gp["instance"][ok], o["primary"]["instance"][ok])
    assert np.abs(g - o["base32"])[ok].max() <= 1e-3
    return gp


def test_no_instances_renders_sky():
    print("[case] edge no-instances", flush=True)
    v, i = _tris_mesh([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]])
    mats = np.array([default_material()])
    ctx, osc = _both(v, i, mats, [(0, 0, 1)], [])
    try:
        gp = _render_both(ctx, osc)
        assert np.all(gp["instance"] == NO_HIT)
        h = ctx.trace_rays(random_rays(1000, seed=1))
        assert np.all(h["instance"] == NO_HIT)
    finally:
        ctx.close()


def test_one_triangle_and_empty_blas():
    print("[case] edge one-triangle + empty BLAS", flush=True)
    v, i = _tris_mesh([[[-1, -1, 0], [1, -1, 0], [0, 1, 0]], [[0, 0, 0], [1, 0, 0], [0, 1, 0]]])
    mats = np.array([default_material(diffuse=(0.2, 0.7, 0.3))])
    # BLAS 0: the first triangle; BLAS 1: zero faces
    ctx, osc = _both(v, i, mats, [(0, 0, 1), (0, 12, 0)], [_instance(1, custom=0), _instance(0, custom=1)])
    try:
        gp = _render_both(ctx, osc)
        assert np.any(gp["instance"] == 1) and not np.any(gp["instance"] == 0)
        rays = random_rays(20_000, seed=2, radius=4.0)
        hg = ctx.trace_rays(rays)
        ho, tie = osc.trace_rays(rays)
        assert len(compare_hits(hg, ho, tie)) == 0
        assert (hg["instance"] != NO_HIT).sum() > 100
    finally:
        ctx.close()


def test_instance_masks_and_cull_masks():
    print("[case] edge instance masks", flush=True)
    v, i = _tris_mesh([[[-1, -1, 0], [1, -1, 0], [0, 1, 0]]])
    mats = np.array([default_material()])
    far = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -2.0]  # the same triangle 2 units behind
    ctx, osc = _both(v, i, mats, [(0, 0, 1)], [_instance(0, mask=0x01, custom=0), _instance(0, far, mask=0x02, custom=1)])
    try:
        rays = np.zeros(64, rgrt.RAY)
        rays["origin"] = np.stack([np.linspace(-0.3, 0.3, 64), np.full(64, -0.2), np.full(64, 3.0)], axis=1)
        rays["dir"] = (0, 0, -1)
        rays["tmin"], rays["tmax"] = 0.001, 100.0
        for cull, expect in [(0xFF, 0), (0x01, 0), (0x02, 1), (0x04, NO_HIT), (0x00, NO_HIT)]:
            hg = ctx.trace_rays(rays, cull_mask=cull)
            ho, _ = osc.trace_rays(rays, cull_mask=cull)
            assert np.all(hg["instance"] == expect), (cull, hg["instance"][:4])
            assert np.array_equal(hg["instance"], ho["instance"]) and np.array_equal(hg["t"], ho["t"])
    finally:
        ctx.close()


def test_mirrored_instance_transform():
    print("[case] edge mirrored instance", flush=True)
    v, i = _tris_mesh([[[-1, -1, 0], [1, -1, 0], [0, 1, 0]], [[0.5, 0.5, 0.3], [1.5, 0.5, 0.3], [0.5, 1.5, 0.3]]])
    mats = np.array([default_material(diffuse=(0.9, 0.4, 0.1), reflectivity=0.3)])
    mirror = [-1, 0, 0, 0.2, 0, 1, 0, 0, 0, 0, 1, 0]  # x -> -x: negative determinant
    ctx, osc = _both(v, i, mats, [(0, 0, 2)], [_instance(0, mirror)])
    try:
        _render_both(ctx, osc)
        rays = random_rays(20_000, seed=3, radius=4.0)
        hg = ctx.trace_rays(rays)
        ho, tie = osc.trace_rays(rays)
        assert len(compare_hits(hg, ho, tie)) == 0
    finally:
        ctx.close()


def test_box_test_keeps_hits_where_child_boxes_meet_far_from_the_node_corner():
    """rgrt_trace_rays finds exactly what FP32 brute force over every
    triangle finds (the oracle's watertight test, no tree) on rays that hit a
    ground grid at x, y ~ 3000 on the seams between child boxes, close to
    their origins, where the box test's plane distances carry an absolute
    rounding error that only the absolute part of its far margin covers."""
    from helpers import seam_scene
    v, idx, inst, rays = seam_scene()
    inst["blas"] = 0
    mats = np.array([default_material()])
    ctx, osc = _both(v, idx, mats, [(0, 0, len(idx) // 3)], list(inst))
    try:
        h = ctx.trace_rays(rays)
        t_b, inst_b, _ = osc.trace_rays_brute32(rays)
        assert (inst_b != NO_HIT).all()
        assert np.array_equal(h["instance"] != NO_HIT, inst_b != NO_HIT)
        assert np.array_equal(h["t"].view(np.uint32), t_b.view(np.uint32))
    finally:
        ctx.close()


def _fan(n=6, seed=3):
    """n triangles (n / 2 quads side by side) tiling [-0.6, 0.6] x [-0.5,
    0.5], each with vertices of its own, distinct vertex normals and vertex 0
    material indices 0 / 1 in turn, then one more triangle (three more
    vertices and indices) that no BLAS covers, so that shifted offset tables
    still address the buffers."""
    rng = np.random.default_rng(seed)
    v = np.zeros(3 * n + 3, VERTEX)
    q = n // 2
    for k in range(n):
        x0 = -0.6 + 1.2 * (k // 2) / q
        x1 = -0.6 + 1.2 * (k // 2 + 1) / q
        tri = [(x0, -0.5, 0.0), (x1, -0.5, 0.0), (x1, 0.5, 0.0)] if k % 2 == 0 else \
              [(x0, -0.5, 0.0), (x1, 0.5, 0.0), (x0, 0.5, 0.0)]
        v["position"][3 * k:3 * k + 3] = tri
        nrm = rng.normal(size=(3, 3)) * 0.3 + (0, 0, 1)
        v["normal"][3 * k:3 * k + 3] = nrm / np.linalg.norm(nrm, axis=1)[:, None]
        v["matIndex"][3 * k] = k % 2
    v["position"][3 * n:] = [(5.0, 5.0, 5.0), (5.1, 5.0, 5.0), (5.0, 5.1, 5.0)]
    v["normal"][3 * n:] = (0.0, 0.0, 1.0)
    return v, np.arange(3 * n + 3, dtype=np.uint32)


def test_shading_records_fall_back_and_follow_vertex_updates():
    """Hits read a per-triangle shading record only when the instance's offset
    table entry matches the vertex offset it was gathered with; otherwise
    (another vertex offset, an index offset that is not a whole triangle)
    the vertices are gathered as the closest-hit stage describes. Normals and
    material indices changed after the build, without a refit, from the host
    and from device memory, are seen by the next frame."""
    print("[case] edge shading records", flush=True)
    from raygun_amd.hiptime import DeviceBuffer
    v, i = _fan()
    red, green, blue = default_material(), default_material(), default_material()
    red["diffuse"] = (0.9, 0.1, 0.1)
    green["diffuse"] = (0.1, 0.8, 0.2)
    blue["diffuse"] = (0.1, 0.2, 0.9)
    mats = np.array([red, green, blue])
    move = lambda x: np.array([[1, 0, 0, x], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)
    instances = [_instance(0, move(-1.3), custom=0, offsets=(0, 0, 0)),   # the record's vertex offset
                 _instance(0, move(0.0), custom=1, offsets=(3, 0, 1)),    # another vertex offset: gathered
                 _instance(0, move(1.3), custom=2, offsets=(0, 1, 0))]    # index offset 1: gathered
    ctx, osc = _both(v, i, mats, [(0, 0, len(i) // 3 - 1)], instances, W=48, H=24)
    buf = None
    try:
        gp = _render_both(ctx, osc, W=48, H=24)
        assert set(np.unique(gp["instance"])) >= {0, 1, 2}
        # normals and materials rewritten from the host, positions kept
        v2 = v.copy()
        v2["normal"] = v["normal"][::-1]
        v2["matIndex"] = 1 - v["matIndex"]
        ctx.update_vertices(0, v2)
        osc.update_vertices(0, v2)
        _render_both(ctx, osc, W=48, H=24)
        # and from device memory
        v3 = v2.copy()
        v3["normal"][:9] = v["normal"][:9]
        v3["matIndex"][:9] = v["matIndex"][:9]
        buf = DeviceBuffer(9 * 32)
        buf.upload(v3[:9])
        ctx.update_vertices_device(0, buf.address, 9)
        ctx.synchronize()
        osc.update_vertices(0, v3[:9])
        _render_both(ctx, osc, W=48, H=24)
    finally:
        ctx.close()
        if buf is not None:
            buf.close()
