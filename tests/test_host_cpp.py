"""The C++ host side (raygun_amd/host: Raytracer / RenderSystem / TopLevelAS
over the C ABI) against the Python mirror on the scene of scene_demo.cpp
(tests/host_scene.py builds the same scene in Python).

CPU: the DFS over the entity tree (hidden and zero-volume subtrees, an entity
without a model, shared meshes), the buffer packing offsets and the frame
constants. GPU: a frame rendered by the C++ program equals, byte for byte, the
frame the Python mirror renders with the same frame constants.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from host_scene import demo_scene
from raygun_amd.records import UBO
from raygun_amd.scene import camera_ubo, pack_models

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "raygun_amd", "host", "scene_demo")


def _demo(*args, **kw):
    if not os.path.exists(DEMO):
        pytest.fail(f"{DEMO} is not built; run raygun_amd/build.py or __graft_entry__.build()")
    return subprocess.run([DEMO, *args], capture_output=True, text=True, timeout=120, **kw)


def _python_side():
    sc = demo_scene()
    models = sc.models()
    v, i, m = pack_models(models)
    handles = {}
    for mo in models:  # handles in first-use order of the distinct meshes
        handles.setdefault(id(mo.mesh), len(handles))
        mo.blas = handles[id(mo.mesh)]
    return sc, (v, i, m), sc.collect_instances()


def test_cpp_host_collects_the_same_instances_and_offsets():
    r = _demo("collect")
    assert r.returncode == 0, r.stderr
    c = json.loads(r.stdout)
    sc, (v, i, m), inst = _python_side()
    assert (c["vertices"], c["indices"], c["materials"]) == (len(v), len(i), len(m))
    # ground, cube1, cube2, cube4: "hidden" (and its child) and the zero-volume "flat" are skipped
    assert len(c["instances"]) == len(inst) == 4
    for k, rec in enumerate(c["instances"]):
        assert rec["customIndex"] == k == int(inst[k]["customIndex"])
        assert rec["mask"] == 0xFF and rec["flags"] == 1
        assert rec["blas"] == int(inst[k]["blas"])
        assert rec["offsets"] == [int(x) for x in inst[k]["offsets"]]
        assert np.array_equal(np.array(rec["transform"], np.float32), inst[k]["transform"])
    ubo = camera_ubo(sc.camera, 320, 180, num_samples=1, max_recursions=3)
    assert np.array_equal(np.array(c["viewInverse"], np.float32), ubo["viewInverse"])
    assert np.array_equal(np.array(c["lightDir"], np.float32), ubo["lightDir"])
    # numpy's LU leaves 1e-16 residues where the exact inverse has zeros
    assert np.allclose(np.array(c["projInverse"], np.float32), ubo["projInverse"], rtol=2e-7, atol=1e-12)


@pytest.mark.gpu
def test_cpp_host_frame_equals_python_mirror_frame(tmp_path):
    out = str(tmp_path / "frame.bin")
    print(f"[case] cpp-host-render 320x180", flush=True)
    r = _demo("render", "320", "180", out)
    assert r.returncode == 0, r.stderr
    cpp = np.fromfile(out, np.uint16).reshape(180, 320, 4)
    ubo = np.fromfile(out + ".ubo", UBO)[0]
    from raygun_amd import rgrt
    sc = demo_scene()
    models = sc.models()
    v, i, m = pack_models(models)
    ctx = rgrt.Context(320, 180, 0)
    try:
        ctx.upload_geometry(v, i)
        ctx.upload_materials(m)
        handles = {}
        for mo in models:
            if id(mo.mesh) not in handles:
                handles[id(mo.mesh)] = ctx.build_blas(mo.mesh.first_vertex, mo.mesh.index_offset_bytes, mo.mesh.num_faces(),
                                                      len(mo.mesh.vertices))
            mo.blas = handles[id(mo.mesh)]
        ctx.build_tlas(sc.collect_instances())
        ctx.set_frame(ubo)
        ctx.render(True)
        py = ctx.read_image(rgrt.IMAGE_FINAL)
    finally:
        ctx.close()
    assert np.array_equal(cpp, py.reshape(180, 320, 4))
    assert np.count_nonzero(cpp) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("in_flight", [1, 3])
def test_cpp_host_graph_frames_equal_eager_frames(tmp_path, in_flight):
    """RenderSystem::renderFrame (one hipGraph per frame, TLAS refits) over
    animated frames equals the eager frames byte for byte, also with three
    frames in flight on contexts sharing the scene; the last frames of both
    runs are identical."""
    print(f"[case] cpp-host-frames in_flight={in_flight}", flush=True)
    out = str(tmp_path / f"frames{in_flight}.ppm")
    r = _demo("frames", "160", "96", "7", out, str(in_flight))
    assert r.returncode == 0, r.stderr
    assert f"{in_flight} in flight, graph == eager" in r.stdout
    if in_flight == 3:
        ref = str(tmp_path / "frames_ref.ppm")
        r1 = _demo("frames", "160", "96", "7", ref, "1")
        assert r1.returncode == 0, r1.stderr
        with open(out, "rb") as a, open(ref, "rb") as b:
            assert a.read() == b.read()
