"""Per-frame scene update: the transform editing helpers (transform.hpp:71-131),
the cached parent transforms (entity.cpp:180-270), the animations
(entity.cpp:300-332) and the main loop's update step (raygun.cpp:122-126),
in the Python mirror and the C++ host side; and the TLAS refit of the oracle,
pinned against a from-scratch build of the same instances.

CPU only.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_py
from host_scene import demo_scene
from raygun_amd import procgen
from raygun_amd.records import INSTANCE, RAY, default_material
from raygun_amd.scene import AnimatableEntity, Entity, ScaleAnimation, Transform, interpolate, quat_to_mat3

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "raygun_amd", "host", "scene_demo")


def test_rotations_follow_glm_closed_forms():
    t = Transform()
    t.rotate(math.pi / 2, (0, 0, 2))  # axis normalised: a quarter turn about +z
    assert np.allclose(t.to_mat4()[:3, :3] @ [1, 0, 0], [0, 1, 0], atol=1e-15)
    e = Transform()
    e.rotate_euler((0, math.pi / 2, 0))  # yaw: +x goes to -z
    assert np.allclose(e.to_mat4()[:3, :3] @ [1, 0, 0], [0, 0, -1], atol=1e-15)
    # quat(euler) for a single axis is the axis-angle quaternion
    p = Transform()
    p.rotate_euler((0.3, 0, 0))
    assert np.allclose(p.rotation, [math.cos(0.15), math.sin(0.15), 0, 0])
    # rotate_around keeps the position (move(-p), rotate, move(p))
    r = Transform((1, 2, 3))
    r.rotate_around((5, 5, 5), (0, 1, 0))
    assert np.allclose(r.position, [1, 2, 3])
    l = Transform((0, 0, 0))
    l.look_at((0, 0, -5))
    assert np.allclose(l.forward(), [0, 0, -1])
    m = interpolate(Transform((0, 0, 0)), Transform((2, 4, 6), scaling=(3, 3, 3)), 1.5)  # clamped
    assert np.allclose(m.position, [2, 4, 6]) and np.allclose(m.scaling, 3)
    inv = Transform((1, 2, 3), (math.cos(0.2), 0, math.sin(0.2), 0), (2, 4, 8)).inverse()
    assert np.allclose(inv.position, [-1, -2, -3]) and np.allclose(inv.scaling, [0.5, 0.25, 0.125])
    assert np.allclose(quat_to_mat3(inv.rotation), quat_to_mat3((math.cos(0.2), 0, math.sin(0.2), 0)).T)


def test_parent_transform_is_cached_and_dropped_when_an_ancestor_moves():
    root = Entity("root")
    a = root.add_child(Entity("a", Transform((1, 0, 0))))
    b = a.add_child(Entity("b", Transform((0, 1, 0))))
    c = b.add_child(Entity("c", Transform((0, 0, 1))))
    assert np.allclose(c.global_transform().position, [1, 1, 1])
    cached = c.parent_transform()
    assert c.parent_transform() is cached  # reused while nothing moves
    a.move((10, 0, 0))
    assert np.allclose(c.global_transform().position, [11, 1, 1])
    a.rotate(math.pi, (0, 1, 0))  # half turn about y: the children's offsets flip in x and z
    assert np.allclose(c.global_transform().position, [11, 1, -1], atol=1e-12)
    b.set_transform(Transform((0, 2, 0)))
    assert np.allclose(c.global_transform().position, [11, 2, -1], atol=1e-12)
    a.scale((2, 2, 2))
    assert np.allclose(c.global_transform().position, [11, 4, -2], atol=1e-12)
    d = Entity("d")
    d.add_child(b)  # re-parenting drops the caches below
    assert np.allclose(c.global_transform().position, [0, 2, 1], atol=1e-12)


def test_scale_animation_runs_and_finishes_once():
    e = AnimatableEntity("e", Transform(scaling=(1, 1, 1)))
    done = []
    e.set_animation(ScaleAnimation(2.0, (1, 1, 1), (3, 5, 7)))
    e.set_animation_finisher(lambda: done.append(1))
    e.update(1.0)
    assert np.allclose(e.transform.scaling, [2, 3, 4]) and e.animation is not None
    e.update(1.0)  # t == 1: still running
    assert np.allclose(e.transform.scaling, [3, 5, 7]) and e.animation is not None and not done
    e.update(0.5)  # t > 1: the last step clamps, then the finisher runs
    assert np.allclose(e.transform.scaling, [3, 5, 7]) and e.animation is None and done == [1]
    e.update(1.0)
    assert done == [1]


def _python_frames(frames):
    sc = demo_scene(animated=True)
    handles = {}
    for mo in sc.models():
        handles.setdefault(id(mo.mesh), len(handles))
        mo.blas = handles[id(mo.mesh)]
    finished = []
    group = cube2 = cube4 = None

    def find(e):
        nonlocal group, cube2, cube4
        if e.name == "group":
            group = e
        elif e.name == "cube2":
            cube2 = e
        elif e.name == "cube4":
            cube4 = e

    sc.root.for_each_entity(find)
    cube2.set_animation_finisher(lambda: finished.append(1))
    out = []
    for _ in range(frames):
        sc.update(0.25)
        group.rotate_euler((0.0, 0.1, 0.0))
        cube4.move((0.05, 0.0, 0.0))
        out.append(sc.collect_instances())
    return out, len(finished)


def test_cpp_host_and_python_mirror_animate_the_same_frames():
    if not os.path.exists(DEMO):
        pytest.fail(f"{DEMO} is not built; run raygun_amd/build.py or __graft_entry__.build()")
    r = subprocess.run([DEMO, "animate", "7"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    c = json.loads(r.stdout)
    py, finished = _python_frames(7)
    assert c["finished"] == finished == 1  # t = 1.25 s ends the one-second animation
    for f, (cf, pf) in enumerate(zip(c["frames"], py)):
        assert len(cf) == len(pf) == 4, f
        for k in range(4):
            assert cf[k]["blas"] == int(pf[k]["blas"])
            assert np.allclose(np.array(cf[k]["transform"], np.float32), pf[k]["transform"], rtol=1e-6, atol=1e-6), (f, k)
    # the DFS order and the BLAS per position do not change between frames
    assert all(np.array_equal(p["blas"], py[0]["blas"]) for p in py)
    # cube2 grew (scale 0.8 -> 1.3) while its parent turned
    s0 = np.linalg.norm(py[0][2]["transform"].reshape(3, 4)[:, 0])
    s6 = np.linalg.norm(py[6][2]["transform"].reshape(3, 4)[:, 0])
    assert abs(s0 - 0.925) < 1e-5 and abs(s6 - 1.3) < 1e-5


def _refit_scene(n, seed):
    a = procgen.uv_sphere(16, 8)
    b = procgen.torus(16, 8)
    v = np.concatenate([a.vertices, b.vertices])
    i = np.concatenate([a.indices, b.indices]).astype(np.uint32)
    osc = oracle_py.OracleScene(v, i, np.array([default_material()]))
    osc.build_blas(0, 0, len(a.indices) // 3)
    osc.build_blas(len(a.vertices), len(a.indices) * 4, len(b.indices) // 3)
    rng = np.random.default_rng(seed)
    inst = np.zeros(n, INSTANCE)
    pos = rng.uniform(-10, 10, size=(n, 3))
    for k in range(n):
        m = np.zeros((3, 4))
        m[:, :3] = np.eye(3) * rng.uniform(0.5, 1.5)
        m[:, 3] = pos[k]
        inst["transform"][k] = m.reshape(-1)
    inst["customIndex"] = np.arange(n)
    inst["mask"] = 0xFF
    inst["blas"] = rng.integers(0, 2, size=n)
    return osc, inst, rng


def _rays(rng, n, centres):
    o = rng.uniform(-20, 20, size=(n, 3))
    d = centres[rng.integers(0, len(centres), size=n)] + rng.uniform(-1, 1, size=(n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, RAY)
    r["origin"], r["dir"], r["tmin"], r["tmax"] = o, d, 0.001, 1e4
    return r


@pytest.mark.parametrize("n", [1, 2, 50])
def test_oracle_tlas_refit_keeps_topology_and_traces_like_a_rebuild(n):
    """After moving every instance, the refitted tree keeps the hierarchy and
    leaf order of the first build, and closest hits through it equal those
    through a tree built from scratch on the moved instances."""
    osc, inst, rng = _refit_scene(n, seed=n)
    osc.build_tlas(inst)
    nodes0, leaf0 = osc.export_tlas()
    moved = inst.copy()
    t = moved["transform"].reshape(n, 3, 4)
    t[:, :, 3] += rng.uniform(-6, 6, size=(n, 3))
    moved["transform"] = t.reshape(n, 12)
    osc.refit_tlas(moved)
    nodes1, leaf1 = osc.export_tlas()
    assert np.array_equal(leaf0, leaf1)
    assert np.array_equal(nodes0["child"], nodes1["child"]) and np.array_equal(nodes0["parent"], nodes1["parent"])
    rays = _rays(rng, 20_000, t[:, :, 3])
    h_refit, tie_r = osc.trace_rays(rays)
    osc2, _, _ = _refit_scene(n, seed=n)
    osc2.build_tlas(moved)
    h_build, tie_b = osc2.trace_rays(rays)
    ok = ~(tie_r | tie_b)
    assert (h_build["instance"] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(h_refit["instance"][ok], h_build["instance"][ok])
    assert np.array_equal(h_refit["primitive"][ok], h_build["primitive"][ok])
    assert np.array_equal(h_refit["t"][ok], h_build["t"][ok])


def test_oracle_tlas_refit_rejects_a_changed_instance_set():
    osc, inst, _ = _refit_scene(10, seed=3)
    osc.build_tlas(inst)
    with pytest.raises(AssertionError):
        osc.refit_tlas(inst[:9])
    other = inst.copy()
    other["blas"][0] ^= 1
    with pytest.raises(AssertionError):
        osc.refit_tlas(other)
