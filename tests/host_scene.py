"""The scene of raygun_amd/host/scene_demo.cpp, built with the Python mirror
(raygun_amd/scene.py): the same meshes, materials, entity tree and camera."""
import math

import numpy as np

from raygun_amd.records import VERTEX, default_material
from raygun_amd.scene import AnimatableEntity, Entity, Mesh, Model, ScaleAnimation, Scene, Transform


def _vtx(p, n):
    v = np.zeros((), VERTEX)
    v["position"] = p
    v["normal"] = n
    return v


def cube():
    verts, idx = [], []
    for axis in range(3):
        for sgn in range(2):
            s = 1.0 if sgn else -1.0
            u, w = (axis + 1) % 3, (axis + 2) % 3
            base = len(verts)
            for k in range(4):
                p = [0.0, 0.0, 0.0]
                n = [0.0, 0.0, 0.0]
                p[axis] = 0.5 * s
                p[u] = 0.5 if k in (1, 2) else -0.5
                p[w] = 0.5 if k >= 2 else -0.5
                n[axis] = s
                verts.append(_vtx(p, n))
            idx += [base + i for i in (0, 1, 2, 0, 2, 3)]
    return Mesh(np.array(verts, VERTEX), np.array(idx, np.uint32), "cube")


def ground():
    up = (0.0, 1.0, 0.0)
    verts = [_vtx(p, up) for p in ((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1))]
    return Mesh(np.array(verts, VERTEX), np.array([0, 2, 1, 0, 3, 2], np.uint32), "ground")


def axis_angle(ax, ay, az, deg):
    l = math.sqrt(ax * ax + ay * ay + az * az)
    h = deg * (math.pi / 180.0) / 2.0
    s = math.sin(h)
    return (math.cos(h), ax / l * s, ay / l * s, az / l * s)


def _entity(name, model, pos, rot=(1, 0, 0, 0), scale=(1, 1, 1)):
    return Entity(name, Transform(pos, rot, scale), model)


def demo_scene(animated=False):
    """animated: cube2 is an AnimatableEntity scaling from 0.8 to 1.3 over one
    second (scene_demo animate)."""
    cube_mesh, ground_mesh = cube(), ground()
    red = Model(cube_mesh, [default_material(diffuse=(0.8, 0.1, 0.1), reflectivity=0.5, roughness=0.3)])
    glass = Model(cube_mesh, [default_material(diffuse=(0.9, 0.9, 1.0), transparency=0.9, ior=1.5)])
    floor = Model(ground_mesh, [default_material(diffuse=(0.6, 0.6, 0.6), reflectivity=0.3, roughness=0.2, effectId=1)])
    sc = Scene()
    sc.root.add_child(_entity("ground", floor, (0, 0, 0), scale=(20, 1, 20)))
    group = sc.root.add_child(_entity("group", None, (0, 0.5, 0), axis_angle(0, 1, 0, 30)))
    group.add_child(_entity("cube1", red, (-1.2, 0, 0)))
    if animated:
        c2 = AnimatableEntity("cube2", Transform((1.2, 0.2, 0), (1, 0, 0, 0), (0.8, 0.8, 0.8)), glass)
        c2.set_animation(ScaleAnimation(1.0, (0.8, 0.8, 0.8), (1.3, 1.3, 1.3)))
        group.add_child(c2)
    else:
        group.add_child(_entity("cube2", glass, (1.2, 0.2, 0), scale=(0.8, 0.8, 0.8)))
    hidden = group.add_child(_entity("hidden", red, (0, 3, 0)))
    hidden.visible = False
    hidden.add_child(_entity("under-hidden", glass, (0, 1, 0)))
    group.add_child(_entity("flat", red, (0, 0, 2), scale=(1, 0, 1)))
    sc.root.add_child(_entity("cube4", red, (0, 1.8, -1), axis_angle(1, 1, 0, 40), (0.6, 0.6, 0.6)))
    sc.camera.fov_deg = 60.0
    sc.camera.look_at((0, 2, 6), (0, 0.5, 0))
    return sc
