"""Scene ingestion (raygun_amd/assets.py): .rgmat.json materials and Collada
models, on files written here and, when the reference's resource directory
is present, on its bundled room / ball example scene.
"""
import json
import math
import os
import textwrap

import numpy as np
import pytest

from raygun_amd import assets
from raygun_amd.records import default_material

REF_RESOURCES = "/root/reference/resources"


def _write(path, obj):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(obj if isinstance(obj, str) else json.dumps(obj))


def test_material_defaults_based_on_and_alt_path(tmp_path):
    root = str(tmp_path)
    _write(f"{root}/materials/base.rgmat.json", {"type": "Material", "diffuse": [0.5, 0.25, 0.125], "ior": 1.3})
    _write(f"{root}/materials/derived.rgmat.json",
           {"type": "Material", "basedOn": "base", "ior": 1.7, "specular": 0.5, "staticFriction": 0.2})
    _write(f"{root}/materials/ui/button.rgmat.json", {"type": "Material", "basedOn": "derived", "effectId": 1})
    res = assets.Resources(root)
    base = res.load_material("base")
    d = default_material()
    assert list(base["diffuse"]) == [0.5, 0.25, 0.125]
    assert base["ior"] == np.float32(1.3)
    for k in ("transparency", "reflectivity", "roughness", "effectId", "rayConsumption", "emission"):
        assert base[k] == d[k], k  # untouched fields keep the gpu_material.def defaults
    der = res.load_material("derived")
    assert list(der["diffuse"]) == [0.5, 0.25, 0.125]  # from the base
    assert der["ior"] == np.float32(1.7)               # overridden
    assert list(der["specular"]) == [0.5, 0.5, 0.5]    # a scalar fills the vector
    ui = res.load_material("ui_button")                # materials/ui_button -> materials/ui/button
    assert ui["effectId"] == 1 and ui["ior"] == np.float32(1.7)
    assert res.load_material("base") is base           # one instance per name


def test_material_errors(tmp_path):
    root = str(tmp_path)
    _write(f"{root}/materials/notmat.rgmat.json", {"type": "Mesh"})
    _write(f"{root}/materials/odd.rgmat.json", {"type": "Material", "shininess": 3})
    res = assets.Resources(root)
    with pytest.raises(assets.AssetError):
        res.load_material("notmat")
    with pytest.raises(assets.AssetError):
        res.load_material("missing")
    with pytest.warns(UserWarning):
        res.load_material("odd")


DAE = textwrap.dedent("""\
    <?xml version="1.0" encoding="utf-8"?>
    <COLLADA xmlns="http://www.collada.org/2005/11/COLLADASchema" version="1.4.1">
      <asset><up_axis>Y_UP</up_axis></asset>
      <library_materials>
        <material id="red-material" name="red"/>
        <material id="blue-material" name="blue"/>
      </library_materials>
      <library_geometries>
        <geometry id="g1" name="g1"><mesh>
          <source id="g1-p"><float_array id="g1-pa" count="15">0 0 0  1 0 0  1 1 0  0 1 0  0 0 1</float_array>
            <technique_common><accessor source="#g1-pa" count="5" stride="3"/></technique_common></source>
          <source id="g1-n"><float_array id="g1-na" count="6">0 0 1  0 1 0</float_array>
            <technique_common><accessor source="#g1-na" count="2" stride="3"/></technique_common></source>
          <vertices id="g1-v"><input semantic="POSITION" source="#g1-p"/></vertices>
          <triangles material="mred" count="1">
            <input semantic="VERTEX" source="#g1-v" offset="0"/><input semantic="NORMAL" source="#g1-n" offset="1"/>
            <p>0 0 1 0 4 1</p>
          </triangles>
          <polylist material="mblue" count="1">
            <input semantic="VERTEX" source="#g1-v" offset="0"/><input semantic="NORMAL" source="#g1-n" offset="1"/>
            <vcount>4</vcount><p>0 0 1 0 2 0 3 0</p>
          </polylist>
        </mesh></geometry>
        <geometry id="g2" name="g2"><mesh>
          <source id="g2-p"><float_array id="g2-pa" count="9">0 0 0  2 0 0  0 2 0</float_array>
            <technique_common><accessor source="#g2-pa" count="3" stride="3"/></technique_common></source>
          <vertices id="g2-v"><input semantic="POSITION" source="#g2-p"/></vertices>
          <triangles material="mred" count="1"><input semantic="VERTEX" source="#g2-v" offset="0"/><p>0 1 2</p></triangles>
        </mesh></geometry>
      </library_geometries>
      <library_visual_scenes><visual_scene id="scene">
        <node id="a" name="A">
          <matrix>2 0 0 1  0 2 0 2  0 0 2 3  0 0 0 1</matrix>
          <instance_geometry url="#g1"><bind_material><technique_common>
            <instance_material symbol="mred" target="#red-material"/><instance_material symbol="mblue" target="#blue-material"/>
          </technique_common></bind_material></instance_geometry>
          <node id="a1" name="A1"><translate>0 5 0</translate>
            <instance_geometry url="#g2"><bind_material><technique_common>
              <instance_material symbol="mred" target="#blue-material"/></technique_common></bind_material></instance_geometry>
          </node>
        </node>
        <node id="b" name="B"><translate>1 0 0</translate><rotate>0 1 0 90</rotate><scale>1 3 1</scale>
          <instance_geometry url="#g2"/>
        </node>
      </visual_scene></library_visual_scenes>
      <scene><instance_visual_scene url="#scene"/></scene>
    </COLLADA>
    """)


@pytest.fixture()
def dae_scene(tmp_path):
    root = str(tmp_path)
    _write(f"{root}/materials/red.rgmat.json", {"type": "Material", "diffuse": [0.9, 0.1, 0.1], "reflectivity": 0.4})
    _write(f"{root}/materials/blue.rgmat.json", {"type": "Material", "diffuse": [0.1, 0.1, 0.9], "transparency": 0.5, "ior": 1.4})
    _write(f"{root}/models/thing.dae", DAE)
    res = assets.Resources(root)
    return res, assets.load_entity("thing", res.entity_path("thing"), res)


def test_collada_nodes_meshes_and_materials(dae_scene):
    res, ent = dae_scene
    assert [c.name for c in ent.children] == ["A", "B"]  # top-level nodes only
    a, b = ent.children
    # A: g1's triangle + its quad as a fan (2 triangles), then the child node's g2
    m = a.model.mesh
    assert m.num_faces() == 1 + 2 + 1
    assert len(m.vertices) == 12  # one vertex per face corner
    assert m.indices.tolist() == list(range(12))  # merged with the offsets of Mesh::merge
    pos = m.vertices["position"]
    assert pos[:3].tolist() == [[0, 0, 0], [1, 0, 0], [0, 0, 1]]
    assert pos[3:9].tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]]
    assert m.vertices["normal"][2].tolist() == [0, 1, 0]
    # matIndex = position of the bound material in the file's list (red 0, blue 1)
    assert m.vertices["matIndex"].tolist() == [0] * 3 + [1] * 6 + [1] * 3
    assert [float(x) for x in a.model.materials["ior"]] == [1.0, np.float32(1.4)]
    assert a.transform.position.tolist() == [1, 2, 3] and np.allclose(a.transform.scaling, 2)
    # B: translate, rotate 90 degrees about y, scale (1, 3, 1)
    assert np.allclose(b.transform.position, [1, 0, 0])
    assert np.allclose(b.transform.scaling, [1, 3, 1])
    q = b.transform.rotation
    assert np.allclose(np.abs(q), [math.cos(math.pi / 4), 0, math.sin(math.pi / 4), 0])
    assert np.allclose(b.transform.to_mat4(), [[0, 0, 1, 1], [0, 3, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1]], atol=1e-12)


def test_collada_scene_traces_on_the_oracle(dae_scene):
    import oracle_py
    from helpers import oracle_scene
    from raygun_amd.procgen import SceneSetup
    from raygun_amd.scene import Scene
    res, ent = dae_scene
    sc = Scene()
    sc.root.add_child(ent)
    sc.camera.fov_deg = 60.0
    sc.camera.look_at((2.0, 4.0, 12.0), (1.0, 3.0, 3.0))
    setup = SceneSetup(sc, 48, 32, max_recursions=3)
    osc, inst, _ = oracle_scene(setup)
    assert len(inst) == 2
    out = osc.render(setup.ubo(), 48, 32)
    assert (out["primary"]["instance"] != 0xFFFFFFFF).sum() > 0


@pytest.mark.skipif(not os.path.isdir(REF_RESOURCES), reason="the reference's resource directory is not present")
def test_bundled_example_scene():
    setup = assets.example_scene(REF_RESOURCES, 64, 36)
    room, ball = setup.scene.root.children
    assert [c.name for c in room.children] == ["Raygun", "ph3_games", "room"]
    assert [c.model.mesh.num_faces() for c in room.children] == [6658, 10428, 548]  # 17,634 in room.dae
    assert ball.model.mesh.num_faces() == 1280
    glass = ball.model.materials[0]
    assert glass["transparency"] == np.float32(0.98) and glass["ior"] == np.float32(1.5)
    names = [float(m["effectId"]) for m in room.children[0].model.materials]
    assert names.count(1.0) == 1  # the floor material carries the grid effect
    from helpers import oracle_scene
    osc, inst, _ = oracle_scene(setup)
    assert len(inst) == 4
    out = osc.render(setup.ubo(), 64, 36)
    assert out["rays"] > 64 * 36


@pytest.mark.gpu
def test_collada_scene_renders_like_the_oracle_on_the_gpu(dae_scene):
    from helpers import gpu_context, oracle_scene
    from raygun_amd import rgrt
    from raygun_amd.procgen import SceneSetup
    from raygun_amd.scene import Scene
    res, ent = dae_scene
    sc = Scene()
    sc.root.add_child(ent)
    sc.camera.fov_deg = 60.0
    sc.camera.look_at((2.0, 4.0, 12.0), (1.0, 3.0, 3.0))
    setup = SceneSetup(sc, 160, 90, max_recursions=4)
    print("[case] collada-scene 160x90", flush=True)
    osc, inst, _ = oracle_scene(setup)
    ctx, _ = gpu_context(setup, flags=rgrt.CFG_FP32_RADIANCE | rgrt.CFG_PRIMARY_AOV)
    try:
        ctx.trace()
        g = ctx.read_image(rgrt.IMAGE_BASE_F32)
        gp = ctx.read_image(rgrt.IMAGE_PRIMARY_HIT)
    finally:
        ctx.close()
    o = osc.render(setup.ubo(), 160, 90)
    ok = ~o["tie"].astype(bool)
    assert np.array_equal(gp["instance"][ok], o["primary"]["instance"][ok])
    assert np.abs(g - o["base32"])[ok].max() <= 1e-3
