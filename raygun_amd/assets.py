"""Scene ingestion: Raygun's material files and Collada models.

- ``load_material``   raygun/material.cpp:57-103 (``.rgmat.json``: ``type`` must be
  "Material"; ``basedOn`` copies a base material first, then every key that
  names a ``gpu_material.def`` parameter overrides it; the physics keys are
  ignored here, other keys are reported) and the lookup of
  raygun/resource_manager.cpp:33-60 (``materials/<name>.rgmat.json`` under the
  resource directory; if absent, the first underscore of that path becomes a
  directory separator, so ``ui_button`` resolves to ``materials/ui/button``;
  one cached instance per name).
- ``load_entity``     raygun/entity.cpp:32-122, the Assimp import the reference
  makes with ``aiProcess_Triangulate``: every top-level node of the visual
  scene becomes a child entity with the node's transform (decomposed into
  translation, rotation, scale as ``Transform(mat4)`` does) and a model whose
  mesh merges the node's meshes and its descendants' (``Mesh::merge``,
  mesh.cpp:55-64); all models share the file's material list, loaded by the
  Collada material names. Polygons are triangulated as fans, every face
  corner is its own vertex (the importer's unindexed layout) and
  ``matIndex`` is the material's position in that list.

Only the Collada subset Blender writes for Raygun's assets is read: float
sources, ``<vertices>`` with a POSITION input, ``<triangles>`` / ``<polylist>``
/ ``<polygons>`` with VERTEX and NORMAL inputs, nodes with a ``<matrix>`` or
``<translate>`` / ``<rotate>`` / ``<scale>``, ``<instance_geometry>`` with material
bindings. The asset's ``up_axis`` is not applied: the importer expresses it
as the root node's transform, which the reference never reads (it takes the
children's own transforms), and Raygun's bundled assets carry Y-up data.
"""
import json
import math
import os
import warnings
import xml.etree.ElementTree as ET

import numpy as np

from .records import MATERIAL, VERTEX, default_material
from .scene import Entity, Mesh, Model, Scene, Transform, mat3_to_quat

_VEC3_PARAMS = ("diffuse", "specular")
_SCALAR_PARAMS = ("transparency", "reflectivity", "roughness", "ior", "effectId", "rayConsumption", "emission")
_PHYSICS_PARAMS = ("staticFriction", "dynamicFriction")


class AssetError(RuntimeError):
    pass


class Resources:
    """A resource directory (``materials/``, ``models/``) with the material
    cache of ResourceManager."""

    def __init__(self, root):
        self.root = os.fspath(root)
        self.materials = {}

    def _path(self, rel):
        if os.path.exists(os.path.join(self.root, rel)):
            return os.path.join(self.root, rel)
        u = rel.find("_")
        if u >= 0:
            rel = os.path.join(rel[:u], rel[u + 1:])
        return os.path.join(self.root, rel)

    def load_material(self, name):
        if name in self.materials:
            return self.materials[name]
        m = _parse_material(self._path(os.path.join("materials", name + ".rgmat.json")), self)
        self.materials[name] = m
        return m

    def entity_path(self, name):
        return self._path(os.path.join("models", name + ".dae"))


def _parse_material(path, resources):
    try:
        with open(path) as f:
            data = json.load(f)
    except OSError as e:
        raise AssetError(f"unable to open material {path}: {e}") from None
    if data.get("type") != "Material":
        raise AssetError(f"not a material: {path}")
    m = np.array(resources.load_material(data["basedOn"]) if "basedOn" in data else default_material(), MATERIAL)
    for key, value in data.items():
        if key in ("type", "basedOn") or key in _PHYSICS_PARAMS:
            continue
        if key in _VEC3_PARAMS:
            v = np.asarray(value, np.float32).reshape(-1)
            m[key] = v[0] if v.size == 1 else v[:3]  # toVec3 accepts a scalar or an array
        elif key in _SCALAR_PARAMS:
            m[key] = value
        else:
            warnings.warn(f"unknown field '{key}' in material {path}")
    return m


# ---------------------------------------------------------------------------
# Collada
# ---------------------------------------------------------------------------
def _tag(e):
    return e.tag.rsplit("}", 1)[-1]


def _children(e, name):
    return [c for c in e if _tag(c) == name]


def _child(e, name):
    for c in e:
        if _tag(c) == name:
            return c
    return None


def _floats(text):
    return np.array(text.split(), np.float64) if text and text.strip() else np.zeros(0)


def _node_matrix(node):
    """Local transform of a <node> (its transform elements in order)."""
    m = np.eye(4)
    for c in node:
        t = _tag(c)
        if t == "matrix":
            m = m @ _floats(c.text).reshape(4, 4)
        elif t == "translate":
            x = np.eye(4)
            x[:3, 3] = _floats(c.text)[:3]
            m = m @ x
        elif t == "scale":
            m = m @ np.diag(list(_floats(c.text)[:3]) + [1.0])
        elif t == "rotate":
            ax, ay, az, deg = _floats(c.text)[:4]
            l = math.sqrt(ax * ax + ay * ay + az * az)
            h = math.radians(deg) / 2.0
            s = math.sin(h) / l
            from .scene import quat_to_mat3
            x = np.eye(4)
            x[:3, :3] = quat_to_mat3((math.cos(h), ax * s, ay * s, az * s))
            m = m @ x
    return m


def transform_from_matrix(m):
    """Transform(mat4): translation, scale = column lengths, rotation of the
    normalised columns (glm::decompose for matrices without shear or
    perspective)."""
    m = np.asarray(m, np.float64)
    t = Transform()
    t.position = m[:3, 3].copy()
    s = np.linalg.norm(m[:3, :3], axis=0)
    r = m[:3, :3] / np.where(s == 0, 1.0, s)[None, :]
    if np.linalg.det(r) < 0:  # a mirror: fold it into the scale
        s[0] = -s[0]
        r[:, 0] = -r[:, 0]
    t.scaling = s
    t.rotation = mat3_to_quat(r)
    return t


class _Collada:
    def __init__(self, path):
        try:
            self.root = ET.parse(path).getroot()
        except (OSError, ET.ParseError) as e:
            raise AssetError(f"unable to load {path}: {e}") from None
        self.ids = {}
        for e in self.root.iter():
            if e.get("id") is not None:
                self.ids[e.get("id")] = e
        lib = _child(self.root, "library_materials")
        self.materials = _children(lib, "material") if lib is not None else []
        self.material_index = {m.get("id"): k for k, m in enumerate(self.materials)}

    def ref(self, url):
        return self.ids.get(url[1:] if url.startswith("#") else url)

    def source(self, url):
        s = self.ref(url)
        if s is not None and _tag(s) == "vertices":  # <vertices> -> its POSITION source
            for i in _children(s, "input"):
                if i.get("semantic") == "POSITION":
                    return self.source(i.get("source"))
        fa = _child(s, "float_array")
        acc = _child(_child(s, "technique_common"), "accessor")
        stride = int(acc.get("stride", "1"))
        return _floats(fa.text).reshape(-1, stride)

    def geometry_meshes(self, geom, bindings):
        """One mesh per primitive group, as the importer splits them."""
        out = []
        mesh = _child(geom, "mesh")
        for prim in mesh:
            kind = _tag(prim)
            if kind not in ("triangles", "polylist", "polygons"):
                continue
            inputs = _children(prim, "input")
            stride = max(int(i.get("offset", "0")) for i in inputs) + 1
            pos = nrm = None
            po = no = 0
            for i in inputs:
                if i.get("semantic") == "VERTEX":
                    pos, po = self.source(i.get("source")), int(i.get("offset", "0"))
                elif i.get("semantic") == "NORMAL":
                    nrm, no = self.source(i.get("source")), int(i.get("offset", "0"))
            if pos is None:
                continue
            if kind == "polygons":
                polys = [np.array(p.text.split(), np.int64).reshape(-1, stride) for p in _children(prim, "p")]
            else:
                p = np.array(_child(prim, "p").text.split(), np.int64).reshape(-1, stride)
                counts = [3] * (len(p) // 3) if kind == "triangles" else [int(x) for x in _child(prim, "vcount").text.split()]
                polys, k = [], 0
                for c in counts:
                    polys.append(p[k:k + c])
                    k += c
            corners = []
            for poly in polys:  # triangle fans
                for j in range(1, len(poly) - 1):
                    corners += [poly[0], poly[j], poly[j + 1]]
            n = len(corners)
            verts = np.zeros(n, VERTEX)
            if n:
                c = np.array(corners)
                verts["position"] = pos[c[:, po], :3]
                if nrm is not None:
                    verts["normal"] = nrm[c[:, no], :3]
            sym = prim.get("material")
            target = bindings.get(sym, sym)
            mi = self.material_index.get(target.lstrip("#") if target else None, 0)
            verts["matIndex"] = mi
            out.append(Mesh(verts, np.arange(n, dtype=np.uint32)))
        return out

    def node_meshes(self, node):
        meshes = []
        for ig in _children(node, "instance_geometry"):
            bindings = {}
            bm = _child(ig, "bind_material")
            if bm is not None:
                for im in bm.iter():
                    if _tag(im) == "instance_material":
                        bindings[im.get("symbol")] = im.get("target")
            geom = self.ref(ig.get("url"))
            if geom is not None:
                meshes += self.geometry_meshes(geom, bindings)
        for child in _children(node, "node"):
            meshes += self.node_meshes(child)
        return meshes


def merge_meshes(meshes, name=""):
    """Mesh::merge over a list: indices shifted by the vertices before them."""
    verts, idx, off = [], [], 0
    for m in meshes:
        verts.append(m.vertices)
        idx.append(m.indices.astype(np.uint64) + off)
        off += len(m.vertices)
    v = np.concatenate(verts) if verts else np.zeros(0, VERTEX)
    i = np.concatenate(idx).astype(np.uint32) if idx else np.zeros(0, np.uint32)
    return Mesh(v, i, name)


def load_entity(name, path, resources=None, load_materials=True):
    """Entity(name, filepath, loadMaterials) (entity.cpp:87-122)."""
    doc = _Collada(path)
    materials = []
    if load_materials:
        if resources is None:
            raise AssetError("load_entity: materials need a resource directory")
        materials = [resources.load_material(m.get("name") or m.get("id")) for m in doc.materials]
    mats = np.array(materials, MATERIAL) if materials else np.array([default_material()], MATERIAL)
    ent = Entity(name)
    scene = None
    vs = _child(doc.root, "library_visual_scenes")
    if vs is not None:
        scene = _child(vs, "visual_scene")
    inst = _child(_child(doc.root, "scene"), "instance_visual_scene") if _child(doc.root, "scene") is not None else None
    if inst is not None and doc.ref(inst.get("url")) is not None:
        scene = doc.ref(inst.get("url"))
    if scene is None:
        return ent
    for node in _children(scene, "node"):
        mesh = merge_meshes(doc.node_meshes(node), node.get("name") or node.get("id") or "")
        child = Entity(node.get("name") or node.get("id") or "", transform_from_matrix(_node_matrix(node)), Model(mesh, mats))
        ent.add_child(child)
    return ent


def example_scene(resources_root, width=1920, height=1080):
    """The bundled example (example/example_scene.cpp:9-23, ball.cpp:10-21,
    ExampleScene::update): the room level, the ball moved to (3, 0, -3) with
    the model of the file's first node, the camera at the ball + (5, 10, 10)
    looking at it. Frame constants as RenderSystem sets them (45 degree field
    of view, 1 sample, 5 recursions)."""
    from .procgen import SceneSetup
    res = Resources(resources_root)
    sc = Scene()
    sc.root.add_child(load_entity("room", res.entity_path("room"), res))
    ball_file = load_entity("Ball", res.entity_path("ball"), res)
    ball = Entity("Ball", Transform(position=(3.0, 0.0, -3.0)), ball_file.children[0].model)
    sc.root.add_child(ball)
    sc.camera.fov_deg = 45.0
    eye = np.array([3.0, 0.0, -3.0]) + np.array([5.0, 10.0, 10.0])
    sc.camera.look_at(eye, (3.0, 0.0, -3.0))
    tris = sum(m.mesh.num_faces() for m in sc.models())
    return SceneSetup(sc, width, height, max_recursions=5, name="example", info={"triangles": tris})
