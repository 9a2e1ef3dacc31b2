"""Seeded procedural meshes and the benchmark scenes built from them.

Meshes are Raygun meshes (VERTEX records, mesh-local uint32 indices, smooth
per-vertex normals, matIndex 0 unless stated). All randomness comes from
numpy's PCG64 seeded generator, so a (config, seed) pair is reproducible.

Scenes:
- ``config1``: UV sphere (radius 1, 128 segments x 64 rings, triangle fans at
  the poles -> 16,128 triangles) on a 2-triangle ground plane at y = -1
  (extent +-50) in ONE mesh / ONE instance, 640x360, vertical FOV 60 deg,
  camera at (0, 0.5, 4) looking at the origin, maxRecursions 1, the
  reference's directional light normalize(0.4, -0.6, -0.8).
- ``config2``: 5 seeded meshes (UV sphere, torus, noise-displaced grid,
  noise-displaced sphere, displaced torus) of ~400k triangles each (2.0M
  unique BLAS triangles), 64 TRS instances uniformly placed in a 100^3 box,
  per-instance materials, 1920x1080, maxRecursions 4.
- ``config5``: 8 seeded noise-displaced meshes (grids, spheres, tori) of ~2M
  triangles each (16M unique BLAS triangles), 256 TRS instances in a 200^3
  box, one instance in five refractive with IOR U[1.3, 1.7], 3840x2160,
  maxRecursions 6, 4 samples per pixel.
"""
import math

import numpy as np

from .records import VERTEX, default_material
from .scene import Camera, Entity, Mesh, Model, Scene, Transform, camera_ubo


# ---------------------------------------------------------------------------
# Meshes
# ---------------------------------------------------------------------------
def _mesh(pos, nrm, idx, mat_index=0, name=""):
    v = np.zeros(len(pos), VERTEX)
    v["position"] = pos.astype(np.float32)
    n = nrm / np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    v["normal"] = n.astype(np.float32)
    v["matIndex"] = mat_index
    return Mesh(v, np.asarray(idx, np.uint32).reshape(-1), name)


def uv_sphere(segments=128, rings=64, radius=1.0, noise=None, mat_index=0):
    """UV sphere; rings-1 latitude circles, single pole vertices joined by
    triangle fans: 2*segments*(rings-2) + 2*segments triangles."""
    lat = np.pi * np.arange(1, rings) / rings  # interior circles
    lon = 2 * np.pi * np.arange(segments) / segments
    st, ct = np.sin(lat)[:, None], np.cos(lat)[:, None]
    sp, cp = np.sin(lon)[None, :], np.cos(lon)[None, :]
    ring = np.stack([st * cp, np.broadcast_to(ct, (rings - 1, segments)), st * sp], axis=-1).reshape(-1, 3)
    dirs = np.concatenate([[[0.0, 1.0, 0.0]], ring, [[0.0, -1.0, 0.0]]])
    r = radius * (1.0 + (noise(dirs) if noise is not None else 0.0))
    pos = dirs * (r[:, None] if np.ndim(r) else r)
    top, bot = 0, len(dirs) - 1

    def vid(i, j):
        return 1 + i * segments + (j % segments)

    tris = []
    j = np.arange(segments)
    tris.append(np.stack([np.full(segments, top), vid(0, j + 1), vid(0, j)], 1))
    for i in range(rings - 2):
        a, b, c, d = vid(i, j), vid(i, j + 1), vid(i + 1, j), vid(i + 1, j + 1)
        tris.append(np.stack([a, b, d], 1))
        tris.append(np.stack([a, d, c], 1))
    tris.append(np.stack([np.full(segments, bot), vid(rings - 2, j), vid(rings - 2, j + 1)], 1))
    idx = np.concatenate(tris)
    nrm = _vertex_normals(pos, idx) if noise is not None else dirs
    return _mesh(pos, nrm, idx, mat_index, "sphere")


def torus(u_seg, v_seg, major=1.0, minor=0.35, noise=None, mat_index=0):
    u = 2 * np.pi * np.arange(u_seg) / u_seg
    v = 2 * np.pi * np.arange(v_seg) / v_seg
    U, V = np.meshgrid(u, v, indexing="ij")
    cx, cz = np.cos(U), np.sin(U)
    n = np.stack([np.cos(V) * cx, np.sin(V), np.cos(V) * cz], -1).reshape(-1, 3)
    rr = minor * (1.0 + (noise(n) if noise is not None else 0.0))
    center = np.stack([major * cx, np.zeros_like(cx), major * cz], -1).reshape(-1, 3)
    pos = center + n * (rr[:, None] if np.ndim(rr) else rr)
    i, j = np.meshgrid(np.arange(u_seg), np.arange(v_seg), indexing="ij")
    a = i * v_seg + j
    b = ((i + 1) % u_seg) * v_seg + j
    c = i * v_seg + (j + 1) % v_seg
    d = ((i + 1) % u_seg) * v_seg + (j + 1) % v_seg
    idx = np.concatenate([np.stack([a, c, b], -1).reshape(-1, 3), np.stack([b, c, d], -1).reshape(-1, 3)])
    nrm = _vertex_normals(pos, idx) if noise is not None else n
    return _mesh(pos, nrm, idx, mat_index, "torus")


def displaced_grid(n, size=2.0, amplitude=0.25, noise=None, mat_index=0):
    """n x n quads on the xz-plane in [-size/2, size/2]^2, heights from noise."""
    g = np.linspace(-size / 2, size / 2, n + 1)
    X, Z = np.meshgrid(g, g, indexing="ij")
    P = np.stack([X, np.zeros_like(X), Z], -1).reshape(-1, 3)
    h = amplitude * (noise(P / size) if noise is not None else 0.0)
    P[:, 1] = h
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = i * (n + 1) + j
    b = (i + 1) * (n + 1) + j
    c = i * (n + 1) + j + 1
    d = (i + 1) * (n + 1) + j + 1
    idx = np.concatenate([np.stack([a, c, b], -1).reshape(-1, 3), np.stack([b, c, d], -1).reshape(-1, 3)])
    return _mesh(P, _vertex_normals(P, idx), idx, mat_index, "grid")


def plane(extent=50.0, y=-1.0, mat_index=0):
    P = np.array([[-extent, y, -extent], [extent, y, -extent], [extent, y, extent], [-extent, y, extent]], np.float64)
    N = np.tile([0.0, 1.0, 0.0], (4, 1))
    idx = np.array([[0, 3, 2], [0, 2, 1]])
    return _mesh(P, N, idx, mat_index, "plane")


def merge(meshes, name=""):
    verts, idx, off = [], [], 0
    for m in meshes:
        verts.append(m.vertices)
        idx.append(m.indices + off)
        off += len(m.vertices)
    return Mesh(np.concatenate(verts), np.concatenate(idx).astype(np.uint32), name)


def _vertex_normals(pos, idx):
    idx = np.asarray(idx).reshape(-1, 3)
    p0, p1, p2 = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    fn = np.cross(p1 - p0, p2 - p0)
    vn = np.zeros_like(pos)
    for k in range(3):
        np.add.at(vn, idx[:, k], fn)
    bad = np.linalg.norm(vn, axis=1) < 1e-30
    vn[bad] = (0.0, 1.0, 0.0)
    return vn


class ValueNoise:
    """Seeded 3-D value noise (trilinear over a random lattice), fractal sum."""

    def __init__(self, seed, octaves=3, base_freq=2.0, lattice=64):
        rng = np.random.default_rng(seed)
        self.table = rng.uniform(-1.0, 1.0, size=(lattice, lattice, lattice))
        self.octaves = octaves
        self.freq = base_freq
        self.L = lattice

    def __call__(self, p):
        p = np.asarray(p, np.float64)
        out = np.zeros(len(p))
        amp, f = 1.0, self.freq
        for _ in range(self.octaves):
            q = p * f + 17.0
            i = np.floor(q).astype(np.int64)
            t = q - i
            t = t * t * (3 - 2 * t)
            acc = 0.0
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        w = (t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1]) * (t[:, 2] if dz else 1 - t[:, 2])
                        acc = acc + w * self.table[(i[:, 0] + dx) % self.L, (i[:, 1] + dy) % self.L, (i[:, 2] + dz) % self.L]
            out += amp * acc
            amp *= 0.5
            f *= 2.0
        return out / 1.75


# ---------------------------------------------------------------------------
# Scenes
# ---------------------------------------------------------------------------
class SceneSetup:
    """A scene plus the render settings it is traced with."""

    def __init__(self, scene, width, height, max_recursions, num_samples=1, name="", seed=0, info=None):
        self.scene = scene
        self.width = width
        self.height = height
        self.max_recursions = max_recursions
        self.num_samples = num_samples
        self.name = name
        self.seed = seed
        self.info = info or {}

    def ubo(self, width=None, height=None):
        return camera_ubo(self.scene.camera, width or self.width, height or self.height,
                          num_samples=self.num_samples, max_recursions=self.max_recursions)


def config1(width=640, height=360, seed=1):
    """Sphere (16,128 tris) + ground plane in one BLAS, one instance."""
    sphere = uv_sphere(128, 64, 1.0, mat_index=0)
    ground = plane(50.0, -1.0, mat_index=1)
    mesh = merge([sphere, ground], "sphere+plane")
    mats = np.array([
        default_material(roughness=0.2, reflectivity=1.0),
        default_material(diffuse=(0.6, 0.6, 0.6), reflectivity=0.3, roughness=0.2, effectId=1),
    ])
    model = Model(mesh, mats)
    sc = Scene()
    sc.camera = Camera(fov_deg=60.0)
    sc.camera.look_at((0.0, 0.5, 4.0), (0.0, 0.0, 0.0))
    sc.root.add_child(Entity("sphere+plane", model=model))
    return SceneSetup(sc, width, height, max_recursions=1, name="config1", seed=seed,
                      info={"triangles": mesh.num_faces(), "instances": 1})


def config2_meshes(seed=2, tris_per_mesh=400_000):
    """The five seeded meshes of config 2, ~tris_per_mesh triangles each."""
    rng = np.random.default_rng(seed)
    noises = [ValueNoise(int(rng.integers(1 << 30)), octaves=3, base_freq=f) for f in (1.5, 2.5, 3.0)]
    # sphere: 2*S*(R-1) triangles with R = S/2
    S = int(round(math.sqrt(tris_per_mesh)))
    S += S % 2
    sphere = uv_sphere(S, S // 2, 1.0)
    # torus: 2*U*V triangles, U = 4V
    V = int(round(math.sqrt(tris_per_mesh / 8)))
    tor = torus(4 * V, V, 1.0, 0.35)
    # grid: 2*n^2 triangles
    n = int(round(math.sqrt(tris_per_mesh / 2)))
    grid = displaced_grid(n, 2.0, 0.3, noise=noises[0])
    bumpy = uv_sphere(S, S // 2, 1.0, noise=lambda d: 0.15 * noises[1](d))
    knot = torus(4 * V, V, 1.0, 0.45, noise=lambda d: 0.3 * noises[2](d))
    return [sphere, tor, grid, bumpy, knot]


def random_quat(rng):
    u1, u2, u3 = rng.uniform(0.0, 1.0, 3)
    return np.array([
        math.sqrt(u1) * math.cos(2 * math.pi * u3),
        math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
        math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
        math.sqrt(u1) * math.sin(2 * math.pi * u3),
    ])


def config2(width=1920, height=1080, seed=2, tris_per_mesh=400_000, num_instances=64, meshes=None):
    """64 TRS instances of 5 seeded meshes in a 100^3 box, per-instance materials."""
    rng = np.random.default_rng(seed + 1000)
    meshes = meshes if meshes is not None else config2_meshes(seed, tris_per_mesh)
    sc = Scene()
    sc.camera = Camera(fov_deg=60.0)
    # 15 units outside the box's +z face, looking in: ~90% of primary rays hit
    sc.camera.look_at((0.0, 5.0, 65.0), (0.0, 0.0, -20.0))
    mats_info = []
    for i in range(num_instances):
        mesh = meshes[int(rng.integers(len(meshes)))]
        refl = float(rng.uniform(0.3, 1.0)) if rng.uniform() < 0.4 else 0.0
        transp = float(rng.uniform(0.5, 0.95)) if rng.uniform() < 0.25 else 0.0
        mat = default_material(
            diffuse=tuple(rng.uniform(0.0, 1.0, 3)),
            roughness=float(rng.uniform(0.0, 1.0)),
            ior=float(rng.uniform(1.0, 1.8)),
            reflectivity=refl,
            transparency=transp,
            effectId=1 if rng.uniform() < 0.1 else 0,
        )
        mats_info.append((refl, transp))
        model = Model(mesh, np.array([mat]))
        t = Transform(position=rng.uniform(-50.0, 50.0, 3), rotation=random_quat(rng),
                      scaling=np.full(3, rng.uniform(6.0, 16.0)))
        sc.root.add_child(Entity(f"inst{i}", transform=t, model=model))
    unique = sum(m.num_faces() for m in meshes)
    inst_tris = sum(e.model.mesh.num_faces() for e in sc.root.children)
    return SceneSetup(sc, width, height, max_recursions=4, name="config2", seed=seed,
                      info={"unique_triangles": unique, "instanced_triangles": inst_tris,
                            "instances": num_instances,
                            "reflective": sum(1 for r, _ in mats_info if r > 0),
                            "transparent": sum(1 for _, t in mats_info if t > 0)})


def config5_meshes(seed=5, tris_per_mesh=2_000_000):
    """Eight seeded noise-displaced meshes of ~tris_per_mesh triangles each:
    displaced grids, bumpy spheres and tori, three of each kind but one."""
    rng = np.random.default_rng(seed)
    S = int(round(math.sqrt(tris_per_mesh)))
    S += S % 2
    V = int(round(math.sqrt(tris_per_mesh / 8)))
    n = int(round(math.sqrt(tris_per_mesh / 2)))
    meshes = []
    for k in range(8):
        noise = ValueNoise(int(rng.integers(1 << 30)), octaves=3, base_freq=float(rng.uniform(1.5, 4.0)))
        amp = float(rng.uniform(0.1, 0.35))
        kind = k % 3
        if kind == 0:
            meshes.append(displaced_grid(n, 2.0, amp, noise=noise))
        elif kind == 1:
            meshes.append(uv_sphere(S, S // 2, 1.0, noise=lambda d, f=noise, a=amp: a * f(d)))
        else:
            meshes.append(torus(4 * V, V, 1.0, 0.4, noise=lambda d, f=noise, a=amp: a * f(d)))
    return meshes


def config5(width=3840, height=2160, seed=5, tris_per_mesh=2_000_000, num_instances=256, meshes=None, num_samples=4):
    """256 TRS instances of 8 seeded displaced meshes in a 200^3 box; one
    instance in five refractive (transparency U[0.5, 0.95], IOR U[1.3, 1.7]),
    maxRecursions 6, num_samples samples per pixel."""
    rng = np.random.default_rng(seed + 1000)
    meshes = meshes if meshes is not None else config5_meshes(seed, tris_per_mesh)
    sc = Scene()
    sc.camera = Camera(fov_deg=60.0)
    sc.camera.look_at((0.0, 10.0, 130.0), (0.0, 0.0, -40.0))
    refractive = 0
    for i in range(num_instances):
        mesh = meshes[int(rng.integers(len(meshes)))]
        is_refr = rng.uniform() < 0.2
        refractive += int(is_refr)
        mat = default_material(
            diffuse=tuple(rng.uniform(0.0, 1.0, 3)),
            roughness=float(rng.uniform(0.0, 1.0)),
            ior=float(rng.uniform(1.3, 1.7)) if is_refr else 1.0,
            reflectivity=float(rng.uniform(0.3, 1.0)) if rng.uniform() < 0.3 else 0.0,
            transparency=float(rng.uniform(0.5, 0.95)) if is_refr else 0.0,
            effectId=1 if rng.uniform() < 0.1 else 0,
        )
        model = Model(mesh, np.array([mat]))
        t = Transform(position=rng.uniform(-100.0, 100.0, 3), rotation=random_quat(rng),
                      scaling=np.full(3, rng.uniform(8.0, 20.0)))
        sc.root.add_child(Entity(f"inst{i}", transform=t, model=model))
    unique = sum(m.num_faces() for m in meshes)
    inst_tris = sum(e.model.mesh.num_faces() for e in sc.root.children)
    return SceneSetup(sc, width, height, max_recursions=6, num_samples=num_samples, name="config5", seed=seed,
                      info={"unique_triangles": unique, "instanced_triangles": inst_tris, "instances": num_instances,
                            "refractive": refractive})
