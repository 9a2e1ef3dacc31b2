"""Shared test helpers: one scene set up in the HIP core and in the oracle."""
import math

import numpy as np

import oracle_py
from raygun_amd import procgen
from raygun_amd.records import HIT, RAY, VERTEX
from raygun_amd.scene import pack_models

NO_HIT = 0xFFFFFFFF


def half(img16):
    return np.asarray(img16, np.uint16).view(np.float16).astype(np.float32)


def oracle_scene(setup):
    """OracleScene + instances for a procgen.SceneSetup."""
    models = setup.scene.models()
    v, i, m = pack_models(models)
    osc = oracle_py.OracleScene(v, i, m)
    blas_of = {}
    for mod in models:
        k = id(mod.mesh)
        if k not in blas_of:
            blas_of[k] = osc.build_blas(mod.mesh.first_vertex, mod.mesh.index_offset_bytes, mod.mesh.num_faces())
        mod.blas = blas_of[k]
    inst = setup.scene.collect_instances()
    osc.build_tlas(inst)
    return osc, inst, (v, i, m)


def gpu_context(setup, width=None, height=None, flags=0):
    """rgrt.Context with the same scene (same BLAS ids as oracle_scene)."""
    from raygun_amd import rgrt
    W, H = width or setup.width, height or setup.height
    ctx = rgrt.Context(W, H, flags)
    models = setup.scene.models()
    v, i, m = pack_models(models)
    ctx.upload_geometry(v, i)
    ctx.upload_materials(m)
    blas_of = {}
    for mod in models:
        k = id(mod.mesh)
        if k not in blas_of:
            blas_of[k] = ctx.build_blas(mod.mesh.first_vertex, mod.mesh.index_offset_bytes, mod.mesh.num_faces(),
                                        len(mod.mesh.vertices))
        mod.blas = blas_of[k]
    inst = setup.scene.collect_instances()
    ctx.build_tlas(inst)
    ctx.set_frame(setup.ubo(W, H))
    return ctx, inst


def random_rays(n, seed, center=(0, 0, 0), radius=3.0, tmin=0.001, tmax=1e4):
    """Seeded rays from a sphere shell towards jittered points near `center`."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center, np.float64)
    o = rng.normal(size=(n, 3))
    o = c + radius * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = c + rng.uniform(-0.6, 0.6, size=(n, 3)) * radius * 0.5
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, RAY)
    r["origin"] = o.astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def seam_scene(n=64, rays=50_000, seed=1, x0=3000.0, z0=100.0):
    """A ground grid of n x n unit quads (2 triangles each) at z = z0 with its
    corner at (x0, x0), and seeded steep rays that hit it on the lines where
    the LBVH splits the grid (x or y = x0 + n/2, n/4, 3n/4, n/8), from 0.05 to
    3 units above. The child boxes of those splits meet there, the hit is
    close to the ray origin, and the origin is far from the nodes' corners
    relative to 1/|d|: the box test's plane distances then carry absolute
    rounding errors (of (corner - o)/d) far above its relative widening.
    Returns vertices, indices, one instance record and the rays."""
    from raygun_amd.records import INSTANCE
    xs = x0 + np.arange(n + 1, dtype=np.float64)
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    v = np.zeros((n + 1) * (n + 1), VERTEX)
    v["position"][:, 0] = gx.ravel()
    v["position"][:, 1] = gy.ravel()
    v["position"][:, 2] = z0
    v["normal"][:, 2] = 1.0
    a = (np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]).ravel()
    quads = np.stack([a, a + n + 1, a + 1, a + 1, a + n + 1, a + n + 2], 1)
    idx = quads.astype(np.uint32).ravel()
    inst = np.zeros(1, INSTANCE)
    inst["transform"][0] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    inst["mask"] = 0xFF
    inst["flags"] = 1
    rng = np.random.default_rng(seed)
    seams = x0 + np.array([n // 2, n // 4, 3 * n // 4, n // 8], np.float64)
    p = np.zeros((rays, 3))
    p[:, 0] = rng.choice(seams, rays)
    p[:, 1] = x0 + rng.uniform(0.5, n - 0.5, rays)
    p[:, 2] = z0
    swap = rng.random(rays) < 0.5
    p[swap, 0], p[swap, 1] = p[swap, 1].copy(), p[swap, 0].copy()
    d = np.stack([rng.normal(0, 0.02, rays), rng.normal(0, 0.02, rays), -np.ones(rays)], 1)
    h = rng.uniform(0.05, 3.0, rays)
    r = np.zeros(rays, RAY)
    r["origin"] = (p - d * h[:, None]).astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = 1e4
    return v, idx, inst, r


def small_config2(width=320, height=180, tris_per_mesh=20_000, seed=2):
    return procgen.config2(width, height, seed=seed, tris_per_mesh=tris_per_mesh)


def compare_hits(h_gpu, h_ref, tie):
    """Indices where the closest-hit identity differs and the oracle saw no tie."""
    same_id = (h_gpu["instance"] == h_ref["instance"]) & ((h_gpu["primitive"] == h_ref["primitive"]) | (h_ref["instance"] == NO_HIT))
    return np.nonzero(~same_id & ~tie)[0]


def sky64(d, L):
    """miss.rmiss:38-74 evaluated in float64 (a second, independent reading);
    a zero direction reads rayDir = 0 (the oracle's definition for the
    zero-direction ray of total internal reflection)."""
    d = np.asarray(d, np.float64)
    L = np.asarray(L, np.float64)
    nd = np.linalg.norm(d)
    rd = d / nd if nd > 0 else np.zeros(3)
    y = abs(d[1] + 1.5) / 3.0
    sun = 1.0 - np.linalg.norm(rd - (-L / np.linalg.norm(L)))
    sun = min(max(sun, 0.0), 2.0)
    glow = min(max(sun, 0.0), 1.0)
    sun = min(max(sun ** 80 * 1000.0, 0.0), 16.0)
    glow = min(max((glow ** 6) ** y, 0.0), 1.0)
    sun *= (y * y) ** (1 / 1.65)
    glow *= (y * y) ** 0.5
    sun += glow
    sun_color = np.array([1.0, 0.6, 0.05]) * sun
    atmosphere = math.sqrt(1.0 - y)
    scatter = 1.0 - min(max((4.0 - L[1]) ** (1 / 15), 0.8), 1.0)
    scatter_color = np.ones(3) * (1 - scatter) + np.array([1.0, 0.3, 0.0]) * 1.5 * scatter
    a = atmosphere / 1.3
    sky = np.array([0.2, 0.4, 0.8]) * (1 - a) + scatter_color * a
    return sun_color + sky
