"""GPU parity of the TLAS build over instance counts on both sides of the
one-workgroup builds (2..256 instances in 256 threads, 257..1024 in 1024)
and of the multi-kernel LBVH above it: binary nodes, leaf order, wide and
compressed nodes identical to the oracle's, and closest hits identical on seeded rays. Every seventh instance
repeats an earlier position (equal Morton codes, so the sort's tie order
matters) and the instances use two BLASes, random rotations, scales and
customIndex permutations.
"""
import numpy as np
import pytest

from helpers import compare_hits
import oracle_py
from raygun_amd import procgen, rgrt
from raygun_amd.records import INSTANCE, RAY, default_material
from raygun_amd.scene import quat_to_mat3

pytestmark = pytest.mark.gpu


def _instances(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-30.0, 30.0, size=(n, 3))
    for i in range(7, n, 7):
        pos[i] = pos[i // 7]
    inst = np.zeros(n, INSTANCE)
    for i in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        m = np.zeros((3, 4))
        m[:, :3] = quat_to_mat3(q) * rng.uniform(0.3, 2.0)
        m[:, 3] = pos[i]
        inst["transform"][i] = m.reshape(-1)
    inst["customIndex"] = rng.permutation(n)
    inst["mask"] = 0xFF
    inst["blas"] = rng.integers(0, 2, size=n)
    return inst


def _rays(n, seed, inst):
    """Rays from a box around the scene towards points near random instances."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-45.0, 45.0, size=(n, 3))
    centre = inst["transform"].reshape(-1, 3, 4)[:, :, 3].astype(np.float64)
    tgt = centre[rng.integers(0, len(inst), size=n)] + rng.uniform(-1.0, 1.0, size=(n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, RAY)
    r["origin"] = o.astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["tmin"] = 0.001
    r["tmax"] = 1e4
    return r


@pytest.fixture(scope="module")
def scene():
    a = procgen.uv_sphere(24, 12)
    b = procgen.torus(24, 12)
    v = np.concatenate([a.vertices, b.vertices])
    i = np.concatenate([a.indices, b.indices]).astype(np.uint32)
    fa, fb = len(a.indices) // 3, len(b.indices) // 3
    mats = np.array([default_material()])
    osc = oracle_py.OracleScene(v, i, mats)
    ob = [osc.build_blas(0, 0, fa), osc.build_blas(len(a.vertices), len(a.indices) * 4, fb)]
    ctx = rgrt.Context(64, 64, 0)
    ctx.upload_geometry(v, i)
    ctx.upload_materials(mats)
    gb = [ctx.build_blas(0, 0, fa, len(a.vertices)), ctx.build_blas(len(a.vertices), len(a.indices) * 4, fb, len(b.vertices))]
    assert ob == gb == [0, 1]
    yield osc, ctx
    ctx.close()


@pytest.mark.parametrize("n", [2, 3, 64, 256, 257, 1024, 1025, 3000])
def test_tlas_matches_oracle(scene, n):
    osc, ctx = scene
    print(f"[case] tlas n={n}", flush=True)
    inst = _instances(n, seed=100 + n)
    ctx.build_tlas(inst)
    osc.build_tlas(inst)
    g_nodes, g_leaf = ctx.export_tlas()
    o_nodes, o_leaf = osc.export_tlas()
    assert np.array_equal(g_leaf, o_leaf)
    for k in ("child", "parent", "lo", "hi"):
        assert np.array_equal(g_nodes[k], o_nodes[k]), k
    gw, ow = ctx.export_tlas_wide(), osc.export_tlas_wide()
    assert len(gw) == len(ow)
    for k in ("child", "lox", "hix", "loy", "hiy", "loz", "hiz"):
        assert np.array_equal(gw[k], ow[k]), k
    assert ctx.export_tlas_wide(compressed=True).tobytes() == osc.export_tlas_wide(compressed=True).tobytes()
    rays = _rays(20_000, seed=n, inst=inst)
    hg = ctx.trace_rays(rays)
    ho, tie = osc.trace_rays(rays)
    assert len(compare_hits(hg, ho, tie)) == 0
    hit = ho["instance"] != 0xFFFFFFFF
    assert hit.mean() > 0.2
    assert np.array_equal(hg["t"][hit & ~tie], ho["t"][hit & ~tie])


def _moved(inst, seed):
    rng = np.random.default_rng(seed)
    out = inst.copy()
    t = out["transform"].reshape(-1, 3, 4)
    t[:, :, 3] += rng.uniform(-8.0, 8.0, size=(len(inst), 3))
    out["transform"] = t.reshape(-1, 12)
    return out


@pytest.mark.parametrize("n", [1, 2, 64, 256, 257, 1024, 1500])
def test_tlas_refit_matches_oracle(scene, n):
    """Build, move every instance, refit on both sides: the same binary, wide
    and compressed trees (hierarchy and leaf order of the build kept) and the
    same closest hits."""
    osc, ctx = scene
    print(f"[case] tlas-refit n={n}", flush=True)
    inst = _instances(n, seed=500 + n)
    ctx.build_tlas(inst)
    osc.build_tlas(inst)
    g0, l0 = ctx.export_tlas()
    moved = _moved(inst, seed=n)
    ctx.refit_tlas(moved)
    osc.refit_tlas(moved)
    g_nodes, g_leaf = ctx.export_tlas()
    o_nodes, o_leaf = osc.export_tlas()
    assert np.array_equal(g_leaf, l0) and np.array_equal(g_nodes["child"], g0["child"])
    assert np.array_equal(g_leaf, o_leaf)
    for k in ("child", "parent", "lo", "hi"):
        assert np.array_equal(g_nodes[k], o_nodes[k]), k
    gw, ow = ctx.export_tlas_wide(), osc.export_tlas_wide()
    assert len(gw) == len(ow)
    for k in ("child", "lox", "hix", "loy", "hiy", "loz", "hiz"):
        assert np.array_equal(gw[k], ow[k]), k
    assert ctx.export_tlas_wide(compressed=True).tobytes() == osc.export_tlas_wide(compressed=True).tobytes()
    rays = _rays(20_000, seed=7 * n, inst=moved)
    hg = ctx.trace_rays(rays)
    ho, tie = osc.trace_rays(rays)
    assert len(compare_hits(hg, ho, tie)) == 0
    hit = ho["instance"] != 0xFFFFFFFF
    assert hit.mean() > 0.2
    assert np.array_equal(hg["t"][hit & ~tie], ho["t"][hit & ~tie])


def test_tlas_refit_rejects_changed_instances(scene):
    osc, ctx = scene
    inst = _instances(10, seed=9)
    ctx.build_tlas(inst)
    with pytest.raises(rgrt.RgrtError):
        ctx.refit_tlas(inst[:9])
    other = inst.copy()
    other["blas"][0] ^= 1
    with pytest.raises(rgrt.RgrtError):
        ctx.refit_tlas(other)
    ctx.refit_tlas(inst)  # the context still works
