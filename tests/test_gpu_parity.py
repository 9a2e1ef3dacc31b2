"""GPU parity: the HIP core against the CPU oracle on the same seeded inputs.

Bars: LBVH nodes identical; closest-hit instance/primitive identical except
where the oracle reports a near-tie (|t1 - t2| <= 1e-6 t); hit t within
1e-5 relative (expected bit-exact: same tree, same watertight arithmetic);
fp32 radiance within 1e-3 absolute; post passes bit-exact given identical
inputs.
"""
import numpy as np
import pytest

from helpers import NO_HIT, compare_hits, gpu_context, half, oracle_scene, random_rays, small_config2
import oracle_py
from raygun_amd import procgen, rgrt

pytestmark = pytest.mark.gpu


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.fixture(scope="module")
def c1():
    setup = procgen.config1()
    osc, inst, bufs = oracle_scene(setup)
    ctx, _ = gpu_context(setup, flags=rgrt.CFG_FP32_RADIANCE | rgrt.CFG_PRIMARY_AOV)
    yield setup, osc, ctx
    ctx.close()


@pytest.fixture(scope="module")
def c2s():
    setup = small_config2()
    osc, inst, bufs = oracle_scene(setup)
    ctx, _ = gpu_context(setup, flags=rgrt.CFG_FP32_RADIANCE | rgrt.CFG_PRIMARY_AOV)
    yield setup, osc, ctx
    ctx.close()


def _nodes_equal(a, b):
    assert len(a) == len(b)
    assert np.array_equal(a["child"], b["child"])
    assert np.array_equal(a["parent"], b["parent"])
    # boxes compared by value (min/max may differ only in the sign of a zero)
    assert np.array_equal(a["lo"], b["lo"])
    assert np.array_equal(a["hi"], b["hi"])


def test_blas_matches_oracle_config1(c1):
    setup, osc, ctx = c1
    _print_case("blas-config1")
    g_nodes, g_leaf = ctx.export_blas(0)
    o_nodes, o_leaf = osc.export_blas(0)
    assert np.array_equal(g_leaf, o_leaf)
    _nodes_equal(g_nodes, o_nodes)


def test_blas_and_tlas_match_oracle_config2_small(c2s):
    setup, osc, ctx = c2s
    _print_case("blas-tlas-config2-small")
    for b in range(5):
        g_nodes, g_leaf = ctx.export_blas(b)
        o_nodes, o_leaf = osc.export_blas(b)
        assert np.array_equal(g_leaf, o_leaf), b
        _nodes_equal(g_nodes, o_nodes)
    g_nodes, g_leaf = ctx.export_tlas()
    o_nodes, o_leaf = osc.export_tlas()
    assert np.array_equal(g_leaf, o_leaf)
    _nodes_equal(g_nodes, o_nodes)


def _wide_equal(a, b):
    assert len(a) == len(b)
    assert np.array_equal(a["child"], b["child"])
    for k in ("lox", "hix", "loy", "hiy", "loz", "hiz"):
        assert np.array_equal(a[k], b[k]), k


def _qwide_equal(a, b):
    assert len(a) == len(b)
    assert a.tobytes() == b.tobytes()


def test_wide_trees_match_oracle(c1, c2s):
    _print_case("wide-trees")
    for (setup, osc, ctx), nblas in ((c1, 1), (c2s, 5)):
        for b in range(nblas):
            _wide_equal(ctx.export_blas_wide(b), osc.export_blas_wide(b))
            _qwide_equal(ctx.export_blas_wide(b, compressed=True), osc.export_blas_wide(b, compressed=True))
        _wide_equal(ctx.export_tlas_wide(), osc.export_tlas_wide())
        _qwide_equal(ctx.export_tlas_wide(compressed=True), osc.export_tlas_wide(compressed=True))


def test_refit_matches_oracle():
    """Deformed vertices + refit on both sides: the same binary, wide and
    compressed trees, and the same closest hits."""
    from test_oracle import _refit_case
    from raygun_amd.records import INSTANCE, default_material
    _print_case("refit", n=400)
    mesh, moved = _refit_case()
    n = mesh.num_faces()
    mats = np.array([default_material()])
    osc = oracle_py.OracleScene(mesh.vertices, mesh.indices, mats)
    ob = osc.build_blas(0, 0, n)
    ctx = rgrt.Context(64, 64, 0)
    try:
        ctx.upload_geometry(mesh.vertices, mesh.indices)
        ctx.upload_materials(mats)
        gb = ctx.build_blas(0, 0, n, len(mesh.vertices))
        osc.update_vertices(0, moved.vertices)
        osc.refit_blas(ob)
        ctx.update_vertices(0, moved.vertices)
        ctx.refit_blas(gb)
        g_nodes, g_leaf = ctx.export_blas(gb)
        o_nodes, o_leaf = osc.export_blas(ob)
        assert np.array_equal(g_leaf, o_leaf)
        _nodes_equal(g_nodes, o_nodes)
        _wide_equal(ctx.export_blas_wide(gb), osc.export_blas_wide(ob))
        _qwide_equal(ctx.export_blas_wide(gb, compressed=True), osc.export_blas_wide(ob, compressed=True))
        inst = np.zeros(1, INSTANCE)
        inst["transform"] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
        inst["mask"] = 0xFF
        inst["blas"] = gb
        ctx.build_tlas(inst)
        inst["blas"] = ob
        osc.build_tlas(inst)
        rays = random_rays(50_000, seed=13, radius=12.0)
        hg = ctx.trace_rays(rays)
        ho, tie = osc.trace_rays(rays)
        assert len(compare_hits(hg, ho, tie)) == 0
    finally:
        ctx.close()


@pytest.mark.parametrize("which", ["config1", "config2s"])
def test_ray_queries_match_oracle(which, c1, c2s):
    setup, osc, ctx = c1 if which == "config1" else c2s
    n = 200_000
    rays = random_rays(n, seed=11, radius=3.0) if which == "config1" else random_rays(n, seed=12, radius=90.0)
    _print_case("ray-queries", scene=which, n=n)
    hg = ctx.trace_rays(rays)
    ho, tie = osc.trace_rays(rays)
    bad = compare_hits(hg, ho, tie)
    assert len(bad) == 0, f"{len(bad)} id mismatches without a tie, first {bad[:5]}"
    both = (hg["instance"] != NO_HIT) & (ho["instance"] != NO_HIT) & ~tie
    assert both.sum() > n // 10
    rel = np.abs(hg["t"][both] - ho["t"][both]) / np.maximum(np.abs(ho["t"][both]), 1e-30)
    assert rel.max() <= 1e-5
    # expected bit-exact (same tree, same arithmetic)
    assert np.mean(hg["t"][both] == ho["t"][both]) > 0.999
    assert np.array_equal(hg["u"][both] == ho["u"][both], np.ones(both.sum(), bool)) or np.mean(hg["u"][both] == ho["u"][both]) > 0.999


def test_cull_mask_zero_never_hits(c1):
    setup, osc, ctx = c1
    rays = random_rays(4096, seed=3)
    h = ctx.trace_rays(rays, cull_mask=0)
    assert np.all(h["instance"] == NO_HIT)


def _render_compare(setup, osc, ctx, W, H, tag):
    ubo = setup.ubo(W, H)
    ctx.resize(W, H)
    ctx.set_frame(ubo)
    _print_case("render", scene=tag, W=W, H=H, maxRec=int(ubo["maxRecursions"]), spp=int(ubo["numSamples"]))
    ctx.trace()
    g = {k: ctx.read_image(i) for k, i in [("base32", rgrt.IMAGE_BASE_F32), ("normal32", rgrt.IMAGE_NORMAL_F32),
                                            ("rough32", rgrt.IMAGE_ROUGH_F32), ("base16", rgrt.IMAGE_BASE),
                                            ("normal16", rgrt.IMAGE_NORMAL), ("rough16", rgrt.IMAGE_ROUGH),
                                            ("primary", rgrt.IMAGE_PRIMARY_HIT)]}
    st = ctx.stats()
    o = osc.render(ubo, W, H)
    tie = o["tie"].astype(bool)
    ok = ~tie
    for k in ("base32", "normal32", "rough32"):
        a, b = g[k], o[k]
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin), k
        d = np.where(fin, np.abs(a - b), 0.0).max(axis=-1)
        assert d[ok].max() <= 1e-3, (k, float(d[ok].max()), np.argwhere(d * ok > 1e-3)[:5])
    pg, po = g["primary"], o["primary"]
    assert np.array_equal(pg["instance"][ok], po["instance"][ok])
    assert np.array_equal(pg["primitive"][ok], po["primitive"][ok])
    hitm = ok & (po["instance"] != NO_HIT)
    assert np.array_equal(pg["t"][hitm], po["t"][hitm])
    return g, o, st


def test_render_config1_matches_oracle(c1):
    setup, osc, ctx = c1
    g, o, st = _render_compare(setup, osc, ctx, 640, 360, "config1")
    if o["tie"].sum() == 0:
        assert st["rays"] == o["rays"]
        assert st["skyLookups"] == o["sky"]


def test_render_config2_small_matches_oracle(c2s):
    setup, osc, ctx = c2s
    g, o, st = _render_compare(setup, osc, ctx, 320, 180, "config2s")
    assert abs(st["rays"] - o["rays"]) <= max(10, o["rays"] // 10000)


@pytest.mark.parametrize("max_rec", [0, 2, 7])
def test_render_recursion_depths(c2s, max_rec):
    setup, osc, ctx = c2s
    old = setup.max_recursions
    setup.max_recursions = max_rec
    try:
        _render_compare(setup, osc, ctx, 160, 90, f"config2s-rec{max_rec}")
    finally:
        setup.max_recursions = old


@pytest.mark.parametrize("spp", [2, 5])
def test_render_ssaa_samples(c1, spp):
    setup, osc, ctx = c1
    old = setup.num_samples
    setup.num_samples = spp
    try:
        _render_compare(setup, osc, ctx, 160, 90, f"config1-spp{spp}")
    finally:
        setup.num_samples = old


@pytest.mark.parametrize("fxaa", [True, False])
def test_post_bit_exact(c1, fxaa):
    setup, osc, ctx = c1
    W, H = 640, 360
    ubo = setup.ubo(W, H)
    ctx.resize(W, H)
    ctx.set_frame(ubo)
    o = osc.render(ubo, W, H)
    _print_case("post", fxaa=fxaa)
    ctx.write_image(rgrt.IMAGE_BASE, o["base16"])
    ctx.write_image(rgrt.IMAGE_NORMAL, o["normal16"])
    ctx.write_image(rgrt.IMAGE_ROUGH, o["rough16"])
    ctx.postprocess(fxaa)
    ref = oracle_py.postprocess(ubo, o["base16"], o["normal16"], o["rough16"], use_fxaa=fxaa)
    assert np.array_equal(ctx.read_image(rgrt.IMAGE_ROUGH_TRANSITIONS), ref["transitions"])
    assert np.array_equal(ctx.read_image(rgrt.IMAGE_ROUGH_COLORS_A), ref["colorsA"])
    assert np.array_equal(ctx.read_image(rgrt.IMAGE_ROUGH_COLORS_B), ref["colorsB"])
    assert np.array_equal(ctx.read_image(rgrt.IMAGE_FINAL), ref["final"])
    assert np.array_equal(ctx.read_image(rgrt.IMAGE_BASE), ref["base"])
