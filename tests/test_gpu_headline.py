"""Parity at the measured sizes: the whole 1920x1080 frame of the 2M-triangle
scene (64 instances of five 400k-triangle meshes) and the config-5 scene
family at 960x540 with 200k triangles per mesh (8 meshes, 256 instances,
maxRecursions 6, 4 spp) against the oracle — every BLAS and the TLAS
byte-equal (binary and compressed wide trees), fp32 radiance within 1e-3 on
every pixel off near-ties, primary instance/primitive/t identical, ray
counts; plus 200k seeded ray queries through the full scene. The deep
trees of these meshes take the traversal stack past its LDS part into the
global spill region, which the reduced scenes never reach."""
import numpy as np
import pytest

from helpers import NO_HIT, compare_hits, gpu_context, oracle_scene, random_rays
from raygun_amd import procgen, rgrt
from test_gpu_parity import _print_case, _qwide_equal, _render_compare

pytestmark = pytest.mark.gpu

FLAGS = rgrt.CFG_FP32_RADIANCE | rgrt.CFG_PRIMARY_AOV


@pytest.fixture(scope="module")
def c2full():
    setup = procgen.config2()
    osc, inst, _ = oracle_scene(setup)
    ctx, _ = gpu_context(setup, flags=FLAGS)
    yield setup, osc, ctx
    ctx.close()


def test_config2_full_trees_match_oracle(c2full):
    setup, osc, ctx = c2full
    _print_case("trees-config2-full")
    for b in range(5):
        g_nodes, g_leaf = ctx.export_blas(b)
        o_nodes, o_leaf = osc.export_blas(b)
        assert len(o_leaf) > 350_000
        assert np.array_equal(g_leaf, o_leaf), b
        assert g_nodes.tobytes() == o_nodes.tobytes(), b
        _qwide_equal(ctx.export_blas_wide(b, compressed=True), osc.export_blas_wide(b, compressed=True))
    g_nodes, g_leaf = ctx.export_tlas()
    o_nodes, o_leaf = osc.export_tlas()
    assert np.array_equal(g_leaf, o_leaf)
    assert g_nodes.tobytes() == o_nodes.tobytes()
    _qwide_equal(ctx.export_tlas_wide(compressed=True), osc.export_tlas_wide(compressed=True))


def test_config2_full_frame_matches_oracle(c2full):
    setup, osc, ctx = c2full
    g, o, st = _render_compare(setup, osc, ctx, 1920, 1080, "config2-full")
    assert o["rays"] > 7_000_000
    assert abs(st["rays"] - o["rays"]) <= max(10, o["rays"] // 100000)


def test_config2_full_ray_queries_match_oracle(c2full):
    setup, osc, ctx = c2full
    n = 200_000
    rays = random_rays(n, seed=31, center=(0.0, 0.0, -20.0), radius=90.0)
    _print_case("ray-queries", scene="config2-full", n=n)
    hg = ctx.trace_rays(rays)
    ho, tie = osc.trace_rays(rays)
    bad = compare_hits(hg, ho, tie)
    assert len(bad) == 0, f"{len(bad)} id mismatches without a tie, first {bad[:5]}"
    both = (hg["instance"] != NO_HIT) & (ho["instance"] != NO_HIT) & ~tie
    assert both.sum() > n // 10
    assert np.array_equal(hg["t"][both], ho["t"][both])


def test_config5_960x540_matches_oracle():
    setup = procgen.config5(960, 540, tris_per_mesh=200_000)
    osc, _, _ = oracle_scene(setup)
    ctx, _ = gpu_context(setup, flags=FLAGS)
    try:
        _print_case("trees-config5-960x540")
        for b in range(8):
            g_nodes, g_leaf = ctx.export_blas(b)
            o_nodes, o_leaf = osc.export_blas(b)
            assert np.array_equal(g_leaf, o_leaf), b
            assert g_nodes.tobytes() == o_nodes.tobytes(), b
        g_nodes, g_leaf = ctx.export_tlas()
        o_nodes, o_leaf = osc.export_tlas()
        assert np.array_equal(g_leaf, o_leaf) and g_nodes.tobytes() == o_nodes.tobytes()
        g, o, st = _render_compare(setup, osc, ctx, 960, 540, "config5-960x540")
        assert abs(st["rays"] - o["rays"]) <= max(10, o["rays"] // 10000)
    finally:
        ctx.close()
