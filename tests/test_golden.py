"""The oracle reproduces the committed golden fixtures (tests/golden/) exactly,
and the HIP core matches them on the GPU."""
import os

import numpy as np
import pytest

import oracle_py
from helpers import NO_HIT, oracle_scene
from raygun_amd import procgen
from raygun_amd.records import HIT, RAY, UBO

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
W, H = 64, 36


def _load():
    g = np.load(os.path.join(HERE, "config1_64x36.npz"))
    r = np.load(os.path.join(HERE, "config1_rays.npz"))
    return g, r


def test_oracle_reproduces_golden_render():
    g, _ = _load()
    setup = procgen.config1(W, H)
    osc, _, _ = oracle_scene(setup)
    ubo = np.frombuffer(g["ubo"].tobytes(), UBO)[0]
    out = osc.render(ubo, W, H, threads=2)
    for k in ("base16", "normal16", "rough16"):
        assert np.array_equal(out[k], g[k]), k
    assert np.array_equal(out["primary"], np.frombuffer(g["primary"].tobytes(), HIT).reshape(H, W))
    post = oracle_py.postprocess(ubo, out["base16"], out["normal16"], out["rough16"], use_fxaa=True)
    assert np.array_equal(post["final"], g["final16"])
    assert np.array_equal(post["colorsA"], g["colorsA16"])
    assert np.array_equal(post["transitions"], g["transitions"])


def test_oracle_reproduces_golden_hits():
    _, r = _load()
    setup = procgen.config1(W, H)
    osc, _, _ = oracle_scene(setup)
    rays = np.frombuffer(r["rays"].tobytes(), RAY)
    hits, tie = osc.trace_rays(rays, threads=2)
    assert np.array_equal(hits, np.frombuffer(r["hits"].tobytes(), HIT))
    t64, inst64, prim64, _ = osc.trace_rays_fp64(rays, threads=2)
    assert np.array_equal(t64, r["t64"]) and np.array_equal(prim64, r["prim64"])


@pytest.mark.gpu
def test_gpu_matches_golden_fixtures():
    from helpers import gpu_context
    from raygun_amd import rgrt
    g, r = _load()
    setup = procgen.config1(W, H)
    ctx, _ = gpu_context(setup, W, H)
    ubo = np.frombuffer(g["ubo"].tobytes(), UBO)[0]
    ctx.set_frame(ubo)
    ctx.trace()
    tie = g["tie"].astype(bool)
    base = ctx.read_image(rgrt.IMAGE_BASE)
    normal = ctx.read_image(rgrt.IMAGE_NORMAL)
    rough = ctx.read_image(rgrt.IMAGE_ROUGH)
    # the fp16 images may differ in the last bit where transcendental
    # functions round differently (log, pow); never by more than 1e-3
    for got, k in ((base, "base16"), (normal, "normal16"), (rough, "rough16")):
        a = got.view(np.float16).astype(np.float32)[~tie]
        b = g[k].view(np.float16).astype(np.float32)[~tie]
        assert np.abs(a - b).max() <= 1e-3 * np.maximum(1.0, np.abs(b)).max()
    rays = np.frombuffer(r["rays"].tobytes(), RAY)
    hits = ctx.trace_rays(rays)
    ref = np.frombuffer(r["hits"].tobytes(), HIT)
    ok = ~r["tie"]
    assert np.array_equal(hits["instance"][ok], ref["instance"][ok])
    hit = ok & (ref["instance"] != NO_HIT)
    assert np.array_equal(hits["primitive"][hit], ref["primitive"][hit])
    assert np.array_equal(hits["t"][hit], ref["t"][hit])
    ctx.close()


def test_config5_scene_family():
    """Config 5's generator: 8 meshes, 256 instances, one in five refractive
    with IOR in [1.3, 1.7], maxRecursions 6, 4 samples; seeded."""
    from raygun_amd import procgen
    a = procgen.config5(64, 36, tris_per_mesh=2_000)
    b = procgen.config5(64, 36, tris_per_mesh=2_000)
    assert a.info == b.info
    assert a.max_recursions == 6 and a.num_samples == 4
    assert a.info["instances"] == 256 and 30 <= a.info["refractive"] <= 75
    mats = [e.model.materials[0] for e in a.scene.root.children]
    for m in mats:
        if m["transparency"] > 0:
            assert 1.3 <= m["ior"] <= 1.7
    assert len({id(e.model.mesh) for e in a.scene.root.children}) == 8
    ubo = a.ubo()
    assert int(ubo["maxRecursions"]) == 6 and int(ubo["numSamples"]) == 4
