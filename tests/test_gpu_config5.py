"""Config 5's scene family (256 instances of displaced meshes, one in five
refractive, maxRecursions 6, 4 samples per pixel) at a reduced size: the HIP
core against the oracle on the same trees — radiance within 1e-3, primary
hit identities and t equal except at the oracle's exact t-ties."""
import pytest

from helpers import gpu_context, oracle_scene
from raygun_amd import procgen, rgrt
from test_gpu_parity import _render_compare

pytestmark = pytest.mark.gpu


def test_render_config5_small_matches_oracle():
    setup = procgen.config5(192, 108, tris_per_mesh=3_000)
    osc, _, _ = oracle_scene(setup)
    ctx, _ = gpu_context(setup, flags=rgrt.CFG_FP32_RADIANCE | rgrt.CFG_PRIMARY_AOV)
    try:
        g, o, st = _render_compare(setup, osc, ctx, 192, 108, "config5s")
        assert abs(st["rays"] - o["rays"]) <= max(10, o["rays"] // 10000)
    finally:
        ctx.close()
