"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

The reference has no golden vectors for this path, so these are the
project's own: seeded inputs and the oracle's outputs, committed so that any
later change of the oracle (or of the scene generators) shows up as a diff.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_py  # noqa: E402
from helpers import oracle_scene, random_rays  # noqa: E402
from raygun_amd import procgen  # noqa: E402

W, H = 64, 36


def make():
    setup = procgen.config1(W, H)
    osc, inst, (v, i, m) = oracle_scene(setup)
    ubo = setup.ubo(W, H)
    out = osc.render(ubo, W, H, threads=1)
    post = oracle_py.postprocess(ubo, out["base16"], out["normal16"], out["rough16"], use_fxaa=True)
    rays = random_rays(2000, seed=21, radius=3.0)
    hits, tie = osc.trace_rays(rays, threads=1)
    t64, inst64, prim64, uv64 = osc.trace_rays_fp64(rays, threads=1)
    np.savez_compressed(
        os.path.join(HERE, "config1_64x36.npz"),
        ubo=np.frombuffer(ubo.tobytes(), np.uint8),
        base16=out["base16"], normal16=out["normal16"], rough16=out["rough16"],
        primary=np.frombuffer(out["primary"].tobytes(), np.uint8),
        tie=out["tie"], rays_count=np.array([out["rays"], out["sky"], out["zero_dir"]], np.int64),
        final16=post["final"], colorsA16=post["colorsA"], transitions=post["transitions"],
    )
    np.savez_compressed(
        os.path.join(HERE, "config1_rays.npz"),
        rays=np.frombuffer(rays.tobytes(), np.uint8),
        hits=np.frombuffer(hits.tobytes(), np.uint8), tie=tie,
        t64=t64, inst64=inst64, prim64=prim64,
    )


if __name__ == "__main__":
    make()
    print("written", os.listdir(HERE))
