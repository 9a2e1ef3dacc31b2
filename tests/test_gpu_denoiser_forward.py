"""The denoiser's forward pass (rgrt_denoiser_infer: every convolution an
implicit GEMM on the bf16 matrix cores) against the numpy oracle
(oracle/unet_oracle.py) with the same bf16 rounding of inputs, weights and
activations: the three used output channels within 1e-2 relative (fp32
accumulation in another order), the padded channels exactly zero; image
widths that are not a multiple of the 16-pixel strips; parameters loaded
with set_params read back unchanged."""
import numpy as np
import pytest

import unet_oracle as U
from raygun_amd import rgrt

pytestmark = pytest.mark.gpu


def _input(seed, n, h, w):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, h, w, 16), np.float32)
    x[..., :11] = rng.uniform(0, 1, size=(n, h, w, 11))
    x = U.bf16(x)
    return x, (x.view(np.uint32) >> 16).astype(np.uint16)


@pytest.mark.parametrize("n,h,w", [(2, 32, 48), (1, 20, 36), (1, 64, 100)])
def test_forward_matches_oracle(n, h, w):
    print(f"[case] denoiser forward n={n} {w}x{h}", flush=True)
    ctx = rgrt.Context(16, 16, 0)
    den = rgrt.Denoiser(ctx, seed=3)
    try:
        p = U.init_params(7)
        den.set_params(p)
        assert np.array_equal(den.get_params(), p)
        x, bits = _input(n * 1000 + h, n, h, w)
        out = den.infer(bits)
        o = (out.astype(np.uint32) << 16).view(np.float32)
        ref = U.Net(p, rounding=True).forward(x.astype(np.float64))["out"]
        assert np.all(o[..., 3:] == 0)
        assert np.all(np.isfinite(o))
        rel = np.linalg.norm(o[..., :3] - ref[..., :3]) / np.linalg.norm(ref[..., :3])
        assert rel < 1e-2, rel
        # per pixel: no pixel far off (a wrong tap or a dropped strip tail
        # would show as whole columns or rows apart)
        err = np.abs(o[..., :3] - ref[..., :3]).max(axis=-1)
        assert err.max() <= 0.05 * np.abs(ref[..., :3]).max() + 1e-3, float(err.max())
    finally:
        den.close()
        ctx.close()


def test_infer_refuses_bad_sizes():
    ctx = rgrt.Context(16, 16, 0)
    den = rgrt.Denoiser(ctx, seed=3)
    try:
        with pytest.raises(rgrt.RgrtError):
            den.infer(np.zeros((1, 18, 32, 16), np.uint16))
    finally:
        den.close()
        ctx.close()
