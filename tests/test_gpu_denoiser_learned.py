"""The learned path end to end on the GPU: training pairs rendered by the
tracer (4-spp stochastic frames, 256-spp means), the network trained on the
GPU (raygun_amd/denoise_train.py in a child process; the library's own
backward kernels are switched off), loaded into the library and applied by
rgrt_denoise (bf16 MFMA inference): on held-out scenes the denoised frames
gain PSNR over the noisy ones against their 256-spp means, which is what the
network adds over the plain tracer at the same sample count. The library's
inference of the trained parameters matches the training network's fp32
forward pass (numpy oracle with bf16 rounding) within bf16 tolerances. Fallback tiles keep the
traced frame exactly."""
import numpy as np
import pytest

import unet_oracle as U
from raygun_amd import denoise as D
from raygun_amd import rgrt

pytestmark = pytest.mark.gpu


def test_trained_network_gains_psnr_on_held_out_scenes():
    print("[case] denoiser: train on 300 traced 64x64 scenes, evaluate 4 held-out", flush=True)
    ctx = rgrt.Context(64, 64, 0)
    den = rgrt.Denoiser(ctx, seed=1)
    try:
        scenes = D.GlossyScenes(ctx, 64, 64)
        first, last, t_render, t_train = D.train_network(ctx, den, scenes, 300, first_seed=100, steps=600, batch=16, lr=1e-3)
        print(f"[case] loss {first:.5f} -> {last:.5f} render {t_render:.1f}s train {t_train:.1f}s", flush=True)
        assert last < 0.5 * first
        gains = []
        for k in range(4):
            ubo = scenes.scene(9000 + k, 4)
            ref = np.clip(D.reference_frame(ctx, ubo, 256, seed=777 + 64 * k), 0, 1)
            noisy = D.noisy_frame(ctx, ubo, 4, seed=99 + k)
            dn, (fb, tiles) = D.denoised_frame(ctx, den, ubo, 4, seed=99 + k, fallback=(1e30, 1e30))
            assert fb == 0 and tiles == 16
            gains.append(D.psnr(dn, ref) - D.psnr(noisy, ref))
        print(f"[case] PSNR gains {np.round(gains, 2).tolist()}", flush=True)
        assert min(gains) > 0.5 and np.mean(gains) > 2.0
        # the library's bf16 inference of these parameters against the
        # oracle's network with the same rounding, on one packed frame
        x, _ = den.dataset_read(0, 1, 64, 64)
        out = den.infer(x)
        xf = (x.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        ref_out = U.Net(den.get_params(), rounding=True).forward(xf)["out"]
        o = (out.astype(np.uint32) << 16).view(np.float32)
        rel = np.linalg.norm(o[..., :3] - ref_out[..., :3]) / np.linalg.norm(ref_out[..., :3])
        # bf16 storage after every layer: the two sides accumulate in another
        # order, so values near a rounding boundary or at zero before a ReLU
        # can land on different sides of it; the denoised colours agree closely
        y, y_ref = xf[..., :3] + o[..., :3], xf[..., :3] + ref_out[..., :3]
        print(f"[case] trained inference vs oracle: rel(correction) {rel:.4f}, max |dy| {np.abs(y - y_ref).max():.4f}",
              flush=True)
        assert rel < 5e-2, rel
        assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < 5e-3
    finally:
        den.close()
        ctx.close()


def test_fallback_tiles_keep_the_traced_frame():
    print("[case] denoiser fallback", flush=True)
    ctx = rgrt.Context(64, 32, 0)
    den = rgrt.Denoiser(ctx, seed=2)
    try:
        scenes = D.GlossyScenes(ctx, 64, 32)
        ubo = scenes.scene(11, 4)
        for _ in D.trace_mean(ctx, ubo, 4, 1, 5):
            pass
        with pytest.raises(rgrt.RgrtError, match="no trained parameters"):
            den.denoise()
        den.set_params(U.init_params(3))
        traced = ctx.read_image(rgrt.IMAGE_BASE).copy()
        fb, tiles = den.denoise(-1.0, -1.0, count=True)  # every tile over the threshold
        assert (fb, tiles) == (8, 8)
        assert np.array_equal(ctx.read_image(rgrt.IMAGE_BASE), traced)
        for _ in D.trace_mean(ctx, ubo, 4, 1, 5):
            pass
        fb, tiles = den.denoise(1e30, 1e30, count=True)
        assert fb == 0
        assert not np.array_equal(ctx.read_image(rgrt.IMAGE_BASE), traced)
        ctx.set_sampling(rgrt.SAMPLING_WHITTED)
        with pytest.raises(rgrt.RgrtError):
            den.denoise()
    finally:
        den.close()
        ctx.close()
