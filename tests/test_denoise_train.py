"""The PyTorch training mirror of the denoiser (raygun_amd/denoise_train.py)
against the numpy network of the oracle (pinned by finite differences in
test_unet_oracle.py): same output, same loss and the same gradient of every
parameter in the library's flat layout, on the CPU in fp32; and a short
training run lowers the loss. CPU only."""
import numpy as np
import pytest

import unet_oracle as U
from raygun_amd import denoise_train as DT

torch = pytest.importorskip("torch")


def _case(seed, n=2, h=8, w=12):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, h, w, 16), np.float32)
    x[..., :11] = rng.uniform(0, 1, size=(n, h, w, 11))
    t = rng.uniform(0, 1, size=(n, h, w, 3)).astype(np.float32)
    return x, t


def test_layout_is_the_oracles():
    assert DT.LAYERS == U.LAYERS
    assert DT.param_layout() == U.param_layout()


def test_forward_loss_and_gradients_match_oracle():
    p = U.init_params(4)
    x, t = _case(1)
    net = DT.UNet(p, "cpu")
    out = net.forward(torch.from_numpy(x))
    ref = U.Net(p.astype(np.float64), rounding=False).forward(x.astype(np.float64))["out"]
    o = out.detach().numpy()
    assert np.abs(o[..., :3] - ref[..., :3]).max() < 1e-4
    assert np.all(o[..., 3:] == 0)
    loss = net.loss(torch.from_numpy(x), torch.from_numpy(t))
    loss.backward()
    l_ref, g_ref, _ = U.Net(p.astype(np.float64), rounding=False).loss_and_grads(x.astype(np.float64), t.astype(np.float64))
    assert abs(float(loss.detach()) - l_ref) <= 1e-5 * l_ref
    g = np.concatenate([np.concatenate([w.grad.numpy().reshape(-1), b.grad.numpy()]) for w, b in zip(net.w, net.b)])
    offs, _ = U.param_layout()
    for l, ((wo, bo), (cin, cout)) in enumerate(zip(offs, U.LAYERS)):
        real_out = 3 if l == len(U.LAYERS) - 1 else cout
        gw = g[wo:bo].reshape(9, cin, cout)[:, :, :real_out]
        rw = g_ref[wo:bo].reshape(9, cin, cout)[:, :, :real_out]
        assert np.linalg.norm(gw - rw) <= 1e-4 * np.linalg.norm(rw) + 1e-9, l
        assert np.linalg.norm(g[bo:bo + real_out] - g_ref[bo:bo + real_out]) <= 1e-4 * np.linalg.norm(g_ref[bo:bo + real_out]) + 1e-9, l


def test_flat_export_round_trips_and_keeps_padding_zero():
    p = U.init_params(5)
    net = DT.UNet(p, "cpu")
    q = net.flat()
    assert np.array_equal(q, p)  # init_params keeps the padded outputs zero already


def test_short_training_lowers_the_loss():
    rng = np.random.default_rng(3)
    n, h, w = 6, 16, 16
    x = np.zeros((n, h, w, 16), np.float32)
    clean = rng.uniform(0.2, 0.8, size=(n, 1, 1, 3)) * np.ones((n, h, w, 3))
    x[..., :3] = clean + rng.normal(0, 0.1, size=(n, h, w, 3))
    x[..., 3:6] = clean
    bits = (DT_bf16_bits(x))
    params, head, tail = DT.train(bits, clean.astype(np.float32), U.init_params(6), steps=120, batch=3, lr=3e-3, device="cpu")
    assert tail < 0.8 * head, (head, tail)
    assert params.shape == U.init_params(6).shape


def DT_bf16_bits(x):
    return (U.bf16(x).view(np.uint32) >> 16).astype(np.uint16)
