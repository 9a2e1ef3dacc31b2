"""Pins the numpy denoiser oracle (oracle/unet_oracle.py): its analytic
gradients in float64 against central finite differences of its own loss, for
parameters of every layer; the bfloat16 rounding against hand-checked values
(ties to even, carries into the exponent); and the rounded form staying close
to the exact one. CPU only."""
import numpy as np
import pytest

import unet_oracle as U


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # bf16 spacing at 1
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -(1.0 + 2 ** -8), 255.5, 3.0e38,
                     1.9921875 + 2 ** -8], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -1.0, 256.0, 2.9999e38, 2.0], np.float64)
    got = U.bf16(vals).astype(np.float64)
    assert got[0] == one and got[1] == 1.0          # tie -> even (1.0)
    assert got[2] == 1.0 + 2 * ulp                  # tie -> even (upper)
    assert got[3] == 1.0 + ulp                      # above the tie -> up
    assert got[4] == -1.0
    assert got[5] == 256.0                          # 255.5 -> 256 (8 significant bits)
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2 ** -8
    assert got[7] == 2.0                            # carry into the exponent
    del want


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    rng = np.random.default_rng(seed)
    N, H, W = 1, 8, 8
    p = U.init_params(seed).astype(np.float64)
    x0 = np.zeros((N, H, W, 16))
    x0[..., :11] = rng.uniform(0, 1, size=(N, H, W, 11))
    target = rng.uniform(0, 1, size=(N, H, W, 3))
    net = U.Net(p, rounding=False)
    loss, g, _ = net.loss_and_grads(x0, target)
    offs, total = U.param_layout()
    checked = 0
    for l, ((wo, bo), (cin, cout)) in enumerate(zip(offs, U.LAYERS)):
        real_in = 11 if l == 0 else cin
        real_out = 3 if l == len(U.LAYERS) - 1 else cout
        picks = [wo + (t * cin + ci) * cout + co for t, ci, co in
                 zip(rng.integers(0, 9, 3), rng.integers(0, real_in, 3), rng.integers(0, real_out, 3))]
        picks.append(bo + int(rng.integers(0, real_out)))
        for i in picks:
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            lp, _, _ = U.Net(pp, rounding=False).loss_and_grads(x0, target)
            lm, _, _ = U.Net(pm, rounding=False).loss_and_grads(x0, target)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (l, i, fd, g[i])
            checked += 1
    assert checked == 4 * len(U.LAYERS)
    # a mistake that drops the skip path's gradient would fail here: the
    # first layer's gradient must not vanish
    assert np.abs(g[offs[0][0]:offs[0][1]]).max() > 0


def test_rounded_network_stays_close_to_exact():
    rng = np.random.default_rng(4)
    p = U.init_params(4)
    x0 = np.zeros((2, 16, 16, 16))
    x0[..., :11] = rng.uniform(0, 1, size=(2, 16, 16, 11))
    target = rng.uniform(0, 1, size=(2, 16, 16, 3))
    l1, g1, a1 = U.Net(p, rounding=True).loss_and_grads(x0, target)
    l0, g0, a0 = U.Net(p, rounding=False).loss_and_grads(U.bf16(x0), target)
    assert abs(l1 - l0) / l0 < 1e-2
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 5e-2
