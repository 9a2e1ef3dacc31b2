"""CPU restatement of the learned denoiser (raygun_amd/csrc/denoise.hip) in
numpy: forward pass, mean-squared-error loss and the gradients of every
parameter. TEST INFRASTRUCTURE ONLY: imported by tests/, never by the
product path.

The network has no counterpart in the reference (it stands in for the
screen-space roughness blur of raygun/compute/rough_blur.h): a 3-level U-Net
with 3x3 convolutions (zero padding 1), ReLU after every layer but the last,
2x2 average pooling, nearest 2x upsampling and skip concatenation
(upsampled channels first):

    x0 (16) -> e1a 16 -> e1b 16 = s1 -> pool -> e2a 32 -> e2b 32 = s2 -> pool
    -> ba 64 -> bb 64 -> up ++ s2 (96) -> d2a 32 -> d2b 32 -> up ++ s1 (48)
    -> d1a 16 -> d1b 16 -> out 16 (first 3 used);  y = x0[:3] + out[:3]
    loss = mean over pixels and the 3 channels of (y - target)^2

Parameters are one flat fp32 vector: per layer the weights [9][cin][cout]
(tap = 3 * (dy + 1) + (dx + 1)) then the biases [cout].

With bf16=True the values are rounded to bfloat16 (round to nearest even)
where the GPU stores them: the packed input, the weights, every activation
and every data gradient, the pooled and the summed values; weight and bias
gradients stay fp32 sums. With bf16=False everything is float64 — that form
is pinned by finite differences (tests/test_unet_oracle.py).
"""
import numpy as np

LAYERS = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16), (16, 16)]
RELU = [1] * 10 + [0]


def param_layout():
    offs, off = [], 0
    for cin, cout in LAYERS:
        offs.append((off, off + 9 * cin * cout))
        off += 9 * cin * cout + cout
    return offs, off


def split_params(p):
    offs, total = param_layout()
    assert p.size == total
    return [(p[w:b].reshape(9, cin, cout), p[b:b + cout]) for (w, b), (cin, cout) in zip(offs, LAYERS)]


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


class Net:
    def __init__(self, params, rounding):
        self.r = (lambda a: bf16(a).astype(np.float64)) if rounding else (lambda a: np.asarray(a, np.float64))
        self.layers = [(self.r(w), np.asarray(b, np.float64)) for w, b in split_params(np.asarray(params, np.float64 if not rounding else np.float32))]

    @staticmethod
    def shift(x, dy, dx):
        """out[y, x] = x[y + dy, x + dx] (zero outside)."""
        n, h, w, c = x.shape
        out = np.zeros_like(x)
        ys, ye = max(0, -dy), min(h, h - dy)
        xs, xe = max(0, -dx), min(w, w - dx)
        out[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
        return out

    def conv(self, x, l, relu=True):
        w, b = self.layers[l]
        y = np.zeros(x.shape[:3] + (w.shape[2],))
        for t in range(9):
            y += self.shift(x, t // 3 - 1, t % 3 - 1) @ w[t]
        y += b
        if relu:
            y = np.maximum(y, 0.0)
        return self.r(y)

    def conv_data_grad(self, dz, l):
        """d input of layer l for the pre-activation gradient dz."""
        w, _ = self.layers[l]
        dx = np.zeros(dz.shape[:3] + (w.shape[1],))
        for t in range(9):
            dx += self.shift(dz, -(t // 3 - 1), -(t % 3 - 1)) @ w[t].T
        return self.r(dx)

    @staticmethod
    def wgrad(x, dz):
        dw = np.zeros((9, x.shape[3], dz.shape[3]))
        for t in range(9):
            xs = Net.shift(x, t // 3 - 1, t % 3 - 1)
            dw[t] = xs.reshape(-1, x.shape[3]).T @ dz.reshape(-1, dz.shape[3])
        return dw

    def pool(self, x):
        n, h, w, c = x.shape
        return self.r(x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4)))

    @staticmethod
    def up(x):
        return x.repeat(2, axis=1).repeat(2, axis=2)

    def forward(self, x0):
        a = {"x0": self.r(x0)}
        a["a1"] = self.conv(a["x0"], 0)
        a["s1"] = self.conv(a["a1"], 1)
        a["p1"] = self.pool(a["s1"])
        a["a2"] = self.conv(a["p1"], 2)
        a["s2"] = self.conv(a["a2"], 3)
        a["p2"] = self.pool(a["s2"])
        a["a3"] = self.conv(a["p2"], 4)
        a["b3"] = self.conv(a["a3"], 5)
        a["c2"] = np.concatenate([self.up(a["b3"]), a["s2"]], axis=3)
        a["a4"] = self.conv(a["c2"], 6)
        a["u2"] = self.conv(a["a4"], 7)
        a["c1"] = np.concatenate([self.up(a["u2"]), a["s1"]], axis=3)
        a["a5"] = self.conv(a["c1"], 8)
        a["u1"] = self.conv(a["a5"], 9)
        a["out"] = self.conv(a["u1"], 10, relu=False)
        return a

    def loss_and_grads(self, x0, target):
        a = self.forward(x0)
        P = x0.shape[0] * x0.shape[1] * x0.shape[2]
        y = a["x0"][..., :3] + a["out"][..., :3]
        d = y - target
        loss = float((d * d).sum() / (3 * P))
        dzo = np.zeros(a["out"].shape)
        dzo[..., :3] = d * (2.0 / (3 * P))
        dzo = self.r(dzo)
        gw = [None] * len(LAYERS)

        def grads(l, x, dz):
            gw[l] = (self.wgrad(x, dz), dz.reshape(-1, dz.shape[3]).sum(axis=0))

        relu_back = lambda dy, y: np.where(y > 0, dy, 0.0)
        grads(10, a["u1"], dzo)
        g = self.conv_data_grad(dzo, 10)
        dz = relu_back(g, a["u1"]); grads(9, a["a5"], dz); g = self.conv_data_grad(dz, 9)
        dz = relu_back(g, a["a5"]); grads(8, a["c1"], dz); g = self.conv_data_grad(dz, 8)
        n, h, w, _ = g.shape
        du2 = self.r(g[..., :32].reshape(n, h // 2, 2, w // 2, 2, 32).sum(axis=(2, 4)))
        ds1x = g[..., 32:]
        dz = relu_back(du2, a["u2"]); grads(7, a["a4"], dz); g = self.conv_data_grad(dz, 7)
        dz = relu_back(g, a["a4"]); grads(6, a["c2"], dz); g = self.conv_data_grad(dz, 6)
        n, h, w, _ = g.shape
        db3 = self.r(g[..., :64].reshape(n, h // 2, 2, w // 2, 2, 64).sum(axis=(2, 4)))
        ds2x = g[..., 64:]
        dz = relu_back(db3, a["b3"]); grads(5, a["a3"], dz); g = self.conv_data_grad(dz, 5)
        dz = relu_back(g, a["a3"]); grads(4, a["p2"], dz); g = self.conv_data_grad(dz, 4)
        ds2 = self.r(self.up(g) * 0.25 + ds2x)
        dz = relu_back(ds2, a["s2"]); grads(3, a["a2"], dz); g = self.conv_data_grad(dz, 3)
        dz = relu_back(g, a["a2"]); grads(2, a["p1"], dz); g = self.conv_data_grad(dz, 2)
        ds1 = self.r(self.up(g) * 0.25 + ds1x)
        dz = relu_back(ds1, a["s1"]); grads(1, a["a1"], dz); g = self.conv_data_grad(dz, 1)
        dz = relu_back(g, a["a1"]); grads(0, a["x0"], dz)
        flat = np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in gw])
        return loss, flat, a


def init_params(seed):
    """He-style random parameters over the real fan-in (11 input channels),
    padded channels zero (the same shape of initialisation as the library's,
    not the same numbers)."""
    rng = np.random.default_rng(seed)
    out = []
    for l, (cin, cout) in enumerate(LAYERS):
        real_in = 11 if l == 0 else cin
        real_out = 3 if l == len(LAYERS) - 1 else cout
        w = rng.normal(0, np.sqrt(2.0 / (9 * real_in)), size=(9, cin, cout))
        w[:, real_in:, :] = 0
        w[:, :, real_out:] = 0
        b = rng.normal(0, 0.05, size=cout)
        b[real_out:] = 0
        out.append(np.concatenate([w.reshape(-1), b]))
    return np.concatenate(out).astype(np.float32)
