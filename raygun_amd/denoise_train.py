"""Training of the learned denoiser on the GPU with PyTorch (fp32), while the
library's own backward kernels are switched off (rgrt_denoiser_train returns
RGRT_ERR_NOT_READY). Inference stays on the library's bf16 MFMA kernels
(rgrt_denoise): the trained parameters are loaded with set_params.

The network is the library's U-Net (csrc/denoise.hip): 11 layers of 3x3
convolutions (zero padding 1), ReLU after all but the last, 2x2 average
pooling, nearest 2x upsampling, skip concatenation with the upsampled
channels first; y = x0[:3] + out[:3], loss = mean squared error of y against
the target over pixels and the 3 channels. Each convolution is computed as
one fp32 GEMM of the 9 shifted inputs side by side against the weights, on
NHWC tensors. Parameters use the library's
flat layout: per layer the weights [9][cin][cout] (tap = 3 (dy + 1) +
(dx + 1)) then the biases [cout].

Training pairs come from the tracer itself (denoise.build_dataset: 4-spp
stochastic frames packed by the library, 256-spp means as targets) and are
read back with Denoiser.dataset_read. PyTorch brings its own HIP runtime, so
training runs in a child process of its own (python -m
raygun_amd.denoise_train DATA.npz OUT.npy ...), never next to the library's
runtime in one process. No counterpart in the reference.
"""
import argparse
import json
import math
import sys

import numpy as np

LAYERS = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16), (16, 16)]


def param_layout():
    offs, off = [], 0
    for cin, cout in LAYERS:
        offs.append((off, off + 9 * cin * cout))
        off += 9 * cin * cout + cout
    return offs, off


def _torch():
    import torch
    return torch


class UNet:
    """The network as fp32 torch tensors on `device` (requires_grad)."""

    def __init__(self, params, device):
        torch = _torch()
        offs, total = param_layout()
        p = np.asarray(params, np.float32)
        assert p.size == total
        self.w, self.b = [], []
        for (wo, bo), (cin, cout) in zip(offs, LAYERS):
            self.w.append(torch.tensor(p[wo:bo].reshape(9, cin, cout), device=device, requires_grad=True))
            self.b.append(torch.tensor(p[bo:bo + cout], device=device, requires_grad=True))
        # padded channels stay zero: the first layer's inputs 11..15 are zero,
        # the last layer's outputs 3..15 are masked out of its gradient
        self.out_mask = torch.zeros(16, device=device)
        self.out_mask[:3] = 1.0

    def parameters(self):
        return self.w + self.b

    def conv(self, x, l, relu=True):
        torch = _torch()
        n, h, w, cin = x.shape
        xp = torch.nn.functional.pad(x, (0, 0, 1, 1, 1, 1))
        wt = self.w[l]
        if l == len(LAYERS) - 1:
            wt = wt * self.out_mask
        # the 9 shifted inputs side by side (tap-major, then input channel),
        # against the weights [9][cin][cout] read as one (9 cin) x cout matrix
        taps = torch.cat([xp[:, ty:ty + h, tx:tx + w, :] for ty in range(3) for tx in range(3)], dim=3)
        y = taps.reshape(-1, 9 * cin) @ wt.reshape(9 * cin, -1)
        b = self.b[l] * self.out_mask if l == len(LAYERS) - 1 else self.b[l]
        y = (y + b).reshape(n, h, w, -1)
        return torch.relu(y) if relu else y

    @staticmethod
    def pool(x):
        n, h, w, c = x.shape
        return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(dim=(2, 4))

    @staticmethod
    def up(x):
        return x.repeat_interleave(2, dim=1).repeat_interleave(2, dim=2)

    def forward(self, x0):
        torch = _torch()
        a1 = self.conv(x0, 0)
        s1 = self.conv(a1, 1)
        a2 = self.conv(self.pool(s1), 2)
        s2 = self.conv(a2, 3)
        a3 = self.conv(self.pool(s2), 4)
        b3 = self.conv(a3, 5)
        a4 = self.conv(torch.cat([self.up(b3), s2], dim=3), 6)
        u2 = self.conv(a4, 7)
        a5 = self.conv(torch.cat([self.up(u2), s1], dim=3), 8)
        u1 = self.conv(a5, 9)
        return self.conv(u1, 10, relu=False)

    def loss(self, x0, target, clamp_target=False):
        """Mean squared error of y = x0[:3] + out[:3] against the target, or
        against the target clamped to [0, 1] (the range the post pass keeps
        and PSNR measures)."""
        out = self.forward(x0)
        if clamp_target:
            target = target.clamp(0.0, 1.0)
        d = x0[..., :3] + out[..., :3] - target
        return (d * d).mean()

    def flat(self):
        out = []
        for w, b in zip(self.w, self.b):
            out.append(w.detach().float().cpu().numpy().reshape(-1))
            out.append(b.detach().float().cpu().numpy())
        p = np.concatenate(out).astype(np.float32)
        offs, _ = param_layout()
        wo, bo = offs[-1]
        w_last = p[wo:bo].reshape(9, 16, 16)
        w_last[:, :, 3:] = 0.0
        p[wo:bo] = w_last.reshape(-1)
        p[bo + 3:bo + 16] = 0.0
        return p


def bf16_bits_to_float(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def train(x0_bits, target, init_params, steps=3000, batch=16, lr=1e-3, seed=0, crop=None, device="cuda", log=None,
          clamp_target=False):
    """Adam on random batches (optionally random crops of `crop` pixels, a
    multiple of 4) of the training set; cosine learning-rate decay. Returns
    (flat fp32 parameters, mean loss of the first and of the last 50 steps)."""
    torch = _torch()
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    X = torch.from_numpy(bf16_bits_to_float(x0_bits)).to(device)
    T = torch.from_numpy(np.ascontiguousarray(target, np.float32)).to(device)
    count, H, W, _ = X.shape
    net = UNet(init_params, device)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    losses = []
    for it in range(steps):
        for group in opt.param_groups:
            group["lr"] = lr * 0.5 * (1.0 + math.cos(math.pi * it / max(1, steps)))
        idx = torch.randint(0, count, (batch,), generator=g)
        xb, tb = X[idx.to(device)], T[idx.to(device)]
        if crop and crop < min(H, W):
            y0 = int(torch.randint(0, (H - crop) // 4 + 1, (1,), generator=g)) * 4
            x0c = int(torch.randint(0, (W - crop) // 4 + 1, (1,), generator=g)) * 4
            xb, tb = xb[:, y0:y0 + crop, x0c:x0c + crop], tb[:, y0:y0 + crop, x0c:x0c + crop]
        loss = net.loss(xb, tb, clamp_target)
        opt.zero_grad()
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
        if log and (it % 200 == 0 or it == steps - 1):
            log(f"[train] step {it} loss {losses[-1]:.6f}")
    head = float(np.mean(losses[:50]))
    tail = float(np.mean(losses[-50:]))
    return net.flat(), head, tail


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the denoiser on a saved training set")
    ap.add_argument("data", help=".npz with x0 (uint16 bf16 bits, N x H x W x 16), target (N x H x W x 3), init (params)")
    ap.add_argument("out", help=".npy for the trained flat parameters")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--crop", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--device", default="cuda")
    ap.add_argument("--clamp-target", action="store_true", help="loss against the target clamped to [0, 1]")
    a = ap.parse_args(argv)
    d = np.load(a.data, allow_pickle=False)
    params, head, tail = train(d["x0"], d["target"], d["init"], steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                               crop=a.crop or None, device=a.device, clamp_target=a.clamp_target,
                               log=lambda m: print(m, file=sys.stderr, flush=True))
    np.save(a.out, params)
    print(json.dumps({"loss_first": head, "loss_last": tail, "steps": a.steps}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
