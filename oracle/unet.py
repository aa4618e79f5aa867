"""The learned denoiser's oracle (test infrastructure only: tests/ and bench.py's checks use it, the library never
does): the U-Net of threert_amd/csrc/unet_net.cpp in numpy, forward and backward.

Two modes:
  exact      float64 throughout: pinned by central finite differences and by a direct loop convolution;
  emulated   the GPU's storage points: activations, pooled maps and output gradients rounded to bfloat16 (round to
             nearest even), input gradients kept in float32, sums in float64 (the GPU's FP32 MFMA accumulation in
             another order: the tests compare with a tolerance).
The network (the reference's denoise::oidn stage, denoise.rs:81-103, has no network of its own in the repository:
OIDN is a third-party library) is this library's, so it is pinned by the mathematics of its parts, not by a
reference run.
"""
from __future__ import annotations

import numpy as np

ARCH = [
    (32, 0, 32, 0), (32, 0, 32, 0), (32, 0, 64, 1), (64, 0, 64, 1), (64, 0, 64, 2), (64, 0, 64, 2),
    (64, 0, 96, 3), (96, 0, 96, 3), (96, 0, 128, 4), (128, 0, 128, 4), (96, 128, 96, 3), (96, 0, 96, 3),
    (64, 96, 64, 2), (64, 0, 64, 2), (64, 64, 64, 1), (64, 0, 64, 1), (32, 64, 32, 0), (32, 0, 32, 0),
    (32, 0, 32, 0),
]
(E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, E4A, E4B, D3A, D3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT) = range(19)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(params: np.ndarray, dtype=np.float64):
    """Per layer (W [cout, 9, cin], b [cout]) from the flat parameters."""
    out, at = [], 0
    for c0, c1, co, _ in ARCH:
        n = co * 9 * (c0 + c1)
        w = params[at:at + n].reshape(co, 9, c0 + c1).astype(dtype)
        at += n
        b = params[at:at + co].astype(dtype)
        at += co
        out.append((w, b))
    return out


def patches(x: np.ndarray) -> np.ndarray:
    """(n, h, w, c) -> (n, h, w, 9, c): the 3x3 neighbourhood, zero outside, tap = 3 * (dy + 1) + (dx + 1)."""
    n, h, w, c = x.shape
    xp = np.zeros((n, h + 2, w + 2, c), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    return np.stack([xp[:, dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)], axis=3)


def up2(x: np.ndarray) -> np.ndarray:
    return x.repeat(2, axis=1).repeat(2, axis=2)


def pool2(x: np.ndarray) -> np.ndarray:
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4)) * 0.25


def conv(x0, x1, w, b):
    """y = conv3x3(concat(x0, up2(x1))) + b, summed in float64."""
    x = x0 if x1 is None else np.concatenate([x0, up2(x1)], axis=3)
    y = np.einsum("nhwkc,okc->nhwo", patches(x.astype(np.float64)), w.astype(np.float64), optimize=True)
    return y + (0.0 if b is None else b)


def conv_input_grad(dy, w):
    """d(concat input) of conv from its output gradient: the flipped, transposed weights over dy."""
    wt = w[:, ::-1, :].transpose(2, 1, 0)  # [cin][9][cout], tap reversed
    return conv(dy, None, wt, None)


def conv_weight_grad(x0, x1, dy):
    x = x0 if x1 is None else np.concatenate([x0, up2(x1)], axis=3)
    dw = np.einsum("nhwo,nhwkc->okc", dy.astype(np.float64), patches(x.astype(np.float64)), optimize=True)
    return dw, dy.astype(np.float64).sum(axis=(0, 1, 2))


def features(color, albedo, normal):
    n, h, w, _ = color.shape
    f = np.zeros((n, h, w, 32), np.float64)
    f[..., 0:3] = np.log1p(np.maximum(color, 0.0))
    f[..., 3:6] = albedo
    f[..., 6:9] = normal * 0.5 + 0.5
    return f


class Net:
    """Forward and backward of the U-Net for a batch (n, h, w, 3) of color / albedo / normal; h, w multiples of 16."""

    def __init__(self, params: np.ndarray, emulate: bool):
        self.emulate = emulate
        self.layers = unpack(params, np.float32 if emulate else np.float64)
        if emulate:
            self.layers = [(bf16(w).astype(np.float64), b.astype(np.float64)) for w, b in self.layers]
        self.q = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if emulate else (lambda a: a)
        self.f32 = (lambda a: a.astype(np.float32).astype(np.float64)) if emulate else (lambda a: a)

    def forward(self, color, albedo, normal):
        q, L = self.q, self.layers
        self.feat = q(features(color, albedo, normal))
        act = {}

        def run(l, x0, x1=None, relu=True):
            y = conv(x0, x1, *L[l])
            return q(np.maximum(y, 0.0)) if relu else self.f32(y)

        self.pool = {}
        act[E0A] = run(E0A, self.feat)
        act[E0B] = run(E0B, act[E0A])
        for k, (a, b, prev) in enumerate(((E1A, E1B, E0B), (E2A, E2B, E1B), (E3A, E3B, E2B), (E4A, E4B, E3B))):
            self.pool[k + 1] = q(pool2(act[prev]))
            act[a] = run(a, self.pool[k + 1])
            act[b] = run(b, act[a])
        for a, b, skip, below in ((D3A, D3B, E3B, E4B), (D2A, D2B, E2B, D3B), (D1A, D1B, E1B, D2B),
                                  (D0A, D0B, E0B, D1B)):
            act[a] = run(a, act[skip], act[below])
            act[b] = run(b, act[a])
        self.y = run(OUT, act[D0B], relu=False)
        self.act = act
        return self.y

    def denoised(self):
        return np.expm1(np.maximum(self.y[..., :3], 0.0))

    def loss_and_grads(self, color, albedo, normal, target):
        """The mean |y - log(1 + target)| over the three color channels, and d(loss)/d(params) (flat, float64)."""
        q, L, act = self.q, self.layers, None
        self.forward(color, albedo, normal)
        act = self.act
        n, h, w, _ = color.shape
        d = self.y[..., :3] - np.log1p(np.maximum(target.astype(np.float64), 0.0))
        if self.emulate:
            d = (self.y[..., :3].astype(np.float32) - np.log1p(np.maximum(target, 0.0)).astype(np.float32)).astype(
                np.float64)
        count = 3.0 * n * h * w
        loss = np.abs(d).sum() / count
        gy = np.zeros((n, h, w, 32))
        gy[..., :3] = np.sign(d) / count
        gy = q(gy)
        grads = [None] * len(L)

        def back(l, x0, x1, dy, need_dx=True):
            grads[l] = conv_weight_grad(x0, x1, dy)
            return self.f32(conv_input_grad(dy, L[l][0])) if need_dx else None

        mask = lambda dx, a: q(np.where(a > 0.0, dx, 0.0))  # noqa: E731
        dx = back(OUT, act[D0B], None, gy)
        g = mask(dx, act[D0B])
        keep = {}
        for a, b, skip, below in ((D0A, D0B, E0B, D1B), (D1A, D1B, E1B, D2B), (D2A, D2B, E2B, D3B),
                                  (D3A, D3B, E3B, E4B)):
            dx = back(b, act[a], None, g)
            g = mask(dx, act[a])
            dx = back(a, act[skip], act[below], g)
            keep[a] = dx
            c0 = ARCH[a][0]
            part = dx[..., c0:]
            nn, hh, ww, cc = part.shape
            summed = self.f32(part.reshape(nn, hh // 2, 2, ww // 2, 2, cc).sum(axis=(2, 4)))
            g = mask(summed, act[below])
        enc = ((E0A, E0B, None, D0A), (E1A, E1B, E0B, D1A), (E2A, E2B, E1B, D2A), (E3A, E3B, E2B, D3A),
               (E4A, E4B, E3B, None))
        for k in range(4, -1, -1):
            a, b, above, _ = enc[k]
            dx = back(b, act[a], None, g)
            g = mask(dx, act[a])
            x_in = self.feat if k == 0 else self.pool[k]
            if k == 0:
                back(a, x_in, None, g, need_dx=False)
                break
            dp = back(a, x_in, None, g)
            dec_above = enc[k - 1][3]
            skip = keep[dec_above][..., :ARCH[above][2]]
            v = self.f32(skip + up2(dp) * 0.25)
            g = mask(v, act[above])
        flat = np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in grads])
        return loss, flat


def direct_conv(x, w, b):
    """A 3x3 zero-padded convolution by explicit loops (the definition, for pinning `conv`)."""
    n, h, wd, c = x.shape
    co = w.shape[0]
    y = np.zeros((n, h, wd, co))
    for i in range(n):
        for yy in range(h):
            for xx in range(wd):
                for t in range(9):
                    sy, sx = yy + t // 3 - 1, xx + t % 3 - 1
                    if 0 <= sy < h and 0 <= sx < wd:
                        y[i, yy, xx] += w[:, t, :] @ x[i, sy, sx]
    return y + b
