"""The learned denoiser: a five-level U-Net trained and run on the GPU (trt_unet_*; the reference's denoise::oidn
stage, denoise.rs:81-103, which hands the image to OIDN).

Inputs per pixel are the accumulator's resolved color, albedo and normal; the network reads log(1 + color), the
albedo and the normal mapped to [0, 1], and predicts log(1 + color). Training minimises the mean absolute error
in that log space against a many-sample render of the same view.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Device, DeviceBuffer, _check, _ptr, lib

# (c0, c1, cout, level) per layer, in parameter order: the encoder e0a .. e4b, the decoder d3a .. d0b, the output.
ARCH = [
    (32, 0, 32, 0), (32, 0, 32, 0), (32, 0, 64, 1), (64, 0, 64, 1), (64, 0, 64, 2), (64, 0, 64, 2),
    (64, 0, 96, 3), (96, 0, 96, 3), (96, 0, 128, 4), (128, 0, 128, 4), (96, 128, 96, 3), (96, 0, 96, 3),
    (64, 96, 64, 2), (64, 0, 64, 2), (64, 64, 64, 1), (64, 0, 64, 1), (32, 64, 32, 0), (32, 0, 32, 0),
    (32, 0, 32, 0),
]
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "e4a", "e4b", "d3a", "d3b", "d2a", "d2b", "d1a",
         "d1b", "d0a", "d0b", "out"]


def layout():
    """(name, weight offset, weight shape (cout, 9, cin), bias offset, cout) per layer in the flat parameters."""
    out, at = [], 0
    for name, (c0, c1, co, _) in zip(NAMES, ARCH):
        n = co * 9 * (c0 + c1)
        out.append((name, at, (co, 9, c0 + c1), at + n, co))
        at += n + co
    return out


def param_count() -> int:
    return sum(co * 9 * (c0 + c1) + co for c0, c1, co, _ in ARCH)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan-in)), zero biases; the output layer's unused channels 3..31 are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(), np.float32)
    for name, wo, shape, bo, co in layout():
        fan_in = shape[1] * shape[2]
        w = rng.standard_normal(shape) * np.sqrt(2.0 / fan_in)
        if name == "out":
            w *= 0.5
            w[3:] = 0.0
        p[wo:wo + w.size] = w.astype(np.float32).ravel()
    return p


def shipped_params() -> np.ndarray:
    """The weights trained by denoise_train.py on 200 seeded scenes (threert_amd/data/unet_weights.npy)."""
    import os

    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_weights.npy"))


def _sig():
    L = lib()
    P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    L.trt_unet_create.argtypes = [P, U64, C.POINTER(C.c_void_p)]
    L.trt_unet_release.argtypes = [P]
    L.trt_unet_param_count.argtypes = [P, C.POINTER(C.c_uint64)]
    L.trt_unet_params_set.argtypes = [P, P, U64]
    L.trt_unet_params_get.argtypes = [P, P, U64]
    L.trt_unet_denoise.argtypes = [P, U32, U32, U32, P, P, P, P]
    L.trt_unet_gradients.argtypes = [P, U32, U32, U32, P, P, P, P, P, P]
    L.trt_unet_train_step.argtypes = [P, U32, U32, U32, P, P, P, P, C.c_float, P]
    L.trt_unet_set_tall_tiles.argtypes = [U32]
    for n in ("trt_unet_create", "trt_unet_release", "trt_unet_param_count", "trt_unet_params_set",
              "trt_unet_params_get", "trt_unet_denoise", "trt_unet_gradients", "trt_unet_train_step",
              "trt_unet_set_tall_tiles"):
        getattr(L, n).restype = I


_sig()


def set_tall_tiles(min_rows: int):
    """Levels at least min_rows tall take 16-pixel-tall convolution tiles (default 128; 2^32 - 1: never)."""
    _check(lib().trt_unet_set_tall_tiles(min_rows))


def _as_nhwc3(a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    return a[None] if a.ndim == 3 else a


class UNet:
    """The network's parameters, optimizer state and buffers on one device, for up to `max_pixels` pixels
    (batch x height x width) a call."""

    def __init__(self, dev: Device, max_pixels: int, params: np.ndarray | None = None, seed: int = 0):
        self.dev = dev
        h = C.c_void_p()
        _check(lib().trt_unet_create(dev.h, int(max_pixels), C.byref(h)))
        self.h = h
        n = C.c_uint64()
        _check(lib().trt_unet_param_count(self.h, C.byref(n)))
        assert n.value == param_count()
        self.set_params(init_params(seed) if params is None else params)
        self._io = {}

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().trt_unet_params_set(self.h, _ptr(p), p.size))

    def params(self) -> np.ndarray:
        p = np.zeros(param_count(), np.float32)
        _check(lib().trt_unet_params_get(self.h, _ptr(p), p.size))
        return p

    def _buf(self, key: str, nbytes: int) -> DeviceBuffer:
        b = self._io.get(key)
        if b is None or b.nbytes < nbytes:
            b = self._io[key] = self.dev.alloc(nbytes)
        return b

    def _upload(self, **arrays):
        out = {}
        for k, a in arrays.items():
            b = self._buf(k, a.nbytes)
            b.upload(a)
            out[k] = b
        return out

    def denoise(self, color, albedo, normal) -> np.ndarray:
        """Host images (h, w, 3) or (n, h, w, 3) in, the denoised color out (same shape)."""
        squeeze = np.asarray(color).ndim == 3
        c, a, nm = _as_nhwc3(color), _as_nhwc3(albedo), _as_nhwc3(normal)
        n, hh, ww, _ = c.shape
        b = self._upload(color=c, albedo=a, normal=nm)
        out = self._buf("out", c.nbytes)
        self.denoise_device(n, hh, ww, b["color"], b["albedo"], b["normal"], out)
        self.dev.sync()
        r = out.download(np.float32, c.size).reshape(c.shape)
        return r[0] if squeeze else r

    def denoise_device(self, n, h, w, color: DeviceBuffer, albedo: DeviceBuffer, normal: DeviceBuffer,
                       out: DeviceBuffer):
        _check(lib().trt_unet_denoise(self.h, n, h, w, color.p, albedo.p, normal.p, out.p))

    def gradients(self, color, albedo, normal, target):
        """(loss, flat FP32 gradient) for host images; waits for the device."""
        c, a, nm, t = (_as_nhwc3(x) for x in (color, albedo, normal, target))
        n, hh, ww, _ = c.shape
        b = self._upload(color=c, albedo=a, normal=nm, target=t)
        loss = np.zeros(1, np.float32)
        g = np.zeros(param_count(), np.float32)
        _check(lib().trt_unet_gradients(self.h, n, hh, ww, b["color"].p, b["albedo"].p, b["normal"].p, b["target"].p,
                                        _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    def train_step_device(self, n, h, w, color: DeviceBuffer, albedo: DeviceBuffer, normal: DeviceBuffer,
                          target: DeviceBuffer, lr: float, want_loss: bool = False):
        loss = np.zeros(1, np.float32)
        _check(lib().trt_unet_train_step(self.h, n, h, w, color.p, albedo.p, normal.p, target.p, float(lr),
                                         _ptr(loss) if want_loss else None))
        return float(loss[0]) if want_loss else None

    def close(self):
        for b in self._io.values():
            b.free()
        self._io = {}
        if self.h:
            lib().trt_unet_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
