"""Denoisers on the GPU.

atrous: the reference's CPU denoiser (denoise.rs:10-74, edge-avoiding a-trous guided by albedo and normal) as HIP
kernels through trt_denoise_atrous.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Device, _check, _ptr, lib


def _sig():
    L = lib()
    P, U32, I = C.c_void_p, C.c_uint32, C.c_int
    L.trt_denoise_atrous.argtypes = [P, U32, U32, P, P, P, P, U32, P]
    L.trt_denoise_atrous.restype = I


_sig()


def atrous_device(dev: Device, w: int, h: int, color, albedo, normal, variance, iterations: int, out):
    """Device buffers (DeviceBuffer): queued on the device's stream."""
    _check(lib().trt_denoise_atrous(dev.h, w, h, color.p, albedo.p, normal.p, variance.p, iterations, out.p))


def atrous(dev: Device, color: np.ndarray, albedo: np.ndarray, normal: np.ndarray, variance: np.ndarray,
           iterations: int = 4) -> np.ndarray:
    """Host images (h, w, 3) float32 and variance (h, w): the denoised colour (h, w, 3)."""
    h, w = color.shape[:2]
    bufs = []
    for a in (color, albedo, normal, variance):
        a = np.ascontiguousarray(a, np.float32)
        b = dev.alloc(a.nbytes)
        b.upload(a)
        bufs.append(b)
    out = dev.alloc(w * h * 12)
    atrous_device(dev, w, h, *bufs, iterations, out)
    res = out.download(np.float32, w * h * 3).reshape(h, w, 3)
    for b in bufs + [out]:
        b.free()
    return res
