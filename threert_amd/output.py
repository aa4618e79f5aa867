"""The output side of a render job: the raw HDR file the reference's `serve` jobs write (main.rs:575-578), 8-bit
images, and the reference's image comparison (compare.rs:13-107: PSNR, SSIM, CIE76 delta E, trail length, luminance
ratio), so that results can be read by the reference's own tools and tests (tools/test-render.ts:46-106 reads the
raw HDR file; `region_mean` is its `region`).

All of this is host code, as it is in the reference. The readback normalisation that comes before it (gpu.rs:4218-4231)
is the library's `trt_accum_resolve`.
"""
from __future__ import annotations

import math

import numpy as np


# ---- the raw HDR file ----------------------------------------------------------------------------------------------
def write_hdr(path, color: np.ndarray) -> None:
    """main.rs:575-578: the image's linear radiance as raw little-endian f32 RGB, rows top to bottom, no header."""
    c = np.asarray(color)
    if c.ndim != 3 or c.shape[2] != 3:
        raise ValueError(f"write_hdr: expected an (h, w, 3) image, got shape {c.shape}")
    np.ascontiguousarray(c, "<f4").tofile(str(path))


def read_hdr(path, width: int, height: int) -> np.ndarray:
    """A raw f32 RGB file of `width` x `height` pixels as an (h, w, 3) float32 array (the size is not in the file: a
    `serve` reply's stats carry it)."""
    px = np.fromfile(str(path), "<f4")
    if px.size != 3 * width * height:
        raise ValueError(f"read_hdr: {path} holds {px.size} floats, not 3 x {width} x {height}")
    return px.reshape(height, width, 3).astype(np.float32)


def region_mean(px: np.ndarray, region) -> np.ndarray:
    """tools/test-render.ts:46-54: the mean RGB of a region [u0, v0, u1, v1] (0..1, v down) of an (h, w, 3) image;
    at least one pixel each way."""
    h, w = px.shape[:2]
    u0, v0, u1, v1 = region
    x0, y0 = math.floor(u0 * w), math.floor(v0 * h)
    x1, y1 = max(x0 + 1, math.ceil(u1 * w)), max(y0 + 1, math.ceil(v1 * h))
    return np.asarray(px, np.float64)[y0:y1, x0:x1].reshape(-1, 3).mean(axis=0)


def luminance(c) -> np.ndarray:
    """Rec. 709 luminance (the weights test-render.ts and compare.rs use)."""
    c = np.asarray(c, np.float64)
    return 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]


# ---- 8-bit images -------------------------------------------------------------------------------------------------
def srgb8(color: np.ndarray, exposure: float = 1.0) -> np.ndarray:
    """A plain display encoding of linear radiance: scaled by `exposure`, clamped to [0, 1], the sRGB transfer curve,
    rounded to 8 bits. It is not the reference's post chain (auto exposure, tone curve, local exposure: post::finish),
    which is outside this library."""
    x = np.clip(np.nan_to_num(np.asarray(color, np.float64) * exposure, nan=0.0, posinf=1.0), 0.0, 1.0)
    s = np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)
    return np.clip(np.floor(s * 255.0 + 0.5), 0, 255).astype(np.uint8)


def write_png(path, rgb8: np.ndarray) -> None:
    from PIL import Image

    Image.fromarray(np.ascontiguousarray(rgb8, np.uint8), "RGB").save(str(path))


def read_rgb8(path) -> np.ndarray:
    """An image file as (h, w, 3) uint8 RGB, as compare.rs's `image::open(..).to_rgb8()` gives it."""
    from PIL import Image

    with Image.open(str(path)) as im:
        return np.asarray(im.convert("RGB"), np.uint8)


# ---- compare.rs ---------------------------------------------------------------------------------------------------
def _round(v: float, d: int) -> float:
    """compare.rs `round`: to `d` decimals, halves away from zero (Rust's f64::round)."""
    p = 10.0 ** d
    a = abs(v * p)
    r = math.floor(a)
    if a - r >= 0.5:
        r += 1.0
    return math.copysign(r, v) / p


def _linear(c8: np.ndarray) -> np.ndarray:
    """compare.rs:28-31: an 8-bit sRGB value to linear."""
    c = c8.astype(np.float64) / 255.0
    return np.where(c <= 0.04045, c / 12.92, np.power((c + 0.055) / 1.055, 2.4))


def lab(c: np.ndarray) -> np.ndarray:
    """compare.rs:228-235: linear sRGB (..., 3) to CIELAB under D65 (the 4-digit sRGB matrix, the CIE 0.008856 /
    7.787 linear segment)."""
    c = np.asarray(c, np.float64)
    r, g, b = c[..., 0], c[..., 1], c[..., 2]
    x = (0.4124 * r + 0.3576 * g + 0.1805 * b) / 0.95047
    y = 0.2126 * r + 0.7152 * g + 0.0722 * b
    z = (0.0193 * r + 0.1192 * g + 0.9505 * b) / 1.08883

    def f(t):
        return np.where(t > 0.008856, np.cbrt(t), 7.787 * t + 16.0 / 116.0)

    fx, fy, fz = f(x), f(y), f(z)
    return np.stack([116.0 * fy - 16.0, 500.0 * (fx - fy), 200.0 * (fy - fz)], axis=-1)


def _gauss11() -> np.ndarray:
    i = np.arange(-5, 6, dtype=np.float64)
    k = np.exp(-(i * i) / (2.0 * 1.5 * 1.5))
    return k / k.sum()


def _blur(src: np.ndarray, k: np.ndarray) -> np.ndarray:
    """compare.rs:246-272: the 11-tap Gaussian along rows, then along columns, edges clamped; taps summed in order."""
    r = len(k) // 2
    h, w = src.shape
    p = np.pad(src, ((0, 0), (r, r)), mode="edge")
    tmp = np.zeros((h, w))
    for j in range(len(k)):
        tmp = tmp + p[:, j:j + w] * k[j]
    p = np.pad(tmp, ((r, r), (0, 0)), mode="edge")
    out = np.zeros((h, w))
    for j in range(len(k)):
        out = out + p[j:j + h, :] * k[j]
    return out


def ssim(ya: np.ndarray, yb: np.ndarray) -> float:
    """compare.rs:243-285: mean SSIM of two luma planes of 0..255 values (11 x 11 Gaussian window, sigma 1.5,
    C1 = (0.01 * 255)^2, C2 = (0.03 * 255)^2)."""
    a, b = np.asarray(ya, np.float64), np.asarray(yb, np.float64)
    k = _gauss11()
    ma, mb = _blur(a, k), _blur(b, k)
    saa, sbb, sab = _blur(a * a, k), _blur(b * b, k), _blur(a * b, k)
    c1, c2 = (0.01 * 255.0) ** 2, (0.03 * 255.0) ** 2
    va, vb, cov = saa - ma * ma, sbb - mb * mb, sab - ma * mb
    s = ((2.0 * ma * mb + c1) * (2.0 * cov + c2)) / ((ma * ma + mb * mb + c1) * (va + vb + c2))
    return float(s.mean())


def _crop(img: np.ndarray, region) -> np.ndarray:
    """compare.rs:19-25: the region in f32 arithmetic, truncated to whole pixels, at least one each way."""
    h, w = img.shape[:2]
    u0, v0, u1, v1 = (np.float32(v) for v in region)
    fw, fh = np.float32(w), np.float32(h)
    x, y = int(u0 * fw), int(v0 * fh)
    cw, ch = max(int((u1 - u0) * fw), 1), max(int((v1 - v0) * fh), 1)
    return img[y:min(y + ch, h), x:min(x + cw, w)]


def _heat(t: np.ndarray) -> np.ndarray:
    """compare.rs:237-242: black, red, yellow, white as t goes 0 to 1."""
    rgb = np.stack([np.clip(t * 3.0, 0, 1), np.clip(t * 3.0 - 1.0, 0, 1), np.clip(t * 3.0 - 2.0, 0, 1)], axis=-1)
    return (rgb * 255.0).astype(np.uint8)


def compare(reference, test, sheet=None, region=None) -> dict:
    """compare.rs:13-107 over two 8-bit RGB images (arrays of (h, w, 3) uint8, or image files): PSNR (dB; 99 when
    equal), SSIM on gamma-encoded luma, mean / p95 / p99 CIE76 delta E in CIELAB, the p95 over rows of each row's
    longest run of pixels with delta E > 10 (`trail_px`), and the test's mean luminance over the reference's.
    `sheet`: a file for reference | test | delta E / 25 as a heat map. The reference resizes a test image of another
    size (Lanczos); here the sizes must match."""
    a = read_rgb8(reference) if not isinstance(reference, np.ndarray) else np.asarray(reference, np.uint8)
    b = read_rgb8(test) if not isinstance(test, np.ndarray) else np.asarray(test, np.uint8)
    if a.shape != b.shape or a.ndim != 3 or a.shape[2] != 3:
        raise ValueError(f"compare: images of {a.shape} and {b.shape}; both must be the same (h, w, 3)")
    if region is not None:
        a, b = _crop(a, region), _crop(b, region)
    h, w = a.shape[:2]
    n = w * h
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.sum(d * d)) / (n * 3)
    psnr = 10.0 * math.log10(255.0 * 255.0 / mse) if mse > 0.0 else 99.0
    ca, cb = _linear(a), _linear(b)
    de = np.sqrt(np.sum((lab(ca) - lab(cb)) ** 2, axis=-1))
    la, lb = float(luminance(ca).sum()), float(luminance(cb).sum())
    wy = np.array([0.299, 0.587, 0.114])
    ya = a[..., 0] * wy[0] + a[..., 1] * wy[1] + a[..., 2] * wy[2]
    yb = b[..., 0] * wy[0] + b[..., 1] * wy[1] + b[..., 2] * wy[2]
    s = ssim(ya, yb)
    flat = np.sort(de.reshape(-1))
    p95 = float(flat[int(n * 0.95)])
    p99 = float(flat[min(int(n * 0.99), n - 1)])
    # each row's longest run of wrong pixels: a pixel's run is its column less the column of the last right one
    col = np.broadcast_to(np.arange(w), (h, w))
    last_right = np.maximum.accumulate(np.where(de > 10.0, -1, col), axis=1)
    runs = np.sort((col - last_right).max(axis=1))
    trail = int(runs[min(int(h * 0.95), h - 1)])
    if sheet is not None:
        out = np.concatenate([a, b, _heat(np.minimum(de / 25.0, 1.0))], axis=1)
        write_png(sheet, out)
    return {"psnr_db": _round(psnr, 2), "ssim": _round(s, 4), "mean_delta_e": _round(float(de.mean()), 2),
            "p95_delta_e": _round(p95, 2), "p99_delta_e": _round(p99, 2), "trail_px": trail,
            "luminance_ratio": _round(lb / max(la, 1e-9), 3)}
