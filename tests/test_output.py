"""The output side: the raw HDR file (main.rs:575-578), 8-bit encoding, and the reference's image comparison
(compare.rs:13-107), pinned by closed forms (PSNR of a known error, SSIM of flat images, published CIELAB values of
the sRGB primaries, the sRGB curve's round trip) and brute force (SSIM's window summed in two dimensions), plus the
readback normalisation (gpu.rs:4218-4231) of the library's resolve on the GPU."""
import math
import struct

import numpy as np
import pytest

from threert_amd import output as O


def test_the_hdr_file_is_raw_little_endian_rgb(tmp_path):
    img = np.arange(18, dtype=np.float32).reshape(2, 3, 3) * 0.5 - 2.0
    p = tmp_path / "x.raw"
    O.write_hdr(p, img)
    assert p.read_bytes() == struct.pack("<18f", *img.reshape(-1).tolist())
    assert (O.read_hdr(p, 3, 2) == img).all()
    with pytest.raises(ValueError):
        O.read_hdr(p, 2, 2)
    with pytest.raises(ValueError):
        O.write_hdr(p, img[..., :2])


def test_region_mean_is_test_renders():
    img = np.zeros((4, 8, 3), np.float32)
    img[1, 2] = (1.0, 2.0, 3.0)
    img[1, 3] = (3.0, 2.0, 1.0)
    # x 2..3 (floor 0.25 * 8 = 2, ceil 0.5 * 8 = 4), y 1 (floor 0.25 * 4 = 1, ceil 0.5 * 4 = 2)
    assert O.region_mean(img, (0.25, 0.25, 0.5, 0.5)).tolist() == [2.0, 2.0, 2.0]
    # an empty region still takes one pixel
    assert O.region_mean(img, (0.25, 0.25, 0.25, 0.25)).tolist() == [1.0, 2.0, 3.0]


def test_srgb_encoding_and_decoding_are_inverse():
    assert O.srgb8(np.array([[[0.5, 0.0, 1.0]]])).tolist() == [[[188, 0, 255]]]  # linear 0.5 is sRGB 188
    assert O.srgb8(np.array([[[np.nan, np.inf, -1.0]]])).tolist() == [[[0, 255, 0]]]
    c = np.arange(256, dtype=np.uint8)
    lin = O._linear(c)
    assert (O.srgb8(np.stack([lin] * 3, -1)[None])[0, :, 0] == c).all()


def test_rounding_is_half_away_from_zero():
    assert O._round(2.5, 0) == 3.0 and O._round(-2.5, 0) == -3.0 and O._round(0.125, 2) == 0.13
    assert O._round(99.0, 2) == 99.0 and O._round(1.23449, 3) == 1.234


def test_cielab_of_the_srgb_primaries():
    """Published D65 CIELAB values of the sRGB primaries, white and mid grey; the 4-digit matrix moves them by less
    than 0.03."""
    for rgb, want in (((255, 0, 0), (53.2408, 80.0925, 67.2032)), ((0, 255, 0), (87.7347, -86.1827, 83.1793)),
                      ((0, 0, 255), (32.2970, 79.1875, -107.8602)), ((255, 255, 255), (100.0, 0.0, 0.0)),
                      ((128, 128, 128), (53.5850, 0.0, 0.0))):
        got = O.lab(O._linear(np.array(rgb, np.uint8)))
        assert np.abs(got - np.array(want)).max() < 0.03, (rgb, got)


def _flat(h, w, v):
    return np.full((h, w, 3), v, np.uint8)


def test_psnr_and_equal_images():
    rng = np.random.default_rng(1)
    a = rng.integers(10, 240, (12, 16, 3)).astype(np.uint8)
    r = O.compare(a, a)
    assert r == {"psnr_db": 99.0, "ssim": 1.0, "mean_delta_e": 0.0, "p95_delta_e": 0.0, "p99_delta_e": 0.0,
                 "trail_px": 0, "luminance_ratio": 1.0}
    # every channel off by 5: MSE 25
    r = O.compare(a, a + 5)
    assert r["psnr_db"] == O._round(10 * math.log10(255.0 ** 2 / 25.0), 2) == 34.15


def test_ssim_of_flat_images_is_its_closed_form():
    """Flat images have no variance or covariance: SSIM = (2 a b + C1) / (a^2 + b^2 + C1)."""
    a, b = _flat(9, 13, 100), _flat(9, 13, 160)
    ya, yb = 100 * 0.299 + 100 * 0.587 + 100 * 0.114, 160 * 0.299 + 160 * 0.587 + 160 * 0.114
    c1 = (0.01 * 255) ** 2
    want = (2 * ya * yb + c1) / (ya * ya + yb * yb + c1)
    assert O.ssim(np.full((9, 13), ya), np.full((9, 13), yb)) == pytest.approx(want, rel=1e-12)
    assert O.compare(a, b)["ssim"] == O._round(want, 4)


def test_ssim_is_the_two_dimensional_window():
    """Brute force: the 11 x 11 Gaussian window summed directly in two dimensions, coordinates clamped at the
    edges, on a small image pair (the separable blur must give the same statistics)."""
    rng = np.random.default_rng(2)
    h, w = 7, 9
    a = rng.uniform(0, 255, (h, w))
    b = np.clip(a + rng.normal(0, 20, (h, w)), 0, 255)
    g = [math.exp(-(i * i) / (2 * 1.5 * 1.5)) for i in range(-5, 6)]
    g = [v / sum(g) for v in g]
    c1, c2 = (0.01 * 255) ** 2, (0.03 * 255) ** 2
    total = 0.0
    for y in range(h):
        for x in range(w):
            m = {"a": 0.0, "b": 0.0, "aa": 0.0, "bb": 0.0, "ab": 0.0}
            for i in range(-5, 6):
                for j in range(-5, 6):
                    yy, xx = min(max(y + i, 0), h - 1), min(max(x + j, 0), w - 1)
                    k = g[i + 5] * g[j + 5]
                    pa, pb = a[yy, xx], b[yy, xx]
                    m["a"] += k * pa
                    m["b"] += k * pb
                    m["aa"] += k * pa * pa
                    m["bb"] += k * pb * pb
                    m["ab"] += k * pa * pb
            va, vb = m["aa"] - m["a"] ** 2, m["bb"] - m["b"] ** 2
            cov = m["ab"] - m["a"] * m["b"]
            total += ((2 * m["a"] * m["b"] + c1) * (2 * cov + c2)) / ((m["a"] ** 2 + m["b"] ** 2 + c1) * (va + vb + c2))
    assert O.ssim(a, b) == pytest.approx(total / (h * w), rel=1e-9)


def test_delta_e_percentiles_take_the_references_ranks():
    """100 pixels, k of them black against white (delta E 100): p95 is the 96th smallest of the sorted list
    (index 95), p99 the 100th."""
    for k, p95, p99 in ((4, 0.0, 100.0), (5, 100.0, 100.0), (1, 0.0, 100.0), (0, 0.0, 0.0)):
        a, b = _flat(10, 10, 255), _flat(10, 10, 255)
        b.reshape(-1, 3)[:k] = 0
        r = O.compare(a, b)
        assert r["p95_delta_e"] == p95 and r["p99_delta_e"] == p99, (k, r)
        assert r["mean_delta_e"] == O._round(k * float(O.lab(np.ones(3))[0]) / 100, 2)


def test_trail_is_the_p95_row_run():
    """Row y holds a run of y wrong pixels (and scattered single ones): sorted runs are 0..19, and the p95 over
    20 rows is index 19."""
    a, b = _flat(20, 40, 255), _flat(20, 40, 255)
    for y in range(20):
        b[y, 5:5 + y] = 0
        if y < 10:
            b[y, 30] = 0  # a lone wrong pixel beside the run
    assert O.compare(a, b)["trail_px"] == 19
    # 40 rows: index 38 of the sorted runs, so the one longest row does not count and the next one does
    a, b = _flat(40, 40, 255), _flat(40, 40, 255)
    b[3, 0:40:2] = 0  # alternating: runs of one
    b[7, 2:32] = 0    # a run of 30
    assert O.compare(a, b)["trail_px"] == 1


def test_luminance_ratio_and_regions(tmp_path):
    a, b = _flat(6, 8, 128), _flat(6, 8, 64)
    lin = O._linear(np.array([128, 64], np.uint8))
    assert O.compare(a, b)["luminance_ratio"] == O._round(lin[1] / lin[0], 3)
    # a region that holds only the differing half
    b2 = a.copy()
    b2[:, 4:] = 64
    assert O.compare(a, b2, region=(0.5, 0.0, 1.0, 1.0))["luminance_ratio"] == O._round(lin[1] / lin[0], 3)
    assert O.compare(a, b2, region=(0.0, 0.0, 0.5, 1.0))["psnr_db"] == 99.0
    # files, and the side-by-side sheet
    O.write_png(tmp_path / "a.png", a)
    O.write_png(tmp_path / "b.png", b2)
    assert (O.read_rgb8(tmp_path / "a.png") == a).all()
    r = O.compare(tmp_path / "a.png", tmp_path / "b.png", sheet=tmp_path / "s.png")
    assert r == O.compare(a, b2)
    s = O.read_rgb8(tmp_path / "s.png")
    assert s.shape == (6, 24, 3) and (s[:, :8] == a).all() and (s[:, 8:16] == b2).all()
    with pytest.raises(ValueError):
        O.compare(a, _flat(6, 9, 0))


@pytest.mark.gpu
def test_resolve_is_the_readback_normalisation(device):
    """trt_accum_resolve against gpu.rs:4218-4231 restated over the raw accumulator (trt_accum_read): color, albedo
    and normal over max(count, 1), variance (E[lum^2] - E[lum]^2)^+ / count; a fresh accumulator resolves to zeros."""
    from threert_amd import render as R

    print("resolve vs the readback normalisation: Cornell 40 x 24, 3 spp", flush=True)
    rs, cam = R.cornell()
    sc = R.upload(device, rs)
    acc = R.Accum(device, 40, 24)
    c, a, n, v = acc.resolve()
    assert not c.any() and not a.any() and not n.any() and not v.any()
    R.render(sc, rs.params(40, 24, cam, seed=2), 0, 3, acc)
    device.sync()
    raw = acc.read().reshape(-1, 20)
    cnt = np.maximum(raw[:, 3], np.float32(1.0))
    p = np.where(np.isfinite(raw), raw, np.float32(0.0))
    mean = p[:, 7] / cnt
    want = (p[:, 0:3] / cnt[:, None], p[:, 4:7] / cnt[:, None], p[:, 8:11] / cnt[:, None],
            np.maximum(p[:, 11] / cnt - mean * mean, np.float32(0.0)) / cnt)
    got = acc.resolve()
    assert (raw[:, 3] == 3).all()
    for g, w in zip(got, want):
        assert (g.reshape(w.shape) == w).all()
    acc.close()
    sc.close()
