"""The a-trous denoiser (denoise.rs:10-74): the oracle pinned by what the filter's definition fixes in closed form
(its impulse response and the dilations of its taps, zero weight across normals at right angles, each edge-stopping
weight read from a single odd pixel, the second iteration's width from the propagated variance), and the GPU's
against the oracle's, bit for bit."""
import math

import numpy as np
import pytest

import oracle


K = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])  # denoise.rs:16 K, as the 5 taps
LW = np.array([0.2126, 0.7152, 0.0722])
S2 = float((K ** 2).sum()) ** 2  # sum of the squared 2-D taps


def guides(h, w, albedo=1.0):
    """Uniform guidance: every normal (0, 0, 1), every albedo grey `albedo`."""
    normal = np.zeros((h, w, 3), np.float32)
    normal[..., 2] = 1.0
    return np.full((h, w, 3), albedo, np.float32), normal


def dilated(step: int) -> np.ndarray:
    k = np.zeros(4 * step + 1)
    k[::step] = K
    return k


def test_an_impulse_spreads_as_the_kernel_and_its_dilations():
    """With uniform guidance and a variance so large that the luminance weight is 1, iteration k is the separable
    kernel [1/16, 1/4, 3/8, 1/4, 1/16] with its taps 2^k apart, renormalised over the taps inside the image: an
    impulse after n iterations is the outer product of the 1-D kernels convolved over k < n (float64 closed form)."""
    h, w = 41, 47
    albedo, normal = guides(h, w)
    var = np.full((h, w), 1e30, np.float32)
    for cy, cx in ((20, 23), (0, 0), (2, 45)):  # centre, a corner, near an edge
        color = np.zeros((h, w, 3), np.float32)
        color[cy, cx] = 1.0
        for iters in (1, 2, 3):
            got = oracle.denoise_atrous(color, albedo, normal, var, iters)[..., 0].astype(np.float64)
            # the closed form: each iteration a separable, border-renormalised filter, in x and in y
            want = np.zeros((h, w))
            want[cy, cx] = 1.0
            for it in range(iters):
                k = dilated(1 << it)
                r = len(k) // 2
                for axis, n in ((1, w), (0, h)):
                    out = np.zeros_like(want)
                    for i in range(n):
                        idx = np.arange(i - r, i + r + 1)
                        ok = (idx >= 0) & (idx < n)
                        wk = k[ok]
                        sl = np.take(want, idx[ok], axis=axis)
                        # renormalised over this axis' taps inside the image (the 2-D sum is their product)
                        val = np.tensordot(sl, wk, axes=([axis], [0])) / wk.sum()
                        if axis == 1:
                            out[:, i] = val
                        else:
                            out[i, :] = val
                    want = out
            assert np.allclose(got, want, rtol=2e-6, atol=1e-9), (cy, cx, iters, np.abs(got - want).max())
    # the interior response is exactly the kernel's outer product
    color = np.zeros((h, w, 3), np.float32)
    color[20, 23] = 1.0
    got = oracle.denoise_atrous(color, albedo, normal, var, 1)[18:23, 21:26, 1]
    assert np.allclose(got, np.outer(K, K), rtol=1e-6)


def test_normals_at_right_angles_do_not_bleed():
    """dot(n, n')^64 with n . n' = 0 is a weight of exactly 0: two halves facing at right angles keep their colours
    (to the rounding of a weighted mean of equal values)."""
    h, w = 24, 32
    albedo, normal = guides(h, w, 0.5)
    normal[:, w // 2:, 2] = 0.0
    normal[:, w // 2:, 0] = 1.0
    rng = np.random.default_rng(4)
    color = np.full((h, w, 3), 0.25, np.float32)
    color[:, w // 2:] = 1.0
    var = rng.random((h, w)).astype(np.float32)
    out = oracle.denoise_atrous(color, albedo, normal, var, 4)
    assert np.allclose(out[:, :w // 2], 0.25, rtol=1e-6) and np.allclose(out[:, w // 2:], 1.0, rtol=1e-6)


def one_odd_pixel(c0, c1, k):
    """The value of a pixel whose taps all carry c0 but one, of tap weight k (of a total 1), carrying c1 with an
    edge-stopping weight: (c0 (1 - k) + c1 k wt) / ((1 - k) + k wt), for wt the tap's weight."""
    return lambda wt: (c0 * (1 - k) + c1 * k * wt) / ((1 - k) + k * wt)


def test_the_edge_stopping_weights_are_their_closed_forms():
    """denoise.rs:48-63: the luminance weight exp(-|dl| / (4 sqrt(var) + 1e-4)) with the variance demodulated by the
    albedo's luminance squared, the normal weight max(n . n', 0)^64 and the albedo weight exp(-|da|^2 / 0.01), each
    read from a pixel next to one odd pixel in a flat image (tap weight 1/4 x 3/8), in float64."""
    h, w = 21, 21
    k = 0.25 * 0.375
    py, px, qy, qx = 10, 10, 10, 11
    # luminance: grey albedo 0.5, variance v: demodulated 4v, so sigma = 4 sqrt(4 v) + 1e-4
    albedo, normal = guides(h, w, 0.5)
    v = 0.01
    color = np.full((h, w, 3), 0.2, np.float32)
    sigma = 4 * math.sqrt(4 * v) + 1e-4
    dl = sigma * 0.8  # demodulated luminance step
    color[py, px] = 0.2 + 0.5 * dl
    out = oracle.denoise_atrous(color, albedo, normal, np.full((h, w), v, np.float32), 1)
    c0, c1 = 0.2 / 0.5, (0.2 + 0.5 * dl) / 0.5
    want = 0.5 * one_odd_pixel(c0, c1, k)(math.exp(-dl / sigma))
    assert abs(out[qy, qx, 1] - want) < 2e-6 * want, (out[qy, qx, 1], want)
    assert abs(out[qy, qx, 1] - 0.5 * one_odd_pixel(c0, c1, k)(math.exp(-dl / (3 * math.sqrt(4 * v))))) > 1e-4
    # normal: the odd pixel's normal turned so that n . n' = 0.5^(1/64): weight 1/2; colours differ, luminance weight 1
    albedo, normal = guides(h, w)
    cos = 0.5 ** (1 / 64)
    normal[py, px] = (math.sqrt(1 - cos * cos), 0.0, cos)
    color = np.full((h, w, 3), 0.3, np.float32)
    color[py, px] = 0.9
    out = oracle.denoise_atrous(color, albedo, normal, np.full((h, w), 1e30, np.float32), 1)
    wn = float(np.float32(cos)) ** 64
    want = one_odd_pixel(0.3, 0.9, k)(wn)
    assert abs(out[qy, qx, 0] - want) < 2e-6 * want and abs(wn - 0.5) < 1e-3
    assert abs(out[qy, qx, 0] - one_odd_pixel(0.3, 0.9, k)(0.5 ** 0.5)) > 1e-3  # an exponent of 32 reads otherwise
    # albedo: the odd pixel's albedo 1 + d on every channel and its colour likewise, so the demodulated colours are
    # equal at one channel's worth; weight exp(-3 d^2 / 0.01)
    albedo, normal = guides(h, w)
    d = 0.05
    albedo[py, px] = 1 + d
    color = np.full((h, w, 3), 0.3, np.float32)
    color[py, px] = 0.6 * (1 + d)  # demodulated 0.6 against 0.3
    out = oracle.denoise_atrous(color, albedo, normal, np.full((h, w), 1e30, np.float32), 1)
    da2 = 3 * float(np.float32(1 + d) - np.float32(1)) ** 2
    want = one_odd_pixel(0.3, 0.6, k)(math.exp(-da2 / 0.01))
    assert abs(out[qy, qx, 2] - want) < 2e-6 * want, (out[qy, qx, 2], want)


def test_the_second_iteration_reads_the_propagated_variance():
    """denoise.rs:24-70 over two iterations, in closed form. A flat image (c) with one odd pixel p (c_p) four columns
    from the left edge, uniform variance v, unit albedo and normals. Iteration 1 (taps 1 apart, sigma1 =
    4 sqrt(v) + 1e-4) changes only p and the pixels within 2 of it, each by the one-odd-pixel form. At the edge the
    variance becomes v x (sum of squared taps) / (sum of taps)^2 over the taps inside the image (vsum / wsum^2, the
    taps summing to less than 1 there). Iteration 2 at q, the edge pixel in p's row (taps 2 apart), widens its
    luminance weight by the 3 x 3 filter of that variance (weights 1/2, 1/4, renormalised at the edge) and reads p and
    the changed pixels 2 columns from it."""
    h, w = 33, 20
    albedo, normal = guides(h, w)
    c, cp, v = 0.2, 1.0, 0.02
    color = np.full((h, w, 3), c, np.float32)
    py, px = 16, 4
    color[py, px] = cp
    out = oracle.denoise_atrous(color, albedo, normal, np.full((h, w), v, np.float32), 2)
    s1 = 4 * math.sqrt(v) + 1e-4
    wl1 = math.exp(-abs(c - cp) / s1)  # grey colours: luminance = value
    k00 = K[2] * K[2]
    after1 = {(0, 0): (cp * k00 + c * (1 - k00) * wl1) / (k00 + (1 - k00) * wl1)}  # p itself: every tap but one odd
    for dx in (-2, 0, 2):
        for dy in (-2, 0, 2):
            if (dx, dy) != (0, 0):  # p's tap weight at p + (dx, dy)
                after1[(dx, dy)] = one_odd_pixel(c, cp, K[2 + dx] * K[2 + dy])(wl1)
    sy = float((K ** 2).sum())  # rows: all five taps inside, summing to 1

    def var1(col):  # the propagated variance at an edge column (its taps all carry c: weights are the taps)
        kx = np.array([K[j + 2] for j in range(-2, 3) if 0 <= col + j < w])
        return v * sy * float((kx ** 2).sum()) / float(kx.sum()) ** 2

    vf = (0.5 * var1(0) + 0.25 * var1(1)) / 0.75  # the 3 x 3 filter at column 0: columns 0 (1/2) and 1 (1/4)

    def second(sigma):
        num = den = 0.0
        for a in range(0, 3):  # columns 0, 2, 4: the taps inside the image
            for b in range(-2, 3):
                val = after1.get((2 * a - 4, 2 * b), c)
                kk = K[a + 2] * K[b + 2]
                wl = math.exp(-abs(c - val) / sigma)
                num += kk * wl * val
                den += kk * wl
        return num / den

    want = second(4 * math.sqrt(vf) + 1e-4)
    got = float(out[py, 0, 0])
    assert abs(got - want) < 5e-6 * want, (got, want)
    # the variance not propagated, or vsum divided by wsum once, reads otherwise
    once = (0.5 * var1(0) * (11 / 16) + 0.25 * var1(1) * (15 / 16)) / 0.75
    for wrong in (4 * math.sqrt(v) + 1e-4, 4 * math.sqrt(once) + 1e-4):
        assert abs(got - second(wrong)) > 1e-4 * want, (got, second(wrong))


def noisy_frame(seed=0, h=40, w=56):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    albedo = np.where((xx > w // 2)[..., None], [0.8, 0.3, 0.2], [0.2, 0.6, 0.7]).astype(np.float32)
    normal = np.zeros((h, w, 3), np.float32)
    normal[..., 1] = np.where(yy > h // 2, 1.0, 0.0)
    normal[..., 2] = np.where(yy > h // 2, 0.0, 1.0)
    clean = albedo * (0.5 + 0.5 * np.sin(xx / 9.0))[..., None]
    color = (clean * rng.exponential(1.0, (h, w, 1))).astype(np.float32)
    variance = (0.05 * rng.random((h, w))).astype(np.float32)
    return color, albedo.astype(np.float32), normal, variance


def test_expf_is_within_two_ulps():
    for x in np.linspace(-80.0, 0.0, 4001, dtype=np.float32)[::7]:
        y = oracle.expf(float(x))
        assert abs(y - math.exp(float(x))) <= 2.5e-7 * math.exp(float(x))


def test_a_flat_image_stays_flat_and_edges_hold():
    h, w = 24, 32
    albedo = np.full((h, w, 3), 0.5, np.float32)
    normal = np.zeros((h, w, 3), np.float32)
    normal[..., 2] = 1.0
    color = np.full((h, w, 3), 0.25, np.float32)
    var = np.zeros((h, w), np.float32)
    out = oracle.denoise_atrous(color, albedo, normal, var, 4)
    assert np.allclose(out, 0.25, rtol=1e-6)
    # two halves facing away from each other do not bleed (dot(n, n')^64 = 0)
    normal[:, w // 2:, 2] = 0.0
    normal[:, w // 2:, 0] = 1.0
    color[:, w // 2:] = 1.0
    out = oracle.denoise_atrous(color, albedo, normal, var, 4)
    assert np.allclose(out[:, : w // 2], 0.25, rtol=1e-6) and np.allclose(out[:, w // 2:], 1.0, rtol=1e-6)


@pytest.mark.gpu
def test_gpu_atrous_is_the_oracles(device):
    from threert_amd import denoise as D

    for seed, (h, w) in [(1, (40, 56)), (2, (33, 70)), (3, (128, 96))]:
        color, albedo, normal, variance = noisy_frame(seed, h, w)
        want = oracle.denoise_atrous(color, albedo, normal, variance, 4)
        got = D.atrous(device, color, albedo, normal, variance, 4)
        assert (got.view(np.uint32) == want.view(np.uint32)).all(), np.abs(got - want).max()
