"""The a-trous denoiser (denoise.rs:10-74): the oracle against an independent float64 restatement and the filter's
invariants; the GPU's against the oracle's, bit for bit."""
import math

import numpy as np
import pytest

import oracle


def numpy_atrous(color, albedo, normal, variance, iterations):
    """denoise.rs in float64, written independently of the C restatement (vectorised over pixels)."""
    h, w = color.shape[:2]
    eps = 1e-3
    am = np.maximum(albedo.astype(np.float64), eps)
    lw = np.array([0.2126, 0.7152, 0.0722])
    cur = color / am
    var = variance / (am @ lw) ** 2
    K = [3 / 8, 1 / 4, 1 / 16]
    nrm = normal.astype(np.float64)
    alb = albedo.astype(np.float64)

    def shifted(a, dx, dy):
        out = np.zeros_like(a)
        ok = np.zeros((h, w), bool)
        ys, yd = (slice(max(dy, 0), h + min(dy, 0)), slice(max(-dy, 0), h + min(-dy, 0)))
        xs, xd = (slice(max(dx, 0), w + min(dx, 0)), slice(max(-dx, 0), w + min(-dx, 0)))
        out[yd, xd] = a[ys, xs]
        ok[yd, xd] = True
        return out, ok

    for it in range(iterations):
        step = 1 << it
        s = np.zeros((h, w))
        ws = np.zeros((h, w))
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = (0.5 if dx == 0 else 0.25) * (0.5 if dy == 0 else 0.25)
                v, ok = shifted(var, dx, dy)
                s += np.where(ok, k * v, 0)
                ws += np.where(ok, k, 0)
        varf = s / ws
        lp = cur @ lw
        sig = 4 * np.sqrt(np.maximum(varf, 0)) + 1e-4
        sm = np.zeros_like(cur)
        wsum = np.zeros((h, w))
        vsum = np.zeros((h, w))
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                cq, ok = shifted(cur, dx * step, dy * step)
                nq, _ = shifted(nrm, dx * step, dy * step)
                aq, _ = shifted(alb, dx * step, dy * step)
                vq, _ = shifted(var, dx * step, dy * step)
                k = K[abs(dx)] * K[abs(dy)]
                wl = np.exp(-np.abs(lp - cq @ lw) / sig)
                wn = np.maximum((nrm * nq).sum(-1), 0) ** 64
                wa = np.exp(-((alb - aq) ** 2).sum(-1) / 0.01)
                wt = np.where(ok, k * wl * wn * wa, 0)
                sm += cq * wt[..., None]
                vsum += wt * wt * vq
                wsum += wt
        pos = wsum > 0
        nxt = np.where(pos[..., None], sm / np.where(pos, wsum, 1)[..., None], cur)
        var = np.where(pos, vsum / np.where(pos, wsum * wsum, 1), var)
        cur = nxt
    return cur * am


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


def test_the_oracle_is_the_filter():
    color, albedo, normal, variance = noisy_frame()
    for iters in (1, 2, 4):
        got = oracle.denoise_atrous(color, albedo, normal, variance, iters)
        want = numpy_atrous(color, albedo, normal, variance, iters)
        assert np.allclose(got, want, rtol=2e-4, atol=1e-6), np.abs(got - want).max()


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
