"""The learned denoiser: its oracle pinned by the mathematics of its parts (the convolution's definition, adjoint
identities, central finite differences of the whole network in float64, bfloat16 rounding), then the GPU's
forward pass, gradients and training held to the oracle.

Tolerances: the GPU stores activations and output gradients in bfloat16 as the emulating oracle does, but sums in
FP32 on MFMA in another order, and a ReLU whose input sits within rounding of zero may switch. Forward: relative
L2 error of the denoised image <= 1e-2. Gradients, per layer: cosine similarity with the emulating oracle >= 0.995
and relative L2 error <= 0.1 (the deepest levels sum over few pixels); and against the exact float64 gradient the
GPU is within twice the error that bfloat16 storage alone gives (the emulating oracle's), plus 5e-3. The loss
within 1e-3 relative."""
import numpy as np
import pytest

from oracle import unet as O
from threert_amd import unet as U


def test_bfloat16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3e38, np.inf, 1e-40], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 2.9980e38, np.inf, 0.0], np.float32)
    got = O.bf16(x)
    assert got[:5].tolist() == want[:5].tolist()
    assert got[6] == np.inf and abs(got[5] / 3e38 - 1) < 2 ** -8 and abs(got[7]) < 2e-40
    assert np.isnan(O.bf16(np.array([np.nan], np.float32))[0])


def test_conv_is_the_definition():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 5, 4))
    w = rng.standard_normal((3, 9, 4))
    b = rng.standard_normal(3)
    assert np.allclose(O.conv(x, None, w, b), O.direct_conv(x, w, b), rtol=1e-12, atol=1e-12)
    # the upsampled second input is nearest-neighbour: each half-size pixel covers a 2x2 block
    x1 = rng.standard_normal((2, 3, 3, 2))
    w2 = rng.standard_normal((3, 9, 6))
    full = np.concatenate([rng.standard_normal((2, 6, 6, 4)), np.zeros((2, 6, 6, 2))], axis=3)
    full[..., 4:] = np.kron(x1, np.ones((1, 2, 2, 1)))
    assert np.allclose(O.conv(full[..., :4], x1, w2, None), O.direct_conv(full, w2, 0.0), rtol=1e-12, atol=1e-12)


def test_the_gradients_of_conv_are_its_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 8, 6, 5))
    w = rng.standard_normal((4, 9, 5))
    y = rng.standard_normal((1, 8, 6, 4))
    lhs = np.sum(O.conv(x, None, w, None) * y)
    assert np.isclose(lhs, np.sum(x * O.conv_input_grad(y, w)), rtol=1e-12)
    dw, db = O.conv_weight_grad(x, None, y)
    assert np.isclose(lhs, np.sum(w * dw), rtol=1e-12)
    assert np.allclose(db, y.sum(axis=(0, 1, 2)))


def _batch(rng, n, h, w):
    color = np.exp(rng.normal(-0.5, 1.2, (n, h, w, 3))).astype(np.float32)
    albedo = rng.uniform(0.05, 0.95, (n, h, w, 3)).astype(np.float32)
    nrm = rng.standard_normal((n, h, w, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=3, keepdims=True)).astype(np.float32)
    target = (color * rng.uniform(0.5, 1.5, (n, h, w, 3))).astype(np.float32)
    return color, albedo, nrm, target


def test_network_gradients_are_central_differences():
    rng = np.random.default_rng(3)
    params = U.init_params(7).astype(np.float64)
    params += rng.normal(0, 0.01, params.size) * (params == 0)  # non-zero biases exercise them too
    color, albedo, nrm, target = _batch(rng, 1, 16, 16)
    net = O.Net(params, emulate=False)
    loss, g = net.loss_and_grads(color, albedo, nrm, target)
    lay = U.layout()
    picks = []
    for name, wo, shape, bo, co in lay:
        n_real = shape[0] if name != "out" else 3
        picks.append(wo + int(rng.integers(0, n_real)) * shape[1] * shape[2] + int(rng.integers(0, shape[1] * shape[2])))
        picks.append(bo + int(rng.integers(0, n_real)))
    eps = 1e-6
    checked = 0
    for i in picks:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        l1 = O.Net(p1, emulate=False).loss_and_grads(color, albedo, nrm, target)[0]
        l2 = O.Net(p2, emulate=False).loss_and_grads(color, albedo, nrm, target)[0]
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
        checked += 1
    assert checked == 2 * len(lay) and loss > 0


# ---- GPU ---------------------------------------------------------------------------------------------------------
def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
def test_gpu_forward_matches_the_oracle(device):
    print("unet forward 2 x 32 x 48", flush=True)
    rng = np.random.default_rng(4)
    color, albedo, nrm, _ = _batch(rng, 2, 32, 48)
    params = U.init_params(5)
    params[U.layout()[-1][3]:U.layout()[-1][3] + 3] = 0.3  # the output's bias lifts most pixels above the clamp
    net = U.UNet(device, 2 * 32 * 48, params)
    got = net.denoise(color, albedo, nrm)
    ref = O.Net(params, emulate=True)
    ref.forward(color, albedo, nrm)
    want = ref.denoised()
    print("forward rel L2", _rel(got, want), flush=True)
    assert _rel(got, want) <= 3e-4  # measured 1.05e-4
    net.close()


@pytest.mark.gpu
def test_tall_tiles_give_the_short_tiles_bits(device):
    """The same input and weights with 16-pixel-tall tiles at the levels 128 pixels tall or more (the default) and
    with 8-pixel tiles everywhere: the forward outputs are the same bits. The gradients agree within 1e-5 relative
    L2: the weight gradients are summed with FP32 atomics, whose order differs from run to run even on one tiling
    (a slip in the tall tiles' halo or strips would move them by orders of magnitude more)."""
    print("unet: tall against short tiles, 1 x 144 x 64 and 2 x 128 x 32", flush=True)
    params = U.init_params(12)
    params[U.layout()[-1][3]:U.layout()[-1][3] + 3] = 0.3
    for n, h, w in ((1, 144, 64), (2, 128, 32)):
        color, albedo, nrm, target = _batch(np.random.default_rng(h), n, h, w)
        out = {}
        for rows in (128, 0xFFFFFFFF):
            U.set_tall_tiles(rows)
            try:
                net = U.UNet(device, n * h * w, params)
                out[rows] = (net.denoise(color, albedo, nrm), net.gradients(color, albedo, nrm, target))
                net.close()
            finally:
                U.set_tall_tiles(128)
        (d0, (l0, g0)), (d1, (l1, g1)) = out[128], out[0xFFFFFFFF]
        assert (d0.view(np.uint32) == d1.view(np.uint32)).all(), (n, h, w)
        assert abs(l0 - l1) <= 1e-6 * abs(l0)
        assert _rel(g0, g1) <= 1e-5, (n, h, w, _rel(g0, g1))


@pytest.mark.gpu
def test_gpu_forward_on_tall_tiles_matches_the_oracle(device):
    """Levels 128 pixels tall or more take 16-pixel-tall tiles (the others 8): the forward pass at 144 x 64, whose
    first level does, within the same bound."""
    print("unet forward 1 x 144 x 64 (tall tiles at the first level)", flush=True)
    rng = np.random.default_rng(11)
    color, albedo, nrm, _ = _batch(rng, 1, 144, 64)
    params = U.init_params(12)
    params[U.layout()[-1][3]:U.layout()[-1][3] + 3] = 0.3
    net = U.UNet(device, 144 * 64, params)
    got = net.denoise(color, albedo, nrm)
    ref = O.Net(params, emulate=True)
    ref.forward(color, albedo, nrm)
    want = ref.denoised()
    print("forward rel L2", _rel(got, want), flush=True)
    assert _rel(got, want) <= 8e-3  # measured 3.96e-3
    net.close()


@pytest.mark.gpu
def test_gpu_gradients_match_the_oracle(device):
    print("unet gradients 2 x 32 x 32 and 1 x 128 x 16 (tall tiles at the first level)", flush=True)
    rng = np.random.default_rng(6)
    color, albedo, nrm, target = _batch(rng, 2, 32, 32)
    params = U.init_params(9)
    _check_gradients(device, params, color, albedo, nrm, target)
    color, albedo, nrm, target = _batch(rng, 1, 128, 16)
    _check_gradients(device, params, color, albedo, nrm, target)


def _check_gradients(device, params, color, albedo, nrm, target):
    n, h, w = color.shape[:3]
    net = U.UNet(device, n * h * w, params)
    loss, g = net.gradients(color, albedo, nrm, target)
    rl, rg = O.Net(params, emulate=True).loss_and_grads(color, albedo, nrm, target)
    _, xg = O.Net(params.astype(np.float64), emulate=False).loss_and_grads(color, albedo, nrm, target)
    print(f"loss gpu {loss:.6f} oracle {rl:.6f}", flush=True)
    assert abs(loss - rl) <= 1e-3 * abs(rl)
    for name, wo, shape, bo, co in U.layout():
        n = shape[0] * shape[1] * shape[2]
        for part, (s, e) in (("w", (wo, wo + n)), ("b", (bo, bo + co))):
            a, b, x = g[s:e], rg[s:e], xg[s:e]
            if np.linalg.norm(b) == 0:
                assert np.linalg.norm(a) == 0, (name, part)
                continue
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
            # bfloat16 storage alone (the emulating oracle) moves the exact gradient by rel(b, x); the GPU may add
            # as much again from its summation order, and no more
            floor = _rel(b, x)
            print(f"{name}.{part}: gpu vs emulated {_rel(a, b):.2e} cos {cos:.5f}; vs exact gpu {_rel(a, x):.2e} "
                  f"emulated {floor:.2e}", flush=True)
            assert cos >= 0.995 and _rel(a, b) <= 0.1, (name, part)
            assert _rel(a, x) <= 2.0 * floor + 5e-3, (name, part)
    net.close()


@pytest.mark.gpu
def test_training_lowers_the_loss(device):
    """Adam steps on one batch whose target is a function of the inputs (the color scaled by its albedo), so the
    loss can fall well below where the network starts."""
    print("unet: 100 Adam steps on one batch", flush=True)
    rng = np.random.default_rng(8)
    color, albedo, nrm, _ = _batch(rng, 4, 32, 32)
    target = (color * (0.25 + albedo)).astype(np.float32)
    net = U.UNet(device, 4 * 32 * 32, U.init_params(1))
    bufs = {}
    for k, a in (("c", color), ("a", albedo), ("n", nrm), ("t", target)):
        bufs[k] = device.alloc(a.nbytes)
        bufs[k].upload(a)
    first = net.train_step_device(4, 32, 32, bufs["c"], bufs["a"], bufs["n"], bufs["t"], 1e-3, want_loss=True)
    for _ in range(98):
        net.train_step_device(4, 32, 32, bufs["c"], bufs["a"], bufs["n"], bufs["t"], 1e-3)
    last = net.train_step_device(4, 32, 32, bufs["c"], bufs["a"], bufs["n"], bufs["t"], 1e-3, want_loss=True)
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.7 * first
    p = net.params()
    assert np.isfinite(p).all()
    for b in bufs.values():
        b.free()
    net.close()


@pytest.mark.gpu
def test_the_trained_denoiser_beats_more_samples(device):
    """The shipped weights on held-out scenes (seeds never trained on) at 384 x 224: from one sample, the U-Net's
    relative RMSE against a 256-sample float64-accumulated reference is under half the input's and under that of a
    plain 4-sample render."""
    from threert_amd import denoise_train as D

    net = U.UNet(device, 384 * 224, U.shipped_params())
    for kind, seed in (("cornell", 20_000), ("soup", 20_001)):
        print(f"held out: {kind} {seed}", flush=True)
        inp, tgt = D.render_set(device, kind, seed, 384, 224, 1, 256)
        color, albedo, normal = inp[0, ..., 0:3], inp[0, ..., 3:6], inp[0, ..., 6:9]
        den = net.denoise(color, albedo, normal)
        four = D.spp_curve(device, kind, seed, 384, 224, spps=(4,))[4]
        e_in, e_den, e_four = D.rel_rmse(color, tgt), D.rel_rmse(den, tgt), D.rel_rmse(four, tgt)
        print(f"  rel RMSE: 1 spp {e_in:.3f}, 4 spp {e_four:.3f}, denoised {e_den:.3f}", flush=True)
        assert e_den < 0.5 * e_in and e_den < e_four
    net.close()
