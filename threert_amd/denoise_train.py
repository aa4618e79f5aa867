"""Training data and training for the learned denoiser, all on the GPU through the library.

Data: seeded scenes from the Cornell-box and clustered-soup generators (the plain and precision workloads) and
small instanced cities; per scene, inputs (color, albedo, normal from the accumulator) read from one render after
1, 2, 4, ... 128 samples, so the network meets every noise level a server would hand it, and one many-sample
target. The target is rendered in launches of `chunk` samples, and the launches' sums are added on the GPU in FP64
(render.reference_fp64, an FP64-accumulated reference).

    python -m threert_amd.denoise_train --scenes 200 --steps 12000 --out weights.npy
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import Device
from . import render as R
from . import scene as S
from . import serve
from .unet import UNet, init_params


def cornell_variant(seed: int):
    """The Cornell box with seeded wall colors, box sizes, places and turns, light size and power, and camera."""
    rng = np.random.default_rng(seed)
    col = lambda lo, hi: tuple(float(x) for x in rng.uniform(lo, hi, 3))  # noqa: E731
    mats = [R.material(col(0.4, 0.85), specular=0.0, roughness=1.0),
            R.material(col(0.05, 0.8), specular=0.0),
            R.material(col(0.05, 0.8), specular=0.0),
            R.material((0.0, 0.0, 0.0), (float(rng.uniform(8, 40)),) * 3, specular=0.0),
            R.material(col(0.2, 0.9), metallic=float(rng.uniform() < 0.4), roughness=float(rng.uniform(0.1, 0.8)))]
    s = float(rng.uniform(0.12, 0.45)) / 2
    q = lambda *p: R.quad(*p)  # noqa: E731
    geo = [q((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), q((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)),
           q((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)), q((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)),
           q((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)),
           q((0.5 - s, 0.999, 0.5 - s), (0.5 + s, 0.999, 0.5 - s), (0.5 + s, 0.999, 0.5 + s), (0.5 - s, 0.999, 0.5 + s)),
           R.box_tris((-0.15, 0.0, -0.15), (0.15, float(rng.uniform(0.3, 0.75)), 0.15)),
           R.box_tris((-0.15, 0.0, -0.15), (0.15, float(rng.uniform(0.15, 0.4)), 0.15))]
    meshes = [R.RMesh(g) for g in geo]
    eye = np.eye(4, dtype=np.float32)
    inst = [R.RInstance(0, eye, [0]), R.RInstance(1, eye, [0]), R.RInstance(2, eye, [0]), R.RInstance(3, eye, [1]),
            R.RInstance(4, eye, [2]), R.RInstance(5, eye, [3]),
            R.RInstance(6, S.trs((float(rng.uniform(0.25, 0.4)), 0.0, float(rng.uniform(0.5, 0.7))),
                                 (0.0, math.radians(rng.uniform(-40, 40)), 0.0), (1, 1, 1)), [4]),
            R.RInstance(7, S.trs((float(rng.uniform(0.6, 0.75)), 0.0, float(rng.uniform(0.25, 0.45))),
                                 (0.0, math.radians(rng.uniform(-40, 40)), 0.0), (1, 1, 1)), [0])]
    rs = R.build(meshes, inst, mats)
    eye_p = (0.5 + rng.uniform(-0.15, 0.15), 0.5 + rng.uniform(-0.15, 0.15), -1.35 + rng.uniform(-0.2, 0.3))
    cam = R.Camera(eye_p, (0.5, 0.45, 0.5), float(rng.uniform(34, 46)))
    return rs, cam, (0.0, 0.0, 0.0)


def scene_of(kind: str, seed: int):
    if kind == "cornell":
        return cornell_variant(seed)
    if kind == "soup":
        n = int(np.exp(np.random.default_rng(seed).uniform(np.log(20_000), np.log(200_000))))
        return serve.soup_scene(n, seed)
    if kind == "bigsoup":  # the serving workload's large scenes: sub-pixel triangles
        n = int(np.exp(np.random.default_rng(seed).uniform(np.log(200_000), np.log(2_000_000))))
        return serve.soup_scene(n, seed)
    rs, cam, sky = R.city(seed=seed, meshes=8, instances=96, extent=80.0)
    return rs, cam, sky


def scene_list(count: int, first_seed: int = 1, kinds=("cornell", "soup", "city")):
    return [(kinds[k % len(kinds)], first_seed + k) for k in range(count)]


def render_set(dev: Device, kind: str, seed: int, w: int, h: int, noisy: int, target_spp: int, chunk: int = 64,
               bounces: int = 4, levels: tuple = ()):
    """(inputs (k, h, w, 9): color, albedo, normal; target (h, w, 3)). The inputs are `noisy` independent 1-sample
    renders, or, given `levels` (ascending sample counts), one render read after each level's samples."""
    rs, cam, sky = scene_of(kind, seed)
    p = rs.params(w, h, cam, seed=seed, max_bounces=bounces, rr_depth=3, sky=sky)
    sc = R.upload(dev, rs)
    inputs = np.zeros((len(levels) or noisy, h, w, 9), np.float32)
    if levels:
        acc = R.Accum(dev, w, h)
        done = 0
        for k, n in enumerate(levels):
            R.render(sc, p, 100_000 + done, n - done, acc)  # sample indices far from the target's
            done = n
            c, a, nm, _ = acc.resolve()
            inputs[k, ..., 0:3], inputs[k, ..., 3:6], inputs[k, ..., 6:9] = c, a, nm
        acc.close()
    for k in range(0 if levels else noisy):
        acc = R.Accum(dev, w, h)
        R.render(sc, p, 100_000 + k, 1, acc)  # sample indices far from the target's
        c, a, n, _ = acc.resolve()
        inputs[k, ..., 0:3], inputs[k, ..., 3:6], inputs[k, ..., 6:9] = c, a, n
        acc.close()
    target = R.reference_fp64(dev, sc, p, w, h, target_spp, chunk).astype(np.float32)
    sc.close()
    return inputs, target


def rel_rmse(x: np.ndarray, ref: np.ndarray) -> float:
    """sqrt(mean((x - ref)^2 / (ref^2 + 1e-2))) over pixels and channels."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2))))


def train(dev: Device, net: UNet, inputs: np.ndarray, targets: np.ndarray, steps: int, lr: float = 1e-3,
          batch: int = 2, log=print, seed: int = 0):
    """inputs (images, noisy, h, w, 9), targets (images, h, w, 3): each step takes `batch` consecutive images of a
    seeded shuffle, with one of their noisy renders. Cosine learning-rate decay to lr / 20."""
    imgs, noisy, h, w, _ = inputs.shape
    rng = np.random.default_rng(seed)
    order = rng.permutation(imgs)
    # device copies, image-major in the shuffled order: color / albedo / normal planes for every noisy render
    px = h * w
    planes = {}
    for name, sl in (("color", slice(0, 3)), ("albedo", slice(3, 6)), ("normal", slice(6, 9))):
        a = np.ascontiguousarray(inputs[order][:, :, :, :, sl].transpose(1, 0, 2, 3, 4))  # (noisy, imgs, h, w, 3)
        planes[name] = dev.alloc(a.nbytes)
        planes[name].upload(a)
    tg = np.ascontiguousarray(targets[order])
    planes["target"] = dev.alloc(tg.nbytes)
    planes["target"].upload(tg)
    import ctypes as C

    from . import _check, lib

    base = {k: v.p.value for k, v in planes.items()}
    losses = []
    t0 = time.perf_counter()
    for step in range(steps):
        i = int(rng.integers(0, imgs - batch + 1))
        k = int(rng.integers(0, noisy))
        off_in = ((k * imgs + i) * px) * 12
        off_t = (i * px) * 12
        cur_lr = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * step / max(steps - 1, 1))))
        want = step % 100 == 0 or step == steps - 1
        loss = np.zeros(1, np.float32)
        _check(lib().trt_unet_train_step(net.h, batch, h, w, C.c_void_p(base["color"] + off_in),
                                         C.c_void_p(base["albedo"] + off_in), C.c_void_p(base["normal"] + off_in),
                                         C.c_void_p(base["target"] + off_t), float(cur_lr),
                                         loss.ctypes.data if want else None))
        if want:
            losses.append((step, float(loss[0])))
            log(f"train: step {step} loss {loss[0]:.4f} lr {cur_lr:.2e} ({time.perf_counter() - t0:.1f} s)")
    dev.sync()
    for v in planes.values():
        v.free()
    return losses


def spp_curve(dev: Device, kind: str, seed: int, w: int, h: int, spps=(1, 4, 16, 64), bounces: int = 4):
    """Color of plain renders at each spp in `spps` (sample indices 200000 on), for the fallback comparison."""
    rs, cam, sky = scene_of(kind, seed)
    p = rs.params(w, h, cam, seed=seed, max_bounces=bounces, rr_depth=3, sky=sky)
    sc = R.upload(dev, rs)
    out = {}
    for n in spps:
        acc = R.Accum(dev, w, h)
        R.render(sc, p, 200_000, n, acc)
        out[n] = acc.resolve()[0]
        acc.close()
    sc.close()
    return out


def evaluate(dev: Device, net: UNet, scenes, w: int = 1280, h: int = 720, target_spp: int = 1024, log=print,
             timing_reps: int = 20, levels: tuple = (1, 4, 16, 64)) -> dict:
    """Per held-out scene: relative RMSE to the FP64-accumulated target of the 1-sample input, the a-trous baseline
    (denoise.rs's stage), the U-Net over one render read at each of `levels` samples (`unet` is its 1-sample
    figure), and plain renders of 4 / 16 / 64 samples; and the U-Net's ms a frame at w x h."""
    from .denoise import atrous

    rows = []
    for kind, seed in scenes:
        inp, tgt = render_set(dev, kind, seed, w, h, 1, target_spp, levels=levels)
        color, albedo, normal = inp[0, ..., 0:3], inp[0, ..., 3:6], inp[0, ..., 6:9]
        den = net.denoise(color, albedo, normal)
        var = np.zeros((h, w), np.float32)  # one sample: the a-trous luminance weight sees no variance estimate
        at = atrous(dev, color, albedo, normal, var)
        curve = spp_curve(dev, kind, seed, w, h)
        row = {"scene": f"{kind}-{seed}", "noisy_1spp": rel_rmse(color, tgt), "atrous": rel_rmse(at, tgt),
               "unet": rel_rmse(den, tgt)}
        for k, n in enumerate(levels[1:], 1):
            row[f"unet_{n}spp"] = rel_rmse(net.denoise(inp[k, ..., 0:3], inp[k, ..., 3:6], inp[k, ..., 6:9]), tgt)
        for n, img in curve.items():
            row[f"plain_{n}spp"] = rel_rmse(img, tgt)
        rows.append(row)
        log(f"eval: {row}")
    # inference time at w x h: inputs resident in HBM, HIP events on the library's stream
    c = np.ascontiguousarray(color)
    bufs = {k: dev.alloc(c.nbytes) for k in ("c", "a", "n", "o")}
    bufs["c"].upload(color)
    bufs["a"].upload(albedo)
    bufs["n"].upload(normal)
    net.denoise_device(1, h, w, bufs["c"], bufs["a"], bufs["n"], bufs["o"])
    dev.sync()
    dev.timer_start()
    for _ in range(timing_reps):
        net.denoise_device(1, h, w, bufs["c"], bufs["a"], bufs["n"], bufs["o"])
    ms = dev.timer_stop() / timing_reps
    for b in bufs.values():
        b.free()
    keys = [k for k in rows[0] if k != "scene"]
    mean = {k: float(np.mean([r[k] for r in rows])) for k in keys}
    return {"scenes": rows, "mean": mean, "ms_per_frame": round(ms, 3), "width": w, "height": h,
            "target_spp": target_spp}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=200)
    ap.add_argument("--width", type=int, default=384)
    ap.add_argument("--height", type=int, default=224)
    ap.add_argument("--noisy", type=int, default=4, help="independent 1-sample inputs a scene (without --levels)")
    ap.add_argument("--levels", default="1,2,4,8,16,32,64,128",
                    help="sample counts of the inputs, read from one render a scene; empty for --noisy")
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default="unet_weights.npy")
    ap.add_argument("--eval-scenes", type=int, default=3)
    ap.add_argument("--eval-width", type=int, default=1280)
    ap.add_argument("--eval-height", type=int, default=720)
    ap.add_argument("--report", default="")
    ap.add_argument("--kinds", default="cornell,soup,city",
                    help="scene generators the training scenes cycle through (cornell, soup, city, bigsoup)")
    args = ap.parse_args()
    dev = Device(0)
    ins, tgs = [], []
    t0 = time.perf_counter()
    levels = tuple(int(x) for x in args.levels.split(",") if x)
    kinds = tuple(x for x in args.kinds.split(",") if x)
    for k, (kind, seed) in enumerate(scene_list(args.scenes, kinds=kinds)):
        i, t = render_set(dev, kind, seed, args.width, args.height, args.noisy, args.target_spp, levels=levels)
        ins.append(i)
        tgs.append(t)
        if k % 10 == 0:
            print(f"data: {k + 1}/{args.scenes} scenes ({time.perf_counter() - t0:.1f} s)", flush=True)
    net = UNet(dev, max(2 * args.width * args.height, args.eval_width * args.eval_height), init_params(0))
    t1 = time.perf_counter()
    losses = train(dev, net, np.stack(ins), np.stack(tgs), args.steps, args.lr, log=lambda s: print(s, flush=True))
    train_s = time.perf_counter() - t1
    np.save(args.out, net.params())
    held_out = scene_list(args.eval_scenes, first_seed=10_000, kinds=kinds)
    rep = evaluate(dev, net, held_out, args.eval_width, args.eval_height, args.target_spp,
                   log=lambda s: print(s, flush=True))
    rep.update({"train_scenes": args.scenes, "train_size": [args.width, args.height],
                "inputs_per_scene": list(levels) if levels else f"{args.noisy} x 1 sample",
                "steps": args.steps, "train_seconds": round(train_s, 1), "losses": losses})
    print("report:", json.dumps(rep), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    net.close()


if __name__ == "__main__":
    main()
