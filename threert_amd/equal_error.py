"""Serving at equal image error: the learned denoiser with plain tracing as fallback, against plain FP32 tracing, on
the serving workload (512 Poisson jobs over 24 seeded scenes of 10k-4M triangles).

The error of an image is its relative RMSE (denoise_train.rel_rmse) against an FP64-accumulated render of 1024
samples. Each of the workload's scenes is calibrated once at 320 x 176 and 4 bounces: after 1, 2, 4, ... 256
samples of one render, the error of the image as it is and of the learned denoiser's output over it. A job asking
for n samples has a plain error e_plain(n) (log-log interpolation between the calibrated points); the denoising
server renders it at the fewest calibrated samples k < n whose denoised error is at most e_plain(n) and denoises it,
or, when no such k exists, renders its n samples plainly (high-sample tracing is the fallback, and its calibrated
error the yardstick). Both servers replay the same arrivals; each job's image error in the denoising run is at most the
plain run's, by calibration.
"""
from __future__ import annotations

import dataclasses
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import Device
from . import render as R
from . import serve
from .denoise_train import rel_rmse
from .unet import UNet, shipped_params

CAL_W, CAL_H = 320, 176
CHECKPOINTS = (1, 2, 4, 8, 16, 32, 64, 128, 256)


def calibrate(dev: Device, net: UNet, built, ref_spp: int = 1024, bounces: int = 4, w: int = CAL_W,
              h: int = CAL_H) -> dict:
    """built = soup_scene's (rs, cam, sky). The plain and the denoised error after each checkpoint's samples (sample
    indices from 200000, away from the reference's)."""
    rs, cam, sky = built
    p = rs.params(w, h, cam, seed=1, max_bounces=bounces, rr_depth=3, sky=sky)
    sc = R.upload(dev, rs)
    ref = R.reference_fp64(dev, sc, p, w, h, ref_spp)
    acc = R.Accum(dev, w, h)
    plain, den, done = {}, {}, 0
    for k in CHECKPOINTS:
        R.render(sc, p, 200_000 + done, k - done, acc)
        done = k
        c, a, n, _ = acc.resolve()
        plain[k] = rel_rmse(c, ref)
        den[k] = rel_rmse(serve.denoise_image(net, c, a, n), ref)
    acc.close()
    sc.close()
    return {"plain": plain, "denoised": den}


def plain_error(cal: dict, spp: int) -> float:
    """The calibrated plain error at `spp` samples: log-log interpolation, 1/sqrt(spp) past the last point."""
    ks = np.array(sorted(cal["plain"]), np.float64)
    es = np.array([cal["plain"][int(k)] for k in ks], np.float64)
    if spp >= ks[-1]:
        return float(es[-1] * np.sqrt(ks[-1] / spp))
    return float(np.exp(np.interp(np.log(max(spp, 1)), np.log(ks), np.log(es))))


def plan(jobs: list[serve.Job], cal: dict) -> list[serve.Job]:
    """The denoising server's jobs: the fewest calibrated samples whose denoised error is at most the job's plain
    error, denoised; or the job as asked."""
    out = []
    for j in jobs:
        c = cal[j.scene]
        target = plain_error(c, j.spp)
        k = next((k for k in CHECKPOINTS if k < j.spp and c["denoised"][k] <= target), None)
        out.append(dataclasses.replace(j, spp=k, denoise=True) if k else j)
    return out


def job_error(c: dict, j: serve.Job) -> float:
    return c["denoised"][j.spp] if j.denoise else plain_error(c, j.spp)


def run(dev: Device, n_jobs: int = 512, rate: float = 40.0, seed: int = 1, log=print,
        policies: tuple = ("equal", "age")) -> dict:
    jobs = serve.workload(n_jobs=n_jobs, rate=rate, seed=seed)
    keys = sorted({j.scene for j in jobs})
    net = UNet(dev, 1920 * 1088, params=shipped_params())
    t0 = time.perf_counter()
    cal = {}
    with ThreadPoolExecutor(max_workers=8) as ex:
        futs = {k: ex.submit(serve.soup_scene, *k) for k in keys}
        for i, k in enumerate(keys):
            cal[k] = calibrate(dev, net, futs[k].result())
            futs[k] = None
            c = cal[k]
            log(f"equal error: calibrated {i + 1}/{len(keys)} {k}: 1/16/256 spp plain {c['plain'][1]:.3f}/"
                f"{c['plain'][16]:.3f}/{c['plain'][256]:.3f}, denoised {c['denoised'][1]:.3f}/{c['denoised'][16]:.3f}/"
                f"{c['denoised'][256]:.3f} ({time.perf_counter() - t0:.0f} s)")
    dn_jobs = plan(jobs, cal)
    runs = {}
    for name, js, unet, policy in [(f"{m}_{pol}", js, un, pol) for pol in policies
                                   for m, js, un in (("plain", jobs, None), ("denoised", dn_jobs, net))]:
        srv = serve.Server(dev, serve.soup_scene, budget_bytes=200 << 30, unet=unet, policy=policy)
        w0 = time.perf_counter()
        res = srv.run(js, prefetch=False,
                      progress=lambda d, a, w, name=name: log(f"equal error: {name} {d} done, {a} rendering, {w} waiting"))
        wall = time.perf_counter() - w0
        srv.close()
        s = serve.summary(res, wall)
        err = [job_error(cal[j.scene], j) for j in js]
        s["mean_rel_rmse"] = round(float(np.mean(err)), 4)
        s["denoised_jobs"] = int(sum(j.denoise for j in js))
        runs[name] = s
        log(f"equal error: {name}: {s}")
    net.close()
    per_job_ok = all(job_error(cal[d.scene], d) <= plain_error(cal[j.scene], j.spp) for d, j in zip(dn_jobs, jobs))
    return {"runs": runs,
            "p99_ratio": {pol: round(runs[f"denoised_{pol}"]["p99_s"] / runs[f"plain_{pol}"]["p99_s"], 4)
                          for pol in policies},
            "each_job_error_at_most_plain": per_job_ok,
            "calibration": {f"{k[0]}:{k[1]}": {m: {str(n): round(e, 4) for n, e in v[m].items()}
                                               for m in ("plain", "denoised")} for k, v in cal.items()},
            "calibration_setup": f"{CAL_W}x{CAL_H}, 4 bounces, reference 1024 samples accumulated in float64"}
