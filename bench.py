"""Closest-hit throughput on one MI355X: the precision workload (3M clustered triangles, 16M incoherent rays).

python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one launch of the 4-wide quantised walk over the whole 16M-ray batch, rays and hits resident in HBM. The
timed region is bracketed by device synchronisation and measured with HIP events on the library's stream (the
stream the kernel runs on). One JSON line is printed on stdout.

Beside the GPU number: the roofline of the walk (algorithmic bytes per launch, counted by a counting launch of the
same walk over the same batch: 64 B per wide node fetched, 48 B per triangle tested, 64 B per unit entered), and
the CPU baseline: the oracle's restatement of the reference's walk (trace_units.wgsl over the reference's binary
node buffer), OpenMP over the host's cores, on a bounded sample of the same rays.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import threert_amd as T  # noqa: E402
from threert_amd import scene as S  # noqa: E402

METRIC = "Mrays/s closest-hit + samples/s at 720p; p99 job latency at equal image error"
HBM_PEAK_GBS = 8000.0
N_TRIS = 3_000_000
N_RAYS = 16 * 1024 * 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(prep, rays: np.ndarray, sample: int) -> dict:
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = min(threads, 16)
    sc = oracle.Scene(*prep.desc())
    sub = rays[:sample].view(oracle.RAY)
    sc.trace_batch(sub[:min(20000, sample)], oracle.WALK_REF, threads=threads)  # warm
    t0 = time.perf_counter()
    sc.trace_batch(sub, oracle.WALK_REF, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(sample / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"the first {sample} of the 16M rays, the oracle's C restatement of trace_units.wgsl over the "
                      f"reference's binary node buffer (-O3, OpenMP, {threads} threads), {dt:.1f} s"}


def bench_render(dev, frames: int) -> dict:
    """Samples/s at 720p: the instanced city (64 seeded meshes x 4096 instances), 1280 x 720, 1 spp a frame,
    3 bounces, sun + constant sky; `frames` frames (one sample each) timed with HIP events."""
    from threert_amd import render as R

    log("bench: building the instanced city")
    t0 = time.perf_counter()
    rs, cam, sky = R.city(seed=1)
    log(f"bench: city ready in {time.perf_counter() - t0:.1f} s ({len(rs.prep.leaves)} unique triangles)")
    w, h = 1280, 720
    p = rs.params(w, h, cam, seed=1, max_bounces=3, rr_depth=3, sky=sky)
    sc = R.upload(dev, rs)
    acc = R.Accum(dev, w, h)
    R.render(sc, p, 0, 1, acc)  # warm
    dev.sync()
    dev.timer_start()
    for f in range(frames):
        R.render(sc, p, 1 + f, 1, acc)
    ms = dev.timer_stop()
    dev.sync()
    ms_frame = ms / frames
    sc.close()
    acc.close()
    sps = w * h / (ms_frame * 1e-3)
    log(f"bench: city 720p {ms_frame:.2f} ms/frame, {sps / 1e6:.1f} Msamples/s")
    return {"metric": "samples/s at 720p", "value": round(sps, 1), "unit": "samples/s",
            "ms_per_frame": round(ms_frame, 3), "frames": frames,
            "workload": "instanced city: 64 seeded meshes (extruded boxes, subdivided spheres, 200-40k triangles) x "
                        "4096 instances, metal-rough GGX, 90,000-lux sun + constant sky, 1280x720, 1 spp a frame, "
                        "3 bounces, megakernel over the 4-wide quantised walk",
            "unique_triangles": int(len(rs.prep.leaves))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--render-frames", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})", file=sys.stderr)
        sys.exit(2)

    log("bench: building the 3M-triangle scene")
    t0 = time.perf_counter()
    prep, lo, hi = S.precision_scene(N_TRIS, seed=1)
    rays = S.incoherent_rays(N_RAYS, lo, hi, seed=1)
    log(f"bench: scene and rays ready in {time.perf_counter() - t0:.1f} s")

    dev = T.Device(0)
    info = dev.info()
    sc = T.Scene(dev, *prep.desc())
    st = sc.stats()
    d_rays = dev.alloc(rays.nbytes)
    d_rays.upload(rays)
    d_hits = dev.alloc(N_RAYS * T.HIT.itemsize)

    reads = sc.reads(d_rays, N_RAYS, d_hits, T.WALK_WIDE)
    log(f"bench: reads per ray {reads['nodes'] / N_RAYS:.1f} nodes, {reads['tris'] / N_RAYS:.1f} tris, "
        f"{reads['bytes'] / N_RAYS:.0f} B")

    for _ in range(args.warmup):
        sc.trace_device(d_rays, N_RAYS, d_hits, T.WALK_WIDE)
    dev.sync()
    log(f"bench: timing {args.steps} steps")
    w0 = time.perf_counter()
    dev.timer_start()
    for _ in range(args.steps):
        sc.trace_device(d_rays, N_RAYS, d_hits, T.WALK_WIDE)
    ms = dev.timer_stop()
    dev.sync()
    wall = time.perf_counter() - w0
    ms_step = ms / args.steps
    mrays = N_RAYS / (ms_step * 1e-3) / 1e6
    hits = d_hits.download(T.HIT, N_RAYS)
    hit_rate = float((hits["tri"] != T.NO_HIT).mean())
    log(f"bench: {ms_step:.3f} ms/step, {mrays:.1f} Mrays/s, wall {wall:.2f} s, hit rate {hit_rate:.3f}")

    achieved = reads["bytes"] / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "closest-hit ray batch: 3M-triangle clustered single-level scene (log-normal areas, "
                        "sigma 1.5), 16,777,216 incoherent rays (origins uniform in the scene box, directions "
                        "uniform on the sphere, PCG32 seed 1), 4-wide 8-bit-quantised BVH",
            "triangles": N_TRIS, "rays": N_RAYS, "hit_rate": round(hit_rate, 4),
            "wide_nodes": st["wide_nodes"], "wide_bytes": st["wide_bytes"], "leaf_bytes": st["leaf_bytes"],
            "device": info["name"], "compute_units": info["compute_units"],
            "reads_per_ray": {"nodes": round(reads["nodes"] / N_RAYS, 2), "tris": round(reads["tris"] / N_RAYS, 2),
                              "bytes": round(reads["bytes"] / N_RAYS, 1)},
            "wall_s_timed": round(wall, 3),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    sc.close()
    d_rays.free()
    d_hits.free()
    if not args.no_render:
        out["render_720p"] = bench_render(dev, args.render_frames)
    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(prep, rays, min(args.cpu_sample, N_RAYS))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
