"""Closest-hit throughput on one MI355X: the precision workload (3M clustered triangles, 16M incoherent rays).

python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one launch of the 4-wide quantised walk over the whole 16M-ray batch, rays and hits resident in HBM. The
timed region is bracketed by device synchronisation and measured with HIP events on the library's stream (the
stream the kernel runs on). One JSON line is printed on stdout.

Beside the GPU number: the roofline of the walk (algorithmic bytes per launch, counted by a counting launch of the
same walk over the same batch: 64 B per wide node fetched, 48 B per triangle tested, 64 B per unit entered), and
the CPU baseline: the oracle's restatement of the reference's walk (trace_units.wgsl over the reference's binary
node buffer), built on this host with -march=native, OpenMP over the host's threads and on one, on a bounded sample
of the same rays. The other workloads follow in the same line: the instanced city at 720p (samples/s, with the path
tracer's own roofline), the Cornell box at 512 x 512 x 16, the 3M-triangle scene at 1080p x 64, the learned
denoiser's inference and held-out error at 1280 x 720, and the serving workload at equal image error (plain tracing
against the denoiser with plain fallback: p50 / p99 latency, mean error).
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
SHADE_BYTES = 12 + 3 * 48 + 160 + 16 + 288  # a shaded hit's gathers: indices, vertices, record, row, material
ACCUM_BYTES = 2 * 48                        # a pixel's colour, albedo and normal sums read and written
N_TRIS = 3_000_000
N_RAYS = 16 * 1024 * 1024


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_trace_wide_precision3M_hbm_traffic.md")


def profiled_traffic():
    """HBM bytes per launch of the walk on this workload, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (calibrated against the copy kernel); None when the summary is missing."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prep, rays: np.ndarray, sample: int, single: int) -> dict:
    """The oracle's restatement of the reference walk (WALK_REF: trace_units.wgsl over the reference's binary node
    buffer), built here for this CPU (-march=native), over `sample` rays on the host's threads (OMP_NUM_THREADS: this
    process's share of the host, else every core) and over `single` rays on one thread."""
    import oracle

    nproc = os.cpu_count() or 1
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or nproc, nproc)
    with oracle.native():
        sc = oracle.Scene(*prep.desc())
        sub = rays[:sample].view(oracle.RAY)
        sc.trace_batch(sub[:min(20000, sample)], oracle.WALK_REF, threads=threads)  # warm
        t0 = time.perf_counter()
        sc.trace_batch(sub, oracle.WALK_REF, threads=threads)
        dt = time.perf_counter() - t0
        t0 = time.perf_counter()
        sc.trace_batch(rays[:single].view(oracle.RAY), oracle.WALK_REF, threads=1)
        dt1 = time.perf_counter() - t0
    model = cpu_model()
    log(f"bench: CPU baseline {sample / dt / 1e6:.3f} Mrays/s on {threads} threads, {single / dt1 / 1e6:.3f} on one "
        f"({model}, nproc {nproc})")
    return {"value": round(sample / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"the first {sample} of the 16M rays, the oracle's C restatement of trace_units.wgsl over the "
                      f"reference's binary node buffer (gcc -O3 -march=native -ffp-contract=off, OpenMP, {threads} "
                      f"threads: this process's share of the host's {nproc}), {dt:.1f} s",
            "single_core": {"value": round(single / dt1 / 1e6, 4), "unit": "Mrays/s", "rays": single,
                            "seconds": round(dt1, 2)},
            "cpu_model": model, "nproc": nproc}


def fp64_check(dev, sc, d_rays, d_hits, hits) -> dict:
    """Every hit of the batch intersected again in double precision on the GPU (trt_trace_check_fp64): the
    distribution of |t - t_FP64| / t_FP64 over the hits, and the disputed rays (FP32 hits the FP64 test misses)."""
    n = len(hits)
    d_t = dev.alloc(8 * n)
    sc.check_fp64(d_rays, n, d_hits, d_t)
    dev.sync()
    t64 = d_t.download(np.float64, n)
    d_t.free()
    hit = hits["tri"] != T.NO_HIT
    fine = hit & (t64 > 0)
    rel = np.abs(hits["t"][fine].astype(np.float64) - t64[fine]) / t64[fine]
    q = np.quantile(rel, [0.5, 0.99, 0.999]) if len(rel) else [0.0, 0.0, 0.0]
    return {"hits": int(hit.sum()), "disputed": int((t64[hit] == -1).sum()), "invalid": int((t64 == -2).sum()),
            "t_rel_err": {"p50": float(q[0]), "p99": float(q[1]), "p999": float(q[2]),
                          "max": float(rel.max()) if len(rel) else 0.0},
            "disputed_rays": np.nonzero(hit & (t64 == -1))[0]}


def watertight_walk(dev, sc, prep, rays, d_rays, d_hits, mt_hits, disputed, launches: int = 5) -> dict:
    """The same batch through the wide walk with the watertight triangle test (TRT_WALK_WATERTIGHT): rays/s, the FP64
    check of its hits, and on the rays whose Moller-Trumbore hit the FP64 test disputes, both tests' answers against
    the double-precision brute force (every triangle, ray placement and Moller-Trumbore in double): the same
    triangle, or one sharing a vertex with it (a ray within rounding of an edge)."""
    import oracle

    sc.trace_device(d_rays, N_RAYS, d_hits, T.WALK_WATERTIGHT)
    dev.sync()
    dev.timer_start()
    for _ in range(launches):
        sc.trace_device(d_rays, N_RAYS, d_hits, T.WALK_WATERTIGHT)
    ms = dev.timer_stop() / launches
    wt = d_hits.download(T.HIT, N_RAYS)
    check = fp64_check(dev, sc, d_rays, d_hits, wt)
    idx = disputed
    f64 = oracle.Scene(*prep.desc()).brute64_batch(rays[idx].view(oracle.RAY),
                                                   threads=int(os.environ.get("OMP_NUM_THREADS", "0")))
    leaves = prep.leaves.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)

    def corners(t):
        v0, e1, e2 = leaves[t]
        return np.stack([v0, v0 + e1, v0 + e2])

    def verdict(h):
        same = touch = 0
        for k, i in enumerate(idx):
            a, b = int(h["tri"][i]), int(f64["tri"][k])
            if b == 0xFFFFFFFF or a == 0xFFFFFFFF:
                same += a == b
                continue
            if a == b:
                same += 1
                continue
            ca, cb = corners(a), corners(b)
            touch += bool((np.abs(ca[:, None] - cb[None]).sum(-1) < 1e-6 * (1 + np.abs(cb).max())).any())
        return {"same_triangle": int(same), "a_neighbour": int(touch), "other": int(len(idx) - same - touch)}

    out = {"mrays_s": round(N_RAYS / (ms * 1e-3) / 1e6, 1), "ms_per_launch": round(ms, 3),
           "fp64_check": {k: v for k, v in check.items() if k != "disputed_rays"},
           "same_hits_as_moller_trumbore": round(float((wt.view(np.uint32).reshape(-1, 6) ==
                                                        mt_hits.view(np.uint32).reshape(-1, 6)).all(1).mean()), 6),
           "on_the_mt_disputed_rays": {"rays": int(len(idx)), "fp64_misses": int((f64["tri"] == 0xFFFFFFFF).sum()),
                                       "moller_trumbore": verdict(mt_hits), "watertight": verdict(wt)}}
    log(f"bench: watertight walk {out}")
    return out


def binary_walk(dev, sc, d_rays, d_hits, wide_hits, launches: int = 3) -> dict:
    """The same batch through the reference's FP32 binary walk (its own 32-byte nodes, 3 node reads an interior step,
    leaves of up to 16 triangles): rays/s, what it reads, and whether its hits are the wide walk's, bit for bit."""
    reads = sc.reads(d_rays, N_RAYS, d_hits, T.WALK_BINARY)
    sc.trace_device(d_rays, N_RAYS, d_hits, T.WALK_BINARY)
    dev.sync()
    dev.timer_start()
    for _ in range(launches):
        sc.trace_device(d_rays, N_RAYS, d_hits, T.WALK_BINARY)
    ms = dev.timer_stop() / launches
    got = d_hits.download(T.HIT, N_RAYS)
    same = bool((got.view(np.uint32) == wide_hits.view(np.uint32)).all())
    return {"mrays_s": round(N_RAYS / (ms * 1e-3) / 1e6, 1), "ms_per_launch": round(ms, 3),
            "reads_per_ray": {"nodes": round(reads["nodes"] / N_RAYS, 2), "tris": round(reads["tris"] / N_RAYS, 2),
                              "records": round(reads["units"] / N_RAYS, 2),
                              "bytes": round(reads["bytes"] / N_RAYS, 1)},
            "hits_equal_to_wide": same}


def hit_agreement(prep, rays: np.ndarray, hits: np.ndarray, sample: int) -> dict:
    """The GPU's hits against the oracle's exact walk (lowest (t, instance, triangle)) over the first `sample` rays:
    the fraction whose whole hit record is the same bits."""
    import oracle

    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1, 16)
    sc = oracle.Scene(*prep.desc())
    want, _ = sc.trace_batch(rays[:sample].view(oracle.RAY), oracle.WALK_EXACT, threads=threads)
    got = hits[:sample].view(np.uint32).reshape(sample, 6)
    same = (got == want.view(np.uint32).reshape(sample, 6)).all(axis=1)
    return {"rays": sample, "same": int(same.sum()), "fraction": float(same.mean())}


def under(dev, how, fn):
    """fn() with the device's render form set to `how` (then back to the default, TRT_RENDER_AUTO)."""
    dev.set_render(how)
    try:
        return fn()
    finally:
        dev.set_render(T.RENDER_AUTO)


def auto_form(leaves: int, paths: int) -> str:
    """The form TRT_RENDER_AUTO takes for a one-scene launch (include/threert.h trt_device_set_render)."""
    return "wavefront" if leaves >= 1024 or paths >= 1 << 22 else "megakernel"


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
    dev.render_ray_counts()  # zero the counters
    dev.timer_start()
    for f in range(frames):
        R.render(sc, p, 1 + f, 1, acc)
    ms = dev.timer_stop()
    dev.sync()
    closest, shadow = dev.render_ray_counts()
    ms_frame = ms / frames
    form = auto_form(len(rs.prep.leaves), w * h)

    def frames_ms():
        R.render(sc, p, 0, 1, acc)
        dev.sync()
        dev.timer_start()
        for f in range(frames):
            R.render(sc, p, 1 + f, 1, acc)
        return round(dev.timer_stop() / frames, 3)

    forms = {"wavefront": under(dev, T.RENDER_WAVEFRONT, frames_ms),
             "megakernel": under(dev, T.RENDER_MEGAKERNEL, frames_ms)}
    dev.render_ray_counts()
    # one counting launch of a frame (not timed): what its walks read and the hits it shaded
    rd = R.render_reads(sc, p, 1, 1, acc)
    sc.close()
    acc.close()
    sps = w * h / (ms_frame * 1e-3)
    walk_b = rd["nodes"] * 64 + rd["tris"] * 48 + rd["units"] * 64
    frame_b = walk_b + rd["shaded"] * SHADE_BYTES + w * h * ACCUM_BYTES
    achieved = frame_b / (ms_frame * 1e-3) / 1e9
    log(f"bench: city 720p {ms_frame:.2f} ms/frame ({form}; ms a frame by form {forms}), {sps / 1e6:.1f} Msamples/s, "
        f"{closest / (ms * 1e-3) / 1e6:.1f} M closest + {shadow / (ms * 1e-3) / 1e6:.1f} M shadow rays/s, "
        f"{frame_b / 1e6:.0f} MB a frame, {achieved:.0f} GB/s")
    kernel = "k_trace_dyn (sharded) + k_wf_shade" if form == "wavefront" else "k_render"
    return {"metric": "samples/s at 720p", "value": round(sps, 1), "unit": "samples/s",
            "ms_per_frame": round(ms_frame, 3), "frames": frames, "form": form, "ms_per_frame_by_form": forms,
            "rays": {"closest": closest, "shadow": shadow,
                     "closest_mrays_s": round(closest / (ms * 1e-3) / 1e6, 1),
                     "shadow_mrays_s": round(shadow / (ms * 1e-3) / 1e6, 1)},
            "roofline": {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                         "bytes_per_frame": int(frame_b),
                         "reads_per_frame": rd, "bytes": "64 B a wide node + 48 B a triangle + 64 B a unit the walks "
                         f"read, {SHADE_BYTES} B a shaded hit (indices 12, 3 vertices 144, record 160, row 16, "
                         f"material 288), {ACCUM_BYTES} B a pixel's accumulator read and written"},
            "workload": "instanced city: 64 seeded meshes (extruded boxes, subdivided spheres, 200-40k triangles) x "
                        "4096 instances, metal-rough GGX, 90,000-lux sun + constant sky, 1280x720, 1 spp a frame, "
                        "3 bounces, from above the skyline; the path tracer over the 4-wide quantised walk",
            "unique_triangles": rs.extras["unique_triangles"],
            "instanced_triangles": rs.extras["instanced_triangles"]}


def bench_cornell(dev, check: bool = False) -> dict:
    """Samples/s on the seeded Cornell box (36 triangles): 512 x 512, 16 spp, 4 bounces, RR from depth 3, clamp
    off, seed 1, one launch for the 16 samples, timed with HIP events. The image's error against a 1024-sample
    FP64-accumulated render (relative RMSE, sqrt(mean((x - ref)^2 / (ref^2 + 0.01)))); with `check`, its parity with
    the oracle's render of the same seed and samples (per-pixel relative RMSE, and the share of pixels whose 20
    accumulator words are the oracle's bits)."""
    from threert_amd import render as R

    rs, cam = R.cornell()
    w = h = 512
    spp = 16
    p = rs.params(w, h, cam, seed=1, max_bounces=4, rr_depth=3)
    sc = R.upload(dev, rs)
    acc = R.Accum(dev, w, h)
    R.render(sc, p, 0, spp, acc)  # warm
    dev.sync()
    acc.clear()
    reps = 5
    dev.timer_start()
    for _ in range(reps):
        R.render(sc, p, 0, spp, acc)
    ms = dev.timer_stop() / reps

    def image_ms():
        R.render(sc, p, 0, spp, acc)
        dev.sync()
        dev.timer_start()
        for _ in range(reps):
            R.render(sc, p, 0, spp, acc)
        return round(dev.timer_stop() / reps, 3)

    forms = {"wavefront": under(dev, T.RENDER_WAVEFRONT, image_ms),
             "megakernel": under(dev, T.RENDER_MEGAKERNEL, image_ms)}
    acc.clear()
    R.render(sc, p, 0, spp, acc)
    raw = acc.read()
    img = raw[..., 0:3] / np.maximum(raw[..., 3:4], 1.0)
    ref = R.reference_fp64(dev, sc, p, w, h, 1024)
    err = float(np.sqrt(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2))))
    sc.close()
    acc.close()
    sps = w * h * spp / (ms * 1e-3)
    form = auto_form(len(rs.prep.leaves), w * h * spp)
    log(f"bench: cornell 512x512x16 {ms:.2f} ms ({form}; by form {forms}), {sps / 1e6:.1f} Msamples/s, "
        f"rel RMSE vs FP64 1024 spp {err:.4f}")
    out = {"metric": "samples/s", "value": round(sps, 1), "unit": "samples/s", "ms_per_image": round(ms, 3),
           "form": form, "ms_per_image_by_form": forms,
           "rel_rmse_vs_fp64_1024spp": round(err, 5),
           "workload": "seeded Cornell box, 36 triangles, 512x512, 16 spp, 4 bounces, RR from depth 3, seed 1"}
    if check:
        import oracle

        t0 = time.perf_counter()
        pr = rs.prep
        want = oracle.RenderScene(pr.nodes, pr.leaves, pr.instances, pr.top_first, pr.units,
                                  **rs.tables()).render(p, 0, spp)
        cw = want[..., 0:3] / np.maximum(want[..., 3:4], 1.0)
        rel = float(np.sqrt(np.mean(((img - cw) / np.maximum(np.abs(cw), 1e-3)) ** 2)))
        same = float((raw.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean())
        out["oracle_parity"] = {"rel_rmse": rel, "pixels_bit_equal": round(same, 6),
                                "oracle_s": round(time.perf_counter() - t0, 2)}
        log(f"bench: cornell against the oracle {out['oracle_parity']}")
    return out


def bench_precision_render(dev) -> dict:
    """Samples/s on the 3M-triangle clustered scene (5 % emissive, sun + sky): 1920 x 1080, 64 spp in 8 launches
    of 8 samples, 4 bounces, RR from depth 3."""
    from threert_amd import render as R
    from threert_amd import serve

    log("bench: building the 3M-triangle render scene")
    rs, cam, sky = serve.soup_scene(N_TRIS, 1)
    w, h, spp, per = 1920, 1080, 64, 8
    p = rs.params(w, h, cam, seed=1, max_bounces=4, rr_depth=3, sky=sky)
    sc = R.upload(dev, rs)
    acc = R.Accum(dev, w, h)
    R.render(sc, p, 0, 1, acc)  # warm
    dev.sync()
    acc.clear()

    def image_ms():
        dev.timer_start()
        for s in range(0, spp, per):
            R.render(sc, p, s, per, acc)
        t = dev.timer_stop()
        dev.sync()
        return t

    ms = image_ms()
    forms = {"wavefront": round(under(dev, T.RENDER_WAVEFRONT, image_ms), 2),
             "megakernel": round(under(dev, T.RENDER_MEGAKERNEL, image_ms), 2)}
    sc.close()
    acc.close()
    sps = w * h * spp / (ms * 1e-3)
    form = auto_form(len(rs.prep.leaves), w * h * per)
    log(f"bench: precision 1080p x 64 spp {ms:.1f} ms ({form}; by form {forms}), {sps / 1e6:.1f} Msamples/s")
    return {"metric": "samples/s", "value": round(sps, 1), "unit": "samples/s", "ms_per_image": round(ms, 2),
            "form": form, "ms_per_image_by_form": forms,
            "workload": "3M clustered triangles (5 % emissive, 150k emitter triangles), sun + constant sky, "
                        "1920x1080, 64 spp (8 launches of 8), 4 bounces, RR from depth 3"}


def tree_build(dev, prep, sc) -> dict:
    """The upload's wide-tree build of the 3M-triangle scene: the library's own mesh tree on the device (the
    default) against one host thread, the same bytes; then collapse and quantisation on the host in both."""
    first = sc.stats()
    dev.set_build(T.BUILD_HOST)
    try:
        host = T.Scene(dev, *prep.desc())
    finally:
        dev.set_build(T.BUILD_AUTO)
    same = all(bool((host.wide_read(k) == sc.wide_read(k)).all()) for k in (0, 1, 2))
    hs = host.stats()
    host.close()
    again = T.Scene(dev, *prep.desc())
    ds = again.stats()
    again.close()
    return {"host_build_ms": round(hs["build_ms"], 1), "device_build_ms": round(ds["build_ms"], 1),
            "device_trees_ms": round(ds["device_build_ms"], 1), "first_device_build_ms": round(first["build_ms"], 1),
            "same_bytes": same, "speedup": round(hs["build_ms"] / ds["build_ms"], 2)}


def bench_copy(dev) -> dict:
    """The library's plain copy kernel over 1 GiB: the HBM bandwidth a streaming kernel reaches, and the
    yardstick for the HBM byte counters."""
    n = 1 << 30
    a, b = dev.alloc(n), dev.alloc(n)
    b.copy_from(a, n)  # warm
    dev.timer_start()
    for _ in range(4):
        b.copy_from(a, n)
    ms = dev.timer_stop() / 4
    a.free()
    b.free()
    gbs = 2 * n / (ms * 1e-3) / 1e9
    log(f"bench: copy 1 GiB {ms:.3f} ms, {gbs:.0f} GB/s read+write")
    return {"bytes": n, "ms": round(ms, 4), "gbs_read_plus_write": round(gbs, 1)}


def bench_serve(dev, n_jobs: int, rate: float) -> dict:
    """The serving workload: `n_jobs` jobs, Poisson arrivals at `rate` per second, seeded clustered scenes of 10k
    to 4M triangles (log-uniform, 24 of them), 256x256 to 1920x1080, 1 to 256 spp, 2 to 6 bounces; continuous
    batching under a 200 GB HBM budget; scenes built on host workers when their first job arrives."""
    from threert_amd import serve

    jobs = serve.workload(n_jobs=n_jobs, rate=rate, seed=1)
    srv = serve.Server(dev, serve.soup_scene, budget_bytes=200 << 30)
    t0 = time.perf_counter()
    res = srv.run(jobs, prefetch=False,
                  progress=lambda d, a, w: log(f"bench: serving {d} done, {a} rendering, {w} waiting"))
    wall = time.perf_counter() - t0
    srv.close()
    s = serve.summary(res, wall)
    s["launches"] = srv.launches
    log(f"bench: serving {s}")
    return s


def bench_serve_wire(dev, n_jobs: int, rate: float) -> dict:
    """The serving workload through the `serve` wire format (threert_amd/wire.py): a client thread writes each job as
    a JSON line at its arrival time into the server's stdin (a pipe), and times each reply line from the moment it
    sent the job. Latency here is the client's: parsing, the reply and the pipe included."""
    import os
    import threading

    from threert_amd import serve, wire

    jobs = serve.workload(n_jobs=n_jobs, rate=rate, seed=1)
    r_fd, w_fd = os.pipe()
    inp, feed = os.fdopen(r_fd, "r"), os.fdopen(w_fd, "w")
    sent, got = {}, []

    class Replies:
        def write(self, text):
            t = time.perf_counter()
            got.extend((t, line) for line in text.splitlines() if line)

        def flush(self):
            pass

    ls = wire.LineServer(dev, inp, Replies(), budget_bytes=200 << 30)
    t0 = time.perf_counter()

    def client():
        for j in jobs:
            wait = t0 + j.arrival - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            line = json.dumps({"id": j.id, "scene": {"kind": "soup", "triangles": j.scene[0], "seed": j.scene[1]},
                               "width": j.width, "height": j.height, "spp": j.spp, "bounces": j.bounces,
                               "seed": j.seed})
            sent[j.id] = time.perf_counter()
            feed.write(line + "\n")
            feed.flush()
        feed.close()

    th = threading.Thread(target=client)
    th.start()
    ls.run(0.0, tick=lambda a, w: log(f"bench: wire {ls.replies - 1} replies, {a} rendering, {w} waiting"))
    th.join()
    wall = time.perf_counter() - t0
    ls.close()
    inp.close()
    replies = [(t, json.loads(line)) for t, line in got]
    done = [(t, r) for t, r in replies if r.get("ok")]
    lat = [t - sent[r["id"]] for t, r in done]
    samples = sum(r["stats"]["width"] * r["stats"]["height"] * r["stats"]["spp"] for _, r in done)
    s = {"jobs": len(done), "failed": len([1 for _, r in replies if r.get("ok") is False]),
         "p50_s": round(serve.percentile(lat, 50), 3), "p99_s": round(serve.percentile(lat, 99), 3),
         "samples": int(samples), "wall_s": round(wall, 3), "samples_per_s": round(samples / wall, 1),
         "ready_line": replies[0][1].get("ready", False) if replies else False}
    log(f"bench: serving over the wire {s}")
    return s


def bench_denoise(dev, w: int = 1280, h: int = 720, reps: int = 20) -> dict:
    """The learned denoiser's inference on one 1280 x 720 frame (the shipped weights; color, albedo and normal
    resident in HBM; HIP events on the library's stream over `reps` frames). Flops: the convolutions'
    sum over layers of 2 x 9 x cin x cout x pixels at the layer's level, against the dense BF16 MFMA peak."""
    from threert_amd import unet as U

    net = U.UNet(dev, w * h, U.shipped_params())
    rng = np.random.default_rng(1)
    img = {k: rng.random((h, w, 3)).astype(np.float32) for k in ("c", "a", "n")}
    bufs = {k: dev.alloc(w * h * 12) for k in ("c", "a", "n", "o")}
    for k in ("c", "a", "n"):
        bufs[k].upload(img[k])
    for _ in range(3):
        net.denoise_device(1, h, w, bufs["c"], bufs["a"], bufs["n"], bufs["o"])
    dev.sync()
    dev.timer_start()
    for _ in range(reps):
        net.denoise_device(1, h, w, bufs["c"], bufs["a"], bufs["n"], bufs["o"])
    ms = dev.timer_stop() / reps
    for b in bufs.values():
        b.free()
    net.close()
    flops = sum(2 * 9 * (c0 + c1) * co * (w >> lv) * (h >> lv) for c0, c1, co, lv in U.ARCH)
    tf = flops / (ms * 1e-3) / 1e12
    log(f"bench: denoise {w}x{h} {ms:.3f} ms, {flops / 1e9:.1f} GFLOP, {tf:.1f} TFLOP/s")
    return {"ms_per_frame": round(ms, 4), "gflop_per_frame": round(flops / 1e9, 2), "tflops": round(tf, 1),
            "frac_of_bf16_peak": round(tf / 2500.0, 4)}


def bench_denoise_eval(dev) -> dict:
    """The learned denoiser at 1280 x 720 (the shipped weights): inference ms a frame and flops against the BF16 MFMA
    peak, and the relative RMSE against 1024-sample FP64-accumulated targets on three held-out scenes (a Cornell
    variant, a clustered soup, a small city; seeds 10000-10002, not trained on), beside plain tracing."""
    from threert_amd import denoise_train as DT
    from threert_amd import unet as U

    inf = bench_denoise(dev)
    net = U.UNet(dev, 1280 * 720, U.shipped_params())
    ev = DT.evaluate(dev, net, DT.scene_list(3, first_seed=10_000), 1280, 720, 1024, log=log, timing_reps=5)
    net.close()
    m = {k: round(v, 4) for k, v in ev["mean"].items()}
    log(f"bench: denoiser held-out mean rel RMSE {m}")
    return {"ms_per_frame": inf["ms_per_frame"], "gflop_per_frame": inf["gflop_per_frame"], "tflops": inf["tflops"],
            "roofline": {"kernel": "k_conv", "bound": "mfma", "achieved": inf["tflops"], "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": inf["frac_of_bf16_peak"], "traffic": None},
            "held_out_rel_rmse": m, "held_out_scenes": [r["scene"] for r in ev["scenes"]],
            "workload": "5-level U-Net (32-128 channels, 3x3 convolutions, BF16 MFMA, FP32 accumulation) on a "
                        "1280x720 frame; error of its output over a 1-sample render (and over 4, 16, 64 samples) "
                        "against a 1024-sample FP64-accumulated render"}


def bench_serve_equal(dev, n_jobs: int, rate: float) -> dict:
    """The serving workload at equal image error (threert_amd/equal_error.py): plain FP32 tracing against the learned
    denoiser with plain tracing as the fallback, equal shares, the same arrivals."""
    from threert_amd import equal_error

    t0 = time.perf_counter()
    r = equal_error.run(dev, n_jobs=n_jobs, rate=rate, log=log, policies=("equal",))
    keep = ("p50_s", "p99_s", "queue_p99_s", "samples_per_s", "wall_s", "mean_rel_rmse", "denoised_jobs")
    out = {"plain": {k: r["runs"]["plain_equal"][k] for k in keep},
           "denoised": {k: r["runs"]["denoised_equal"][k] for k in keep},
           "p99_ratio": r["p99_ratio"]["equal"], "p50_ratio": round(r["runs"]["denoised_equal"]["p50_s"] /
                                                                    r["runs"]["plain_equal"]["p50_s"], 4),
           "each_job_error_at_most_plain": r["each_job_error_at_most_plain"],
           "calibration": r["calibration_setup"], "seconds": round(time.perf_counter() - t0, 1),
           "workload": f"{n_jobs} render jobs, Poisson arrivals at {rate}/s, 24 seeded scenes of 10k-4M triangles, "
                       "256x256-1920x1080, 1-256 spp, 2-6 bounces, 200 GB budget, continuous batching"}
    log(f"bench: serving at equal error {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=N_RAYS)
    ap.add_argument("--agree-sample", type=int, default=N_RAYS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--render-frames", type=int, default=10)
    ap.add_argument("--no-denoise", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--cpu-single", type=int, default=1 << 20)
    ap.add_argument("--build-only", action="store_true",
                    help="measure only the upload's wide-tree build of the 3M-triangle scene (device and host)")
    ap.add_argument("--denoise-only", action="store_true",
                    help="measure only the learned denoiser's inference on a 1280x720 frame")
    ap.add_argument("--serve", action="store_true", help="measure the serving workload instead (p50/p99 latency)")
    ap.add_argument("--serve-wire", action="store_true",
                    help="the serving workload through the JSON-lines wire format, latency timed by the client")
    ap.add_argument("--serve-jobs", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-equal", action="store_true",
                    help="the serving workload twice at equal image error: plain tracing, then the learned denoiser "
                         "with plain tracing as fallback (p99 latency of each)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.serve_equal:
        from threert_amd import equal_error

        dev = T.Device(0)
        r = equal_error.run(dev, n_jobs=args.serve_jobs, rate=args.serve_rate, log=log)
        best = min(r["p99_ratio"], key=lambda pol: r["runs"][f"denoised_{pol}"]["p99_s"])
        d = r["runs"][f"denoised_{best}"]
        print(json.dumps({"metric": "p99 job latency at equal image error", "value": d["p99_s"], "unit": "s",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
                          "policy": best, "vs_plain_p99": r["p99_ratio"][best],
                          "config": {"workload": f"{args.serve_jobs} render jobs, Poisson arrivals at "
                                                 f"{args.serve_rate}/s, seeded scenes of 10k-4M triangles, "
                                                 "256x256-1920x1080, 1-256 spp, 2-6 bounces, 200 GB budget; "
                                                 "each job at the fewest samples whose calibrated denoised error is "
                                                 "at most its plain error, denoised on the device, plain tracing "
                                                 "otherwise"},
                          "equal_error": r}), flush=True)
        return
    if args.denoise_only:
        dev = T.Device(0)
        d = bench_denoise(dev)
        print(json.dumps({"metric": "denoiser inference, 1280x720", "value": d["ms_per_frame"], "unit": "ms",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "bf16", "data": "synthetic",
                          "config": {"workload": "19-layer U-Net, 3x3 convolutions on BF16 MFMA, one 1280x720 frame"},
                          "roofline": {"bound": "mfma", "achieved": d["tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                                       "frac": d["frac_of_bf16_peak"], "traffic": None},
                          "denoise": d}), flush=True)
        return
    if args.serve_wire:
        dev = T.Device(0)
        s = bench_serve_wire(dev, args.serve_jobs, args.serve_rate)
        print(json.dumps({"metric": "p99 job latency (serving over the wire format)", "value": s["p99_s"],
                          "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": f"{args.serve_jobs} render jobs as JSON lines, Poisson arrivals at "
                                                 f"{args.serve_rate}/s, seeded scenes of 10k-4M triangles, "
                                                 "256x256-1920x1080, 1-256 spp, 2-6 bounces, 200 GB budget"},
                          "serving": s}), flush=True)
        return
    if args.serve:
        dev = T.Device(0)
        s = bench_serve(dev, args.serve_jobs, args.serve_rate)
        print(json.dumps({"metric": "p99 job latency (serving)", "value": s["p99_s"], "unit": "s", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": f"{args.serve_jobs} render jobs, Poisson arrivals at "
                                                 f"{args.serve_rate}/s, seeded scenes of 10k-4M triangles, "
                                                 "256x256-1920x1080, 1-256 spp, 2-6 bounces, 200 GB budget"},
                          "serving": s}), flush=True)
        return

    log("bench: building the 3M-triangle scene")
    t0 = time.perf_counter()
    prep, lo, hi = S.precision_scene(N_TRIS, seed=1)
    rays = S.incoherent_rays(N_RAYS, lo, hi, seed=1)
    log(f"bench: scene and rays ready in {time.perf_counter() - t0:.1f} s")

    dev = T.Device(0)
    info = dev.info()
    sc = T.Scene(dev, *prep.desc())
    st = sc.stats()
    build = tree_build(dev, prep, sc)
    log(f"bench: tree build {build}")
    if args.build_only:
        print(json.dumps({"metric": "wide-tree build of a 3M-triangle mesh", "value": build["device_build_ms"],
                          "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": "trt_scene_upload's wide-tree build: 3M-triangle clustered mesh",
                                     "tree_build": build}}), flush=True)
        return
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
    each = []  # the same launches one at a time, for the median of at least 5 timed runs
    for _ in range(max(5, min(args.steps, 20))):
        dev.timer_start()
        sc.trace_device(d_rays, N_RAYS, d_hits, T.WALK_WIDE)
        each.append(dev.timer_stop())
    ms_median = float(np.median(each))
    mrays = N_RAYS / (ms_step * 1e-3) / 1e6
    hits = d_hits.download(T.HIT, N_RAYS)
    hit_rate = float((hits["tri"] != T.NO_HIT).mean())
    log(f"bench: {ms_step:.3f} ms/step, {mrays:.1f} Mrays/s, wall {wall:.2f} s, hit rate {hit_rate:.3f}")
    # the boundary's host-buffer entry (trt_trace_rays): the caller's rays and hits cross PCIe each call
    sc.trace(rays[:1024])
    t0 = time.perf_counter()
    host_hits = sc.trace(rays)
    host_ms = (time.perf_counter() - t0) * 1e3
    host_buffers = {"ms": round(host_ms, 2), "mrays_s": round(N_RAYS / (host_ms * 1e-3) / 1e6, 1),
                    "hits_equal": bool((host_hits.view(np.uint32) == hits.view(np.uint32)).all())}
    log(f"bench: host buffers (PCIe both ways) {host_buffers}")
    fp64 = fp64_check(dev, sc, d_rays, d_hits, hits)
    disputed = fp64.pop("disputed_rays")
    log(f"bench: FP64 check {fp64}")
    watertight = watertight_walk(dev, sc, prep, rays, d_rays, d_hits, hits, disputed)
    binary = binary_walk(dev, sc, d_rays, d_hits, hits)
    log(f"bench: binary walk {binary}")

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
            "ms_median_single_launch": round(ms_median, 4),
            "mrays_s_median": round(N_RAYS / (ms_median * 1e-3) / 1e6, 2),
            "fp64_check": fp64,
            "watertight_walk": watertight,
            "binary_walk": binary,
            "tree_build": build,
            "host_buffers_pcie": host_buffers,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic()},
        "cpu_baseline": None,
    }
    sc.close()
    d_rays.free()
    d_hits.free()
    out["config"]["copy_1gib"] = bench_copy(dev)
    if not args.no_render:
        out["render_720p"] = bench_render(dev, args.render_frames)
        out["render_cornell_512"] = bench_cornell(dev, check=not args.no_cpu)
        out["render_1080p_precision"] = bench_precision_render(dev)
    if not args.no_denoise:
        out["denoise_720p"] = bench_denoise_eval(dev)
    if not args.no_serve:
        out["serving_equal_error"] = bench_serve_equal(dev, args.serve_jobs, args.serve_rate)
    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(prep, rays, min(args.cpu_sample, N_RAYS), min(args.cpu_single, N_RAYS))
        out["config"]["hit_id_agreement"] = hit_agreement(prep, rays, hits, min(args.agree_sample, N_RAYS))
        log(f"bench: hit-ID agreement {out['config']['hit_id_agreement']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
