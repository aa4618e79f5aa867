"""The reference's `serve` wire format (main.rs:345-398) over the continuous-batching server: one JSON job a line on
stdin, one JSON reply a line on stdout, and a `{"ready": true, "adapter": ..., "init_ms": ...}` line first.

    python -m threert_amd.wire [--budget-gb 200] [--denoiser] < jobs.jsonl

A job is `{id, scene, spp?, width?, height?, bounces?, seed?, camera?, denoise?, hdr?, out?}`:
- `scene` says what to render, in place of the reference's `game` bundle (this library has no script host):
  `{"kind": "cornell"}`, `{"kind": "soup", "triangles": n, "seed": s}` (clustered triangles under a sun and a sky), or
  `{"kind": "city", "seed": s, "meshes": m, "instances": k}` (the instanced city);
- `spp` (default 64, as tools/test-render.ts asks), `width` x `height` (1280 x 720), `bounces` (4), `seed` (1: the
  sampler's seed, Params.size.w);
- `camera`: `{"position": [x, y, z], "target": [x, y, z], "fov": degrees}` (main.rs:138-151);
- `denoise`: true sends the image through the learned denoiser (the server must be started with `--denoiser`);
- `hdr`: a file for the image's raw f32 RGB radiance (main.rs:575-578); `out`: an 8-bit PNG of it (a plain sRGB
  encoding at exposure 1, not the reference's post chain).

The reply is `{id, ok: true, stats}` or `{id, ok: false, error}`. Jobs render together, so replies come in the order
jobs finish, each with its id; a client that waits for each reply before it sends the next job (as
tools/test-render.ts does) sees them in its own order. Keys of the reference's jobs that ask for what this library
does not render (a game bundle, live frames, game state, overrides, presets, scene time, approximations,
reconstruction, scale, signal files) get `ok: false` with an error that starts "unsupported", which test-render.ts
counts as unsupported rather than failed.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import threading
import time
from collections import deque

from . import output as O
from . import render as R
from . import serve as V

UNSUPPORTED = ("game", "live", "state", "overrides", "preset", "time", "approx", "reconstruction", "scale", "signals")
LIMITS = {"spp": (1, 1 << 20), "width": (1, 16384), "height": (1, 16384), "bounces": (0, 32), "seed": (0, 2**32 - 1)}
DENOISE_PIXELS = 1920 * 1088  # the denoising step's buffers (the U-Net's and the server's)
SCENE_LIMITS = {"triangles": (1, 50_000_000), "meshes": (2, 256), "instances": (1, 1 << 20), "seed": (0, 2**31 - 1)}


class JobError(ValueError):
    """A job line the server refuses: the reply carries the message."""


def make_scene(kind: str, *args):
    """A scene key (as `parse_job` makes it) to (RenderScene, Camera, sky radiance)."""
    if kind == "cornell":
        rs, cam = R.cornell()
        return rs, cam, (0.0, 0.0, 0.0)
    if kind == "soup":
        return V.soup_scene(*args)
    if kind == "city":
        seed, meshes, instances = args
        return R.city(seed, meshes, instances)
    raise JobError(f"unknown scene kind {kind!r}")


def _int(v: dict, key: str, default: int, limits=LIMITS) -> int:
    x = v.get(key, default)
    if x is None:
        x = default
    if isinstance(x, bool) or not isinstance(x, (int, float)) or x != int(x):
        raise JobError(f"{key} must be a whole number, got {x!r}")
    lo, hi = limits[key]
    if not lo <= int(x) <= hi:
        raise JobError(f"{key} must be in [{lo}, {hi}], got {int(x)}")
    return int(x)


def _vec3(c: dict, key: str) -> tuple:
    a = c.get(key)
    if not (isinstance(a, list) and len(a) == 3 and all(isinstance(x, (int, float)) and not isinstance(x, bool)
                                                        and math.isfinite(x) for x in a)):
        raise JobError(f"camera: {key} must be [x, y, z]")
    return tuple(float(x) for x in a)


def scene_key(s) -> tuple:
    if not isinstance(s, dict) or "kind" not in s:
        raise JobError("job: a scene ({\"kind\": \"cornell\" | \"soup\" | \"city\", ...}) is required")
    kind = s["kind"]
    if kind == "cornell":
        return ("cornell",)
    if kind == "soup":
        return ("soup", _int(s, "triangles", 100_000, SCENE_LIMITS), _int(s, "seed", 1, SCENE_LIMITS))
    if kind == "city":
        return ("city", _int(s, "seed", 1, SCENE_LIMITS), _int(s, "meshes", 64, SCENE_LIMITS),
                _int(s, "instances", 4096, SCENE_LIMITS))
    raise JobError(f"unknown scene kind {kind!r}")


def parse_job(line: str, seq: int):
    """One job line to (id, Job, outputs {hdr, out}); raises JobError with the reply's message. `seq` is the job's
    number on this connection (the server's own key for it; the reply echoes the client's `id`)."""
    try:
        v = json.loads(line)
    except json.JSONDecodeError as e:
        raise JobError(f"job: not JSON ({e.msg} at column {e.colno})") from None
    if not isinstance(v, dict):
        raise JobError("job: a JSON object is required")
    jid = v.get("id")
    for k in UNSUPPORTED:
        if v.get(k) is not None and not (k == "game" and "scene" in v):
            raise JobError(f"unsupported: `{k}` jobs need the reference's {'script host' if k != 'live' else 'real-time renderer'}"
                           f"; this library renders `scene` jobs offline")
    key = scene_key(v.get("scene"))
    cam = None
    if v.get("camera") is not None:
        c = v["camera"]
        if not isinstance(c, dict):
            raise JobError("camera: an object {position, target, fov?} is required")
        pos, target = _vec3(c, "position"), _vec3(c, "target")
        if pos == target:
            raise JobError("camera: position == target")
        fov = c.get("fov", 40.0 if key[0] != "city" else 55.0)
        if isinstance(fov, bool) or not isinstance(fov, (int, float)) or not 0.0 < fov < 180.0:
            raise JobError(f"camera: fov must be in (0, 180) degrees, got {fov!r}")
        cam = R.Camera(pos, target, float(fov))
    denoise = v.get("denoise")
    if denoise is not None and not isinstance(denoise, bool):
        raise JobError("denoise must be true or false")
    outs = {}
    for k in ("hdr", "out"):
        if v.get(k) is not None:
            if not isinstance(v[k], str) or not v[k]:
                raise JobError(f"{k} must be a file path")
            outs[k] = v[k]
    job = V.Job(seq, key, _int(v, "width", 1280), _int(v, "height", 720), _int(v, "spp", 64), _int(v, "bounces", 4),
                seed=_int(v, "seed", 1), denoise=bool(denoise), keep=bool(outs), camera=cam)
    return jid, job, outs


class LineServer:
    """Reads job lines from `inp` on a thread of its own, feeds them to a `serve.Server` as they arrive and writes
    each reply to `out` when its job finishes."""

    def __init__(self, dev, inp, out, unet=None, budget_bytes: int = 200 << 30, launch_samples: float = 4e7,
                 workers: int = 8):
        self.dev, self.inp, self.out = dev, inp, out
        self.srv = V.Server(dev, make_scene, budget_bytes, launch_samples, keep_images=False, workers=workers,
                            unet=unet)
        self.lock = threading.Lock()
        self.cv = threading.Condition()
        self.queue: deque = deque()
        self.eof = False
        self.jobs: dict = {}  # seq -> (client id, Job, outputs)
        self.replies = 0
        self.adapter = "unknown"

    def reply(self, obj: dict) -> None:
        with self.lock:
            self.out.write(json.dumps(obj) + "\n")
            self.out.flush()
            self.replies += 1

    def _read(self) -> None:
        seq = 0
        for line in iter(self.inp.readline, ""):
            if not line.strip():
                continue
            try:
                jid, job, outs = parse_job(line, seq)
            except JobError as e:
                try:
                    jid = json.loads(line).get("id")
                except Exception:
                    jid = None
                self.reply({"id": jid, "ok": False, "error": str(e)})
                continue
            if job.denoise and self.srv.unet is None:
                self.reply({"id": jid, "ok": False, "error": "unsupported: denoise needs a server started with a denoiser"})
                continue
            if job.denoise and (job.width + (-job.width) % 16) * (job.height + (-job.height) % 16) > DENOISE_PIXELS:
                self.reply({"id": jid, "ok": False, "error": "denoise: images of at most 1920 x 1088 pixels"})
                continue
            self.jobs[seq] = (jid, job, outs)
            seq += 1
            with self.cv:
                self.queue.append((time.perf_counter(), job))
                self.cv.notify()
        with self.cv:
            self.eof = True
            self.cv.notify()

    def _arrivals(self, t: float) -> list:
        with self.cv:
            got = list(self.queue)
            self.queue.clear()
        for received, job in got:
            job.arrival = max(received - self.srv.t0, 0.0)
        return [j for _, j in got]

    def _closed(self) -> bool:
        with self.cv:
            return self.eof and not self.queue

    def _idle(self, t: float) -> None:
        with self.cv:
            if not self.queue and not self.eof:
                self.cv.wait(0.05)

    def _finished(self, r: V.Result) -> None:
        jid, job, outs = self.jobs.pop(r.id)
        if r.error is not None:
            self.reply({"id": jid, "ok": False, "error": r.error})
            return
        t = time.perf_counter()
        try:
            if "hdr" in outs:
                O.write_hdr(outs["hdr"], r.image)
            if "out" in outs:
                O.write_png(outs["out"], O.srgb8(r.image))
        except OSError as e:
            self.reply({"id": jid, "ok": False, "error": f"writing the image: {e}"})
            return
        write_ms = (time.perf_counter() - t) * 1000.0
        rs = self.srv.store.host[job.scene].result()[0]
        render_s = max(r.finished - r.admitted, 1e-9)
        stats = {"scene": "-".join(str(k) for k in job.scene), "out": outs.get("out"), "hdr": outs.get("hdr"),
                 "width": job.width, "height": job.height, "spp": job.spp, "bounces": job.bounces,
                 "triangles": int(rs.prep.leaves.shape[0]), "instances": int(rs.prep.units),
                 "emitters": int(rs.emitter_tris.shape[0]), "lights": int(rs.lights.shape[0]),
                 "denoise": "unet" if r.denoised else "none", "adapter": self.adapter,
                 "samples_per_sec": job.width * job.height * job.spp / render_s,
                 "timings_ms": {"queue_ms": round((r.admitted - r.arrival) * 1000.0, 1),
                                "render_ms": round(render_s * 1000.0, 1), "write_ms": round(write_ms, 1),
                                "total_ms": round((r.finished - r.arrival) * 1000.0 + write_ms, 1)}}
        self.reply({"id": jid, "ok": True, "stats": stats})

    def run(self, init_ms: float = 0.0, tick=None, every: float = 5.0) -> None:
        """The ready line, then jobs until the input ends and every job has its reply; `tick(active, waiting)` at
        most every `every` seconds."""
        info = self.dev.info()
        self.adapter = f"{info['name']} (gfx{info['gcn_arch']}, {info['compute_units']} CUs)"
        self.reply({"ready": True, "adapter": self.adapter, "init_ms": round(init_ms, 1)})
        reader = threading.Thread(target=self._read, daemon=True)
        reader.start()
        try:
            self.srv.serve(self._arrivals, self._closed, self._finished, self._idle, tick, every)
        finally:
            reader.join(timeout=1.0)

    def close(self) -> None:
        self.srv.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--denoiser", action="store_true", help="load the shipped U-Net weights for `denoise` jobs")
    ap.add_argument("--launch-samples", type=float, default=4e7)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    from . import Device
    from . import unet as U

    dev = Device(0)
    net = U.UNet(dev, DENOISE_PIXELS, U.shipped_params()) if a.denoiser else None
    ls = LineServer(dev, sys.stdin, sys.stdout, unet=net, budget_bytes=int(a.budget_gb * (1 << 30)),
                    launch_samples=a.launch_samples)
    try:
        ls.run((time.perf_counter() - t0) * 1000.0)
    finally:
        ls.close()
        if net is not None:
            net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
