"""A render server with continuous batching (the serving counterpart of the reference's `serve` loop,
main.rs:345-398, which renders one JSON job at a time, rebuilding and re-uploading its scene each time).

Jobs arrive at any time. Each is admitted when the HBM budget has room for its accumulator (and its scene, if that
scene is not resident yet); admitted jobs share every launch (`trt_render_batch`): each launch carries a slice of
samples of every active job, so a job that arrives while others render joins at the next launch instead of waiting
for them to finish. Scenes stay resident in HBM keyed by their description and are evicted least-recently-used
when the budget needs their room; accumulators live from admission to the job's last sample. A job's image is the
one a lone render of it gives, bit for bit: each pixel adds its samples in sample order whatever else shares the
launch.
"""
from __future__ import annotations

import ctypes as C
import heapq
import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import Device, _check, _ptr, lib
from . import render as R


class _Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", C.c_void_p), ("first_sample", C.c_uint32), ("spp", C.c_uint32),
                ("acc", C.c_void_p)]


def _sig():
    L = lib()
    L.trt_render_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    L.trt_render_batch.restype = C.c_int


_sig()


def render_batch(dev: Device, jobs: list[tuple]):
    """jobs: (Scene, params (156 u32), first_sample, spp, Accum) each; one launch, queued on the stream."""
    arr = (_Job * len(jobs))()
    keep = []
    for k, (sc, p, first, spp, acc) in enumerate(jobs):
        p = np.ascontiguousarray(p, np.uint32)
        keep.append(p)
        arr[k] = _Job(sc.h.value if isinstance(sc.h, C.c_void_p) else sc.h, p.ctypes.data, first, spp,
                      acc.handle.value if isinstance(acc.handle, C.c_void_p) else acc.handle)
    _check(lib().trt_render_batch(dev.h, arr, len(jobs)))


@dataclass
class Job:
    id: int
    scene: tuple          # the scene's description (a key for the store; `make_scene(*scene)` builds it)
    width: int
    height: int
    spp: int
    bounces: int
    arrival: float = 0.0  # seconds from the start of the run
    seed: int = 1
    denoise: bool = False  # the finished image goes through the server's learned denoiser
    keep: bool = False     # the finished image comes back to the host (also when the server keeps no images)
    camera: object = None  # a render.Camera in place of the scene's own


@dataclass
class Result:
    id: int
    arrival: float
    admitted: float
    finished: float
    samples: int
    image: np.ndarray | None = None
    denoised: bool = False
    error: str | None = None  # the job failed (its scene could not be built or uploaded); it rendered nothing

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Active:
    job: Job
    entry: "_Resident"
    acc: R.Accum
    params: np.ndarray
    done: int = 0
    admitted: float = 0.0


@dataclass
class _Resident:
    scene: object
    rs: R.RenderScene
    cam: R.Camera
    sky: tuple
    bytes: int
    users: int = 0


class SceneStore:
    """Scenes resident in HBM by key, least recently used first out. A worker pool builds each scene on the host and
    uploads it (the library builds its wide trees during the upload, seconds for millions of triangles), so the
    serving loop seldom waits for a build. A worker uploads only while the upload fits what `headroom()` says is
    left of the budget (the bytes it reserves count at once, `pending`); past that it leaves the upload to the
    scene's first admission, which makes room first."""

    def __init__(self, dev: Device, make_scene, workers: int = 4, headroom=None):
        self.dev = dev
        self.make_scene = make_scene
        self.headroom = headroom  # bytes of the budget not yet taken by admitted jobs (None: no limit)
        self.resident: "OrderedDict[tuple, _Resident]" = OrderedDict()
        self.host: dict = {}
        self.pending = 0  # bytes of scenes uploaded by the workers (or being uploaded) and not yet resident
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.lock = threading.Lock()

    def _build(self, key):
        """[rs, cam, sky, the upload or None, the bytes it holds in `pending`]"""
        rs, cam, sky = self.make_scene(*key)
        b = scene_bytes(rs)
        with self.lock:
            ok = self.headroom is None or self.pending + b <= self.headroom()
            if ok:
                self.pending += b
        if not ok:
            return [rs, cam, sky, None, 0]
        try:
            return [rs, cam, sky, R.upload(self.dev, rs), b]
        except Exception:
            with self.lock:
                self.pending -= b
            raise

    def prefetch(self, key):
        with self.lock:
            if key not in self.host:
                self.host[key] = self.pool.submit(self._build, key)

    def ready(self, key) -> bool:
        return key in self.resident or (key in self.host and self.host[key].done())

    def bytes_of(self, key) -> int:
        if key in self.resident:
            return 0
        rs = self.host[key].result()[0]
        return scene_bytes(rs)

    def pending_except(self, key) -> int:
        """`pending` without the bytes `key`'s own upload holds there."""
        with self.lock:
            own = 0
            f = self.host.get(key)
            if f is not None and f.done() and f.exception() is None and key not in self.resident:
                own = f.result()[4]
            return self.pending - own

    def acquire(self, key) -> _Resident:
        if key in self.resident:
            e = self.resident.pop(key)
            self.resident[key] = e
        else:
            built = self.host[key].result()
            rs, cam, sky, sc = built[:4]
            if sc is None:  # never uploaded, or dropped since: upload now
                sc = R.upload(self.dev, rs)
            with self.lock:
                self.pending -= built[4]
                built[3], built[4] = None, 0  # the upload now belongs to the resident entry
            e = _Resident(sc, rs, cam, sky, scene_bytes(rs))
            self.resident[key] = e
        e.users += 1
        return e

    def drop_uploads(self, need: int, keep) -> int:
        """Frees workers' uploads not yet resident (other than `keep`'s) until `need` bytes are free; their scenes
        are uploaded again at admission. Returns the bytes freed."""
        freed = 0
        with self.lock:
            for key, f in self.host.items():
                if freed >= need:
                    break
                if key == keep or key in self.resident or not f.done() or f.exception() is not None:
                    continue
                built = f.result()
                if built[3] is not None:
                    built[3].close()
                    freed += built[4]
                    self.pending -= built[4]
                    built[3], built[4] = None, 0
        return freed

    def release(self, e: _Resident):
        e.users -= 1

    def evict_for(self, need: int, used: int, budget: int) -> int:
        """Frees idle scenes, least recently used first, until `need` more bytes fit; returns bytes freed."""
        freed = 0
        for key in list(self.resident):
            if used - freed + need <= budget:
                break
            e = self.resident[key]
            if e.users == 0:
                e.scene.close()
                freed += e.bytes
                del self.resident[key]
        return freed

    def resident_bytes(self) -> int:
        return sum(e.bytes for e in self.resident.values())

    def close(self):
        self.pool.shutdown(wait=True)
        for e in self.resident.values():
            e.scene.close()
        self.resident.clear()
        for f in self.host.values():
            if f.done() and f.exception() is None and f.result()[3] is not None:
                f.result()[3].close()  # built and uploaded, never used
        self.host.clear()
        self.pending = 0


ACCUM_BYTES = 80  # records.wgsl Pixel: five vec4 per pixel


def denoise_image(net, color: np.ndarray, albedo: np.ndarray, normal: np.ndarray) -> np.ndarray:
    """The learned denoiser over an image of any size: edge-padded to multiples of 16 (the network's four poolings),
    denoised on the device, cropped back."""
    h, w = color.shape[:2]
    ph, pw = (-h) % 16, (-w) % 16
    pad = lambda x: np.pad(x, ((0, ph), (0, pw), (0, 0)), mode="edge")  # noqa: E731
    return net.denoise(pad(color), pad(albedo), pad(normal))[:h, :w]


def params_of(e: _Resident, j: Job) -> np.ndarray:
    """The job's Params: its size, seed and bounces over the scene's camera (or the job's) and sky (RR from depth 3)."""
    return e.rs.params(j.width, j.height, j.camera or e.cam, seed=j.seed, max_bounces=j.bounces, rr_depth=3,
                       sky=e.sky)


def scene_bytes(rs: R.RenderScene) -> int:
    p = rs.prep
    b = p.nodes.nbytes + p.leaves.nbytes + p.instances.nbytes
    b += sum(a.nbytes for a in rs.tables().values() if isinstance(a, np.ndarray))
    # the wide tree is at most twice the binary buffer; the leaves are copied once more in the wide trees' order
    return int(b + p.nodes.nbytes * 2 + p.leaves.nbytes)


class Server:
    """Continuous batching over one device. `run(jobs)` replays jobs at their arrival times (real time) and returns
    one Result per job."""

    def __init__(self, dev: Device, make_scene, budget_bytes: int = 200 << 30, launch_samples: float = 4e7,
                 keep_images: bool = False, max_jobs: int = 256, workers: int = 8, unet=None, policy: str = "equal"):
        self.dev = dev
        self.unet = unet  # a threert_amd.unet.UNet for jobs with `denoise` set
        if policy not in ("equal", "age"):
            raise ValueError(f"unknown policy {policy!r}")
        self.policy = policy
        self._dn = None  # device buffers of the denoising step
        self.budget = budget_bytes
        self.used = 0  # bytes of admitted jobs' accumulators and of resident scenes (kept across serve calls, as they are)
        self.store = SceneStore(dev, make_scene, workers, headroom=lambda: self.budget - self.used)
        self.launch_samples = launch_samples  # path samples one launch aims at (all jobs together)
        self.keep_images = keep_images
        self.max_jobs = max_jobs
        self.launches = 0
        self.samples = 0
        self.t0 = time.perf_counter()  # the clock arrivals are measured on: reset when a run starts

    def _denoise(self, a: _Active):
        """The finished job's image through the learned denoiser, all on the device: the accumulator resolved into
        edge-padded buffers (sides multiples of 16), the network, one wait; the image comes back to the host only
        when the server keeps images."""
        w, h = a.job.width, a.job.height
        pw, ph = w + (-w) % 16, h + (-h) % 16
        nbytes = 12 * pw * ph
        if self._dn is None or self._dn[0].nbytes < nbytes:
            if self._dn is not None:
                for b in self._dn:
                    b.free()
            self._dn = [self.dev.alloc(max(nbytes, 12 * 1920 * 1088)) for _ in range(4)]
        c, al, nm, out = self._dn
        a.acc.resolve_device(pw, ph, c, al, nm)
        self.unet.denoise_device(1, ph, pw, c, al, nm, out)
        self.dev.sync()
        if not (self.keep_images or a.job.keep):
            return None
        return out.download(np.float32, 3 * pw * ph).reshape(ph, pw, 3)[:h, :w]

    def _slices(self, active: list[_Active], now: float) -> list[int]:
        """Passes each active job takes this launch: its share of the launch's samples in whole passes over its
        pixels, at least one. `equal`: the same share for every job; `age`: shares in proportion to the time each job
        has waited since it arrived, so old jobs (large ones, mostly) catch up."""
        if self.policy == "age":
            ages = [max(now - a.job.arrival, 1e-3) for a in active]
            tot = sum(ages)
            shares = [self.launch_samples * g / tot for g in ages]
        else:
            shares = [self.launch_samples / max(len(active), 1)] * len(active)
        out = []
        for a, share in zip(active, shares):
            px = a.job.width * a.job.height
            k = int(max(1, share // px))
            out.append(min(k, a.job.spp - a.done))
        return out

    def run(self, jobs: list[Job], prefetch: bool = True, progress=None, every: float = 5.0) -> list[Result]:
        """`prefetch` starts every scene's host build at once (else at its first job's arrival); `progress(done,
        active, waiting)` is called at most every `every` seconds."""
        jobs = sorted(jobs, key=lambda j: j.arrival)
        if prefetch:
            for j in jobs:
                self.store.prefetch(j.scene)
        pending = deque(jobs)
        results: list[Result] = []

        def arrivals(t):
            out = []
            while pending and pending[0].arrival <= t:
                out.append(pending.popleft())
            return out

        def idle(t):
            time.sleep(max(0.0, min(pending[0].arrival - t, 0.002)) if pending else 0.001)

        def tick(active, waiting):
            if progress:
                progress(len(results), active, waiting)

        self.serve(arrivals, lambda: not pending, results.append, idle, tick, every)
        return sorted(results, key=lambda r: r.id)

    def serve(self, arrivals, closed, emit, idle, tick=None, every: float = 5.0) -> None:
        """The serving loop over a source of jobs: `arrivals(t)` returns the jobs that have arrived by `t` seconds
        (their `arrival` set), `closed()` says no more will come, `idle(t)` waits a little when nothing is active;
        `emit(result)` gets each job's Result as it finishes (or fails); `tick(active, waiting)` at most every
        `every` seconds. Returns when the source is closed and every job has finished."""
        waiting: deque[Job] = deque()
        active: list[_Active] = []
        self.t0 = time.perf_counter()
        now = lambda: time.perf_counter() - self.t0  # noqa: E731
        told = 0.0
        while True:
            t = now()
            for j in arrivals(t):
                self.store.prefetch(j.scene)
                waiting.append(j)
            if not (waiting or active) and closed():
                break
            if tick and t - told >= every:
                told = t
                tick(len(active), len(waiting))
            # admission in arrival order among jobs whose scene is built; the first that does not fit the budget
            # stops admission, so a large job is not overtaken for ever by small ones
            for j in list(waiting):
                if len(active) >= self.max_jobs:
                    break
                if not self.store.ready(j.scene):
                    continue
                try:
                    need = j.width * j.height * ACCUM_BYTES + self.store.bytes_of(j.scene)
                except Exception as e:  # the scene's build or upload failed: the job fails, the server goes on
                    waiting.remove(j)
                    emit(Result(j.id, j.arrival, now(), now(), 0, None, False, f"{type(e).__name__}: {e}"))
                    continue
                # the workers' uploads of other scenes count against the budget too
                pending = self.store.pending_except(j.scene)
                if self.used + pending + need > self.budget:
                    self.used -= self.store.evict_for(need + pending, self.used, self.budget)
                    if self.used + pending + need > self.budget:
                        pending -= self.store.drop_uploads(self.used + pending + need - self.budget, j.scene)
                    if self.used + pending + need > self.budget:
                        break
                waiting.remove(j)
                accum = j.width * j.height * ACCUM_BYTES
                try:  # an upload here (a scene dropped or never uploaded) may fail: the job fails
                    e = self.store.acquire(j.scene)
                except Exception as ex:
                    emit(Result(j.id, j.arrival, now(), now(), 0, None, False, f"{type(ex).__name__}: {ex}"))
                    continue
                self.used += need - accum  # the scene, resident now if it was not
                try:
                    acc = R.Accum(self.dev, j.width, j.height)
                except Exception as ex:
                    self.store.release(e)
                    emit(Result(j.id, j.arrival, now(), now(), 0, None, False, f"{type(ex).__name__}: {ex}"))
                    continue
                self.used += accum
                active.append(_Active(j, e, acc, params_of(e, j), 0, now()))
            if not active:
                idle(now())
                continue
            ks = self._slices(active, now())
            try:
                render_batch(self.dev, [(a.entry.scene, a.params, a.done, k, a.acc) for a, k in zip(active, ks)])
                self.dev.sync()
            except Exception as ex:  # the launch failed (or a walk stopped and reported): its jobs fail, serving goes on
                for a in active:
                    emit(Result(a.job.id, a.job.arrival, a.admitted, now(), 0, None, False,
                                f"{type(ex).__name__}: {ex}"))
                    a.acc.close()
                    self.used -= a.job.width * a.job.height * ACCUM_BYTES
                    self.store.release(a.entry)
                active = []
                continue
            self.launches += 1
            still = []
            for a, k in zip(active, ks):
                a.done += k
                self.samples += k * a.job.width * a.job.height
                if a.done >= a.job.spp:
                    keep = self.keep_images or a.job.keep
                    img = None
                    if a.job.denoise:
                        img = self._denoise(a)
                    elif keep:
                        img = a.acc.resolve()[0]
                    emit(Result(a.job.id, a.job.arrival, a.admitted, now(), a.job.spp * a.job.width * a.job.height,
                                img if keep else None, a.job.denoise))
                    a.acc.close()
                    self.used -= a.job.width * a.job.height * ACCUM_BYTES
                    self.store.release(a.entry)
                else:
                    still.append(a)
            active = still

    def close(self):
        self.store.close()
        for b in self._dn or []:
            b.free()
        self._dn = None


# ---- the serving workload ----------------------------------------------------------------------------------------
def workload(n_jobs: int = 512, rate: float = 40.0, seed: int = 1, n_scenes: int = 24, min_tris: int = 10_000,
             max_tris: int = 4_000_000, scale: float = 1.0) -> list[Job]:
    """Poisson arrivals at `rate` jobs/s; scenes seeded procedural, log-uniform in [min_tris, max_tris] triangles
    (`n_scenes` distinct ones, each job picks one); resolution from 256x256 to 1920x1080 and spp from 1 to 256,
    log-uniform; bounces 2-6. `scale` shrinks resolution and spp for quick runs."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(min_tris), np.log(max_tris), n_scenes)).astype(int)
    gaps = rng.exponential(1.0 / rate, n_jobs)
    t = np.cumsum(gaps) - gaps[0]
    jobs = []
    for k in range(n_jobs):
        s = int(rng.integers(0, n_scenes))
        f = rng.uniform(0, 1)  # 256x256 .. 1920x1080, log-uniform along the path between them
        w = int(round(np.exp(np.log(256) + f * (np.log(1920) - np.log(256))) * scale))
        h = int(round(np.exp(np.log(256) + f * (np.log(1080) - np.log(256))) * scale))
        spp = max(1, int(round(np.exp(rng.uniform(0, np.log(256))) * scale)))
        jobs.append(Job(k, (int(sizes[s]), 1000 + s), max(8, w), max(8, h), spp, int(rng.integers(2, 7)),
                        float(t[k]), seed=k + 1))
    return jobs


def summary(results: list[Result], wall: float) -> dict:
    """Latency percentiles (arrival to result), queueing (arrival to admission) and aggregate samples/s."""
    lat = [r.latency for r in results]
    queue = [r.admitted - r.arrival for r in results]
    samples = sum(r.samples for r in results)
    return {"jobs": len(results), "p50_s": round(percentile(lat, 50), 3), "p99_s": round(percentile(lat, 99), 3),
            "queue_p50_s": round(percentile(queue, 50), 3), "queue_p99_s": round(percentile(queue, 99), 3),
            "samples": int(samples), "wall_s": round(wall, 3), "samples_per_s": round(samples / wall, 1)}


def soup_scene(n_tris: int, seed: int):
    """A seeded procedural scene of `n_tris` clustered triangles (the precision scene's generator), 5 % emissive,
    Lambert/GGX materials in four sections, under a sun and a constant sky; a camera outside the clusters' box."""
    from . import scene as S

    tris, emissive = S.clustered_soup(n_tris, seed, extent=100.0, median_area=0.02)
    rng = np.random.default_rng(seed)
    sections = np.where(emissive > 0, 3, rng.integers(0, 3, n_tris)).astype(np.uint32)
    mats = [R.material((0.7, 0.7, 0.7), roughness=0.9), R.material((0.8, 0.5, 0.3), metallic=1.0, roughness=0.3),
            R.material((0.3, 0.6, 0.8), roughness=0.5, specular=0.5),
            R.material((0, 0, 0), (float(np.mean(emissive[emissive > 0])) if (emissive > 0).any() else 10.0,) * 3)]
    rs = R.build([R.RMesh(tris, sections=sections)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0, 1, 2, 3])],
                 mats, [R.directional_light((-0.3, -1.0, -0.2), (3.0, 3.0, 3.0))])
    cam = R.Camera((0.0, 40.0, -160.0), (0.0, 0.0, 0.0), 50.0)
    return rs, cam, (0.3, 0.4, 0.6)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else float("nan")
