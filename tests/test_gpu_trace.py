"""Ray batches on the GPU against the oracle (parity tests proper; they run through the C ABI).

Closest hits must be the oracle's bit for bit (t, triangle, u, v, instance, section), for the walk over the
reference's own binary node buffer and for the library's 4-wide quantised walk alike: the answer is the lowest
(t, instance, triangle), which no tree changes. Any-hit modes must agree on hit or miss; the binary walk, being the
oracle's walk step for step, must also report the same triangle."""
import sys

import numpy as np
import pytest

import oracle
import threert_amd as T
from threert_amd import scene as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rays_at(prep: S.Prepared, n: int, seed: int, mode: int) -> np.ndarray:
    """Half the rays aimed at a unit's centre, half anywhere (units.rs:1396-1401)."""
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, T.RAY)
    rays["o"] = ((rng.random((n, 3)) - 0.5) * 60).astype(np.float32)
    d = rng.normal(size=(n, 3))
    if prep.units:
        tops = prep.nodes[prep.top_first:prep.top_first + 2 * prep.units - 1]
        leaves = tops[tops["count"] > 0]
        centres = (leaves["bmin"] + leaves["bmax"]) * 0.5
        pick = centres[rng.integers(0, len(centres), n)] + (rng.random((n, 3)) - 0.5)
        aim = pick - rays["o"]
        d[::2] = aim[::2]
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays["d"] = d.astype(np.float32)
    rays["tmax"] = T.INF
    rays["mode"] = mode
    return rays


def same(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return (a.view(np.uint32).reshape(len(a), 6) == b.view(np.uint32).reshape(len(b), 6)).all(axis=1)


def check_modes(dev, prep, rays_n=3000, seed=0, what=""):
    sc_gpu = T.Scene(dev, *prep.desc())
    sc_cpu = oracle.Scene(*prep.desc())
    for mode in (T.TRACE_CLOSEST, T.TRACE_SHADOW, T.TRACE_ANY):
        rays = rays_at(prep, rays_n, seed + mode, mode)
        want, _ = sc_cpu.trace_batch(rays.view(oracle.RAY), oracle.WALK_EXACT)
        for walk in (T.WALK_BINARY, T.WALK_WIDE):
            got = sc_gpu.trace(rays, walk)
            if mode == T.TRACE_CLOSEST or walk == T.WALK_BINARY:
                bad = ~same(got, want)
                assert not bad.any(), (what, mode, walk, np.nonzero(bad)[0][:5], got[bad][:3], want[bad][:3])
            else:
                assert ((got["tri"] != T.NO_HIT) == (want["tri"] != T.NO_HIT)).all(), (what, mode, walk)
                hit = got["tri"] != T.NO_HIT
                assert (got["t"][hit] < rays["tmax"][hit]).all()
    sc_gpu.close()


@pytest.mark.parametrize("seed", range(8))
def test_instanced_scenes_match_the_oracle(device, seed):
    say(f"instanced scene seed {seed}")
    prep = S.instanced_scene(seed, meshes=5, instances=40)
    check_modes(device, prep, 4000, seed, f"seed {seed}")


def test_one_mesh_one_leaf_and_flat_quads(device):
    """Units whose tree is one leaf (the top tree's root a leaf too), and two triangles of a quad in one plane."""
    say("one-leaf units")
    quad = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[1, 0, 0], [1, 1, 0], [0, 1, 0]]], np.float32)
    prep = S.prepare([S.Mesh(quad)], [S.Placement(0, S.translation([0.0, 0.0, 5.0]))])
    check_modes(device, prep, 2000, 11, "one quad")
    prep = S.prepare([S.Mesh(quad), S.Mesh(quad * 3)],
                     [S.Placement(0, S.translation([0.0, 0.0, 5.0])), S.Placement(1, np.eye(4, dtype=np.float32)),
                      S.Placement(0, S.trs([1, 2, 3], [0.5, 0.1, 0.2], [2, 1, 0.5]), cast_shadow=False)])
    check_modes(device, prep, 2000, 12, "quads")


def test_ties_go_to_the_lowest_instance_and_triangle(device):
    """Two units on top of each other (the same mesh twice at one place): every hit is a tie, won by record 0;
    rays through a quad's shared edge tie between its two triangles."""
    say("ties")
    quad = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[1, 0, 0], [1, 1, 0], [0, 1, 0]]], np.float32)
    w = S.translation([0.0, 0.0, 5.0])
    prep = S.prepare([S.Mesh(quad)], [S.Placement(0, w), S.Placement(0, w)])
    n = 2000
    rays = np.zeros(n, T.RAY)
    rng = np.random.default_rng(3)
    rays["o"] = np.c_[rng.random(n), rng.random(n), np.zeros(n)].astype(np.float32)
    rays["d"] = np.array([0, 0, 1], np.float32)
    rays["o"][: n // 4, 1] = 1.0 - rays["o"][: n // 4, 0]  # on the diagonal both triangles share
    rays["tmax"] = T.INF
    sc = T.Scene(device, *prep.desc())
    want, _ = oracle.Scene(*prep.desc()).trace_batch(rays.view(oracle.RAY), oracle.WALK_EXACT)
    brute = oracle.Scene(*prep.desc()).brute_batch(rays.view(oracle.RAY), oracle.WALK_EXACT)
    assert same(want, brute).all()
    hit = want["tri"] != T.NO_HIT
    assert hit.sum() > n // 2 and (want["inst"][hit] == 0).all()
    for walk in (T.WALK_BINARY, T.WALK_WIDE):
        got = sc.trace(rays, walk)
        assert same(got, want).all(), walk
    sc.close()


def test_edge_cases(device):
    """An empty batch; a scene without units; tmax shorter than the hit; tmax past 3e38; axis-aligned rays; rays
    from a vertex."""
    say("edge cases")
    prep = S.instanced_scene(3, meshes=3, instances=10)
    sc = T.Scene(device, *prep.desc())
    assert len(sc.trace(np.zeros(0, T.RAY))) == 0
    rays = rays_at(prep, 1000, 5, 0)
    full = sc.trace(rays)
    hit = full["tri"] != T.NO_HIT
    short = rays.copy()
    short["tmax"][hit] = full["t"][hit]
    got = sc.trace(short)
    assert (got["tri"][hit] == T.NO_HIT).all() and (got["t"][hit] == full["t"][hit]).all()
    far = rays.copy()
    far["tmax"] = np.float32(np.inf)
    assert same(sc.trace(far), full).all()
    axis = np.zeros(600, T.RAY)
    for k, a in enumerate(np.eye(3, dtype=np.float32)):
        for s, sign in enumerate((1, -1)):
            sl = slice(100 * (2 * k + s), 100 * (2 * k + s + 1))
            axis["o"][sl] = rays["o"][:100]
            axis["d"][sl] = a * sign
    axis["tmax"] = T.INF
    want, _ = oracle.Scene(*prep.desc()).trace_batch(axis.view(oracle.RAY))
    for walk in (T.WALK_BINARY, T.WALK_WIDE):
        assert same(sc.trace(axis, walk), want).all()
    vert = np.zeros(300, T.RAY)
    vert["o"] = prep.leaves[:300, 0, :3]
    vert["d"] = rays["d"][:300]
    vert["tmax"] = T.INF
    want, _ = oracle.Scene(*prep.desc()).trace_batch(vert.view(oracle.RAY))
    for walk in (T.WALK_BINARY, T.WALK_WIDE):
        assert same(sc.trace(vert, walk), want).all()
    sc.close()
    empty = S.Prepared(np.repeat(S.FILL, 1), np.zeros((1, 3, 4), np.float32), np.zeros(1, T.INSTANCE), 0, 0)
    sc = T.Scene(device, *empty.desc())
    got = sc.trace(rays[:10])
    assert (got["tri"] == T.NO_HIT).all() and (got["t"] == rays["tmax"][:10]).all()
    sc.close()


def test_precision_scene_subset_matches_the_oracle(device):
    """A 200k-triangle clustered scene, 400k incoherent rays (PCG32 seed 1): every closest hit is the oracle's."""
    say("precision subset")
    prep, lo, hi = S.precision_scene(200_000, seed=2)
    rays = S.incoherent_rays(400_000, lo, hi, seed=1)
    want, _ = oracle.Scene(*prep.desc()).trace_batch(rays.view(oracle.RAY), oracle.WALK_EXACT)
    sc = T.Scene(device, *prep.desc())
    for walk in (T.WALK_WIDE, T.WALK_BINARY):
        got = sc.trace(rays, walk)
        bad = ~same(got, want)
        assert not bad.any(), (walk, bad.sum(), got[bad][:3], want[bad][:3])
    assert (want["tri"] != T.NO_HIT).mean() > 0.03
    sc.close()


def test_full_size_walks_agree(device):
    """3M triangles, 16M incoherent rays: the wide walk's hits are the binary walk's, every one (two trees, one
    answer), and a 40k-ray sample is the oracle's."""
    say("full size: building 3M triangles")
    prep, lo, hi = S.precision_scene(3_000_000, seed=1)
    rays = S.incoherent_rays(16 * 1024 * 1024, lo, hi, seed=1)
    sc = T.Scene(device, *prep.desc())
    say("full size: tracing")
    wide = sc.trace(rays, T.WALK_WIDE)
    binary = sc.trace(rays, T.WALK_BINARY)
    bad = ~same(wide, binary)
    assert not bad.any(), (bad.sum(), wide[bad][:3], binary[bad][:3])
    again = sc.trace(rays, T.WALK_WIDE)
    assert same(again, wide).all()
    pick = np.random.default_rng(0).choice(len(rays), 40_000, replace=False)
    want, _ = oracle.Scene(*prep.desc()).trace_batch(rays[pick].view(oracle.RAY), oracle.WALK_EXACT)
    assert same(wide[pick], want).all()
    sc.close()


def fp64_distances(prep: S.Prepared, rays: np.ndarray, hits: np.ndarray) -> np.ndarray:
    """What trt_trace_check_fp64 states, in numpy float64: the ray placed in the hit's record (o - translation, then
    the normal rows unless the record is a pure shift) and the hit triangle's v0, e1, e2 tested by Moller-Trumbore,
    each sum in the kernel's order. 0 for no hit, -1 where the double-precision test misses."""
    out = np.zeros(len(rays))
    hit = hits["tri"] != T.NO_HIT
    idx = np.nonzero(hit)[0]
    rec = prep.instances[hits["inst"][idx]]
    o = rays["o"][idx].astype(np.float64) - rec["m"][:, :, 3].astype(np.float64)
    d = rays["d"][idx].astype(np.float64)
    rows = rec["n"][:, :, :3].astype(np.float64)
    shifted = (rec["n"][:, 1, 3].view(np.uint32) & 2) != 0

    def place(v):
        return (rows[:, 0] * v[:, 0:1] + rows[:, 1] * v[:, 1:2]) + rows[:, 2] * v[:, 2:3]

    o = np.where(shifted[:, None], o, place(o))
    d = np.where(shifted[:, None], d, place(d))
    lv = prep.leaves[hits["tri"][idx]].astype(np.float64)
    v0, e1, e2 = lv[:, 0, :3], lv[:, 1, :3], lv[:, 2, :3]

    def dot(a, b):
        return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]

    def cross(a, b):
        return np.stack([a[:, 1] * b[:, 2] - b[:, 1] * a[:, 2], a[:, 2] * b[:, 0] - b[:, 2] * a[:, 0],
                         a[:, 0] * b[:, 1] - b[:, 0] * a[:, 1]], axis=1)

    pv = cross(d, e2)
    det = dot(e1, pv)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / det
        tv = o - v0
        u = dot(tv, pv) * inv
        qv = cross(tv, e1)
        v = dot(d, qv) * inv
        t = dot(e2, qv) * inv
        ok = (det != 0) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t > 0)
    out[idx] = np.where(ok, t, -1.0)
    return out


def test_fp64_check_is_the_double_precision_test(device):
    """trt_trace_check_fp64 against its statement in numpy float64, bit for bit, on rotated and scaled instances;
    and the walk's FP32 distances within FP32 rounding of the FP64 ones."""
    prep = S.instanced_scene(3, meshes=5, instances=40)
    rays = rays_at(prep, 6000, 3, T.TRACE_CLOSEST)
    sc = T.Scene(device, *prep.desc())
    d_r, d_h, d_t = device.alloc(rays.nbytes), device.alloc(len(rays) * T.HIT.itemsize), device.alloc(8 * len(rays))
    d_r.upload(rays)
    sc.trace_device(d_r, len(rays), d_h, T.WALK_WIDE)
    sc.check_fp64(d_r, len(rays), d_h, d_t)
    device.sync()
    hits = d_h.download(T.HIT, len(rays))
    t64 = d_t.download(np.float64, len(rays))
    for b in (d_r, d_h, d_t):
        b.free()
    sc.close()
    want = fp64_distances(prep, rays, hits)
    hit = hits["tri"] != T.NO_HIT
    say(f"{hit.sum()} hits, {(t64 == -1).sum()} disputed")
    assert hit.sum() > 1000
    assert (t64[~hit] == 0).all()
    assert (t64.view(np.uint64) == want.view(np.uint64)).all(), np.nonzero(t64 != want)[0][:5]
    fine = hit & (t64 > 0)
    rel = np.abs(hits["t"][fine].astype(np.float64) - t64[fine]) / t64[fine]
    assert rel.max() < 1e-4, rel.max()
    assert (t64[hit] == -1).mean() < 0.01


def test_threads_share_a_device(device):
    """The ABI is thread-safe per device: four threads, each uploading its own scene (trees built on the device)
    and tracing batches on it while the others do, get the hits a lone thread gets."""
    import threading

    say("threads sharing the device")
    preps = [S.instanced_scene(20 + k, meshes=4, instances=25, tris_per_mesh=(100, 30000)) for k in range(4)]
    rays = [rays_at(p, 20000, 40 + k, T.TRACE_CLOSEST) for k, p in enumerate(preps)]
    want = []
    for p, r in zip(preps, rays):
        sc = T.Scene(device, *p.desc())
        want.append(sc.trace(r))
        sc.close()
    got, errors = [None] * 4, []

    def work(k):
        try:
            sc = T.Scene(device, *preps[k].desc())
            for _ in range(3):
                got[k] = sc.trace(rays[k])
            sc.close()
        except Exception as e:  # reported below
            errors.append((k, e))

    device.set_build(T.BUILD_DEVICE)
    try:
        ts = [threading.Thread(target=work, args=(k,)) for k in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=120)
    finally:
        device.set_build(T.BUILD_AUTO)
    assert not errors, errors
    for k in range(4):
        assert got[k] is not None and same(got[k], want[k]).all(), k
