"""The watertight triangle test (TRT_WALK_WATERTIGHT; Woop, Benthin and Wald, JCGT 2013) beside the reference's
Moller-Trumbore: the oracle's walk equals its brute force with the same test; rays aimed at the shared edges and
vertices of a grid whose vertices are exact in float never slip between two triangles (Moller-Trumbore's misses on
the same rays are counted beside it); the test's answer agrees with a double-precision brute force where it hits; and
on the GPU the wide walk's watertight hits are the oracle's, bit for bit."""
import numpy as np
import pytest

import oracle
import threert_amd as T
from threert_amd import scene as S

N = 64  # the grid's squares a side
NO_TRI = 0xFFFFFFFF


def grid_scene(tilt: bool):
    """N x N unit squares at integer x, z in one mesh (two triangles a square, split along alternating diagonals), so
    every stored edge vector is exact and neighbours share their edges bit for bit; `tilt` places the mesh by an
    instance rotated off the axes (the ray is then placed in its space by the record's rows)."""
    i, j = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    i, j = i.ravel(), j.ravel()
    p00 = np.stack([i, np.zeros_like(i), j], 1).astype(np.float32)
    p10, p11, p01 = p00 + [1, 0, 0], p00 + [1, 0, 1], p00 + [0, 0, 1]
    flip = ((i + j) % 2 == 1)[:, None, None]
    a = np.where(flip, np.stack([p00, p10, p11], 1), np.stack([p00, p10, p01], 1))
    b = np.where(flip, np.stack([p00, p11, p01], 1), np.stack([p10, p11, p01], 1))
    tris = np.concatenate([a, b]).astype(np.float32)
    world = S.trs((-N / 2, 0.0, -N / 2), (0.3, 0.2, 0.1), (1.0, 1.0, 1.0)) if tilt else S.translation((-N / 2, 0, -N / 2))
    prep = S.prepare([S.Mesh(tris)], [S.Placement(0, world)])
    return prep, np.asarray(world, np.float64)


def edge_rays(world, n: int, seed: int) -> np.ndarray:
    """Rays from random points well above the grid to grid vertices, points on square edges and points on the
    squares' diagonals (the edges two triangles share), all strictly inside the grid."""
    rng = np.random.default_rng(seed)
    kind = rng.integers(0, 3, n)
    x = rng.integers(1, N, n).astype(np.float64)
    z = rng.integers(1, N, n).astype(np.float64)
    f = rng.random(n)
    x = np.where(kind == 1, x + f, x)          # along a square's edge
    z = np.where(kind == 2, z + f, z)          # along the next square edge
    diag = rng.random(n) < 0.3
    x = np.where(diag, np.floor(x) + f, x)     # on a diagonal of the square
    z = np.where(diag, np.floor(z) + f, z)
    target = np.stack([x, np.zeros(n), z], 1) @ world[:3, :3].T + world[:3, 3]
    o = target + np.stack([rng.uniform(-40, 40, n), rng.uniform(5, 60, n), rng.uniform(-40, 40, n)], 1) @ \
        world[:3, :3].T
    d = target - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.zeros(n, oracle.RAY)
    rays["o"], rays["d"], rays["tmax"] = o.astype(np.float32), d.astype(np.float32), 3e38
    return rays


@pytest.mark.parametrize("tilt", [False, True])
def test_no_ray_slips_between_the_grids_triangles(tilt):
    prep, world = grid_scene(tilt)
    sc = oracle.Scene(*prep.desc())
    rays = edge_rays(world, 100_000, seed=7 + tilt)
    wt, _ = sc.trace_batch(rays, oracle.WALK_WT, threads=8)
    mt, _ = sc.trace_batch(rays, oracle.WALK_EXACT, threads=8)
    wt_miss, mt_miss = int((wt["tri"] == NO_TRI).sum()), int((mt["tri"] == NO_TRI).sum())
    print(f"tilt {tilt}: watertight misses {wt_miss}, Moller-Trumbore misses {mt_miss} of {len(rays)}")
    assert wt_miss == 0
    # where both hit, the same plane: t within float rounding; on the same triangle the same barycentrics (u weighs
    # v1, v weighs v2 in both tests)
    both = (wt["tri"] != NO_TRI) & (mt["tri"] != NO_TRI)
    assert np.allclose(wt["t"][both], mt["t"][both], rtol=1e-5)
    same = both & (wt["tri"] == mt["tri"])
    assert same.mean() > 0.5
    assert np.allclose(wt["u"][same], mt["u"][same], atol=1e-4) and np.allclose(wt["v"][same], mt["v"][same], atol=1e-4)


def test_the_watertight_walk_is_its_brute_force():
    for seed in (1, 2):
        prep = S.instanced_scene(seed, meshes=4, instances=30)
        sc = oracle.Scene(*prep.desc())
        rng = np.random.default_rng(seed)
        rays = np.zeros(4000, oracle.RAY)
        rays["o"] = ((rng.random((4000, 3)) - 0.5) * 60).astype(np.float32)
        d = rng.normal(size=(4000, 3))
        rays["d"] = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
        rays["tmax"] = 3e38
        walk, _ = sc.trace_batch(rays, oracle.WALK_WT, threads=8)
        brute = sc.brute_batch(rays, oracle.WALK_WT, threads=8)
        assert (walk.view(np.uint32) == brute.view(np.uint32)).all()
        assert (walk["tri"] != NO_TRI).mean() > 0.05


def test_where_the_tests_disagree_double_precision_decides():
    """On the tilted grid's edge rays, each float test's answer against the double-precision brute force: FP64 hits
    the grid with every ray; the watertight test hits every one at FP64's distance (to float rounding), on FP64's
    triangle or on a neighbour that shares a vertex with it (a ray through an edge touches both; FP64 names the
    lowest); Moller-Trumbore misses rays FP64 hits."""
    prep, world = grid_scene(True)
    sc = oracle.Scene(*prep.desc())
    rays = edge_rays(world, 20_000, seed=3)
    wt, _ = sc.trace_batch(rays, oracle.WALK_WT, threads=8)
    mt, _ = sc.trace_batch(rays, oracle.WALK_EXACT, threads=8)
    f64 = sc.brute64_batch(rays, threads=8)
    assert (f64["tri"] != NO_TRI).all() and (wt["tri"] != NO_TRI).all()
    assert np.allclose(wt["t"], f64["t"], rtol=1e-5)
    leaves = prep.leaves.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    corners = np.stack([leaves[:, 0], leaves[:, 0] + leaves[:, 1], leaves[:, 0] + leaves[:, 2]], 1)  # exact here
    a, b = corners[wt["tri"]], corners[f64["tri"].astype(np.int64)]
    touch = (np.abs(a[:, :, None, :] - b[:, None, :, :]).sum(-1) == 0).any(axis=(1, 2))
    assert touch.all(), int((~touch).sum())
    mt_miss = mt["tri"] == NO_TRI
    print(f"watertight names FP64's own triangle on {(wt['tri'] == f64['tri']).mean():.3f} of the rays, a neighbour "
          f"on the rest; Moller-Trumbore misses {int(mt_miss.sum())} rays FP64 hits")
    assert mt_miss.sum() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["grid", "instances", "precision"])
def test_gpu_watertight_walk_is_the_oracles(device, which):
    print(f"watertight: {which}", flush=True)
    if which == "grid":
        prep, world = grid_scene(True)
        rays = edge_rays(world, 100_000, seed=9)
    elif which == "instances":
        prep = S.instanced_scene(5, meshes=4, instances=30)
        rng = np.random.default_rng(5)
        rays = np.zeros(20_000, oracle.RAY)
        rays["o"] = ((rng.random((20_000, 3)) - 0.5) * 60).astype(np.float32)
        d = rng.normal(size=(20_000, 3))
        rays["d"] = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
        rays["tmax"] = 3e38
    else:
        prep, lo, hi = S.precision_scene(200_000, seed=2)
        rays = S.incoherent_rays(200_000, lo, hi, seed=4).view(oracle.RAY)
    want, _ = oracle.Scene(*prep.desc()).trace_batch(rays, oracle.WALK_WT, threads=8)
    sc = T.Scene(device, *prep.desc())
    got = sc.trace(rays.view(T.RAY), T.WALK_WATERTIGHT)
    sc.close()
    bad = ~(got.view(np.uint32).reshape(-1, 6) == want.view(np.uint32).reshape(-1, 6)).all(axis=1)
    assert not bad.any(), (which, int(bad.sum()), got[bad][:3], want[bad][:3])
    if which == "grid":
        assert (got["tri"] != T.NO_HIT).all()
