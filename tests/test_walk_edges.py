"""The exact walk finds what brute force finds on rays through triangle edges on flat boxes: the city's extruded
boxes have axis-aligned faces, so a mesh tree's box around a face's triangles is flat, and a ray through an edge of
one of them can be missed by a float slab test that the triangle test accepts (u or v exactly 0). The walks widen
each box before testing it; this holds the oracle's exact walk to brute force on such rays. No GPU here."""
import numpy as np

import oracle
from threert_amd import render as R


def aimed_at_edges(rs, n: int, seed: int) -> np.ndarray:
    """Rays from random points above the scene to points on triangle edges (v0 + s e2: u = 0), in world space."""
    rng = np.random.default_rng(seed)
    prep = rs.prep
    rays = np.zeros(n, oracle.RAY)
    inst = rng.integers(1, len(prep.instances), n)  # not the ground plane
    rec = prep.instances[inst]
    # a triangle of the record's mesh: its leaves are [first, first + mesh size); pick within the mesh's range
    firsts = np.unique(prep.instances["info"][:, 0])
    ends = {int(f): int(e) for f, e in zip(firsts, list(firsts[1:]) + [len(prep.leaves)])}
    tri = np.array([int(f) + int(rng.integers(0, ends[int(f)] - int(f))) for f in rec["info"][:, 0]])
    leaf = prep.leaves[tri]
    v0, e2 = leaf[:, 0, :3].astype(np.float64), leaf[:, 2, :3].astype(np.float64)
    p_model = v0 + rng.random((n, 1)) * e2
    m = rec["m"].astype(np.float64)  # 3 rows of [M | t]
    p = np.einsum("nij,nj->ni", m[:, :, :3], p_model) + m[:, :, 3]
    o = p + np.stack([rng.uniform(-30, 30, n), rng.uniform(20, 60, n), rng.uniform(-30, 30, n)], 1)
    d = p - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays["o"] = o.astype(np.float32)
    rays["d"] = d.astype(np.float32)
    rays["tmax"] = 3e38
    return rays


def test_rays_through_edges_of_flat_boxes_find_the_brute_force_hit():
    rs, _, _ = R.city(seed=3, meshes=6, instances=40, extent=60.0)
    sc = oracle.Scene(*rs.prep.desc())
    rays = aimed_at_edges(rs, 4000, seed=1)
    walk, _ = sc.trace_batch(rays, oracle.WALK_EXACT, threads=8)
    brute = sc.brute_batch(rays, oracle.WALK_EXACT, threads=8)
    same = (walk.view(np.uint32).reshape(-1, 6) == brute.view(np.uint32).reshape(-1, 6)).all(axis=1)
    assert same.all(), (np.nonzero(~same)[0][:5], (~same).sum())
    assert (walk["tri"] != 0xFFFFFFFF).mean() > 0.9
    edge = (walk["u"] == 0) | (walk["v"] == 0)
    assert edge.sum() > 0  # some rays do hit on an edge


def test_the_city_ray_the_binary_box_test_missed():
    """A bounce ray of the 1280 x 720 city render (pixel 349, 360, sample 0), recorded bit for bit: its closest hit is
    on an edge (u = 0) of a wall triangle at t = 18.30, which the slab test of the triangle's flat box rejects
    without a margin (the unwidened walk reports a wall 24.78 away)."""
    rs, _, _ = R.city(seed=1)
    ray = np.zeros(1, oracle.RAY)
    ray.view(np.uint32)[:] = [0x4249DA89, 0x4221E1D0, 0xC31D758D, 0x7F61B1E6, 0x3E8DD8C8, 0xBF03FA9B, 0xBF4F933C, 0]
    sc = oracle.Scene(*rs.prep.desc())
    exact, _ = sc.trace_batch(ray, oracle.WALK_EXACT)
    ref, _ = sc.trace_batch(ray, oracle.WALK_REF)
    assert exact.view(np.uint32).tolist() == [0x419264E3, 0xEA913, 0x0, 0x3F7226E6, 0x758, 0x0]
    assert ref["t"][0] > exact["t"][0]  # the reference's walk, boxes as they are, misses it
