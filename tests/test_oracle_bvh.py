"""The tree builders: the reference's own scenarios (bvh.rs:405-490, units.rs:1536-1610) on the oracle, and the
library's host builders held to the oracle's bytes on the same inputs."""
import math

import numpy as np
import pytest

import oracle
import threert_amd as T
from tests.lcg import Rng, Stream
from threert_amd import scene as S

INF32 = np.float32(np.inf)


def single_unit(tris: np.ndarray, tree: dict) -> oracle.Scene:
    """One unit, identity placement, whose mesh tree is `tree` (a bare tree walked as the kernel walks units)."""
    nodes = np.concatenate([tree["nodes"], np.array([((-1e30,) * 3, 0, (1e30,) * 3, 1)], dtype=oracle.NODE)])
    order = tree["order"]
    leaves = S.leaves_of(tris[order].astype(np.float32), np.zeros(len(order), np.uint32), order)
    rec = S.instance_record(np.eye(4, dtype=np.float32), 0)
    oracle.unit_record(rec, 0, True)
    return oracle.Scene(nodes, leaves, rec, len(tree["nodes"]), 1)


def garbage_cases():
    rnd = Stream(11)

    def p(s):
        return rnd.p(s)

    good = []
    for _ in range(600):
        c = p(20.0)
        good.append([c + p(1.0), c + p(1.0), c + p(1.0)])
    good = np.array(good, np.float32)
    nan = np.array([np.nan, 0, 0], np.float32)
    inf = np.array([0, np.inf, 0], np.float32)
    mixed = good.copy()
    for k in range(0, len(mixed), 7):
        mixed[k][k % 3] = nan if k % 2 == 0 else inf
    allbad = np.array([[nan, [0, 0, 0], [1, 1, 1]], [[0, 0, 0], inf, [1, 1, 1]],
                       [[-np.inf] * 3, [0, 0, 0], [1, 1, 1]]], np.float32)
    limit = []
    for k in range(200):
        c = np.float32(3.0e38) * np.float32(1.0 if k % 2 == 0 else -1.0)
        c = np.array([c, c, c], np.float32)
        limit.append([c, c * np.float32(0.5), np.zeros(3, np.float32)])
    limit = np.array(limit, np.float32)
    thousand = np.repeat(good[:1], 1000, axis=0)
    needle = []
    for k in range(500):
        c = np.array([np.float32(k) * np.float32(1e-3), 0, 0], np.float32)
        needle.append([c, c + np.array([0, 1e-4, 0], np.float32), c + np.array([0, 0, 1e-4], np.float32)])
    needle = np.array(needle, np.float32)
    return [("mixed", mixed), ("all bad", allbad), ("near the limit", limit),
            ("one triangle a thousand times", thousand), ("a needle", needle)]


@pytest.mark.parametrize("what,tris", garbage_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_garbage_in_makes_a_tree_the_kernel_walks(what, tris):
    """bvh.rs:409-441: every tree passes check_tree, items that are not finite come last and in no leaf, and the
    kernel's walk ends inside its buffers."""
    b = oracle.build_tris(tris, 48)
    bad = [i for i in range(len(tris)) if not np.isfinite(tris[i]).all()]
    good_n = len(tris) - len(bad)
    leaves = b["nodes"][b["nodes"]["count"] > 0]
    if good_n:
        assert int((leaves["left_or_first"] + leaves["count"]).max()) <= good_n, what
        assert oracle.check_tree(b["nodes"], 48, good_n) == b["depth"], what
    assert sorted(b["order"][good_n:].tolist()) == bad, what
    # the library's builder: the same bytes
    lb = T.build_tris(tris, 48)
    assert lb["nodes"].tobytes() == b["nodes"].tobytes() and (lb["order"] == b["order"]).all(), what
    if good_n == 0:
        return
    sc = single_unit(np.nan_to_num(tris, posinf=0, neginf=0), b)
    for k in range(200):
        o = np.array([k - 100.0, 7.0, 3.0], np.float32)
        d = np.array([0.3, -1.0, 0.2], np.float32) - o * np.float32(0.001)
        d = (d / np.float32(np.linalg.norm(d))).astype(np.float32)
        sc.trace(o, d, variant=oracle.WALK_REF)


def test_a_tree_is_no_deeper_than_its_stack():
    """bvh.rs:446-458: boxes spaced in powers of two make SAH deep; held, the depth is capped, and with room the
    bytes are the unheld tree's."""
    tris = []
    for k in range(120):
        c = np.array([np.float32(2.0) ** np.float32(k), 0, 0], np.float32)
        tris.append([c, c + np.array([0, 1, 0], np.float32), c + np.array([0, 0, 1], np.float32)])
    tris = np.array(tris, np.float32)
    free = oracle.build_tris(tris, 1000)
    assert free["depth"] > 20
    for cap in (8, 12, 20):
        held = oracle.build_tris(tris, cap)
        assert held["depth"] <= cap
        assert oracle.check_tree(held["nodes"], cap, len(tris)) == held["depth"]
        lb = T.build_tris(tris, cap)
        assert lb["nodes"].tobytes() == held["nodes"].tobytes()
    roomy = oracle.build_tris(tris, free["depth"])
    assert roomy["depth"] == free["depth"] and (roomy["order"] == free["order"]).all()
    assert roomy["nodes"].tobytes() == free["nodes"].tobytes()
    assert T.build_tris(tris, 1000)["nodes"].tobytes() == free["nodes"].tobytes()


def test_the_tree_over_nothing_is_one_leaf():
    """bvh.rs:63, 82-84: no triangle is the EMPTY node (a leaf of one item without a box)."""
    b = oracle.build_tris(np.zeros((0, 3, 3), np.float32))
    assert len(b["nodes"]) == 1 and b["nodes"][0]["count"] == 1 and b["nodes"][0]["left_or_first"] == 0
    assert T.build_tris(np.zeros((0, 3, 3), np.float32))["nodes"].tobytes() == b["nodes"].tobytes()


@pytest.mark.parametrize("n", [1, 2, 4, 5, 17, 1000])
def test_the_kernel_walks_every_tree_within_its_buffers(n):
    """bvh.rs:467-490 (seed 7): 2000 random rays and rays along the axes end inside the buffers, and the walk's
    closest hit is brute force's."""
    rnd = Stream(7)
    tris = []
    for _ in range(n):
        c = rnd.p(20.0)
        tris.append([c + rnd.p(1.0), c + rnd.p(1.0), c + rnd.p(1.0)])
    tris = np.array(tris, np.float32)
    b = oracle.build_tris(tris, 48)
    assert T.build_tris(tris, 48)["nodes"].tobytes() == b["nodes"].tobytes()
    sc = single_unit(tris, b)
    rays = []
    for _ in range(2000):
        o = rnd.p(60.0)
        d = rnd.p(2.0)
        nd = np.float32(np.linalg.norm(d))
        d = d / nd if nd > 0 else np.array([1, 0, 0], np.float32)
        rays.append((o, d))
    for a in ([1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [1, -1, 0]):
        a = np.array(a, np.float32) / np.float32(np.linalg.norm(a))
        rays += [(np.zeros(3, np.float32), a), (-a * 5, a), (a * 5, -a), (np.array([0, 4.5, 0.001], np.float32), -a)]
    rr = np.zeros(len(rays), oracle.RAY)
    rr["o"] = [r[0] for r in rays]
    rr["d"] = [r[1] for r in rays]
    rr["tmax"] = oracle.INF
    walk, reads = sc.trace_batch(rr, oracle.WALK_REF, threads=1)
    brute = sc.brute_batch(rr, oracle.WALK_REF, threads=1)
    assert (walk["t"].view(np.uint32) == brute["t"].view(np.uint32)).all()
    assert reads["mesh_verts"] > 0  # the rays reach the leaves


def test_the_builders_keep_their_depths():
    """units.rs:1536-1575: a top tree of 1000 units has 1999 nodes, depth 10..19; by surface area it costs a ray
    less than the median tree (5 % less; 10 % with walls among lamps); the library's builders agree."""
    rng = Rng(5)
    boxes = []
    for k in range(1000):
        t = rng.v(100.0)
        w = np.eye(4, dtype=np.float32)
        w[:3, 3] = t
        boxes.append(oracle.unit_box(w.T.copy(), np.full(3, -0.25, np.float32), np.full(3, 0.25, np.float32)))
    boxes = np.array(boxes, np.float32)
    top = oracle.top_tree(boxes, np.arange(1000, dtype=np.uint32))
    assert len(top["nodes"]) == 1999
    depth = oracle.check_top(top["nodes"], 1000)
    assert 10 <= depth <= 19
    assert T.top_tree(boxes, np.arange(1000, dtype=np.uint32))["nodes"].tobytes() == top["nodes"].tobytes()
    sah, median = oracle.top_cost(top["nodes"]), oracle.top_cost(oracle.build_median(boxes, 1)["nodes"])
    assert sah < 0.95 * median
    mixed = boxes.copy()
    mixed[::10, :3] -= np.array([12, 6, 0], np.float32)
    mixed[::10, 3:] += np.array([12, 6, 0], np.float32)
    sah = oracle.top_cost(oracle.build_sah(mixed, 19)["nodes"])
    median = oracle.top_cost(oracle.build_median(mixed, 1)["nodes"])
    assert sah < 0.9 * median
    assert T.build_sah(mixed, 19)["nodes"].tobytes() == oracle.build_sah(mixed, 19)["nodes"].tobytes()
    assert T.build_median(mixed, 1)["nodes"].tobytes() == oracle.build_median(mixed, 1)["nodes"].tobytes()


def test_sah_depth_holds_whatever_the_boxes():
    """units.rs:1576-1608: as many boxes as the depth allows, one level more, plenty; boxes in one place, a chain
    each further than the rest together, two clumps, anywhere: the tree is no deeper than asked."""
    rng = Rng(5)
    for rnd in range(200):
        n = 1 + rng.below(300)
        least = math.ceil(math.log2(n)) if n > 1 else 0
        max_depth = least + [0, 0, 1, 2, 40][rnd % 5]
        kind = rnd % 4
        boxes = np.zeros((n, 6), np.float32)
        for k in range(n):
            if kind == 0:
                c = np.zeros(3, np.float32)
            elif kind == 1:
                c = np.array([np.float32(2.0) ** np.float32(min(k, 100)), 0, 0], np.float32)
            elif kind == 2:
                c = rng.v(1.0) + (np.float32(100) if k % 7 == 0 else np.float32(0))
            else:
                c = rng.v(100.0)
            boxes[k, :3] = c - np.float32(0.5)
            boxes[k, 3:] = c + np.float32(0.5)
        b = oracle.build_sah(boxes, max_depth)
        d = oracle.check_tree(b["nodes"], max_depth, n, 1)
        assert isinstance(d, int) and d <= max_depth, (rnd, n, max_depth, d)
        assert len(b["nodes"]) == 2 * n - 1
        assert T.build_sah(boxes, max_depth)["nodes"].tobytes() == b["nodes"].tobytes()


def test_mesh_trees_fall_back_to_median_splits():
    """units.rs:231-246: a mesh whose SAH tree is deeper than 45 is built again by median splits (leaves of 4)."""
    tris = []
    for k in range(2000):
        c = np.array([np.float32(1.5) ** np.float32(k % 200) + np.float32(k) * np.float32(1e-3), 0, 0], np.float32)
        tris.append([c, c + np.array([0, 1, 0], np.float32), c + np.array([0, 0, 1], np.float32)])
    tris = np.array(tris, np.float32)
    m = oracle.mesh_tree(tris)
    assert m["depth"] <= 45
    assert isinstance(oracle.check_tree(m["nodes"], 45, len(tris)), int)
    lm = T.mesh_tree(tris)
    assert lm["nodes"].tobytes() == m["nodes"].tobytes() and (lm["order"] == m["order"]).all()
    assert np.allclose(lm["lo"], m["lo"]) and np.allclose(lm["hi"], m["hi"])


def test_mesh_trees_refuse_what_the_reference_refuses():
    """units.rs:234-239: no triangles, or a vertex that is not finite: no tree."""
    with pytest.raises(ValueError):
        oracle.mesh_tree(np.zeros((0, 3, 3), np.float32))
    bad = np.zeros((1, 3, 3), np.float32)
    bad[0, 1, 2] = np.nan
    with pytest.raises(ValueError):
        oracle.mesh_tree(bad)
    with pytest.raises(T.ThreeRTError):
        T.mesh_tree(bad)
    with pytest.raises(T.ThreeRTError):
        T.mesh_tree(np.zeros((0, 3, 3), np.float32))


def test_check_tree_refuses_broken_trees():
    """units.rs:1610 (the checks refuse broken trees): a child before its parent, a node reached twice, a leaf
    outside the items, a box that is not finite, a tree too deep."""
    tris = Rng(3).tris(40, 5.0, 1.0)
    t = oracle.build_tris(tris, 48)
    nodes = t["nodes"].copy()
    assert isinstance(oracle.check_tree(nodes, 48, 40), int)
    bad = nodes.copy()
    bad[0]["left_or_first"] = 0
    assert isinstance(oracle.check_tree(bad, 48, 40), str)
    bad = nodes.copy()
    leaf = np.nonzero(bad["count"] > 0)[0][0]
    bad[leaf]["left_or_first"] = 39
    bad[leaf]["count"] = 4
    assert isinstance(oracle.check_tree(bad, 48, 40), str)
    bad = nodes.copy()
    inner = np.nonzero(bad["count"] == 0)[0][0]
    bad[inner]["bmin"][0] = np.nan
    assert isinstance(oracle.check_tree(bad, 48, 40), str)
    assert isinstance(oracle.check_tree(nodes, 1, 40), str)


def test_unit_records_carry_root_and_flags():
    """units.rs:329-342: n0.w the root's bits; n1.w keeps INST_* bits and adds UNIT_SHADOW / UNIT_SHIFTED; a
    translation alone is SHIFTED."""
    for world, shifted in [(S.translation([1, 2, 3]), True), (S.trs([1, 2, 3], [0.3, 0, 0], [1, 1, 1]), False)]:
        for cast in (True, False):
            a = S.instance_record(world, 7)
            a["n"][0, 1, 3] = np.array([4 | 1 | 256], np.uint32).view(np.float32)[0]
            b = a.copy()
            oracle.unit_record(a, 1234, cast)
            T.unit_record(b, 1234, cast)
            assert a.tobytes() == b.tobytes()
            assert a["n"][0, 0, 3].view(np.uint32) == 1234
            bits = int(a["n"][0, 1, 3].view(np.uint32))
            assert bits == 4 | (1 if cast else 0) | (2 if shifted else 0)
