"""The light tree (ltree.rs, pathtrace.wgsl:489-618): the library's builder held to the reference's own tests and
its LightTree::check invariants, the half floats to numpy's binary16, and the oracle's emitter_draw / emitter_pmf
to each other, to a sum of one and to draw frequencies; a render with the tree converges to the render without it.
No GPU here."""
import math

import numpy as np
import pytest

import oracle
import threert_amd as T
from threert_amd import render as R
from threert_amd import scene as S

LEAF = 0x80000000
PATH_BITS = 26


def tri(i: int, x: float):
    """ltree.rs:466-468 `tri`: a right triangle at x of area 0.5 and its centroid."""
    v = np.array([[x, 0, 0], [x + 1, 0, 0], [x, 1, 0]], np.float32)
    c = np.array([x + np.float32(1.0 / 3.0), np.float32(1.0 / 3.0), 0.0], np.float32)
    return v, c


def words(t) -> np.ndarray:
    return t.nodes.view(np.uint32)[:, 3]


def check(t, entries: int) -> int:
    """ltree.rs:162-233 LightTree::check, restated: from node 0 every interior node's children are after it and
    inside the tree, every node is reached once and no deeper than 26, the leaves together hold every entry once, and
    every entry's code leads level by level to the leaf that holds it. Returns the depth."""
    w = words(t)
    n = len(w)
    seen = np.zeros(n, bool)
    held = np.zeros(entries, np.int64)
    stack, deepest = [(0, 0)], 0
    while stack:
        i, depth = stack.pop()
        assert not seen[i], f"node {i} reached twice"
        seen[i] = True
        assert depth <= PATH_BITS
        deepest = max(deepest, depth)
        if w[i] & LEAF:
            first, count = int(w[i] & 0x1FFFFFFF), int((w[i] >> 29) & 3) + 1
            assert first + count <= entries
            held[first:first + count] += 1
        else:
            left = int(w[i])
            assert i < left and left + 1 < n, f"node {i}'s children {left}, {left + 1}"
            stack += [(left, depth + 1), (left + 1, depth + 1)]
    assert seen.all(), f"node {int(np.argmin(seen))} not reached"
    assert (held == 1).all()
    for k, c in enumerate(t.code):
        depth = int(c >> PATH_BITS) - 1
        assert 0 <= depth <= PATH_BITS
        i = 0
        for level in range(depth):
            assert not w[i] & LEAF, f"entry {k}'s code goes past leaf {i}"
            i = int(w[i]) + int((c >> level) & 1)
        assert w[i] & LEAF and int(w[i] & 0x1FFFFFFF) <= k < int(w[i] & 0x1FFFFFFF) + int((w[i] >> 29) & 3) + 1
    return deepest


def sphere(node) -> tuple[np.ndarray, float]:
    h = node[:2].view(np.uint32)
    half = np.array([h[0] & 0xFFFF, h[0] >> 16, h[1] & 0xFFFF, h[1] >> 16], np.uint16).view(np.float16)
    return half[:3].astype(np.float64), float(half[3])


def build(n: int, place, weight):
    vs, cs = zip(*(tri(i, place(i)) for i in range(n)))
    return T.build_light_tree(np.stack(vs), np.stack(cs), np.array([weight(i) for i in range(n)], np.float32)), \
        np.stack(vs)


def test_paths_lead_to_the_leaves_holding_their_entries_and_powers_sum():
    """ltree.rs:495-533 on the library's builder: 1000 emitters along x, weights 1 + i."""
    n = 1000
    t, vs = build(n, lambda i: 2.0 * i, lambda i: 1.0 + i)
    assert abs(t.power - sum(1.0 + i for i in range(n))) < 1.0
    assert sorted(t.order.tolist()) == list(range(n))
    depth = check(t, n)
    assert depth <= math.ceil(math.log2(n))
    # a node's power is its emitters' weights; its sphere holds their triangles
    w = words(t)

    def under(i):
        if w[i] & LEAF:
            return [int(w[i] & 0x1FFFFFFF)]
        return under(int(w[i])) + under(int(w[i]) + 1)

    for i in (0, 1, 2, 5, 17, len(w) - 1):
        ks = under(i)
        want = sum(1.0 + int(t.order[k]) for k in ks)
        assert abs(t.nodes[i, 2] - want) <= 1e-2 * want
        c, r = sphere(t.nodes[i])
        for k in ks:
            d = np.linalg.norm(vs[t.order[k]].astype(np.float64) - c, axis=1)
            assert (d <= r).all(), (i, k, d, r)


def test_few_emitters_make_no_tree():
    """ltree.rs:536-543 (MIN_STATIC = 256), and emitters without weight are left out."""
    assert build(100, float, lambda i: 1.0)[0] is None
    assert build(255, float, lambda i: 1.0)[0] is None
    assert build(256, float, lambda i: 1.0)[0] is not None
    t = build(400, float, lambda i: 0.0 if i % 3 else 1.0)[0]  # 134 weigh something
    assert t is None
    t = build(900, float, lambda i: 0.0 if i % 3 == 0 else 1.0)[0]
    assert len(t.order) == 600 and not (t.order % 3 == 0).any()
    check(t, 600)


def test_garbage_in_makes_a_tree_the_kernel_reads():
    """ltree.rs:474-492: triangles that are not finite, centroids that are not, weights near f32's limit (the
    reference's moving emitters are static here: this builder has no moving subtree)."""
    n = 700
    vs, cs = zip(*(tri(i, float(i % 50)) for i in range(n)))
    vs, cs = np.stack(vs), np.stack(cs)
    for k in range(0, n, 9):
        vs[k, 0, k % 3] = np.nan if k % 2 == 0 else np.inf
        cs[k] = (np.nan, 0.0, 0.0)
    w = np.array([3.0e38 if i % 11 == 0 else 1.0 for i in range(n)], np.float32)
    t = T.build_light_tree(vs, cs, w)
    check(t, n)


def test_sphere_centres_are_binary16_roundings():
    """encode (ltree.rs:118-131): a leaf's centre is its triangle box's f32 centre rounded to binary16 (numpy's
    conversion, round to nearest even), and its radius is no smaller than the box's half diagonal plus the rounding."""
    rng = np.random.default_rng(3)
    n = 600
    vs = (rng.normal(size=(n, 3, 3)) * rng.choice([1e-3, 1.0, 300.0, 6e4], size=(n, 1, 1))).astype(np.float32)
    cs = vs.mean(axis=1).astype(np.float32)
    t = T.build_light_tree(vs, cs, np.ones(n, np.float32))
    w = words(t)
    checked = 0
    for i in range(len(w)):
        if not w[i] & LEAF:
            continue
        v = vs[t.order[int(w[i] & 0x1FFFFFFF)]]
        lo, hi = v.min(axis=0), v.max(axis=0)
        c32 = np.float32(0.5) * (lo + hi)
        with np.errstate(over="ignore"):
            c16 = c32.astype(np.float16)
        h = t.nodes[i, :2].view(np.uint32)
        assert np.array([h[0] & 0xFFFF, h[0] >> 16, h[1] & 0xFFFF], np.uint16).tolist() == c16.view(np.uint16).tolist()
        _, r = sphere(t.nodes[i])
        need = 0.5 * np.linalg.norm((hi - lo).astype(np.float64)) + np.linalg.norm(c16.astype(np.float64) - c32)
        if np.isfinite(need) and need < 65504:
            assert r >= need * (1 - 1e-6), (i, r, need)
            checked += 1
    assert checked > 300


def lit_floor(n_emitters: int = 600, seed: int = 5, tree: bool = True):
    """A 10 x 10 Lambert floor (albedo 0.5) under n small emitter triangles at heights 0.5-3, radiance 1-20."""
    rng = np.random.default_rng(seed)
    floor = R.quad((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5))
    centre = np.stack([rng.uniform(-5, 5, n_emitters), rng.uniform(0.5, 3.0, n_emitters),
                       rng.uniform(-5, 5, n_emitters)], 1)
    tris = (centre[:, None, :] + rng.normal(size=(n_emitters, 3, 3)) * 0.08).astype(np.float32)
    sec = rng.integers(0, 4, n_emitters).astype(np.uint32)
    mats = [R.material((0.5, 0.5, 0.5), specular=0.0)] + [R.material((0, 0, 0), (e, e, e), specular=0.0)
                                                          for e in (1.0, 4.0, 9.0, 20.0)]
    meshes = [R.RMesh(floor), R.RMesh(tris, sections=sec)]
    eye = np.eye(4, dtype=np.float32)
    rs = R.build(meshes, [R.RInstance(0, eye, [0]), R.RInstance(1, eye, [1, 2, 3, 4])], mats, light_tree=tree)
    return rs, R.Camera((0.0, 7.0, -7.0), (0.0, 0.0, 0.0), 50.0)


def oracle_of(rs):
    p = rs.prep
    return oracle.RenderScene(p.nodes, p.leaves, p.instances, p.top_first, p.units, **rs.tables())


def test_the_emitter_table_follows_the_tree():
    """lib.rs:419-481: the tree's emitters first in the CDF in leaf order, each triangle's code its path, the header's
    y the weight outside the tree (here none), P.ltree = (-, entries, tail, RIS_OFFLINE)."""
    rs, cam = lit_floor()
    n = len(rs.emitter_tris)
    assert rs.tree_entries == n and len(rs.light_tree) == 2 * n - 1
    order = rs.emitters[:n, 0].view(np.uint32)
    assert sorted(order.tolist()) == list(range(n))
    codes = rs.emitter_tris.view(np.uint32)[:, 11]
    w = rs.light_tree.view(np.uint32)[:, 3]
    for k in range(0, n, 7):  # entry k's emitter's code leads to the leaf of entry k
        c, i = int(codes[order[k]]), 0
        for level in range((c >> PATH_BITS) - 1):
            i = int(w[i]) + ((c >> level) & 1)
        assert w[i] & LEAF and int(w[i] & 0x1FFFFFFF) == k
    p = rs.params(32, 32, cam)
    assert p[85] == n and p.view(np.float32)[87] == 6.0
    header = rs.emitters[n + (65537 + 3) // 4]
    assert 0.0 <= header[1] < 1e-5
    assert np.isclose(rs.light_tree[0, 2], header[0], rtol=1e-5)


def test_draws_and_pmf_agree_and_sum_to_one():
    """emitter_draw and emitter_pmf (pathtrace.wgsl:547-618) in the oracle: with no hemisphere (n = 0) the pmf over
    every emitter sums to 1; at a floor point the pmf is what the draw returns for what it draws, and 200,000 draws
    land on each emitter as often as its pmf says (within 5 standard errors), missing as often as 1 - sum."""
    rs, cam = lit_floor()
    orc = oracle_of(rs)
    p = rs.params(32, 32, cam)
    n = len(rs.emitter_tris)
    x, up = np.array([0.3, 0.0, -1.2], np.float32), np.array([0.0, 1.0, 0.0], np.float32)
    pm0 = np.array([orc.emitter_pmf(p, x, np.zeros(3, np.float32), e) for e in range(n)], np.float64)
    assert abs(pm0.sum() - 1.0) < 1e-5 and (pm0 > 0).all()
    pm = np.array([orc.emitter_pmf(p, x, up, e) for e in range(n)], np.float64)
    assert 0.5 < pm.sum() <= 1.0 + 1e-5
    order = rs.emitters[:n, 0].view(np.uint32)
    rng = np.random.default_rng(11)
    draws = 200_000
    hits = np.zeros(n, np.int64)
    misses = 0
    for k in range(draws):
        entry, pr = orc.emitter_draw(p, x, up, rng.random(3).astype(np.float32))
        if entry < 0:
            misses += 1
            continue
        e = int(order[entry])
        hits[e] += 1
        if k < 2000:
            assert abs(pr - pm[e]) <= 1e-5 * pm[e], (e, pr, pm[e])
    freq = hits / draws
    se = np.sqrt(pm * (1 - pm) / draws)
    big = pm > 2e-4
    assert (np.abs(freq[big] - pm[big]) <= 5 * se[big]).all()
    assert hits[pm == 0].sum() == 0
    miss = 1.0 - pm.sum()
    assert abs(misses / draws - miss) <= 5 * math.sqrt(max(miss * (1 - miss), 1e-12) / draws) + 1e-6


def test_a_render_with_the_tree_converges_to_the_render_without_it():
    """The tree changes which emitters are sampled, not the expectation: the lit floor rendered by the oracle with the
    tree and with the CDF alone (lib.rs's THREERT_LTREE=0), 1024 samples a pixel, agree in mean luminance within 4
    standard errors (float64 over the pixels; about 0.4 % of the mean), and both are brighter than nothing."""
    means = []
    for tree in (True, False):
        rs, cam = lit_floor(tree=tree)
        orc = oracle_of(rs)
        p = rs.params(40, 30, cam, max_bounces=2, rr_depth=2)
        acc = orc.render(p, 0, 1024, threads=8).astype(np.float64)
        cnt = acc[..., 3]
        lum = acc[..., 7] / cnt
        var = np.maximum(acc[..., 11] / cnt - lum * lum, 0.0) / cnt
        means.append((lum.mean(), math.sqrt(var.sum()) / lum.size))
    (a, sa), (b, sb) = means
    assert a > 0.05 and b > 0.05
    assert abs(a - b) <= 4 * math.hypot(sa, sb), means
