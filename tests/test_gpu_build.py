"""The library's own mesh trees and their 4-wide nodes built on the device (trt_device_set_build(TRT_BUILD_DEVICE),
and by default for large scenes) against the same built on the host (TRT_BUILD_HOST): the wide nodes, the leaves
in the trees' order and the units must be the same bytes. The host build is the reference for this: it is the tree every other test walks against
the oracle's hits. Cases: instanced scenes, the reference's garbage-in scenarios (bvh.rs:409-441: non-finite
corners, coordinates near the float limit, one triangle a thousand times, a needle), a 200k-triangle clustered
scene, and the Cornell box and a sunlit city as the renderer uploads them."""
import sys

import numpy as np
import pytest

import threert_amd as T
from threert_amd import render as R
from threert_amd import scene as S
from tests.test_oracle_bvh import garbage_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def both(dev, upload):
    """(host-built scene, device-built scene) from `upload()`."""
    try:
        dev.set_build(T.BUILD_HOST)
        h = upload()
        dev.set_build(T.BUILD_DEVICE)
        d = upload()
    finally:
        dev.set_build(T.BUILD_AUTO)
    return h, d


def same_trees(h, d, what, host_meshes=0):
    sh, sd = h.stats(), d.stats()
    say(f"  {what}: build host {sh['build_ms']:.1f} ms, device {sd['build_ms']:.1f} ms (trees on the device "
        f"{sd['device_build_ms']:.1f} ms), meshes device {sd['device_meshes']} host {sd['host_meshes']}")
    assert sh["device_meshes"] == 0
    assert sd["host_meshes"] == host_meshes, (what, sd)
    for which, name in ((0, "wide nodes"), (1, "leaves in tree order"), (2, "units")):
        a, b = h.wide_read(which), d.wide_read(which)
        assert a.size == b.size and (a == b).all(), (what, name, int(np.count_nonzero(a != b)))
    h.close()
    d.close()


@pytest.mark.parametrize("seed", range(4))
def test_instanced_scenes_build_the_same_on_the_device(device, seed):
    prep = S.instanced_scene(seed, meshes=6, instances=30, tris_per_mesh=(1, 3000))
    say(f"instanced seed {seed}: {len(prep.leaves)} triangles")
    same_trees(*both(device, lambda: T.Scene(device, *prep.desc())), f"seed {seed}")


def test_garbage_in_builds_the_same_on_the_device(device):
    """Each of the reference's garbage scenarios as a mesh of its own, all in one scene (so one device build takes
    them together). The reference's mesh builder refuses corners that are not finite, so those two cases ("mixed",
    "all bad") are made from finite meshes of the same size whose leaves are then spoiled as the scenario spoils
    its corners (a caller's leaves are not checked against its tree). Meshes whose every split would cost more than
    a float holds (coordinates near 3e38), or that have no finite triangle, are left to the host, which splits
    them at the median."""
    # (near the limit at half the scenario's 3e38, so that the unit's padded box stays finite for the top tree)
    cases = [(w, t * np.float32(0.5) if w == "near the limit" else t) for w, t in garbage_cases()]
    finite = [np.nan_to_num(t, nan=0.0, posinf=1.0, neginf=-1.0) for _, t in cases]
    prep = S.prepare([S.Mesh(t) for t in finite],
                     [S.Placement(k, S.translation([3.0 * k, 0.0, 0.0])) for k in range(len(cases))])
    leaves = prep.leaves.copy()
    for k, (what, tris) in enumerate(cases):
        if np.isfinite(tris).all():
            continue
        first, order = prep.mesh_first[k], prep.mesh_order[k]
        bad = ~np.isfinite(tris[order]).all(axis=(1, 2))
        leaves[first + np.nonzero(bad)[0], 0, 0] = np.float32(np.nan)
        say(f"  {what}: {int(bad.sum())} of {len(tris)} leaves spoiled")
    prep.leaves = leaves
    say(f"garbage: {[(w, len(t)) for w, t in cases]}")
    h, d = both(device, lambda: T.Scene(device, *prep.desc()))
    host = d.stats()["host_meshes"]
    assert 1 <= host <= 2  # "all bad" (nothing finite), and "near the limit" if its splits overflow
    same_trees(h, d, "garbage", host)


def test_a_clustered_scene_builds_the_same_on_the_device(device):
    prep, _, _ = S.precision_scene(200_000, seed=2)
    say("clustered 200k")
    same_trees(*both(device, lambda: T.Scene(device, *prep.desc())), "clustered 200k")


def test_render_scenes_build_the_same_on_the_device(device):
    rs, _ = R.cornell()
    say("cornell")
    same_trees(*both(device, lambda: R.upload(device, rs)), "cornell")
    rs, _, _ = R.city(3, meshes=12, instances=300)
    say("city 12 meshes x 300")
    same_trees(*both(device, lambda: R.upload(device, rs)), "city")
