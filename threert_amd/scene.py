"""Scenes as the reference's prepare step lays them out for software traversal, and the seeded generators.

`prepare` follows lib.rs:530-594 / 836-913 (prepare_as with units): every mesh gets a tree (units.rs:231
mesh_tree) placed side by side in the node buffer (units.rs:532-595 Layout: mesh range first, the top tree after
it), its triangles appended in the tree's order as 48-byte leaves (indexed.rs:23 leaf), every instance a unit with
a 160-byte record (scene.rs:239 Instance::gpu, units.rs:329 unit_record), and the top tree built over the units'
world boxes (units.rs:288 UnitBox::new, units.rs:303 top_tree). The trees come from the library's host builders.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import INSTANCE, NODE, RAY, mesh_tree, top_tree, unit_box, unit_record

FILL = np.array([((0, 0, 0), 0, (0, 0, 0), 1)], dtype=NODE)  # units.rs:43 FILL


@dataclass
class Mesh:
    tris: np.ndarray                 # (n, 3, 3) float32, model space
    sections: np.ndarray | None = None  # (n,) u32 section of each triangle (its material row); 0 when None


@dataclass
class Placement:
    mesh: int
    world: np.ndarray                # (4, 4) float32, row-major: p_world = world @ (p, 1)
    cast_shadow: bool = True
    object: int = 0
    bits: int = 0                    # scene.rs Instance::bits: INST_CUTOUT (16) | INST_GLASS (32) of its rows


@dataclass
class Prepared:
    nodes: np.ndarray
    leaves: np.ndarray               # (n, 3, 4) float32
    instances: np.ndarray
    top_first: int
    units: int
    mesh_first: list = field(default_factory=list)   # each mesh's first leaf
    mesh_root: list = field(default_factory=list)    # each mesh's tree root
    mesh_order: list = field(default_factory=list)   # each mesh's tree order
    depths: list = field(default_factory=list)

    def desc(self):
        return self.nodes, self.leaves, self.instances, self.top_first, self.units


def leaves_of(tris: np.ndarray, sections: np.ndarray, index: np.ndarray) -> np.ndarray:
    """indexed.rs:23-26 leaf: (v0, section bits), (v1 - v0, index bits), (v2 - v0, 0)."""
    n = len(tris)
    out = np.zeros((n, 3, 4), np.float32)
    out[:, 0, :3] = tris[:, 0]
    out[:, 1, :3] = tris[:, 1] - tris[:, 0]
    out[:, 2, :3] = tris[:, 2] - tris[:, 0]
    out[:, 0, 3] = sections.astype(np.uint32).view(np.float32)
    out[:, 1, 3] = index.astype(np.uint32).view(np.float32)
    return out


def section_rank(sections: np.ndarray) -> np.ndarray:
    """Each triangle's index in its section, in the mesh's triangle order (the leaf's second id)."""
    sections = np.asarray(sections, np.uint32)
    rank = np.zeros(len(sections), np.uint32)
    for s in np.unique(sections):
        idx = np.nonzero(sections == s)[0]
        rank[idx] = np.arange(len(idx), dtype=np.uint32)
    return rank


def instance_record(world: np.ndarray, first: int, obj: int = 0, bits: int = 0) -> np.ndarray:
    """scene.rs:239-246 Instance::gpu: m rows, normal rows (M3^-1)^T with the bits words, info, p rows.
    (The inverse is taken in f64 and rounded; glam's f32 inverse is not reproduced.)"""
    w = np.asarray(world, np.float64)
    r = np.zeros(1, INSTANCE)
    r["m"][0] = w[:3, :].astype(np.float32)
    nm = np.linalg.inv(w[:3, :3]).T
    r["n"][0, :, :3] = nm.astype(np.float32)
    mirrored = 4 if np.linalg.det(w[:3, :3]) < 0 else 0
    r["n"][0, 1, 3] = np.array([mirrored | (bits & (16 | 32))], np.uint32).view(np.float32)[0]
    r["info"][0] = (first, obj, 0, 0)
    r["p"][0] = r["m"][0]
    return r


def prepare(meshes: list[Mesh], placements: list[Placement], mesh_room: int | None = None) -> Prepared:
    trees, firsts, roots, orders, depths = [], [], [], [], []
    leaf_parts = []
    first = 0
    for m in meshes:
        tris = np.ascontiguousarray(m.tris, np.float32).reshape(-1, 3, 3)
        t = mesh_tree(tris)
        order = t["order"]
        sec = np.zeros(len(tris), np.uint32) if m.sections is None else np.asarray(m.sections, np.uint32)
        leaf_parts.append(leaves_of(tris[order], sec[order], section_rank(sec)[order]))
        trees.append(t)
        firsts.append(first)
        orders.append(order)
        depths.append(t["depth"])
        first += len(tris)
    mesh_nodes = sum(len(t["nodes"]) for t in trees)
    cap = mesh_nodes if mesh_room is None else max(mesh_room, mesh_nodes)
    units = len(placements)
    top_cap = 2 * max(1, units) - 1
    nodes = np.repeat(FILL, cap + top_cap)
    at = 0
    for t in trees:
        roots.append(at)
        nodes[at:at + len(t["nodes"])] = t["nodes"]
        at += len(t["nodes"])
    instances = np.zeros(max(1, units), INSTANCE)
    boxes = np.zeros((units, 6), np.float32)
    for k, p in enumerate(placements):
        rec = instance_record(p.world, firsts[p.mesh], p.object, p.bits)
        unit_record(rec, roots[p.mesh], p.cast_shadow)
        instances[k] = rec[0]
        t = trees[p.mesh]
        boxes[k] = unit_box(np.asarray(p.world, np.float32).T.copy(), t["lo"], t["hi"])
    top_first = cap
    if units:
        top = top_tree(boxes, np.arange(units, dtype=np.uint32))
        nodes[top_first:top_first + len(top["nodes"])] = top["nodes"]
    leaves = np.concatenate(leaf_parts) if leaf_parts else np.zeros((0, 3, 4), np.float32)
    return Prepared(nodes, leaves, instances, top_first, units, firsts, roots, orders, depths)


# ---- transforms ---------------------------------------------------------------------------------------------------
def translation(t) -> np.ndarray:
    w = np.eye(4, dtype=np.float32)
    w[:3, 3] = t
    return w


def trs(t, axis_angles, s) -> np.ndarray:
    """Translation, rotation (XYZ Euler angles, radians), scale: world = T R S."""
    ax, ay, az = axis_angles
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    w = np.eye(4)
    w[:3, :3] = rx @ ry @ rz @ np.diag(s)
    w[:3, 3] = t
    return w.astype(np.float32)


# ---- PCG32 (O'Neill, pcg32_random_r / pcg32_srandom_r), vectorised over independent streams ---------------------
_MULT = np.uint64(6364136223846793005)


class PCG32:
    def __init__(self, seed: int, streams: np.ndarray):
        streams = np.asarray(streams, np.uint64)
        self.inc = (streams << np.uint64(1)) | np.uint64(1)
        self.state = np.zeros_like(streams)
        self.next()
        self.state = self.state + np.uint64(seed)
        self.next()

    def next(self) -> np.ndarray:
        old = self.state
        self.state = old * _MULT + self.inc
        xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
        rot = (old >> np.uint64(59)).astype(np.uint32)
        return (xs >> rot) | (xs << ((np.uint32(32) - rot) & np.uint32(31)))

    def uniform(self) -> np.ndarray:
        """[0, 1) with 24 bits, float32."""
        return (self.next() >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / (1 << 24))


def incoherent_rays(n: int, lo, hi, seed: int = 1, first: int = 0) -> np.ndarray:
    """Closest-hit rays: origins uniform in the box [lo, hi], directions uniform on the sphere (z = 1 - 2u,
    phi = 2 pi v), tmax = 3e38. Ray i draws from PCG32(seed, stream first + i): u_ox, u_oy, u_oz, u, v."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    g = PCG32(seed, np.arange(first, first + n, dtype=np.uint64))
    rays = np.zeros(n, RAY)
    for k in range(3):
        rays["o"][:, k] = lo[k] + g.uniform() * (hi[k] - lo[k])
    u = g.uniform().astype(np.float64)
    v = g.uniform().astype(np.float64)
    z = 1.0 - 2.0 * u
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * v
    rays["d"][:, 0] = (r * np.cos(phi)).astype(np.float32)
    rays["d"][:, 1] = (r * np.sin(phi)).astype(np.float32)
    rays["d"][:, 2] = z.astype(np.float32)
    rays["tmax"] = np.float32(3.0e38)
    rays["mode"] = 0
    return rays


# ---- seeded scenes ------------------------------------------------------------------------------------------------
def random_triangles(rng: np.random.Generator, n: int, spread: float, size: float) -> np.ndarray:
    c = (rng.random((n, 1, 3), dtype=np.float32) - 0.5) * spread
    return (c + (rng.random((n, 3, 3), dtype=np.float32) - 0.5) * size).astype(np.float32)


def clustered_soup(n_tris: int, seed: int = 1, clusters: int | None = None, extent: float = 100.0,
                   median_area: float = 0.01, sigma: float = 1.5, emissive_share: float = 0.05):
    """The precision scene: `n_tris` triangles in clustered meshes, areas log-normal (median `median_area`,
    sigma `sigma`), a share of them emissive at 1-200 W/m^2. Returns (tris (n,3,3), emissive (n,) float32)."""
    rng = np.random.default_rng(seed)
    if clusters is None:
        clusters = max(1, n_tris // 1000)
    centres = (rng.random((clusters, 3)) - 0.5) * extent
    radii = 10.0 ** rng.uniform(-0.3, 0.7, clusters)
    which = rng.integers(0, clusters, n_tris)
    cen = centres[which] + rng.normal(size=(n_tris, 3)) * (0.5 * radii[which])[:, None]
    shape = rng.normal(size=(n_tris, 3, 3))
    e1 = shape[:, 1] - shape[:, 0]
    e2 = shape[:, 2] - shape[:, 0]
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)
    target = median_area * np.exp(sigma * rng.normal(size=n_tris))
    scale = np.sqrt(target / np.maximum(area, 1e-12))
    shape = shape - shape.mean(axis=1, keepdims=True)
    tris = (cen[:, None, :] + shape * scale[:, None, None]).astype(np.float32)
    emissive = np.where(rng.random(n_tris) < emissive_share, rng.uniform(1.0, 200.0, n_tris), 0.0).astype(np.float32)
    return tris, emissive


def precision_scene(n_tris: int = 3_000_000, seed: int = 1) -> tuple[Prepared, np.ndarray, np.ndarray]:
    """A single-level scene (one unit, identity placement) over `clustered_soup`. Returns the prepared scene and
    its world box (lo, hi)."""
    tris, _ = clustered_soup(n_tris, seed)
    prep = prepare([Mesh(tris)], [Placement(0, np.eye(4, dtype=np.float32))])
    lo = tris.reshape(-1, 3).min(axis=0)
    hi = tris.reshape(-1, 3).max(axis=0)
    return prep, lo, hi


def instanced_scene(seed: int, meshes: int = 4, instances: int = 20, tris_per_mesh=(1, 300)):
    """Random meshes placed by random transforms (a third translation only, some identity), as the reference's
    random-edit test builds its scenes (units.rs:1391-1405 pose)."""
    rng = np.random.default_rng(seed)
    ms = []
    for _ in range(meshes):
        n = int(rng.integers(tris_per_mesh[0], tris_per_mesh[1] + 1))
        if rng.random() < 0.2:
            tris = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[1, 0, 0], [1, 1, 0], [0, 1, 0]]], np.float32)
        else:
            tris = random_triangles(rng, n, 4.0, 1.5)
        ms.append(Mesh(tris, sections=rng.integers(0, 3, len(tris)).astype(np.uint32)))
    ps = []
    for _ in range(instances):
        kind = rng.random()
        if kind < 0.05:
            w = np.eye(4, dtype=np.float32)
        elif kind < 0.35:
            w = translation((rng.random(3) - 0.5) * 30.0)
        else:
            w = trs((rng.random(3) - 0.5) * 30.0, rng.random(3) * 6.0, 0.3 + 2.0 * rng.random(3))
        ps.append(Placement(int(rng.integers(0, meshes)), w, bool(rng.random() < 0.8)))
    return prepare(ms, ps)
