"""The oracle: ctypes bindings to liboracle.so, the plain-C CPU restatement of ThreeRT's software path-trace pass.

Test infrastructure only. It is imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
always as the checker (or the timed CPU restatement), never by the threert_amd product.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NODE = np.dtype([("bmin", "<f4", 3), ("left_or_first", "<u4"), ("bmax", "<f4", 3), ("count", "<u4")])
INSTANCE = np.dtype([("m", "<f4", (3, 4)), ("n", "<f4", (3, 4)), ("info", "<u4", 4), ("p", "<f4", (3, 4))])
RAY = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("mode", "<u4")])
HIT = np.dtype([("t", "<f4"), ("tri", "<u4"), ("u", "<f4"), ("v", "<f4"), ("inst", "<u4"), ("part", "<u4")])
HIT64 = np.dtype([("t", "<f8"), ("u", "<f8"), ("v", "<f8"), ("tri", "<u4"), ("inst", "<u4")])
READS = np.dtype([("top_nodes", "<u8"), ("mesh_nodes", "<u8"), ("mesh_verts", "<u8"), ("records", "<u8"),
                  ("steps", "<u8"), ("deepest", "<u8")])

WALK_REF = 0
WALK_EXACT = 1
WALK_WT = 2  # EXACT with the watertight triangle test
INF = np.float32(3.0e38)
NO_HIT = 0xFFFFFFFF


class _Tree(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint64), ("order", C.POINTER(C.c_uint32)),
                ("order_count", C.c_uint64), ("depth", C.c_uint32), ("lo", C.c_float * 3), ("hi", C.c_float * 3)]


class _Scene(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint64), ("leaves", C.c_void_p),
                ("leaf_count", C.c_uint64), ("instances", C.c_void_p), ("instance_count", C.c_uint64),
                ("top", C.c_uint32), ("units", C.c_uint32)]


def build() -> str:
    r = subprocess.run(["make", "-B", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the oracle failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


NATIVE_PATH = os.path.join(HERE, "_native", "liboracle.so")


@contextlib.contextmanager
def native():
    """The oracle built for the CPU this runs on (-march=native, contraction still off: the same bits), compiled here
    and used for the calls inside the block: the CPU baseline's copy."""
    global _lib, LIB_PATH
    r = subprocess.run(["make", "-B", "-C", HERE, "native"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the native oracle failed:\n{r.stdout}\n{r.stderr}")
    saved = (_lib, LIB_PATH)
    _lib, LIB_PATH = None, NATIVE_PATH
    try:
        yield lib()
    finally:
        _lib, LIB_PATH = saved


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, U64, U32, I = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
        for name, args in [("or_build_tris", [P, U64, U32, P]), ("or_build_median", [P, U64, U32, P]),
                           ("or_build_sah", [P, U64, U32, P]), ("or_mesh_tree", [P, U64, P]),
                           ("or_top_tree", [P, P, U64, P])]:
            getattr(L, name).argtypes = args
            getattr(L, name).restype = I
        L.or_tree_free.argtypes = [P]
        L.or_check_tree.argtypes = [P, U64, U32, U64, U32, C.c_char_p, U64]
        L.or_check_tree.restype = I
        L.or_check_top.argtypes = [P, U64, U32, C.c_char_p, U64]
        L.or_check_top.restype = I
        L.or_unit_box.argtypes = [P, P, P, P]
        L.or_unit_record.argtypes = [P, U32, I]
        L.or_top_cost.argtypes = [P, U64]
        L.or_top_cost.restype = C.c_double
        L.or_ray_aabb.argtypes = [P, P, P, P, C.c_float]
        L.or_ray_aabb.restype = C.c_float
        L.or_intersect_tri.argtypes = [P, P, P, C.c_float, P]
        L.or_inv.argtypes = [P, P]
        L.or_trace.argtypes = [P, P, P, C.c_float, U32, U32, U64, P, P]
        L.or_trace.restype = I
        L.or_brute.argtypes = [P, P, P, C.c_float, U32, U32, P]
        L.or_brute.restype = I
        L.or_trace_batch.argtypes = [P, P, U64, U32, I, P, P]
        L.or_trace_batch.restype = C.c_int64
        L.or_brute_batch.argtypes = [P, P, U64, U32, I, P]
        L.or_brute_batch.restype = C.c_int64
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _take_tree(t: _Tree):
    nodes = np.frombuffer(C.string_at(t.nodes, t.node_count * 32), dtype=NODE).copy()
    order = np.ctypeslib.as_array(t.order, shape=(t.order_count,)).copy() if t.order_count else np.zeros(0, np.uint32)
    out = {"nodes": nodes, "order": order.astype(np.uint32), "depth": int(t.depth),
           "lo": np.array(t.lo[:], np.float32), "hi": np.array(t.hi[:], np.float32)}
    lib().or_tree_free(C.byref(t))
    return out


def build_tris(tris: np.ndarray, max_depth: int = 48):
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    t = _Tree()
    if lib().or_build_tris(_ptr(tris), len(tris), max_depth, C.byref(t)):
        raise ValueError(f"{len(tris)} items do not fit a tree {max_depth} deep")
    return _take_tree(t)


def build_median(boxes: np.ndarray, max_leaf: int):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    t = _Tree()
    lib().or_build_median(_ptr(boxes), len(boxes), max_leaf, C.byref(t))
    return _take_tree(t)


def build_sah(boxes: np.ndarray, max_depth: int):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    t = _Tree()
    lib().or_build_sah(_ptr(boxes), len(boxes), max_depth, C.byref(t))
    return _take_tree(t)


def mesh_tree(tris: np.ndarray):
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    t = _Tree()
    if lib().or_mesh_tree(_ptr(tris), len(tris), C.byref(t)):
        raise ValueError("no mesh tree (no triangles, or a vertex that is not finite)")
    return _take_tree(t)


def top_tree(boxes: np.ndarray, records: np.ndarray):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    records = np.ascontiguousarray(records, dtype=np.uint32)
    t = _Tree()
    if lib().or_top_tree(_ptr(boxes), _ptr(records), len(boxes), C.byref(t)):
        raise ValueError("no top tree")
    return _take_tree(t)


def check_tree(nodes: np.ndarray, max_depth: int, items: int, max_leaf: int = 16):
    nodes = np.ascontiguousarray(nodes, dtype=NODE)
    msg = C.create_string_buffer(256)
    d = lib().or_check_tree(_ptr(nodes), len(nodes), max_depth, items, max_leaf, msg, 256)
    return d if d >= 0 else msg.value.decode()


def check_top(nodes: np.ndarray, records: int):
    nodes = np.ascontiguousarray(nodes, dtype=NODE)
    msg = C.create_string_buffer(256)
    d = lib().or_check_top(_ptr(nodes), len(nodes), records, msg, 256)
    return d if d >= 0 else msg.value.decode()


def top_cost(nodes: np.ndarray) -> float:
    nodes = np.ascontiguousarray(nodes, dtype=NODE)
    return float(lib().or_top_cost(_ptr(nodes), len(nodes)))


def unit_box(world_colmajor: np.ndarray, lo, hi) -> np.ndarray:
    w = np.ascontiguousarray(world_colmajor, dtype=np.float32).reshape(16)
    lo = np.ascontiguousarray(lo, dtype=np.float32)
    hi = np.ascontiguousarray(hi, dtype=np.float32)
    out = np.zeros(6, np.float32)
    lib().or_unit_box(_ptr(w), _ptr(lo), _ptr(hi), _ptr(out))
    return out


def unit_record(rec: np.ndarray, root: int, cast_shadow: bool) -> None:
    """In place on a one-element INSTANCE array."""
    assert rec.dtype == INSTANCE and rec.flags.c_contiguous
    lib().or_unit_record(_ptr(rec), root, 1 if cast_shadow else 0)


def ray_aabb(o, inv, bmin, bmax, tmax) -> float:
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, inv, bmin, bmax)]
    return float(np.float32(lib().or_ray_aabb(*[_ptr(x) for x in a], C.c_float(tmax))))


def intersect_tri(o, d, leaf: np.ndarray, tmax) -> np.ndarray:
    o = np.ascontiguousarray(o, dtype=np.float32)
    d = np.ascontiguousarray(d, dtype=np.float32)
    leaf = np.ascontiguousarray(leaf, dtype=np.float32).reshape(12)
    out = np.zeros(3, np.float32)
    lib().or_intersect_tri(_ptr(o), _ptr(d), _ptr(leaf), C.c_float(tmax), _ptr(out))
    return out


class Scene:
    """The kernel's buffers (nodes, leaves, records, P.motion.z/w) as the oracle walks them."""

    def __init__(self, nodes: np.ndarray, leaves: np.ndarray, instances: np.ndarray, top: int, units: int):
        self.nodes = np.ascontiguousarray(nodes, dtype=NODE)
        self.leaves = np.ascontiguousarray(leaves, dtype=np.float32).reshape(-1, 3, 4)
        self.instances = np.ascontiguousarray(instances, dtype=INSTANCE)
        self._s = _Scene(_ptr(self.nodes), len(self.nodes), _ptr(self.leaves), len(self.leaves),
                         _ptr(self.instances), len(self.instances), top, units)

    def trace(self, o, d, tmax=INF, mode=0, variant=WALK_EXACT):
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        hit = np.zeros(1, HIT)
        reads = np.zeros(1, READS)
        rc = lib().or_trace(C.byref(self._s), _ptr(o), _ptr(d), C.c_float(tmax), mode, variant, 0, _ptr(hit),
                            _ptr(reads))
        if rc:
            raise RuntimeError("the walk read outside its buffers or did not stop")
        return hit[0], reads[0]

    def brute(self, o, d, tmax=INF, mode=0, variant=WALK_EXACT):
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        hit = np.zeros(1, HIT)
        if lib().or_brute(C.byref(self._s), _ptr(o), _ptr(d), C.c_float(tmax), mode, variant, _ptr(hit)):
            raise RuntimeError("brute force read outside its buffers")
        return hit[0]

    def trace_batch(self, rays: np.ndarray, variant=WALK_EXACT, threads=0):
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), HIT)
        reads = np.zeros(1, READS)
        failed = lib().or_trace_batch(C.byref(self._s), _ptr(rays), len(rays), variant, threads, _ptr(hits),
                                      _ptr(reads))
        if failed:
            raise RuntimeError(f"{failed} walks read outside their buffers or did not stop")
        return hits, reads[0]

    def brute64_batch(self, rays: np.ndarray, threads=0) -> np.ndarray:
        """The closest hit in double precision by brute force (or_brute64_batch): HIT64 records, t -1 and tri
        NO_TRI for a miss; ties to the lowest (instance, triangle)."""
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), HIT64)
        L = lib()
        L.or_brute64_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_void_p]
        L.or_brute64_batch.restype = C.c_int64
        if L.or_brute64_batch(C.byref(self._s), _ptr(rays), len(rays), threads, _ptr(hits)):
            raise RuntimeError("FP64 brute force read outside its buffers")
        return hits

    def brute_batch(self, rays: np.ndarray, variant=WALK_EXACT, threads=0):
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), HIT)
        failed = lib().or_brute_batch(C.byref(self._s), _ptr(rays), len(rays), variant, threads, _ptr(hits))
        if failed:
            raise RuntimeError(f"{failed} brute-force rays read outside their buffers")
        return hits


# ---- the offline path loop (render.c) -----------------------------------------------------------------------------
class _RenderScene(C.Structure):
    _fields_ = [("geo", _Scene), ("vertices", C.c_void_p), ("vertex_count", C.c_uint64), ("indices", C.c_void_p),
                ("tri_count", C.c_uint64), ("rows", C.c_void_p), ("row_count", C.c_uint64),
                ("materials", C.c_void_p), ("material_count", C.c_uint64), ("lights", C.c_void_p),
                ("light_count", C.c_uint32), ("emitters", C.c_void_p), ("emitter_vec4s", C.c_uint64),
                ("emitter_tris", C.c_void_p), ("emitter_tri_count", C.c_uint64), ("objects", C.c_void_p),
                ("object_count", C.c_uint64), ("energy", C.c_void_p), ("light_tree", C.c_void_p),
                ("light_tree_count", C.c_uint64), ("glass_energy", C.c_void_p), ("tex_pages", C.c_void_p),
                ("tex_side", C.c_uint32), ("tex_layers", C.c_uint32), ("tex_info", C.c_void_p),
                ("tex_info_vec4s", C.c_uint64), ("env", C.c_void_p), ("env_w", C.c_uint32), ("env_h", C.c_uint32),
                ("env_cdf", C.c_void_p), ("env_cdf_count", C.c_uint64), ("vcolors", C.c_void_p),
                ("vcolor_words", C.c_uint64)]


class RenderScene:
    """The kernel's buffers plus the shading tables, as the oracle's path loop reads them."""

    def __init__(self, nodes, leaves, instances, top, units, vertices, indices, rows, materials, lights, emitters,
                 emitter_tris, objects, energy, light_tree=None, glass_energy=None, tex_pages=None, tex_side=0,
                 tex_layers=0, tex_info=None, env=None, env_cdf=None, vcolors=None):
        self.geo = Scene(nodes, leaves, instances, top, units)
        if light_tree is None:
            light_tree = np.zeros((0, 4), np.float32)
        if glass_energy is None:
            glass_energy = np.zeros((0, 4), np.float32)
        elif np.asarray(glass_energy).shape != (1024, 4):
            raise ValueError("glass_energy must be the (1024, 4) float32 table")
        self.keep = [np.ascontiguousarray(x) for x in (vertices, indices, rows, materials, lights, emitters,
                                                       emitter_tris, objects, energy, light_tree)]
        self.keep.append(np.ascontiguousarray(glass_energy, np.float32))
        if tex_info is None:
            tex_pages, tex_info = np.zeros(0, np.uint8), np.zeros((0, 4), np.float32)
        else:
            layer = sum((tex_side >> k) ** 2 * 4 for k in range(5))
            if np.asarray(tex_pages).nbytes != layer * tex_layers:
                raise ValueError("tex_pages must hold tex_layers layers of 5 RGBA8 levels of tex_side >> level")
        self.keep += [np.ascontiguousarray(tex_pages, np.uint8).reshape(-1),
                      np.ascontiguousarray(tex_info, np.float32).reshape(-1, 4)]
        # env: (h, w, 4) uint16 half-float texels; env_cdf: h + w h floats
        if env is None:
            env, env_cdf = np.zeros((0, 0, 4), np.uint16), np.zeros(0, np.float32)
        else:
            env = np.asarray(env)
            if env.dtype != np.uint16 or env.ndim != 3 or env.shape[2] != 4:
                raise ValueError("env must be (h, w, 4) uint16 half-float texels")
            if np.asarray(env_cdf).size != env.shape[0] * (1 + env.shape[1]):
                raise ValueError("env_cdf must hold h + w h floats")
        self.keep += [np.ascontiguousarray(env, np.uint16), np.ascontiguousarray(env_cdf, np.float32).reshape(-1)]
        # vcolors: 3 RGBA8 words a triangle in the leaves' order
        self.keep.append(np.zeros(0, np.uint32) if vcolors is None else np.ascontiguousarray(vcolors, np.uint32).reshape(-1))
        v, i, r, m, li, e, et, o, en, lt, ge, tp, ti, ev, ec, vc = self.keep
        self._s = _RenderScene(self.geo._s, _ptr(v), len(v), _ptr(i), len(i), _ptr(r), len(r), _ptr(m), len(m),
                               _ptr(li), len(li), _ptr(e), len(e), _ptr(et), len(et), _ptr(o), len(o) // 3,
                               _ptr(en), _ptr(lt) if len(lt) else None, len(lt), _ptr(ge) if len(ge) else None,
                               _ptr(tp) if len(ti) else None, tex_side, tex_layers, _ptr(ti) if len(ti) else None,
                               len(ti), _ptr(ev) if ev.size else None, ev.shape[1], ev.shape[0],
                               _ptr(ec) if ev.size else None, len(ec), _ptr(vc) if vc.size else None, len(vc))
        L = lib()
        L.or_render_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int]
        L.or_render_sample.restype = C.c_int64
        L.or_emitter_draw.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_int64)]
        L.or_emitter_draw.restype = C.c_float
        L.or_emitter_pmf.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
        L.or_emitter_pmf.restype = C.c_float
        L.or_glass_energy.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_int]
        L.or_glass_energy.restype = C.c_float
        L.or_glass_share.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float]
        L.or_glass_share.restype = C.c_float
        L.or_tex.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_void_p]
        L.or_tex.restype = C.c_int
        L.or_env.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_env.restype = C.c_int

    def emitter_draw(self, params: np.ndarray, x, n, u) -> tuple[int, float]:
        """pathtrace.wgsl:547-585 emitter_draw at x (normal n) for uniforms u[3]: (CDF entry, p); entry -1: none."""
        p = np.ascontiguousarray(params, np.uint32)
        xs, ns, us = (np.ascontiguousarray(a, np.float32) for a in (x, n, u))
        e = C.c_int64()
        pr = lib().or_emitter_draw(C.byref(self._s), _ptr(p), _ptr(xs), _ptr(ns), _ptr(us), C.byref(e))
        if pr < 0:
            raise RuntimeError("emitter_draw read outside its tables")
        return int(e.value), float(pr)

    def emitter_pmf(self, params: np.ndarray, x, n, ei: int) -> float:
        """pathtrace.wgsl:590-618 emitter_pmf: the probability emitter_draw at x (normal n) picks emitter ei."""
        p = np.ascontiguousarray(params, np.uint32)
        xs, ns = (np.ascontiguousarray(a, np.float32) for a in (x, n))
        pr = lib().or_emitter_pmf(C.byref(self._s), _ptr(p), _ptr(xs), _ptr(ns), ei)
        if pr < 0:
            raise RuntimeError("emitter_pmf read outside its tables")
        return float(pr)

    def glass_energy(self, nov: float, roughness: float, ior: float, leaving: bool) -> float:
        """pathtrace.wgsl:1582-1603 glass_energy: R + T of the rough interface, interpolated in the table."""
        if len(self.keep[10]) == 0:
            raise ValueError("the scene has no glass table")
        r = lib().or_glass_energy(C.byref(self._s), nov, roughness, ior, int(bool(leaving)))
        if r < 0:
            raise RuntimeError("glass_energy read outside its table")
        return float(r)

    def tex(self, tex_id: int, u: float, v: float, lod: float) -> np.ndarray:
        """pathtrace.wgsl:1114-1117 tex: texture tex_id at (u, v), lod = log2 of the footprint in its UV units."""
        out = np.zeros(4, np.float32)
        if lib().or_tex(C.byref(self._s), tex_id, u, v, lod, _ptr(out)) < 0:
            raise RuntimeError("tex read outside its tables")
        return out

    def env(self, params: np.ndarray, d, u=(0.5, 0.5, 0.5, 0.5)) -> tuple[np.ndarray, float, np.ndarray]:
        """pathtrace.wgsl:1895-1941 under `params`: env_radiance and env_pdf along d, and sample_env's direction for the
        uniforms u[4]."""
        p = np.ascontiguousarray(params, np.uint32)
        ds, us = (np.ascontiguousarray(a, np.float32) for a in (d, u))
        rad, smp, pdf = np.zeros(3, np.float32), np.zeros(3, np.float32), C.c_float()
        if lib().or_env(C.byref(self._s), _ptr(p), _ptr(ds), _ptr(us), _ptr(rad), C.byref(pdf), _ptr(smp)) < 0:
            raise RuntimeError("the environment's reads fell outside its tables")
        return rad, float(pdf.value), smp

    def glass_share(self, params: np.ndarray, material: int, cos_d: float) -> float:
        """pathtrace.wgsl:930-970 glass_at's share for `material` at |cos| cos_d, under `params`."""
        p = np.ascontiguousarray(params, np.uint32)
        r = lib().or_glass_share(C.byref(self._s), _ptr(p), material, cos_d)
        if r < 0:
            raise RuntimeError("glass_share read outside its tables")
        return float(r)

    def render(self, params: np.ndarray, first_sample: int, spp: int, accum: np.ndarray | None = None,
               rows=None, threads=0) -> np.ndarray:
        """Samples first..first+spp-1 of rows [y0, y1) added to `accum` (h, w, 20), sample after sample."""
        p = np.ascontiguousarray(params, np.uint32).copy()
        w, h = int(p[16]), int(p[17])
        if accum is None:
            accum = np.zeros((h, w, 20), np.float32)
        y0, y1 = (0, h) if rows is None else rows
        for s in range(first_sample, first_sample + spp):
            p[18] = s
            failed = lib().or_render_sample(C.byref(self._s), _ptr(p), y0, y1, _ptr(accum), threads)
            if failed:
                raise RuntimeError(f"{failed} walks read outside their buffers")
        return accum


def log2(x: float) -> float:
    """The kernel's log2 (Cephes log2f, render.c or_log2)."""
    L = lib()
    L.or_log2.argtypes, L.or_log2.restype = [C.c_float], C.c_float
    return float(L.or_log2(x))


def exp2(x: float) -> float:
    """The kernel's exp2 (Cephes exp2f, render.c or_exp2)."""
    L = lib()
    L.or_exp2.argtypes, L.or_exp2.restype = [C.c_float], C.c_float
    return float(L.or_exp2(x))


def rng4_at(pixel_seed: int, index: int, d: int) -> np.ndarray:
    """sample.rs:78-82 rng4 (the oracle's restatement of sample.wgsl)."""
    L = lib()
    L.or_rng4_at.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    out = np.zeros(4, np.float32)
    L.or_rng4_at(pixel_seed, index, d, _ptr(out))
    return out


def sobol_dirs() -> np.ndarray:
    L = lib()
    L.or_sobol_dirs.argtypes = [C.c_void_p]
    out = np.zeros(96, np.uint32)
    L.or_sobol_dirs(_ptr(out))
    return out.reshape(3, 32)


def hash_hp(x: int) -> int:
    L = lib()
    L.or_hash_hp.argtypes = [C.c_uint32]
    L.or_hash_hp.restype = C.c_uint32
    return int(L.or_hash_hp(x))


def denoise_atrous(color, albedo, normal, variance, iterations: int = 4) -> np.ndarray:
    """denoise.rs:10-74 restated (render.c's sibling denoise.c)."""
    L = lib()
    L.or_denoise_atrous.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_uint32, C.c_void_p]
    h, w = color.shape[:2]
    a = [np.ascontiguousarray(x, np.float32) for x in (color, albedo, normal, variance)]
    out = np.zeros((h, w, 3), np.float32)
    L.or_denoise_atrous(w, h, *[_ptr(x) for x in a], iterations, _ptr(out))
    return out


def expf(x: float) -> float:
    L = lib()
    L.or_expf.argtypes = [C.c_float]
    L.or_expf.restype = C.c_float
    return float(L.or_expf(C.c_float(x)))


def sincos(x: float):
    L = lib()
    L.or_sincos.argtypes = [C.c_float, C.c_void_p, C.c_void_p]
    s, c = C.c_float(), C.c_float()
    L.or_sincos(C.c_float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def atan2(y: float, x: float) -> float:
    """The kernel's atan2 (Cephes atanf, render.c or_atan2)."""
    L = lib()
    L.or_atan2.argtypes, L.or_atan2.restype = [C.c_float, C.c_float], C.c_float
    return float(L.or_atan2(y, x))


def acos(x: float) -> float:
    """The kernel's acos (Cephes asinf, render.c or_acos)."""
    L = lib()
    L.or_acos.argtypes, L.or_acos.restype = [C.c_float], C.c_float
    return float(L.or_acos(x))
