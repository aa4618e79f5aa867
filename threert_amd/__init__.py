"""threert_amd: an MI355X-native (gfx950 HIP) core for ThreeRT's software path-trace pass.

This module is the host-side mirror of the reference's interface for the path, over the library's C ABI
(include/threert.h, libthreert.so built in place by threert_amd/build.py):

  Device           gpu.rs:2113 Gpu::new                    (one HIP device and its stream)
  Scene            gpu.rs:2613 Gpu::upload                 (the prepared node / leaf / record buffers in HBM)
  Scene.trace      pathtrace.wgsl:764 trace, units.rs:1019 Kernel::trace, for a batch of rays
  mesh_tree, top_tree, unit_box, unit_record               units.rs builders (lib.rs:836-913 unit_trees)

Every call goes to the HIP library; there is no CPU path. Importing works without a GPU (the builders are host
code); creating a Device needs one and raises when there is none or the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libthreert.so")

# Byte formats (include/threert.h).
NODE = np.dtype([("bmin", "<f4", 3), ("left_or_first", "<u4"), ("bmax", "<f4", 3), ("count", "<u4")])
INSTANCE = np.dtype([("m", "<f4", (3, 4)), ("n", "<f4", (3, 4)), ("info", "<u4", 4), ("p", "<f4", (3, 4))])
RAY = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("mode", "<u4")])
HIT = np.dtype([("t", "<f4"), ("tri", "<u4"), ("u", "<f4"), ("v", "<f4"), ("inst", "<u4"), ("part", "<u4")])
assert NODE.itemsize == 32 and INSTANCE.itemsize == 160 and RAY.itemsize == 32 and HIT.itemsize == 24

TRT_OK, TRT_E_INVALID, TRT_E_OOM, TRT_E_DEVICE, TRT_E_GUARD = 0, 1, 2, 3, 4
TRACE_CLOSEST, TRACE_SHADOW, TRACE_ANY = 0, 1, 2
WALK_WIDE, WALK_BINARY = 0, 1
WALK_WATERTIGHT = 2  # the 4-wide walk with the watertight triangle test (Woop, Benthin and Wald 2013)
BUILD_DEVICE, BUILD_HOST, BUILD_AUTO = 0, 1, 2  # trt_device_set_build
RENDER_AUTO, RENDER_MEGAKERNEL, RENDER_WAVEFRONT = 0, 1, 2  # trt_device_set_render
NO_HIT = 0xFFFFFFFF
INF = np.float32(3.0e38)

EXPORTS = [
    "trt_last_error", "trt_device_create", "trt_device_release", "trt_device_info_get", "trt_device_sync",
    "trt_mem_alloc", "trt_mem_free", "trt_mem_upload", "trt_mem_download", "trt_timer_start", "trt_timer_stop",
    "trt_build_tris", "trt_build_mesh_tree", "trt_build_sah", "trt_build_median", "trt_build_top_tree",
    "trt_unit_box", "trt_unit_record", "trt_tree_free", "trt_scene_upload", "trt_scene_release",
    "trt_scene_stats_get", "trt_trace_rays", "trt_trace_rays_device", "trt_trace_reads",
    "trt_accum_create", "trt_accum_release", "trt_accum_clear", "trt_accum_read", "trt_accum_resolve", "trt_render",
    "trt_denoise_atrous", "trt_render_batch", "trt_mem_copy", "trt_trace_check_fp64", "trt_render_ray_counts",
    "trt_accum_resolve_device", "trt_accum_add_fp64", "trt_device_set_build", "trt_device_set_render", "trt_scene_wide_read",
    "trt_unet_create", "trt_unet_release", "trt_unet_param_count", "trt_unet_params_set", "trt_unet_params_get",
    "trt_unet_denoise", "trt_unet_gradients", "trt_unet_train_step", "trt_trace_rays_device_noalloc",
    "trt_build_light_tree", "trt_light_tree_free", "trt_render_reads", "trt_unet_set_tall_tiles",
]


class ThreeRTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _Tree(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint64), ("order", C.POINTER(C.c_uint32)),
                ("order_count", C.c_uint64), ("depth", C.c_uint32), ("lo", C.c_float * 3), ("hi", C.c_float * 3)]


class _SceneDesc(C.Structure):
    """trt_scene_desc: the trees, then the shading tables (NULL for a scene that only traces rays)."""
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint64), ("leaves", C.c_void_p),
                ("leaf_count", C.c_uint64), ("instances", C.c_void_p), ("instance_count", C.c_uint64),
                ("top_first", C.c_uint32), ("unit_count", C.c_uint32),
                ("vertices", C.c_void_p), ("vertex_count", C.c_uint64), ("indices", C.c_void_p),
                ("index_count", C.c_uint64), ("rows", C.c_void_p), ("row_count", C.c_uint64),
                ("materials", C.c_void_p), ("material_count", C.c_uint64), ("lights", C.c_void_p),
                ("light_count", C.c_uint64), ("emitters", C.c_void_p), ("emitter_vec4s", C.c_uint64),
                ("emitter_tris", C.c_void_p), ("emitter_tri_count", C.c_uint64), ("objects", C.c_void_p),
                ("object_count", C.c_uint64), ("energy", C.c_void_p), ("light_tree", C.c_void_p),
                ("light_tree_vec4s", C.c_uint64), ("glass_energy", C.c_void_p), ("tex_pages", C.c_void_p),
                ("tex_side", C.c_uint32), ("tex_layers", C.c_uint32), ("tex_info", C.c_void_p),
                ("tex_info_vec4s", C.c_uint64), ("env", C.c_void_p), ("env_width", C.c_uint32),
                ("env_height", C.c_uint32), ("env_cdf", C.c_void_p), ("env_cdf_count", C.c_uint64),
                ("vcolors", C.c_void_p), ("vcolor_words", C.c_uint64)]


class _LightTree(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint64), ("order", C.c_void_p), ("code", C.c_void_p),
                ("entry_count", C.c_uint64), ("power", C.c_float)]


class _SceneStats(C.Structure):
    _fields_ = [("wide_nodes", C.c_uint64), ("wide_bytes", C.c_uint64), ("binary_bytes", C.c_uint64),
                ("leaf_bytes", C.c_uint64), ("unit_bytes", C.c_uint64), ("max_stack", C.c_uint32),
                ("units", C.c_uint32), ("device_bytes", C.c_uint64), ("build_ms", C.c_float),
                ("device_build_ms", C.c_float), ("device_meshes", C.c_uint32), ("host_meshes", C.c_uint32)]


class _DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("compute_units", C.c_uint32), ("wave_size", C.c_uint32),
                ("hbm_bytes", C.c_uint64), ("clock_khz", C.c_uint32), ("gcn_arch", C.c_uint32)]


_lib = None


def lib():
    """The loaded libthreert.so. Raises when it is not built: the product has no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ThreeRTError(TRT_E_DEVICE, f"{LIB_PATH} is not built (run threert_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        P, U64, U32, I = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
        L.trt_last_error.restype = C.c_char_p
        sig = {
            "trt_device_create": [I, C.POINTER(C.c_void_p)], "trt_device_release": [P],
            "trt_device_info_get": [P, P], "trt_device_sync": [P],
            "trt_mem_alloc": [P, U64, C.POINTER(C.c_void_p)], "trt_mem_free": [P, P],
            "trt_mem_upload": [P, P, P, U64], "trt_mem_download": [P, P, P, U64], "trt_mem_copy": [P, P, P, U64],
            "trt_timer_start": [P], "trt_timer_stop": [P, C.POINTER(C.c_float)],
            "trt_build_tris": [P, U64, U32, P], "trt_build_mesh_tree": [P, U64, P],
            "trt_build_sah": [P, U64, U32, P], "trt_build_median": [P, U64, U32, P],
            "trt_build_top_tree": [P, P, U64, P],
            "trt_scene_upload": [P, P, C.POINTER(C.c_void_p)], "trt_scene_release": [P],
            "trt_scene_stats_get": [P, P], "trt_trace_rays": [P, P, U64, P, U32],
            "trt_trace_rays_device": [P, P, U64, P, U32], "trt_trace_rays_device_noalloc": [P, P, U64, P, U32],
            "trt_trace_reads": [P, P, U64, P, U32, P],
            "trt_trace_check_fp64": [P, P, P, U64, P],
            "trt_render_ray_counts": [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
            "trt_device_set_build": [P, U32], "trt_device_set_render": [P, U32], "trt_scene_wide_read": [P, U32, P, U64],
            "trt_build_light_tree": [P, P, P, U64, P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I
        L.trt_unit_box.argtypes = [P, P, P, P]
        L.trt_unit_box.restype = None
        L.trt_unit_record.argtypes = [P, U32, I]
        L.trt_unit_record.restype = None
        L.trt_tree_free.argtypes = [P]
        L.trt_tree_free.restype = None
        L.trt_light_tree_free.argtypes = [P]
        L.trt_light_tree_free.restype = None
        _lib = L
    return _lib


def _check(rc: int):
    if rc != TRT_OK:
        raise ThreeRTError(rc, lib().trt_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _take_tree(t: _Tree):
    nodes = np.frombuffer(C.string_at(t.nodes, t.node_count * 32), dtype=NODE).copy()
    order = (np.ctypeslib.as_array(t.order, shape=(t.order_count,)).copy() if t.order_count
             else np.zeros(0, np.uint32))
    out = {"nodes": nodes, "order": order.astype(np.uint32), "depth": int(t.depth),
           "lo": np.array(t.lo[:], np.float32), "hi": np.array(t.hi[:], np.float32)}
    lib().trt_tree_free(C.byref(t))
    return out


# ---- host builders (units.rs / bvh.rs) --------------------------------------------------------------------------
def build_tris(tris: np.ndarray, max_depth: int = 48):
    """bvh.rs:81 build_tris."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    t = _Tree()
    _check(lib().trt_build_tris(_ptr(tris), len(tris), max_depth, C.byref(t)))
    return _take_tree(t)


def mesh_tree(tris: np.ndarray):
    """units.rs:231 mesh_tree: {'nodes', 'order', 'depth', 'lo', 'hi'}."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    t = _Tree()
    _check(lib().trt_build_mesh_tree(_ptr(tris), len(tris), C.byref(t)))
    return _take_tree(t)


def build_sah(boxes: np.ndarray, max_depth: int):
    """units.rs:140 build_sah over boxes (n x 6)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    t = _Tree()
    _check(lib().trt_build_sah(_ptr(boxes), len(boxes), max_depth, C.byref(t)))
    return _take_tree(t)


def build_median(boxes: np.ndarray, max_leaf: int):
    """units.rs:84 build_median over boxes (n x 6)."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    t = _Tree()
    _check(lib().trt_build_median(_ptr(boxes), len(boxes), max_leaf, C.byref(t)))
    return _take_tree(t)


def top_tree(boxes: np.ndarray, records: np.ndarray):
    """units.rs:303 top_tree: the leaves' left_or_first are records."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    records = np.ascontiguousarray(records, dtype=np.uint32)
    t = _Tree()
    _check(lib().trt_build_top_tree(_ptr(boxes), _ptr(records), len(boxes), C.byref(t)))
    return _take_tree(t)


class LightTree:
    """ltree.rs LightTree as the kernel and the emitter table read it: nodes (vec4s), the tree's emitters in leaf
    order, each entry's path code, the root's power."""

    def __init__(self, nodes: np.ndarray, order: np.ndarray, code: np.ndarray, power: float):
        self.nodes, self.order, self.code, self.power = nodes, order, code, power


def build_light_tree(verts: np.ndarray, centroids: np.ndarray, weights: np.ndarray):
    """ltree.rs:409-459 build over emitter triangles (world vertices n x 3 x 3, centroids n x 3, weights n): a
    LightTree, or None below 256 emitters that weigh something (the CDF alone)."""
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    centroids = np.ascontiguousarray(centroids, np.float32).reshape(-1, 3)
    weights = np.ascontiguousarray(weights, np.float32).reshape(-1)
    n = len(weights)
    assert len(verts) == n and len(centroids) == n
    t = _LightTree()
    _check(lib().trt_build_light_tree(_ptr(verts), _ptr(centroids), _ptr(weights), n, C.byref(t)))
    try:
        if t.node_count == 0:
            return None
        nodes = np.ctypeslib.as_array(C.cast(t.nodes, C.POINTER(C.c_float)), (t.node_count, 4)).copy()
        order = np.ctypeslib.as_array(C.cast(t.order, C.POINTER(C.c_uint32)), (t.entry_count,)).copy()
        code = np.ctypeslib.as_array(C.cast(t.code, C.POINTER(C.c_uint32)), (t.entry_count,)).copy()
        return LightTree(nodes, order, code, float(np.float32(t.power)))
    finally:
        lib().trt_light_tree_free(C.byref(t))


def unit_box(world_colmajor: np.ndarray, lo, hi) -> np.ndarray:
    """units.rs:288 UnitBox::new: (lo xyz, hi xyz) of a model box under a column-major world matrix."""
    w = np.ascontiguousarray(world_colmajor, dtype=np.float32).reshape(16)
    lo = np.ascontiguousarray(lo, dtype=np.float32)
    hi = np.ascontiguousarray(hi, dtype=np.float32)
    out = np.zeros(6, np.float32)
    lib().trt_unit_box(_ptr(w), _ptr(lo), _ptr(hi), _ptr(out))
    return out


def unit_record(rec: np.ndarray, root: int, cast_shadow: bool) -> None:
    """units.rs:329 unit_record, in place on a one-element INSTANCE array."""
    assert rec.dtype == INSTANCE and rec.flags.c_contiguous
    lib().trt_unit_record(_ptr(rec), root, 1 if cast_shadow else 0)


# ---- device and scene --------------------------------------------------------------------------------------------
class Device:
    """One HIP device and the stream every kernel of the library runs on (gpu.rs:2113 Gpu::new)."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        _check(lib().trt_device_create(ordinal, C.byref(h)))
        self.h = h

    def info(self) -> dict:
        i = _DeviceInfo()
        _check(lib().trt_device_info_get(self.h, C.byref(i)))
        return {"name": i.name.decode(), "compute_units": i.compute_units, "wave_size": i.wave_size,
                "hbm_bytes": i.hbm_bytes, "clock_khz": i.clock_khz, "gcn_arch": i.gcn_arch}

    def sync(self):
        _check(lib().trt_device_sync(self.h))

    def render_ray_counts(self) -> tuple:
        """(closest, shadow) rays the render launches traced since the last call; waits for the device."""
        c, sh = C.c_uint64(), C.c_uint64()
        _check(lib().trt_render_ray_counts(self.h, C.byref(c), C.byref(sh)))
        return int(c.value), int(sh.value)

    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def set_render(self, how: int):
        """How later render launches run their paths: RENDER_AUTO (default: the wavefront for scenes of 1,024
        triangles or more or launches of 4,194,304 paths or more, else the megakernel), RENDER_MEGAKERNEL or
        RENDER_WAVEFRONT (bounce by bounce over queued rays). The same bits either way."""
        _check(lib().trt_device_set_render(self.h, how))

    def set_build(self, where: int):
        """Where later uploads build the library's own mesh trees: BUILD_AUTO (default: the device from 65,536
        mesh triangles), BUILD_DEVICE or BUILD_HOST."""
        _check(lib().trt_device_set_build(self.h, where))

    def timer_start(self):
        _check(lib().trt_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_float()
        _check(lib().trt_timer_stop(self.h, C.byref(ms)))
        return float(ms.value)

    def close(self):
        if self.h:
            lib().trt_device_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory (HBM) owned by the caller: inputs kept resident between launches."""

    def __init__(self, dev: Device, nbytes: int):
        self.dev, self.nbytes = dev, int(nbytes)
        p = C.c_void_p()
        _check(lib().trt_mem_alloc(dev.h, self.nbytes, C.byref(p)))
        self.p = p

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().trt_mem_upload(self.dev.h, self.p, _ptr(a), a.nbytes))

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().trt_mem_download(self.dev.h, _ptr(out), self.p, out.nbytes))
        return out

    def copy_from(self, src: "DeviceBuffer", nbytes: int):
        """Queues a device-to-device copy of `nbytes` (a multiple of 16) from `src`."""
        assert nbytes <= self.nbytes and nbytes <= src.nbytes
        _check(lib().trt_mem_copy(self.dev.h, self.p, src.p, nbytes))

    def free(self):
        if self.p:
            lib().trt_mem_free(self.dev.h, self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    """A prepared scene resident in HBM (gpu.rs:2613 Gpu::upload): the reference's node / leaf / record buffers
    and the library's 4-wide quantised tree built from them."""

    def __init__(self, dev: Device, nodes: np.ndarray, leaves: np.ndarray, instances: np.ndarray, top_first: int,
                 units: int):
        self.dev = dev
        nodes = np.ascontiguousarray(nodes, dtype=NODE)
        leaves = np.ascontiguousarray(leaves, dtype=np.float32).reshape(-1, 3, 4)
        instances = np.ascontiguousarray(instances, dtype=INSTANCE)
        desc = _SceneDesc(_ptr(nodes), len(nodes), _ptr(leaves), len(leaves), _ptr(instances), len(instances),
                          top_first, units)
        h = C.c_void_p()
        _check(lib().trt_scene_upload(dev.h, C.byref(desc), C.byref(h)))
        self.h = h

    def stats(self) -> dict:
        s = _SceneStats()
        _check(lib().trt_scene_stats_get(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _SceneStats._fields_}

    def trace(self, rays: np.ndarray, walk: int = WALK_WIDE) -> np.ndarray:
        """Hits for host rays (RAY records): the lowest (t, instance, triangle) for closest rays."""
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), HIT)
        _check(lib().trt_trace_rays(self.h, _ptr(rays), len(rays), _ptr(hits), walk))
        return hits

    def trace_device(self, rays: DeviceBuffer, n: int, hits: DeviceBuffer, walk: int = WALK_WIDE):
        """Queues a batch whose rays are already in HBM; Device.sync() waits."""
        _check(lib().trt_trace_rays_device(self.h, rays.p, n, hits.p, walk))

    def trace_device_noalloc(self, rays: DeviceBuffer, n: int, hits: DeviceBuffer, walk: int = WALK_WIDE):
        """trace_device without any device allocation (trt_trace_rays_device_noalloc): the walk's counter and
        spilled stack in the GPU's fixed scratch. Device.sync() waits."""
        _check(lib().trt_trace_rays_device_noalloc(self.h, rays.p, n, hits.p, walk))

    def reads(self, rays: DeviceBuffer, n: int, hits: DeviceBuffer, walk: int = WALK_WIDE) -> dict:
        """What the walk reads over a batch (units.rs Reads): node fetches, triangle tests, unit records, and the
        algorithmic bytes they come to in the walk's own formats."""
        out = np.zeros(3, np.uint64)
        _check(lib().trt_trace_reads(self.h, rays.p, n, hits.p, walk, _ptr(out)))
        node_b, unit_b = (64, 64) if walk == WALK_WIDE else (32, 160)
        nodes, tris, units = (int(x) for x in out)
        return {"nodes": nodes, "tris": tris, "units": units,
                "bytes": nodes * node_b + tris * 48 + units * unit_b}

    def wide_read(self, which: int) -> np.ndarray:
        """One of the wide-tree buffers as bytes (trt_scene_wide_read): 0 nodes, 1 leaves in the trees' order,
        2 units."""
        st = self.stats()
        size = {0: st["wide_bytes"], 1: st["leaf_bytes"], 2: st["unit_bytes"]}[which]
        out = np.zeros(size, np.uint8)
        _check(lib().trt_scene_wide_read(self.h, which, _ptr(out) if size else None, size))
        return out

    def check_fp64(self, rays: DeviceBuffer, n: int, hits: DeviceBuffer, t64: DeviceBuffer):
        """Queues the FP64 re-intersection of a batch's hits (trt_trace_check_fp64): t64 receives n doubles."""
        _check(lib().trt_trace_check_fp64(self.h, rays.p, hits.p, n, t64.p))

    def close(self):
        if self.h:
            lib().trt_scene_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
