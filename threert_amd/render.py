"""Render scenes: the shading tables the offline pass reads, the Params block, and the seeded workloads.

The tables follow the reference's layouts (the prepare step of lib.rs:530-785 for the features used here):
  vertices   indexed.rs:10-13, 48 bytes: (position, u), (normal, v), (tangent, hand); a triangle's corners here
             are its own (not shared between triangles: the kernel reads them through the indices either way)
  indices    three a triangle, in the leaves' (tree) order, counted from the mesh's first vertex (record info.z)
  rows       scene.rs:99-135 Row: (material, section's first triangle, its first emitter or NO_EMITTER, 0), one a
             section of the instance's mesh, from record info.w
  materials  scene.rs:480-491 GpuMaterial, 288 bytes
  lights     scene.rs:495-501 GpuLight, 80 bytes (directional: type 0, emission direction, cos of the half angle)
  emitters   lib.rs:451-488 emitter_cdf: (emitter bits, cumulative weight, weight, area) a triangle, the 65,537-entry
             guide table four to a vec4, then the header (total weight, 0, 0, 0)
  emitter triangles  lib.rs:143-157 GpuEmitterTri, world space, 64 bytes
  energy     lib.rs:1026-1104 specular_albedo / energy_lut, 512 vec4
  params     gpu.rs:129-185 Params, 39 vec4
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import (INSTANCE, Device, Scene, ThreeRTError, TRT_E_INVALID, _check, _ptr, lib)
from . import scene as S
from . import build_light_tree

NO_EMITTER = 0xFFFFFFFF
EMITTER_GUIDE = 65536
RIS_OFFLINE = 6  # ltree.rs:45: light-tree candidates resampled at camera hits
OBJECT_STRIDE = 3
INST_GLASS = 32  # scene.rs:263: an instance with glass rows
ROW_GLASS = 1 << 25  # scene.rs:103: a row's class bit for glass (mesh_class 2)


def material(base=(0.8, 0.8, 0.8), emissive=(0.0, 0.0, 0.0), metallic=0.0, roughness=1.0, specular=0.5,
             clearcoat=0.0, clearcoat_roughness=0.0, transmission=0.0, ior=1.5, base_map=-1, normal_map=-1,
             roughness_map=-1, metalness_map=-1, emissive_map=-1, transmission_map=-1, normal_scale=1.0,
             repeat=(1.0, 1.0), offset=(0.0, 0.0), unlit=False, vertex_colors=False) -> np.ndarray:
    """scene.rs:480-491 GpuMaterial (72 words; scene.rs:1238-1262 build_material): base colour, emission, (metallic,
    roughness, specular, normal scale); the texture slots (tex0: base colour (sRGB), normal, roughness (green),
    metalness (blue); tex1.x emissive (sRGB); uv.z the transmission map's slot + 1 (red); -1 none), the uv repeat and
    offset; the layer (scene.rs:1257): transmission (glass, clamped to [0, 1]), index of refraction, clearcoat and its
    roughness (both clamped to [0, 1]); `unlit`: three.js MeshBasicMaterial (tex1.z MAT_UNLIT 4: its colour to camera
    rays and mirror-like paths, lit by nothing, lighting nothing); `vertex_colors`: the mesh's vertex colours multiply
    the base colour (uv.w)."""
    m = np.zeros(72, np.float32)
    m[0:3] = base
    m[3] = 1.0
    m[4:7] = emissive
    m[8:12] = (metallic, roughness, specular, normal_scale)
    m[12:20] = np.array([base_map, normal_map, roughness_map, metalness_map, emissive_map, -1, 4 if unlit else 0, 0],
                        np.int32).view(np.float32)
    m[20:24] = (repeat[0], repeat[1], float(transmission_map) + 1.0, 1.0 if vertex_colors else 0.0)
    m[24:26] = offset
    m[28:32] = (min(max(transmission, 0.0), 1.0), ior, min(max(clearcoat, 0.0), 1.0),
                min(max(clearcoat_roughness, 0.0), 1.0))
    m[64:72] = np.full(8, -1, np.int32).view(np.float32)
    return m


def directional_light(direction, lux, diameter_deg=0.5357, cast_shadow=True) -> np.ndarray:
    """scene.rs:495-501 GpuLight for a directional light: `direction` is where the light travels."""
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    li = np.zeros(20, np.float32)
    li[4:7] = d
    li[7] = math.cos(math.radians(diameter_deg) * 0.5)
    li[8:11] = lux
    li[15] = 1.0 if cast_shadow else 0.0
    li[17] = 1.0
    return li


def spot_cone(angle: float, penumbra: float) -> tuple[float, float]:
    """scene.rs:1309-1313 spot_cone: the outer cone's cosine and 1 / (inner cosine - outer cosine), the difference no
    less than 0.001; the inner cone angle x (1 - penumbra)."""
    a = np.float32(min(max(angle, 0.0), math.pi))
    co = np.float32(math.cos(a))
    ci = np.float32(math.cos(np.float32(a * np.float32(1.0 - min(max(penumbra, 0.0), 1.0)))))
    return float(co), float(np.float32(1.0) / np.float32(max(ci - co, np.float32(1e-3))))


def point_light(position, candela, radius=0.0, range_=0.0, cast_shadow=True, specular_scale=1.0) -> np.ndarray:
    """scene.rs:1332-1341 GpuLight of a point light: candela rgb, a sphere of `radius` for soft shadows, `range_` the
    distance its windowed attenuation reaches (0: none)."""
    li = np.zeros(20, np.float32)
    li[0:4] = (*position, 1.0)
    li[4:8] = (0.0, 0.0, -1.0, radius)
    li[8:11] = candela
    li[11] = range_
    li[15] = 1.0 if cast_shadow else 0.0
    li[17] = specular_scale
    return li


def spot_light(position, direction, candela, angle, penumbra=0.0, radius=0.0, range_=0.0, cast_shadow=True,
               specular_scale=1.0) -> np.ndarray:
    """scene.rs:1342-1353 GpuLight of a spot light: `direction` its axis, `angle` its outer half angle (radians),
    `penumbra` the share of it over which it fades."""
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    li = point_light(position, candela, radius, range_, cast_shadow, specular_scale)
    li[3] = 2.0
    li[4:7] = d
    li[12:14] = spot_cone(angle, penumbra)
    return li


def rect_light(center, normal, width_axis, width, height, nits, range_=0.0, barn_door_deg=90.0, cast_shadow=True,
               specular_scale=1.0) -> np.ndarray:
    """scene.rs:1355-1367 GpuLight of a hidden rect light (UE RectLight): a one-sided Lambertian rectangle of
    luminance `nits` facing `normal`, `width` along `width_axis` and `height` across it, cut beyond the barn-door
    angle."""
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    ux = np.asarray(width_axis, np.float64)
    ux = ux / np.linalg.norm(ux) * (width * 0.5)
    li = np.zeros(20, np.float32)
    li[0:4] = (*center, 3.0)
    li[4:8] = (*n, height * 0.5)
    li[8:11] = nits
    li[11] = range_
    li[12:15] = ux
    li[15] = 1.0 if cast_shadow else 0.0
    li[17] = specular_scale
    li[18] = math.cos(math.radians(min(max(barn_door_deg, 0.0), 90.0)))
    return li


def light_order(lights: list) -> list:
    """lib.rs:367-380 light_order, as the kernel needs it: directional lights first, then local lights that cast no
    shadow, then shadowed ones (the reference also sorts each class spatially for its light groups, which only skip
    work; the kernel here reads no groups and the order within a class is kept)."""
    cls = lambda li: 0 if li[3] == 0.0 else (1 if li[15] < 0.5 else 2)  # noqa: E731
    return sorted(range(len(lights)), key=lambda i: cls(lights[i]))


# ---- lib.rs:1026-1104 specular_albedo / energy_lut, in f64 ------------------------------------------------------
def specular_albedo(mu: float, alpha: float, nr: int = 48, nphi: int = 8) -> tuple[float, float]:
    mu = min(max(mu, 1e-4), 1.0)
    if alpha < 1e-6:
        fc = (1.0 - mu) ** 5
        return 1.0 - fc, fc
    vx, vz = math.sqrt(max(1.0 - mu * mu, 0.0)), mu
    k = math.sqrt(alpha * alpha * vx * vx + vz * vz)
    hx, hz = alpha * vx / k, vz / k
    s = 0.5 * (1.0 + hz)
    inv_g1 = 0.5 * (1.0 + math.sqrt(1.0 + alpha * alpha * (1.0 - mu * mu) / (mu * mu)))
    i = np.arange(nr, dtype=np.float64)
    t = (i + 0.5) / nr
    r = np.sqrt(1.0 - (1.0 - t) * (1.0 - t))
    w = 2.0 * (1.0 - t)
    j = np.arange(nphi, dtype=np.float64)
    phi = math.pi * ((j + 0.5) / nphi - 0.5)
    r2, phi2 = np.meshgrid(r, phi, indexing="ij")
    w2 = np.repeat(w[:, None], nphi, axis=1)
    p1 = r2 * np.cos(phi2)
    p2 = (1.0 - s) * np.sqrt(np.maximum(1.0 - p1 * p1, 0.0)) + s * r2 * np.sin(phi2)
    pz = np.sqrt(np.maximum(1.0 - p1 * p1 - p2 * p2, 0.0))
    m0 = alpha * (pz * hx - p2 * hz)
    m1 = alpha * p1
    m2 = np.maximum(p2 * hx + pz * hz, 0.0)
    ln = np.sqrt(m0 * m0 + m1 * m1 + m2 * m2)
    voh = (vx * m0 + vz * m2) / ln
    nol = 2.0 * voh * m2 / ln - vz
    ok = (nol > 0.0) & (voh > 0.0)
    g2 = 2.0 * mu * nol / (nol * (mu * (1.0 - alpha) + alpha) + mu * (nol * (1.0 - alpha) + alpha))
    fc = (1.0 - voh) ** 5
    a = np.where(ok, w2 * (1.0 - fc) * g2, 0.0).sum()
    b = np.where(ok, w2 * fc * g2, 0.0).sum()
    kk = inv_g1 / (nr * nphi)
    return float(a * kk), float(b * kk)


_LUT = None


def energy_lut() -> np.ndarray:
    global _LUT
    if _LUT is None:
        n = 32
        vals = []
        for iy in range(n):
            r = iy / (n - 1)
            for ix in range(n):
                x = ix / (n - 1)
                a, b = specular_albedo(x * x, r * r, 48, 8)
                vals += [a, b]
        _LUT = np.array(vals, np.float32).reshape(512, 4)
    return _LUT


# ---- lib.rs:1186-1254 glass_energy / glass_energy_lut, in f64 ---------------------------------------------------
GLASS_E = (16, 16, 8)                 # nodes over sqrt(cos), roughness, sqrt((ior - 1) / (ior + 1)) (pathtrace.wgsl)
GLASS_E_ZMAX = 1.0 / math.sqrt(2.0)   # ior 3


def fresnel_dielectric(c, eta):
    """Unpolarised Fresnel reflectance, `c` the cosine on the incident side, `eta` = n_t / n_i; 1 at total internal
    reflection (Walter et al. 2007, eq. 22)."""
    c = np.asarray(c, np.float64)
    g2 = eta * eta - 1.0 + c * c
    g = np.sqrt(np.maximum(g2, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (g - c) / (g + c)
        b = (c * (g + c) - 1.0) / (c * (g - c) + 1.0)
        f = 0.5 * a * a * (1.0 + b * b)
    return np.where(g2 <= 0.0, 1.0, f)


def glass_energy(mu: float, alpha: float, eta: float, n: int = 24) -> float:
    """R + T of a rough dielectric interface (GGX, single scattering) toward a view at cosine `mu`, without the tint:
    the mean over n x n stratified visible normals (Heitz 2018) of G2 / G1 for the ray reflected or refracted by the
    Fresnel reflectance at the normal; reflection with the kernel's approximate joint visibility, refraction with the
    height-correlated Smith term (Walter eq. 40). `alpha` = roughness^2, `eta` = n(far) / n(view)."""
    a2 = alpha * alpha

    def lam(c):
        c = np.minimum(np.abs(c), 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            return 0.5 * (-1.0 + np.sqrt(1.0 + a2 * (1.0 - c * c) / (c * c)))

    v0, v2 = math.sqrt(max(1.0 - mu * mu, 0.0)), mu
    k = math.sqrt(alpha * alpha * v0 * v0 + v2 * v2)
    vh0, vh2 = alpha * v0 / k, v2 / k
    t20, t22 = -vh2, vh0
    lv, s = float(lam(mu)), 0.5 * (1.0 + vh2)
    i = np.arange(n, dtype=np.float64)
    r = np.sqrt((i + 0.5) / n)[:, None]
    phi = (2.0 * math.pi * (i + 0.5) / n)[None, :]
    p1 = r * np.cos(phi)
    p2 = (1.0 - s) * np.sqrt(np.maximum(1.0 - p1 * p1, 0.0)) + s * r * np.sin(phi)
    pz = np.sqrt(np.maximum(1.0 - p1 * p1 - p2 * p2, 0.0))
    n0, n1, n2 = p2 * t20 + pz * vh0, p1, p2 * t22 + pz * vh2
    m0, m1, m2 = alpha * n0, alpha * n1, np.maximum(n2, 0.0)
    ln = np.sqrt(m0 * m0 + m1 * m1 + m2 * m2)
    m0, m1, m2 = m0 / ln, m1 / ln, m2 / ln
    vm = v0 * m0 + v2 * m2
    f = fresnel_dielectric(vm, eta)
    nol = 2.0 * vm * m2 - v2
    g2r = 2.0 * mu * nol / np.maximum(nol * (mu * (1.0 - alpha) + alpha) + mu * (nol * (1.0 - alpha) + alpha), 1e-7)
    refl = np.where(nol > 0.0, f * g2r * (1.0 + lv), 0.0)
    c2 = 1.0 - (1.0 - vm * vm) / (eta * eta)
    lz = (vm / eta - np.sqrt(np.maximum(c2, 0.0))) * m2 - v2 / eta
    ok = (f < 1.0) & (c2 > 0.0) & (lz < 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        tr = np.where(ok, (1.0 - f) * (1.0 + lv) / (1.0 + lv + lam(lz)), 0.0)
    return float((refl + tr).sum() / (n * n))


_GLASS_LUT = None


def glass_energy_lut() -> np.ndarray:
    """glass_energy over the table's nodes, entering then leaving (ior, then its inverse), four values a vec4: the
    kernel's glass energy table (1024 vec4)."""
    global _GLASS_LUT
    if _GLASS_LUT is None:
        nm, nr, nz = GLASS_E
        vals = []
        for side in range(2):
            for iz in range(nz):
                z = max(iz / (nz - 1) * GLASS_E_ZMAX, 0.02)
                ior = (1.0 + z * z) / (1.0 - z * z)
                for ir in range(nr):
                    r = max(ir / (nr - 1), 0.02)
                    for im in range(nm):
                        y = im / (nm - 1)
                        e = glass_energy(max(y * y, 1e-3), r * r, ior if side == 0 else 1.0 / ior, 24)
                        vals.append(min(max(e, 0.05), 1.0))
        _GLASS_LUT = np.array(vals, np.float32).reshape(-1, 4)
    return _GLASS_LUT


# ---- the render scene -------------------------------------------------------------------------------------------
@dataclass
class RMesh:
    tris: np.ndarray                   # (n, 3, 3) model space
    normals: np.ndarray | None = None  # (n, 3, 3) per corner; None: the triangle's own
    sections: np.ndarray | None = None  # (n,) section of each triangle
    uvs: np.ndarray | None = None      # (n, 3, 2) per corner; None: zeros
    tangents: np.ndarray | None = None  # (n, 3, 4) per corner (xyz, hand); None: zeros (hand 0: derived from the uvs)
    colors: np.ndarray | None = None   # (n, 3, 3) linear vertex colours per corner (three.js vertexColors); None: white


# ---- texpack.rs: texture pages ----------------------------------------------------------------------------------
TEX_GUTTER, TEX_MIPS, TEX_ALIGN, TEX_MAX_PAGE, TEX_MAX_DIM = 16, 5, 16, 4096, 2048


@dataclass
class Texture:
    rgba: np.ndarray   # (h, w, 4) uint8, as stored (sRGB-encoded colour maps are decoded by the kernel)
    wrap: bool = True  # repeat (else clamp) outside [0, 1]


def _content_size(w: int, h: int) -> tuple[int, int]:
    """texpack.rs:48-60 content_size."""
    cw, ch = min(w, TEX_MAX_DIM), min(h, TEX_MAX_DIM)
    k = min(np.float32(cw) / np.float32(w), np.float32(ch) / np.float32(h))
    cw, ch = max(int(np.float32(w) * k), 1), max(int(np.float32(h) * k), 1)
    if cw >= 512 and cw & (cw - 1) == 0:
        cw -= 2 * TEX_GUTTER
    if ch >= 512 and ch & (ch - 1) == 0:
        ch -= 2 * TEX_GUTTER
    return cw, ch


def _mip_chain(level0: np.ndarray) -> list:
    """texpack.rs:396-412 mip_chain: 2 x 2 boxes of the stored values, rounded ((a + b + c + d + 2) / 4)."""
    chain = [level0]
    for _ in range(1, TEX_MIPS):
        p = chain[-1].astype(np.uint32)
        h, w = max(p.shape[0] // 2, 1), max(p.shape[1] // 2, 1)
        s = p[0:2 * h:2, 0:2 * w:2] + p[0:2 * h:2, 1:2 * w:2] + p[1:2 * h:2, 0:2 * w:2] + p[1:2 * h:2, 1:2 * w:2]
        chain.append(((s + 2) // 4).astype(np.uint8))
    return chain


def pack_textures(textures: list) -> tuple[np.ndarray, int, int, np.ndarray]:
    """texpack.rs:298-393 pack and Atlas::info (72-86): every texture shelf-packed into square RGBA8 layers with
    16-texel gutters of wrapped (or clamped) texels, tallest first, each layer's 2 x 2 box mip chain. Returns (pages:
    per layer levels 0..4 one after the other, side, layers, info: 2 vec4 a texture). Textures whose size the
    reference's packer changes (wider than 2048, or a power of two from 512) are refused: its resize
    (texpack.rs:415 resize_triangle) is not restated here; the library takes any pages."""
    src = []
    for k, tx in enumerate(textures):
        img = np.ascontiguousarray(tx.rgba, np.uint8)
        if img.ndim != 3 or img.shape[2] != 4:
            raise ValueError(f"texture {k}: rgba must be (h, w, 4) uint8")
        h, w = img.shape[:2]
        if _content_size(w, h) != (w, h):
            raise ValueError(f"texture {k}: {w} x {h} would be resized by the reference's packer (not restated)")
        src.append(img)
    up = lambda v: (v + TEX_ALIGN - 1) // TEX_ALIGN * TEX_ALIGN  # noqa: E731
    padded = [(up(i.shape[1] + 2 * TEX_GUTTER), up(i.shape[0] + 2 * TEX_GUTTER)) for i in src]
    order = sorted(range(len(src)), key=lambda i: (-padded[i][1], -padded[i][0]))
    area = sum(pw * ph for pw, ph in padded)
    widest = max((max(p) for p in padded), default=1)
    side = max(1 << (widest - 1).bit_length(), 256)
    while side < TEX_MAX_PAGE and side * side < area:
        side *= 2
    place = [(0, 0, 0)] * len(src)
    layer = x = y = shelf = 0
    for i in order:
        pw, ph = padded[i]
        if x + pw > side:
            x, y, shelf = 0, y + shelf, 0
        if y + ph > side:
            layer, x, y, shelf = layer + 1, 0, 0, 0
        place[i] = (layer, x, y)
        x += pw
        shelf = max(shelf, ph)
    pages = []
    for lyr in range(layer + 1):
        page = np.zeros((side, side, 4), np.uint8)
        for i in (j for j in range(len(src)) if place[j][0] == lyr):
            _, px, py = place[i]
            pw, ph = padded[i]
            h, w = src[i].shape[:2]
            if textures[i].wrap:
                rows, cols = np.arange(ph) - TEX_GUTTER, np.arange(pw) - TEX_GUTTER
                rows, cols = rows % h, cols % w
            else:
                rows = np.clip(np.arange(ph) - TEX_GUTTER, 0, h - 1)
                cols = np.clip(np.arange(pw) - TEX_GUTTER, 0, w - 1)
            page[py:py + ph, px:px + pw] = src[i][rows[:, None], cols[None, :]]
        pages.append(np.concatenate([lv.reshape(-1) for lv in _mip_chain(page)]))
    s = np.float32(side)
    info = np.zeros((max(len(src), 1) * 2, 4), np.float32)
    for i, (lyr, px, py) in enumerate(place):
        h, w = src[i].shape[:2]
        info[2 * i] = (np.float32(px + TEX_GUTTER) / s, np.float32(py + TEX_GUTTER) / s, np.float32(w) / s,
                       np.float32(h) / s)
        info[2 * i + 1] = (lyr, 1.0 if textures[i].wrap else 0.0, w, h)
    return np.concatenate(pages), side, layer + 1, info


@dataclass
class RInstance:
    mesh: int
    world: np.ndarray
    materials: list                    # material index a section of the mesh
    cast_shadow: bool = True


@dataclass
class Camera:
    position: tuple
    target: tuple
    fov_deg: float = 40.0
    up: tuple = (0.0, 1.0, 0.0)


@dataclass
class RenderScene:
    prep: S.Prepared
    vertices: np.ndarray
    indices: np.ndarray
    rows: np.ndarray
    materials: np.ndarray
    lights: np.ndarray
    n_directional: int
    emitters: np.ndarray
    emitter_tris: np.ndarray
    objects: np.ndarray
    energy: np.ndarray
    total_weight: float
    extras: dict = field(default_factory=dict)
    # ltree.rs LightTree nodes (vec4s; none: the CDF alone), its entry count (the CDF's first entries) and the share
    # of the emitters' weight outside it (the table header's y)
    light_tree: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    tree_entries: int = 0
    # lib.rs glass_energy_lut (1024 vec4) when a material has glass (gpu.rs features_of: layer.x > 0); None: no glass
    glass_energy: np.ndarray | None = None
    # texpack.rs pages (per layer, levels 0..4), their side and layers, and the info table; None: no textures
    tex_pages: np.ndarray | None = None
    tex_side: int = 0
    tex_layers: int = 0
    tex_info: np.ndarray | None = None
    # an environment map: RGBA float texels (h, w, 4), row 0 the top (+y), and lib.rs build_env_cdf's h + w h floats;
    # None: a constant sky (the Params' colour)
    env: np.ndarray | None = None
    env_cdf: np.ndarray | None = None
    # lib.rs gpu_vcolors: (triangles, 3) RGBA8 sRGB words in the leaves' order; None: no vertex colours
    vcolors: np.ndarray | None = None

    def env_half(self) -> np.ndarray | None:
        """gpu.rs:3552-3561 write_env_texels: the texels the kernel reads, half floats of min(v, 65000)."""
        if self.env is None:
            return None
        return np.fmin(self.env, np.float32(65000.0)).astype(np.float16).view(np.uint16)

    def tables(self) -> dict:
        return {"vertices": self.vertices, "indices": self.indices, "rows": self.rows, "materials": self.materials,
                "lights": self.lights, "emitters": self.emitters, "emitter_tris": self.emitter_tris,
                "objects": self.objects, "energy": self.energy, "light_tree": self.light_tree,
                "glass_energy": self.glass_energy, "tex_pages": self.tex_pages, "tex_side": self.tex_side,
                "tex_layers": self.tex_layers, "tex_info": self.tex_info, "env": self.env_half(),
                "env_cdf": self.env_cdf, "vcolors": self.vcolors}

    def params(self, width: int, height: int, camera: Camera, seed: int = 1, max_bounces: int = 4, rr_depth: int = 3,
               clamp: float = 1e9, energy_compensation: float = 0.0, sky=(0.0, 0.0, 0.0), env_nee: bool = True,
               sample: int = 0, lens_radius: float = 0.0, focus: float = 1.0, background=None) -> np.ndarray:
        """The 624-byte Params block (gpu.rs:129-185, lib.rs:713-756 for a camera); lens_radius > 0 makes the camera a
        thin lens focused at `focus` along the view direction (cam_right.w, cam_up.w; pathtrace.wgsl:4017-4023);
        `background`: the radiance camera rays see where nothing is (three.js scene.background, Params.background =
        (rgb, 1), the reference's F_BACKGROUND)."""
        p = np.zeros(156, np.uint32)
        f = p.view(np.float32)
        pos = np.asarray(camera.position, np.float64)
        fwd = np.asarray(camera.target, np.float64) - pos
        fwd /= np.linalg.norm(fwd)
        right = np.cross(fwd, np.asarray(camera.up, np.float64))
        right /= np.linalg.norm(right)
        up = np.cross(right, fwd)
        f[0:3] = pos
        f[3] = math.tan(math.radians(camera.fov_deg) * 0.5)
        f[4:7] = fwd
        f[7] = width / height
        f[8:11] = right
        f[11] = lens_radius
        f[12:15] = up
        f[15] = focus
        p[16:20] = (width, height, sample, seed)
        n_em = len(self.emitter_tris)
        p[20:24] = (len(self.lights), n_em, max_bounces, rr_depth)
        f[24:28] = (0.0, 0.0, 0.0, -1.0)
        f[28:32] = (clamp, self.total_weight, energy_compensation, 0.0)
        f[32:35] = sky
        env_on = 1 if (env_nee and (0.2126 * sky[0] + 0.7152 * sky[1] + 0.0722 * sky[2]) > 0) else 0
        if self.env is not None:  # lib.rs:1004-1009 SkyBaked::write: a map, its size; sampled unless env_nee is off
            f[24:27] = (1.0, self.env.shape[1], self.env.shape[0])
            f[32:35] = 0.0
            env_on = 1 if env_nee else 0
        # lib.rs:725 counts2: the directional lights, the sky's sampling, -, the first shadowed local light
        first_shadowed = next((k for k in range(len(self.lights)) if self.lights[k, 3] != 0.0 and
                               self.lights[k, 15] >= 0.5), len(self.lights))
        p[36:40] = (self.n_directional, env_on, 0, first_shadowed)
        if background is not None:
            f[96:100] = (*background, 1.0)
        f[68] = 0.0
        f[70] = np.uint32(self.prep.top_first).view(np.float32)
        f[71] = np.uint32(self.prep.units).view(np.float32)
        f[76] = self.total_weight
        if self.tree_entries:  # lib.rs:673 Params::ltree: -, its entries (bits), the tail's share, RIS_OFFLINE
            p[85] = self.tree_entries
            f[86] = self.emitters[len(self.emitter_tris) + (EMITTER_GUIDE + 1 + 3) // 4, 1]
            f[87] = float(RIS_OFFLINE)
        return p


def tri_color_words(colors: np.ndarray) -> np.ndarray:
    """scene.rs:2324-2334 tri_colors: (n, 3, 3) linear colours to (n, 3) RGBA8 words, each channel clamped to [0, 1],
    sRGB-encoded (post.rs:526-532 linear_to_srgb) and rounded half away from zero, alpha 255."""
    c = np.clip(np.asarray(colors, np.float32), np.float32(0.0), np.float32(1.0))
    s = np.where(c < np.float32(0.0031308), c * np.float32(12.92),
                 np.float32(1.055) * np.power(c, np.float32(1.0 / 2.4)) - np.float32(0.055)).astype(np.float32)
    q = np.floor(s.astype(np.float64) * 255.0 + 0.5).astype(np.uint32)
    return (q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16) | np.uint32(0xff << 24)).astype(np.uint32)


def build_env_cdf(texels: np.ndarray) -> np.ndarray:
    """lib.rs:916-942 build_env_cdf, in f32 as the reference sums: per row the running sum of luminance (clamped at
    0) x sin(theta at the row's centre) + 1e-12, divided by the row's total; then the rows' totals' running share.
    Returns the h row entries followed by the h x w conditional ones."""
    t = np.asarray(texels, np.float32)
    h, w = t.shape[:2]
    f32 = np.float32
    sin_t = np.sin((np.arange(h, dtype=f32) + f32(0.5)) / f32(h) * f32(math.pi)).astype(f32)
    lum = f32(0.2126) * t[..., 0] + f32(0.7152) * t[..., 1] + f32(0.0722) * t[..., 2]
    term = (np.maximum(lum, f32(0.0)) * sin_t[:, None] + f32(1e-12)).astype(f32)
    cond = np.cumsum(term, axis=1, dtype=f32)
    rows = cond[:, -1].copy()
    cond = (cond / rows[:, None]).astype(f32)
    total = np.cumsum(rows, dtype=f32)[-1]
    marg = np.cumsum((rows / total).astype(f32), dtype=f32)
    return np.concatenate([marg, cond.reshape(-1)]).astype(f32)


def build(meshes: list[RMesh], instances: list[RInstance], materials: list[np.ndarray],
          lights: list[np.ndarray] = (), light_tree: bool = True, textures: list = (), env=None) -> RenderScene:
    """Lays out a scene as prepare does: trees and records (scene.prepare), then vertices, indices, rows, emitters.
    With 256 emitters or more that weigh something, a light tree over them (lib.rs:650-676, on unless THREERT_LTREE=0:
    here `light_tree=False`), whose emitters come first in the CDF in its leaf order. Glass (a material with
    transmission) is a class of its rows (scene.rs:2472 mesh_class): an instance whose sections are all glass casts no
    shadow (scene.rs:127 rows_cast) and carries INST_GLASS (rows_bits); glass beside other sections of a shadow-casting
    instance (the reference's F_IGLASS) is refused."""
    mats = np.stack(materials).astype(np.float32)
    placements = []
    for k, i in enumerate(instances):
        glass = [bool(mats[m, 28] > 0.0) for m in i.materials]
        cast = i.cast_shadow and not all(glass)
        if any(glass) and cast:
            raise ValueError(f"instance {k}: glass beside other sections of a shadow-casting instance is not "
                             "supported (the reference's F_IGLASS)")
        placements.append(S.Placement(i.mesh, i.world, cast, 0, INST_GLASS if any(glass) else 0))
    prep = S.prepare([S.Mesh(m.tris, m.sections) for m in meshes], placements)
    # vertices: three a triangle, in the mesh's original order; indices in the tree order
    verts, vfirst, idx_parts = [], [], []
    at = 0
    for k, m in enumerate(meshes):
        tris = np.asarray(m.tris, np.float32).reshape(-1, 3, 3)
        n = len(tris)
        if m.normals is None:
            fn = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]).astype(np.float64)
            fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
            nrm = np.repeat(fn[:, None, :], 3, axis=1).astype(np.float32)
        else:
            nrm = np.asarray(m.normals, np.float32).reshape(-1, 3, 3)
        v = np.zeros((n * 3, 12), np.float32)
        v[:, 0:3] = tris.reshape(-1, 3)
        v[:, 4:7] = nrm.reshape(-1, 3)
        if m.uvs is not None:  # indexed.rs Vertex: (position, u), (normal, v), (tangent, hand)
            uv = np.asarray(m.uvs, np.float32).reshape(-1, 2)
            v[:, 3], v[:, 7] = uv[:, 0], uv[:, 1]
        if m.tangents is not None:
            v[:, 8:12] = np.asarray(m.tangents, np.float32).reshape(-1, 4)
        verts.append(v)
        vfirst.append(at)
        at += n * 3
        order = prep.mesh_order[k].astype(np.uint32)
        idx_parts.append((order[:, None] * 3 + np.arange(3, dtype=np.uint32)[None, :]).astype(np.uint32))
    vertices = np.concatenate(verts) if verts else np.zeros((0, 12), np.float32)
    indices = np.concatenate(idx_parts) if idx_parts else np.zeros((0, 3), np.uint32)
    # lib.rs:601 gpu_vcolors: scene.rs:2324-2334 tri_colors' words in the leaves' order, white (scene.rs WHITE) for a
    # mesh without colours; none when no mesh has any
    vcolors = None
    if any(m.colors is not None for m in meshes):
        parts = []
        for k, m in enumerate(meshes):
            n = len(np.asarray(m.tris).reshape(-1, 3, 3))
            if m.colors is None:
                parts.append(np.full((n, 3), 0xffffffff, np.uint32))
                continue
            order = prep.mesh_order[k]
            parts.append(tri_color_words(np.asarray(m.colors, np.float32).reshape(n, 3, 3))[order])
        vcolors = np.concatenate(parts).astype(np.uint32)
    # rows, emitters
    rows, em_tris, entries, centres = [], [], [], []
    n_em_so_far = 0
    lum = lambda e: 0.2126 * e[0] + 0.7152 * e[1] + 0.0722 * e[2]  # noqa: E731
    for r, inst in enumerate(instances):
        m = meshes[inst.mesh]
        n = len(m.tris)
        sec = np.zeros(n, np.uint32) if m.sections is None else np.asarray(m.sections, np.uint32)
        nsec = int(sec.max()) + 1 if n else 1
        rfirst = len(rows)
        prep.instances["info"][r, 2] = vfirst[inst.mesh]
        prep.instances["info"][r, 3] = rfirst
        mesh_first = prep.mesh_first[inst.mesh]
        order = prep.mesh_order[inst.mesh]
        pos_of = np.empty(n, np.int64)
        pos_of[order] = np.arange(n)
        w = np.asarray(inst.world, np.float64)
        tris64 = np.asarray(m.tris, np.float64).reshape(-1, 3, 3)
        for s in range(nsec):
            mi = inst.materials[s] if s < len(inst.materials) else inst.materials[-1]
            first_em = NO_EMITTER
            if lum(mats[mi][4:7]) > 0.0:
                first_em = n_em_so_far
                members = np.nonzero(sec == s)[0]  # in the section's order (its rank)
                pw = np.einsum("ij,tcj->tci", w[:3, :3], tris64[members]) + w[:3, 3]
                p32 = pw.astype(np.float32)
                area = 0.5 * np.linalg.norm(np.cross(pw[:, 1] - pw[:, 0], pw[:, 2] - pw[:, 0]), axis=1)
                rec_em = np.zeros((len(members), 16), np.uint32)
                fl = rec_em.view(np.float32)
                fl[:, 0:3] = p32[:, 0]
                rec_em[:, 3] = mesh_first + pos_of[members]
                fl[:, 4:7] = p32[:, 1] - p32[:, 0]
                rec_em[:, 7] = r
                fl[:, 8:11] = p32[:, 2] - p32[:, 0]
                rec_em[:, 11] = np.arange(first_em, first_em + len(members), dtype=np.uint32)
                rec_em[:, 12] = s
                em_tris.append(rec_em)
                a32 = area.astype(np.float32)
                entries.append((a32 * np.float32(lum(mats[mi][4:7])), a32))
                centres.append((p32[:, 0] + p32[:, 1] + p32[:, 2]) / np.float32(3.0))  # lib.rs:210 placed_emitter
                n_em_so_far += len(members)
            rows.append((mi | (ROW_GLASS if mats[mi, 28] > 0.0 else 0), 0, first_em, 0))
    rows = np.array(rows, np.uint32).reshape(-1, 4)
    n_em = n_em_so_far
    emitter_tris = (np.concatenate(em_tris).view(np.float32).reshape(-1, 16) if em_tris
                    else np.zeros((0, 16), np.float32))
    wgt = np.concatenate([e[0] for e in entries]).astype(np.float32) if n_em else np.zeros(0, np.float32)
    areas = np.concatenate([e[1] for e in entries]) if n_em else np.zeros(0, np.float32)
    # lib.rs:650-676: the light tree over the emitters (GpuEmitterTri::vertices for its bounds, the placed centroids)
    tree = None
    if light_tree and n_em:
        v0 = emitter_tris[:, 0:3]
        tv = np.stack([v0, v0 + emitter_tris[:, 4:7], v0 + emitter_tris[:, 8:11]], axis=1)
        tree = build_light_tree(tv, np.concatenate(centres), wgt)
    # lib.rs:419-441 emitter_order / emitter_codes: the tree's emitters in leaf order, then the rest in list order;
    # each triangle's code its path (tree) or its entry in the CDF
    if tree is not None:
        placed = np.zeros(n_em, bool)
        placed[tree.order] = True
        order = np.concatenate([tree.order, np.nonzero(~placed)[0].astype(np.uint32)]).astype(np.uint32)
        codes = np.empty(n_em, np.uint32)
        codes[order] = np.arange(n_em, dtype=np.uint32)
        codes[tree.order] = tree.code
        emitter_tris.view(np.uint32)[:, 11] = codes
    else:
        order = np.arange(n_em, dtype=np.uint32)
    # lib.rs:451-481 emitter_cdf: f32 running sum in the CDF's order (np.cumsum accumulates in order, in f32)
    cdf = np.zeros((n_em, 4), np.float32)
    if n_em:
        cdf[:, 0] = order.view(np.float32)
        cdf[:, 1] = np.cumsum(wgt[order], dtype=np.float32)
        cdf[:, 2] = wgt[order]
        cdf[:, 3] = areas[order]
        total = np.float32(cdf[-1, 1])
    else:
        total = np.float32(0.0)
    if total > 0:
        cdf[:, 1] = (cdf[:, 1] / total).astype(np.float32)
    # the guide: for u = k / 65536, the first entry whose cumulative weight is not below u
    us = np.arange(EMITTER_GUIDE + 1, dtype=np.float32) / np.float32(EMITTER_GUIDE)
    guide = np.searchsorted(cdf[:, 1], us, side="left").astype(np.uint32)
    g4 = np.zeros(((EMITTER_GUIDE + 1 + 3) // 4) * 4, np.uint32)
    g4[:EMITTER_GUIDE + 1] = guide
    header = np.array([[total, 0, 0, 0]], np.float32)
    if tree is not None and total > 0:  # lib.rs:672-674: the share of the weight outside the tree
        header[0, 1] = max((total - np.float32(tree.power)) / total, np.float32(0.0))
    emitters = np.concatenate([cdf, g4.view(np.float32).reshape(-1, 4), header]) if n_em else np.zeros((0, 4),
                                                                                                         np.float32)
    objects = np.zeros((OBJECT_STRIDE, 4), np.float32)
    objects[0, 3] = 1.0
    lights = np.stack([lights[k] for k in light_order(lights)]).astype(np.float32) if len(lights) else \
        np.zeros((0, 20), np.float32)
    n_dir = int((lights[:, 3] == 0.0).sum())
    rs = RenderScene(prep, vertices, indices, rows, mats, lights, n_dir, emitters, emitter_tris, objects,
                     energy_lut(), float(total))
    if tree is not None:
        rs.light_tree = tree.nodes.astype(np.float32)
        rs.tree_entries = len(tree.order)
    if np.any(mats[:, 28] > 0.0):
        rs.glass_energy = glass_energy_lut()
    if len(textures):
        rs.tex_pages, rs.tex_side, rs.tex_layers, rs.tex_info = pack_textures(list(textures))
    rs.vcolors = vcolors
    if env is not None:  # an equirectangular map (h, w, 3 or 4): lib.rs SkyBaked's EnvMap and its CDF
        e = np.asarray(env, np.float32)
        if e.ndim != 3 or e.shape[2] not in (3, 4) or min(e.shape[:2]) < 1:
            raise ValueError("env must be an (h, w, 3 or 4) array of texels")
        rs.env = np.concatenate([e[..., :3], np.ones(e.shape[:2] + (1,), np.float32)], axis=2) if e.shape[2] == 3 \
            else np.ascontiguousarray(e)
        rs.env_cdf = build_env_cdf(rs.env)
    return rs


# ---- device side -----------------------------------------------------------------------------------------------
from . import _SceneDesc as _RenderDesc  # noqa: E402  (trt_scene_desc with its shading tables)


def upload(dev: Device, rs: RenderScene) -> Scene:
    """gpu.rs:2613 Gpu::upload with the shading tables."""
    p = rs.prep
    glass = np.zeros((0, 4), np.float32) if rs.glass_energy is None else rs.glass_energy
    keep = [np.ascontiguousarray(x) for x in (p.nodes, p.leaves, p.instances, rs.vertices, rs.indices, rs.rows,
                                              rs.materials, rs.lights, rs.emitters, rs.emitter_tris, rs.objects,
                                              rs.energy, rs.light_tree, np.asarray(glass, np.float32))]
    nodes, leaves, inst, verts, idx, rows, mats, lights, ems, emt, objs, energy, ltree, gl = keep
    if gl.size and gl.shape != (1024, 4):
        raise ValueError("glass_energy must be the (1024, 4) float32 table")
    null = lambda a: _ptr(a) if a.size else None  # noqa: E731
    desc = _RenderDesc(_ptr(nodes), len(nodes), _ptr(leaves), len(leaves), _ptr(inst), len(inst), p.top_first,
                       p.units, null(verts), len(verts), null(idx), len(idx), null(rows), len(rows), null(mats),
                       len(mats), null(lights), len(lights), null(ems), len(ems), null(emt), len(emt), null(objs),
                       len(objs) // OBJECT_STRIDE, _ptr(energy), null(ltree), len(ltree), null(gl))
    if rs.tex_info is not None:
        pages = np.ascontiguousarray(rs.tex_pages, np.uint8)
        info = np.ascontiguousarray(rs.tex_info, np.float32)
        keep += [pages, info]
        desc.tex_pages, desc.tex_side, desc.tex_layers = _ptr(pages), rs.tex_side, rs.tex_layers
        desc.tex_info, desc.tex_info_vec4s = _ptr(info), len(info)
    if rs.vcolors is not None:
        vc = np.ascontiguousarray(rs.vcolors, np.uint32)
        keep.append(vc)
        desc.vcolors, desc.vcolor_words = _ptr(vc), vc.size
    if rs.env is not None:
        env = np.ascontiguousarray(rs.env, np.float32)
        cdf = np.ascontiguousarray(rs.env_cdf, np.float32)
        keep += [env, cdf]
        desc.env, desc.env_width, desc.env_height = _ptr(env), env.shape[1], env.shape[0]
        desc.env_cdf, desc.env_cdf_count = _ptr(cdf), len(cdf)
    sc = Scene.__new__(Scene)
    sc.dev = dev
    h = C.c_void_p()
    _check(lib().trt_scene_upload(dev.h, C.byref(desc), C.byref(h)))
    sc.h = h
    return sc


class Accum:
    """records.wgsl Pixel accumulator in HBM (gpu.rs:2388 accum_buffer)."""

    def __init__(self, dev: Device, width: int, height: int):
        self.dev, self.w, self.h = dev, width, height
        h = C.c_void_p()
        _check(lib().trt_accum_create(dev.h, width, height, C.byref(h)))
        self.handle = h

    def clear(self):
        _check(lib().trt_accum_clear(self.handle))

    def read(self) -> np.ndarray:
        out = np.zeros((self.h, self.w, 20), np.float32)
        _check(lib().trt_accum_read(self.handle, _ptr(out)))
        return out

    def resolve(self):
        c = np.zeros((self.h, self.w, 3), np.float32)
        a = np.zeros_like(c)
        n = np.zeros_like(c)
        v = np.zeros((self.h, self.w), np.float32)
        _check(lib().trt_accum_resolve(self.handle, _ptr(c), _ptr(a), _ptr(n), _ptr(v)))
        return c, a, n, v

    def add_fp64(self, total):
        """trt_accum_add_fp64: this accumulator's color sums and count added to the FP64 image `total` (a device
        buffer of width x height x 4 doubles), queued on the stream."""
        _check(lib().trt_accum_add_fp64(self.handle, total.p))

    def resolve_device(self, pw: int, ph: int, color, albedo, normal):
        """trt_accum_resolve_device: the means into device buffers of pw x ph x 3 floats (queued on the stream)."""
        _check(lib().trt_accum_resolve_device(self.handle, pw, ph, color.p, albedo.p, normal.p))

    def close(self):
        if self.handle:
            lib().trt_accum_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def reference_fp64(dev: Device, scene: Scene, params: np.ndarray, width: int, height: int, spp: int,
                   chunk: int = 64) -> np.ndarray:
    """A converged reference image: samples 0 .. spp - 1 rendered in launches of `chunk` samples, each launch's FP32
    sums added once into an FP64 image on the device (trt_accum_add_fp64); the mean color (height, width, 3) in
    float64."""
    total = dev.alloc(32 * width * height)
    total.upload(np.zeros(4 * width * height, np.float64))
    acc = Accum(dev, width, height)
    for s in range(0, spp, chunk):
        acc.clear()
        render(scene, params, s, min(chunk, spp - s), acc)
        acc.add_fp64(total)
    out = total.download(np.float64, 4 * width * height).reshape(height, width, 4)
    acc.close()
    total.free()
    return out[..., 0:3] / np.maximum(out[..., 3:4], 1.0)


def render(scene: Scene, params: np.ndarray, first_sample: int, spp: int, acc: Accum):
    """gpu.rs:4116 Gpu::render's sample loop: samples first..first+spp-1 into `acc`, queued on the stream."""
    params = np.ascontiguousarray(params, np.uint32)
    if params.size != 156:
        raise ThreeRTError(TRT_E_INVALID, "params must be 156 words")
    _check(lib().trt_render(scene.h, _ptr(params), first_sample, spp, acc.handle))


def render_reads(scene: Scene, params: np.ndarray, first_sample: int, spp: int, acc: Accum) -> dict:
    """render as one counting launch (trt_render_reads), blocking: what its walks read (64-byte wide nodes,
    48-byte triangles, 64-byte units) and the hits it shaded."""
    params = np.ascontiguousarray(params, np.uint32)
    out = np.zeros(4, np.uint64)
    _check(lib().trt_render_reads(scene.h, _ptr(params), first_sample, spp, acc.handle, _ptr(out)))
    return {"nodes": int(out[0]), "tris": int(out[1]), "units": int(out[2]), "shaded": int(out[3])}


def _sig():
    L = lib()
    P, U32, I = C.c_void_p, C.c_uint32, C.c_int
    L.trt_accum_create.argtypes = [P, U32, U32, C.POINTER(C.c_void_p)]
    L.trt_accum_release.argtypes = [P]
    L.trt_accum_clear.argtypes = [P]
    L.trt_accum_read.argtypes = [P, P]
    L.trt_accum_resolve.argtypes = [P, P, P, P, P]
    L.trt_accum_resolve_device.argtypes = [P, U32, U32, P, P, P]
    L.trt_accum_add_fp64.argtypes = [P, P]
    L.trt_render.argtypes = [P, P, U32, U32, P]
    L.trt_render_reads.argtypes = [P, P, U32, U32, P, P]
    for n in ("trt_accum_create", "trt_accum_release", "trt_accum_clear", "trt_accum_read", "trt_accum_resolve",
              "trt_accum_resolve_device", "trt_accum_add_fp64", "trt_render", "trt_render_reads"):
        getattr(L, n).restype = I


_sig()


# ---- seeded workloads --------------------------------------------------------------------------------------------
def quad(p0, p1, p2, p3) -> np.ndarray:
    """Two triangles (p0 p1 p2), (p0 p2 p3)."""
    return np.array([[p0, p1, p2], [p0, p2, p3]], np.float32)


def box_tris(lo, hi) -> np.ndarray:
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    P = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2)]
    return np.concatenate([quad(P[a], P[b], P[c], P[d]) for a, b, c, d in faces])


def cornell(seed: int = 1):
    """The plain-method scene: a unit-cube room of five Lambertian walls (0.73 grey; red 0.65/0.05/0.05; green
    0.12/0.45/0.15), two boxes rotated +-18 degrees, a 0.25 x 0.25 emissive ceiling quad (radiance 17), 36 triangles.
    Returns (RenderScene, Camera)."""
    mats = [material((0.73, 0.73, 0.73), specular=0.0), material((0.65, 0.05, 0.05), specular=0.0),
            material((0.12, 0.45, 0.15), specular=0.0), material((0.0, 0.0, 0.0), (17.0, 17.0, 17.0), specular=0.0)]
    floor = quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0))
    ceiling = quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1))
    back = quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1))
    left = quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1))
    right = quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0))
    light = quad((0.375, 0.999, 0.375), (0.625, 0.999, 0.375), (0.625, 0.999, 0.625), (0.375, 0.999, 0.625))
    tall = box_tris((-0.15, 0.0, -0.15), (0.15, 0.6, 0.15))
    short = box_tris((-0.15, 0.0, -0.15), (0.15, 0.3, 0.15))
    meshes = [RMesh(x) for x in (floor, ceiling, back, left, right, light, tall, short)]
    eye = np.eye(4, dtype=np.float32)
    inst = [RInstance(0, eye, [0]), RInstance(1, eye, [0]), RInstance(2, eye, [0]), RInstance(3, eye, [1]),
            RInstance(4, eye, [2]), RInstance(5, eye, [3]),
            RInstance(6, S.trs((0.33, 0.0, 0.62), (0.0, math.radians(18.0), 0.0), (1, 1, 1)), [0]),
            RInstance(7, S.trs((0.67, 0.0, 0.35), (0.0, math.radians(-18.0), 0.0), (1, 1, 1)), [0])]
    rs = build(meshes, inst, mats)
    cam = Camera((0.5, 0.5, -1.35), (0.5, 0.5, 0.5), 40.0)
    return rs, cam


def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    tris = np.array([[v[a], v[b], v[c]] for a, b, c in f], np.float64)
    tris /= np.linalg.norm(tris, axis=2, keepdims=True)
    for _ in range(subdiv):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = (a + b), (b + c), (c + a)
        ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        tris = np.concatenate([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1),
                               np.stack([ab, bc, ca], 1)])
    return tris.astype(np.float32), tris.astype(np.float32)


def extruded_box(rng: np.random.Generator, n_target: int) -> np.ndarray:
    """A box whose six faces are k x k grids (12 k^2 triangles, about n_target), w x h x d: a building."""
    k = max(1, int(round(math.sqrt(n_target / 12.0))))
    w, d, h = rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.5), rng.uniform(1.0, 6.0)
    g = np.linspace(0.0, 1.0, k + 1)
    i, j = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    u0, u1, v0, v1 = g[i].ravel(), g[i + 1].ravel(), g[j].ravel(), g[j + 1].ravel()
    parts = []
    for axis in range(3):
        for side in (0.0, 1.0):
            corners = []
            for uu, vv in ((u0, v0), (u1, v0), (u1, v1), (u0, v1)):
                p = np.zeros((k * k, 3))
                p[:, axis] = side
                p[:, (axis + 1) % 3] = uu
                p[:, (axis + 2) % 3] = vv
                corners.append(p)
            a, b, c, e = corners
            q = np.stack([np.stack([a, b, c], 1), np.stack([a, c, e], 1)], 1).reshape(-1, 3, 3)
            if side == 0.0:
                q = q[:, ::-1]
            parts.append(q)
    tris = np.concatenate(parts)
    tris = (tris - np.array([0.5, 0.0, 0.5])) * np.array([w, h, d])
    return tris.astype(np.float32)


def uv_sphere(n_target: int) -> tuple[np.ndarray, np.ndarray]:
    """A unit latitude-longitude sphere of r rings and 2r segments (4 r^2 - 4 r triangles, about n_target), with its
    vertex normals: a subdivided sphere of any size."""
    r = max(2, int(round(math.sqrt(n_target / 4.0))))
    th = np.linspace(0.0, math.pi, r + 1)
    ph = np.linspace(0.0, 2 * math.pi, 2 * r + 1)
    P = np.stack([np.sin(th)[:, None] * np.cos(ph)[None, :], np.repeat(np.cos(th)[:, None], 2 * r + 1, 1),
                  np.sin(th)[:, None] * np.sin(ph)[None, :]], -1)
    i, j = np.meshgrid(np.arange(r), np.arange(2 * r), indexing="ij")
    a, b, c, d = P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]
    top = np.stack([a, c, d], 2)[1:]        # (the first ring's top triangles are degenerate at the pole)
    bottom = np.stack([a, b, c], 2)[:-1]    # (the last ring's bottom ones likewise)
    tris = np.concatenate([top.reshape(-1, 3, 3), bottom.reshape(-1, 3, 3)])
    return tris.astype(np.float32), tris.astype(np.float32)


def city(seed: int = 1, meshes: int = 64, instances: int = 4096, extent: float = 400.0):
    """The instanced city: a ground plane and `meshes` - 1 seeded base meshes (extruded boxes and subdivided spheres,
    200-40k triangles: 200 + 39,800 u^1.5, a mean near 16k, so 64 meshes hold about 1M unique triangles) placed
    `instances` times by random TRS (about 60M instanced at 4096), metal-rough GGX (roughness U[0.05, 1], metallic
    Bernoulli(0.3)), a 90,000-lux sun and a constant sky. Returns (RenderScene, Camera, sky)."""
    rng = np.random.default_rng(seed)
    ms = [RMesh(quad((-extent, 0, -extent), (-extent, 0, extent), (extent, 0, extent), (extent, 0, -extent)))]
    for k in range(meshes - 1):
        n = int(200 + 39800 * rng.uniform() ** 1.5)
        if k % 2 == 0:
            ms.append(RMesh(extruded_box(rng, n)))
        else:
            t, nrm = uv_sphere(n)
            ms.append(RMesh(t, normals=nrm))
    mats, inst = [material((0.5, 0.5, 0.5), roughness=0.8)], [RInstance(0, np.eye(4, dtype=np.float32), [0])]
    for i in range(instances - 1):
        m = int(rng.integers(1, meshes))
        s = rng.uniform(2.0, 12.0)
        pos = ((rng.random() - 0.5) * 2 * extent * 0.9, 0.0 if m % 2 == 1 else s * 0.9, (rng.random() - 0.5) * 2 * extent * 0.9)
        w = S.trs(pos, (0.0, rng.uniform(0, 2 * math.pi), 0.0), (s, s * rng.uniform(0.6, 1.8), s))
        base = rng.uniform(0.2, 0.9, 3)
        mats.append(material(tuple(base), metallic=float(rng.random() < 0.3), roughness=float(rng.uniform(0.05, 1.0))))
        inst.append(RInstance(m, w, [len(mats) - 1]))
    sun = directional_light((-0.4, -0.8, -0.45), (90000.0, 88000.0, 82000.0))
    rs = build(ms, inst, mats, [sun])
    sizes = [len(m.tris) for m in ms]
    rs.extras.update(unique_triangles=int(sum(sizes)), instanced_triangles=int(sum(sizes[i.mesh] for i in inst)))
    # the camera above the skyline (the tallest boxes reach about 130 m), off to one side, looking at the centre: the
    # top of each instance bounded by its mesh box's transformed corners
    boxes = []
    for m in ms:
        t = np.asarray(m.tris, np.float64).reshape(-1, 3)
        lo, hi = t.min(axis=0), t.max(axis=0)
        boxes.append(np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])]))
    top = max(float((boxes[i.mesh] @ np.asarray(i.world, np.float64)[:3, :3].T
                     + np.asarray(i.world, np.float64)[:3, 3])[:, 1].max()) for i in inst)
    cam = Camera((-extent * 0.625, max(0.45 * extent, 1.1 * top), -extent * 0.75), (0.0, 0.0, 0.0), 55.0)
    return rs, cam, (6000.0, 8000.0, 12000.0)
