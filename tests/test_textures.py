"""Textures: the pages (texpack.rs), the kernel's trilinear read (pathtrace.wgsl:1102-1117 tex / tex_level,
srgb_to_linear), the level from the ray cone and the triangle's texel density (mesh_uv / mesh_density / mesh_lod
1119-1140, the path loop's cone 4165-4167), and the maps of `surface` (colour, roughness, metalness, emissive, normal,
transmission: 1344-1425), glass_at (950-956) and emitter_radiance (2269-2283), in the oracle.

Pinned by:
- log2 / exp2 (the kernel's, Cephes) within 1.5 ulp of float64 over their ranges;
- the packer's layout against texpack.rs's rules (gutters of wrapped or clamped texels, 2 x 2 box mips rounded to
  nearest, the info rows), and the read against the sampler's definition (gpu.rs:2690: repeat, linear, linear): a
  texel centre reads the texel, half-way reads the mean, u + 1 reads u (wrap), a clamped texture's edge, level 0.5
  the mean of levels 0 and 1;
- the reference's analytic texture render tests (tests/render/*.ts): texture-srgb-decode (sRGB 188 reads 0.5029),
  texture-level-instance (a magnified checker on a scaled instance reads 0 and 1, not the 0.5 of a level five too
  high), material-glass-transmission-map (the mapped half of a clear slab passes 92.3 % of a wall), and
  shadow-glass-map-instance with a sun for its point light (a pane's white square passes 0.96^2);
- closed forms for the other maps: a metalness map's blue and a roughness map's green multiply the factors (the
  same bits as the factor alone), an emissive map scales the emission by its decoded colour, a normal map tilted by
  theta toward the tangent lights a plane under a sun along its normal by cos theta, and toward a sun off the
  normal as the tilted normal's cosine (explicit tangents; the bitangent is cross(normal, tangent) x hand, glTF 2.0
  section 3.7.2.1). The tangent derived from the uvs where the vertices have none (hit_tangent's hand rule) has no
  fixture of the reference's: its handedness is parity unpinned beyond the oracle's restatement.
"""
import math

import numpy as np
import pytest

import oracle
from threert_amd import render as R
from threert_amd import scene as S


def oracle_scene(rs):
    p = rs.prep
    return oracle.RenderScene(p.nodes, p.leaves, p.instances, p.top_first, p.units, **rs.tables())


def ulps(got, want):
    return abs(float(got) - want) / float(np.spacing(np.float32(abs(want))))


def test_log2_and_exp2_within_ulps():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0.5, 2.0, 2000), np.exp(rng.uniform(-80, 80, 3000)),
                         rng.uniform(1 - 1e-3, 1 + 1e-3, 500)]).astype(np.float32)
    assert max(ulps(oracle.log2(float(x)), math.log2(float(x))) for x in xs if x != 1.0) <= 1.5
    assert oracle.log2(1.0) == 0.0 and oracle.log2(8.0) == 3.0 and oracle.log2(0.0) == -math.inf
    ys = np.concatenate([rng.uniform(-20, 20, 3000), rng.uniform(-126, 127, 2000)]).astype(np.float32)
    assert max(ulps(oracle.exp2(float(y)), 2.0 ** float(y)) for y in ys) <= 1.5
    assert oracle.exp2(3.0) == 8.0 and oracle.exp2(-1.0) == 0.5


# ---- the pages and the read -----------------------------------------------------------------------------------
def quad_scene(textures, mat_kw, tris=None, uvs=None):
    tris = R.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)) if tris is None else tris
    if uvs is None:
        uvs = np.array([[[0, 1], [1, 1], [1, 0]], [[0, 1], [1, 0], [0, 0]]], np.float32)
    rs = R.build([R.RMesh(tris, uvs=uvs)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0])],
                 [R.material(**mat_kw)], textures=textures)
    return rs


def test_the_pages_hold_each_texture_with_its_gutter_and_mips():
    """texpack.rs pack: the content at (x + 16, y + 16) of its padded rect, the 16-texel gutter filled by wrapped
    (wrap) or clamped texels, rects of multiples of 16, the tallest first; each level the 2 x 2 box of the one above,
    (a + b + c + d + 2) // 4; info (u0, v0, du, dv) in page UV and (layer, wrap, width, height)."""
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, (20, 37, 4), dtype=np.uint8)
    b = rng.integers(0, 256, (64, 8, 4), dtype=np.uint8)
    pages, side, layers, info = R.pack_textures([R.Texture(a, True), R.Texture(b, False)])
    assert side == 256 and layers == 1 and info.shape == (4, 4)
    lv0 = pages[:side * side * 4].reshape(side, side, 4)
    # b pads to 48 x 96 and goes first, at (0, 0); a (37 x 20) pads to 80 x 64 and follows it on the shelf
    for k, (img, wrap) in enumerate(((a, True), (b, False))):
        u0, v0, du, dv = info[2 * k]
        x0, y0 = int(round(u0 * side)) - 16, int(round(v0 * side)) - 16
        h, w = img.shape[:2]
        assert (du, dv) == (np.float32(w) / np.float32(side), np.float32(h) / np.float32(side))
        assert tuple(info[2 * k + 1]) == (0.0, 1.0 if wrap else 0.0, w, h)
        assert x0 % 16 == 0 and y0 % 16 == 0
        rect = lv0[y0:y0 + h + 32, x0:x0 + w + 32]
        assert np.array_equal(rect[16:16 + h, 16:16 + w], img)
        ys, xs = np.arange(-16, h + 16), np.arange(-16, w + 16)
        if wrap:
            want = img[(ys % h)[:, None], (xs % w)[None, :]]
        else:
            want = img[np.clip(ys, 0, h - 1)[:, None], np.clip(xs, 0, w - 1)[None, :]]
        assert np.array_equal(rect, want)
    assert int(round(info[2][1] * side)) - 16 == 0  # b, the taller, is first on the shelf
    # the mip chain
    at = side * side * 4
    lv1 = pages[at:at + (side // 2) ** 2 * 4].reshape(side // 2, side // 2, 4).astype(np.int64)
    q = lv0.astype(np.int64)
    box = (q[0::2, 0::2] + q[0::2, 1::2] + q[1::2, 0::2] + q[1::2, 1::2] + 2) // 4
    assert np.array_equal(lv1, box)
    assert len(pages) == sum((side >> k) ** 2 * 4 for k in range(5))


def test_the_read_is_the_samplers():
    """Repeat addressing, linear within and between levels: at a texel centre the texel (as unorm), half-way
    between two the mean, u + 1 the same as u on a wrapped texture, past the edge of a clamped one its edge texel,
    level 0.5 the mean of levels 0 and 1."""
    img = np.zeros((8, 8, 4), np.uint8)
    img[..., 0] = (np.arange(8) * 30)[None, :]      # red ramps along u
    img[..., 1] = (np.arange(8) * 20)[:, None]      # green along v
    img[..., 2] = 200
    img[..., 3] = 255
    noise = np.random.default_rng(2).integers(0, 256, (8, 8, 4), dtype=np.uint8)
    rs = quad_scene([R.Texture(img, True), R.Texture(img, False), R.Texture(noise, True)], {"base_map": 0})
    sc = oracle_scene(rs)
    lod0 = -0.5 * math.log2(64) - 2  # well below level 0: the level clamps to 0
    for i in range(8):
        for j in range(8):
            got = sc.tex(0, (i + 0.5) / 8, (j + 0.5) / 8, lod0)
            assert np.allclose(got, img[j, i] / 255.0, atol=2e-6), (i, j, got)
    got = sc.tex(0, 3.0 / 8, 2.5 / 8, lod0)  # between texels 2 and 3 along u
    assert np.allclose(got[0], (img[2, 2, 0] + img[2, 3, 0]) / 2 / 255.0, atol=2e-6)
    assert np.allclose(sc.tex(0, 0.3 + 1.0, 0.6 - 2.0, lod0), sc.tex(0, 0.3, 0.6, lod0), atol=2e-6)
    assert np.allclose(sc.tex(1, 1.7, 0.5 / 8, lod0)[0], img[0, 7, 0] / 255.0, atol=2e-6)  # clamped to the edge
    assert np.allclose(sc.tex(1, -0.4, 0.5 / 8, lod0)[0], img[0, 0, 0] / 255.0, atol=2e-6)
    # levels: the texture's own size is added to the lod (8 x 8: 3), then 0.5 is half-way between 0 and 1 (a noise
    # texture: a ramp's levels agree, the box mips keeping a linear function)
    l0 = sc.tex(2, 0.37, 0.41, -3.0)
    l1 = sc.tex(2, 0.37, 0.41, -2.0)
    lh = sc.tex(2, 0.37, 0.41, -2.5)
    assert np.allclose(lh, (l0 + l1) / 2, atol=3e-6) and not np.allclose(l0, l1, atol=1e-3)


# ---- renders against the reference's texture tests ---------------------------------------------------------
def flat_texture(v, size=8):
    px = np.full((size, size, 4), v, np.uint8)
    px[..., 3] = 255
    return px


def region_mean(acc, x0, y0, x1, y1):
    c = acc[..., :3] / np.maximum(acc[..., 3:4], 1)
    h, w = c.shape[:2]
    return c[int(y0 * h):int(math.ceil(y1 * h)), int(x0 * w):int(math.ceil(x1 * w))].reshape(-1, 3).mean(axis=0)


def test_srgb_maps_decode_to_linear():
    """texture-srgb-decode.ts: a map of sRGB 188 on a white Lambert plane under a 1-nit sky reads 0.5029 +- 2 %."""
    plane = R.quad((-1000, 0, -1000), (-1000, 0, 1000), (1000, 0, 1000), (1000, 0, -1000))
    uvs = np.array([[[0, 0], [0, 1], [1, 1]], [[0, 0], [1, 1], [1, 0]]], np.float32) * 1000
    rs = R.build([R.RMesh(plane, uvs=uvs)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0])],
                 [R.material((1, 1, 1), specular=0.0, roughness=1.0, base_map=0)],
                 textures=[R.Texture(flat_texture(188))])
    cam = R.Camera((0.0, 2.0, 2.0), (0.0, 0.0, 0.0), 30.0)
    p = rs.params(32, 18, cam, max_bounces=4, rr_depth=8, sky=(1.0, 1.0, 1.0))
    m = region_mean(oracle_scene(rs).render(p, 0, 64), 0.45, 0.45, 0.55, 0.55)
    assert np.all(np.abs(m - 0.5029) < 0.02 * 0.5029), m


def checker(alpha=False):
    px = np.zeros((64, 64, 4), np.uint8)
    on = ((np.arange(64)[:, None] >> 3) + (np.arange(64)[None, :] >> 3)) % 2 == 0
    px[..., :3] = np.where(on, 255, 0)[..., None]
    px[..., 3] = 255
    return px


def test_the_level_follows_the_instanced_triangle():
    """texture-level-instance.ts: a 2 m quad whose mesh is 1/32 of it, scaled 32 times by its instance, with an 8 x 8
    checker of 64 x 64 texels, straight on from 3 m (a square 14 pixels wide: magnified): inside a black square 0
    and inside a white one 1. From the mesh's own size the level would be five too high and both would read 0.5."""
    s = 1.0 / 32
    tris = R.quad((-s, -s, 0), (s, -s, 0), (s, s, 0), (-s, s, 0))
    rs = R.build([R.RMesh(tris, uvs=np.array([[[0, 1], [1, 1], [1, 0]], [[0, 1], [1, 0], [0, 0]]], np.float32))],
                 [R.RInstance(0, S.trs((0, 0, 0), (0, 0, 0), (32, 32, 32)), [0])],
                 [R.material((1, 1, 1), specular=0.0, roughness=1.0, base_map=0)],
                 textures=[R.Texture(checker())])
    cam = R.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 30.0)
    w, h = 160, 90  # the reference test's 16:9; a square is 0.25 m, 1/8 of the quad
    p = rs.params(w, h, cam, max_bounces=2, rr_depth=8, sky=(1.0, 1.0, 1.0))
    acc = oracle_scene(rs).render(p, 0, 16)
    c = acc[..., :3] / np.maximum(acc[..., 3:4], 1)
    # the two squares just above the centre, left and right of it (the image's centre is the quad's)
    tan = math.tan(math.radians(15.0))
    px = lambda x: w / 2 + x / (3.0 * tan) * (h / 2)  # noqa: E731  (metres on the quad -> pixels)
    py = lambda y: h / 2 - y / (3.0 * tan) * (h / 2)  # noqa: E731
    left = c[int(py(0.2)):int(py(0.05)), int(px(-0.2)):int(px(-0.05))].reshape(-1, 3).mean(axis=0)
    right = c[int(py(0.2)):int(py(0.05)), int(px(0.05)):int(px(0.2))].reshape(-1, 3).mean(axis=0)
    lo, hi = sorted((float(left.mean()), float(right.mean())))
    assert lo < 0.03 * hi and abs(hi - 1.0) < 0.03, (left, right)


def test_a_transmission_map_masks_the_glass():
    """material-glass-transmission-map.ts: the clear slab before a 100-nit wall with a transmission map of 1 in its
    top half and 0 below: the top passes (1 - 0.04)^2 / (1 - 0.04^2) of the wall (+- 2 %); the bottom, a white
    diffuse box lit from behind only, reads 0."""
    T = (1 - 0.04) ** 2 / (1 - 0.04 ** 2)
    px = np.zeros((64, 1, 4), np.uint8)
    px[:32] = 255
    px[32:, :, 3] = 255
    wall = R.quad((-5, -5, -2), (5, -5, -2), (5, 5, -2), (-5, 5, -2))
    slab = R.box_tris((-0.5, -0.5, -0.025), (0.5, 0.5, 0.025))
    # v from the slab's position: 0 at its top, 1 at its bottom (row 0 of the map is the top); u along x
    uvs = np.stack([np.stack([(t[:, 0] + 0.5), (0.5 - t[:, 1])], axis=1) for t in slab]).astype(np.float32)
    rs = R.build([R.RMesh(wall), R.RMesh(slab, uvs=uvs)],
                 [R.RInstance(0, np.eye(4, dtype=np.float32), [0]), R.RInstance(1, np.eye(4, dtype=np.float32), [1])],
                 [R.material((0, 0, 0), emissive=(100.0, 100.0, 100.0)),
                  R.material((1, 1, 1), metallic=0.0, roughness=0.0, transmission=1.0, ior=1.5, transmission_map=0)],
                 light_tree=False, textures=[R.Texture(px, wrap=False)])
    cam = R.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 10.0)
    p = rs.params(48, 27, cam, max_bounces=16, rr_depth=16)
    acc = oracle_scene(rs).render(p, 0, 32)
    top = region_mean(acc, 0.45, 0.15, 0.55, 0.3)
    bottom = region_mean(acc, 0.45, 0.7, 0.55, 0.85)
    assert np.all(np.abs(top - 100 * T) < 0.02 * 100 * T), top
    assert np.all(bottom < 0.5), bottom


def test_a_pane_map_colours_the_sun_shadow():
    """shadow-glass-map-instance.ts with a sun (1000 lux, straight down) for its point light: a clear pane (ior 1.5,
    smooth, 5 cm) whose colour map is white in a square at its middle (24 of 64 texels) and black around, its mesh
    32 times the pane and its instance scaled 1/32; the albedo-0.5 plane under its middle reads 0.5 E / pi x 0.96^2
    (+- 2 %): both surfaces pass their white square at the pane's own level. Five levels too high the square is
    grey."""
    px = np.zeros((64, 64, 4), np.uint8)
    px[20:44, 20:44, :3] = 255
    px[..., 3] = 255
    k = 32.0
    plane = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    pane = R.box_tris((-0.5 * k, -0.025 * k, -0.5 * k), (0.5 * k, 0.025 * k, 0.5 * k))
    uvs = np.stack([np.stack([t[:, 0] / k + 0.5, t[:, 2] / k + 0.5], axis=1) for t in pane]).astype(np.float32)
    rs = R.build([R.RMesh(plane), R.RMesh(pane, uvs=uvs)],
                 [R.RInstance(0, np.eye(4, dtype=np.float32), [0]),
                  R.RInstance(1, S.trs((0, 1, 0), (0, 0, 0), (1 / k, 1 / k, 1 / k)), [1])],
                 [R.material((0.5, 0.5, 0.5), specular=0.0),
                  R.material((1, 1, 1), metallic=0.0, roughness=0.0, transmission=1.0, ior=1.5, base_map=0)],
                 [R.directional_light((0.0, -1.0, 0.0), 1000.0)], textures=[R.Texture(px)])
    cam = R.Camera((0.0, 0.6, 3.0), (0.0, 0.0, 0.0), 4.0)
    p = rs.params(16, 16, cam, max_bounces=1, rr_depth=8)
    m = region_mean(oracle_scene(rs).render(p, 0, 16), 0.4, 0.4, 0.6, 0.6)
    want = 0.5 * 1000 / math.pi * 0.96 ** 2
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


# ---- the other maps -------------------------------------------------------------------------------------------
def sunlit_plane(mat_kw, textures, sun=(0.0, -1.0, 0.0), tangents=None, lux=1000.0):
    """A 20 m Lambert plane facing +y (u along +x, v along +z) under one sun, seen from above; direct light only."""
    plane = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    uvs = np.stack([np.stack([(t[:, 0] + 10) / 20, (t[:, 2] + 10) / 20], axis=1) for t in plane]).astype(np.float32)
    rs = R.build([R.RMesh(plane, uvs=uvs, tangents=tangents)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0])],
                 [R.material(**mat_kw)], [R.directional_light(sun, lux, diameter_deg=0.1)], textures=textures)
    cam = R.Camera((0.0, 5.0, 0.01), (0.0, 0.0, 0.0), 20.0)
    return rs, rs.params(16, 16, cam, max_bounces=1, rr_depth=8)


def test_roughness_and_metalness_maps_multiply_their_factors():
    """A metalness map reads its blue, a roughness map its green: a metallic-1 plane under a map of blue 0 renders
    the bits of the metallic-0 plane, and roughness 1 under a map of green 64 those of roughness 64 / 255."""
    px = flat_texture(0)
    px[..., 1] = 64
    px[..., 0] = 255
    texs = [R.Texture(px)]
    base = {"base": (0.7, 0.6, 0.5), "specular": 0.5}
    rs_a, p = sunlit_plane({**base, "metallic": 1.0, "metalness_map": 0, "roughness": 0.5}, texs,
                           sun=(0.3, -1.0, 0.2))
    rs_b, _ = sunlit_plane({**base, "metallic": 0.0, "roughness": 0.5}, texs, sun=(0.3, -1.0, 0.2))
    a, b = oracle_scene(rs_a).render(p, 0, 4), oracle_scene(rs_b).render(p, 0, 4)
    assert np.array_equal(a[..., :3].view(np.uint32), b[..., :3].view(np.uint32))
    # (the filter's weights may move a constant texel by an ulp: the roughness agrees to float rounding)
    rough = float(np.float32(64) * np.float32(1.0 / 255.0))
    rs_a, p = sunlit_plane({**base, "roughness": 1.0, "roughness_map": 0}, texs, sun=(0.3, -1.0, 0.2))
    rs_b, _ = sunlit_plane({**base, "roughness": rough}, texs, sun=(0.3, -1.0, 0.2))
    a, b = oracle_scene(rs_a).render(p, 0, 4), oracle_scene(rs_b).render(p, 0, 4)
    assert np.allclose(a[..., :3], b[..., :3], rtol=1e-5, atol=1e-6)
    rs_c, _ = sunlit_plane({**base, "roughness": 1.0}, texs, sun=(0.3, -1.0, 0.2))
    assert not np.allclose(oracle_scene(rs_c).render(p, 0, 4)[..., :3], b[..., :3], rtol=1e-3)


def test_an_emissive_map_scales_the_emission():
    """An emissive plane (factor 2) under a map of sRGB 188, seen directly, reads 2 x 0.5029 = 1.0058 (+- 0.5 %)."""
    rs, p = sunlit_plane({"base": (0, 0, 0), "emissive": (2.0, 2.0, 2.0), "specular": 0.0, "emissive_map": 0},
                         [R.Texture(flat_texture(188))], lux=0.0)
    m = region_mean(oracle_scene(rs).render(p, 0, 4), 0.25, 0.25, 0.75, 0.75)
    assert np.all(np.abs(m - 2 * 0.502886) < 0.005 * 1.0058), m


def normal_texel(n):
    """A normal map texel (8 bits) for the unit tangent-space normal n, and the normal it decodes to."""
    q = np.round((np.asarray(n) + 1.0) / 2.0 * 255.0).astype(np.uint8)
    return q, q.astype(np.float64) / 255.0 * 2.0 - 1.0


@pytest.mark.parametrize("theta_deg,scale", [(0.0, 1.0), (30.0, 1.0), (50.0, 0.5)])
def test_a_normal_map_tilts_the_shading_normal(theta_deg, scale):
    """A constant normal map tilted by theta toward the tangent (explicit vertex tangents along +x, hand 1): under a
    sun along the plane's normal a white Lambert plane reads E / pi x cos of the decoded normal's tilt (its x and y
    scaled by the material's normal scale), and under a sun 45 degrees off the normal toward +x, E / pi x the tilted
    normal's cosine with it (+- 0.5 %). A flat map leaves the plane as it is."""
    th = math.radians(theta_deg)
    q, dec = normal_texel((math.sin(th), 0.0, math.cos(th)))
    px = np.zeros((8, 8, 4), np.uint8)
    px[..., :3] = q
    px[..., 3] = 255
    tan = np.tile(np.array([1.0, 0.0, 0.0, 1.0], np.float32), (2, 3, 1))
    mat = {"base": (1, 1, 1), "specular": 0.0, "normal_map": 0, "normal_scale": scale}
    # the shading normal in world space: tangent +x, bitangent cross(n, t) * hand = -z, normal +y
    nw = np.array([dec[0] * scale, dec[2], -dec[1] * scale])
    nw /= np.linalg.norm(nw)
    for sun in ((0.0, -1.0, 0.0), (-math.sqrt(0.5), -math.sqrt(0.5), 0.0)):
        rs, p = sunlit_plane(mat, [R.Texture(px)], sun=sun, tangents=tan)
        m = region_mean(oracle_scene(rs).render(p, 0, 8), 0.25, 0.25, 0.75, 0.75)
        l = -np.asarray(sun)
        want = 1000.0 / math.pi * max(float(nw @ l), 0.0)
        assert np.all(np.abs(m - want) < 0.005 * want + 1e-3), (sun, m, want)


def test_the_maps_leave_scenes_without_textures_alone():
    """A scene with pages but materials naming no texture renders the bits of the same scene without pages."""
    rs, cam = R.cornell()
    p = rs.params(24, 24, cam, seed=3, max_bounces=4, rr_depth=3)
    want = oracle_scene(rs).render(p, 0, 2)
    rs.tex_pages, rs.tex_side, rs.tex_layers, rs.tex_info = R.pack_textures([R.Texture(checker())])
    got = oracle_scene(rs).render(p, 0, 2)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


# ---- GPU parity ---------------------------------------------------------------------------------------------------
def textured_garden(light_tree: bool = False):
    """Every map on the megakernel: a floor under a repeated checker (colour), a box with roughness and metalness
    maps, a sphere with a normal map and tangents derived from its uvs, a second with vertex tangents, an emissive
    panel of 288 triangles under an emissive map (a light tree over them with light_tree), a glass pane with a colour
    map and a glass slab with a transmission map, a sun and a sky."""
    rng = np.random.default_rng(11)
    noise = rng.integers(0, 256, (32, 48, 4), dtype=np.uint8)
    normals = np.zeros((16, 16, 4), np.uint8)
    yy, xx = np.mgrid[0:16, 0:16]
    nx, ny = 0.4 * np.sin(xx * 0.8), 0.4 * np.cos(yy * 0.6)
    nz = np.sqrt(1 - nx * nx - ny * ny)
    normals[..., 0], normals[..., 1], normals[..., 2] = [np.round((c + 1) / 2 * 255) for c in (nx, ny, nz)]
    normals[..., 3] = 255
    stripes = np.zeros((64, 1, 4), np.uint8)
    stripes[::8] = 255
    stripes[..., 3] = 255
    textures = [R.Texture(checker()), R.Texture(noise), R.Texture(normals), R.Texture(stripes, wrap=False)]
    ground = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    guv = np.stack([np.stack([t[:, 0] / 4, t[:, 2] / 4], axis=1) for t in ground]).astype(np.float32)
    box = R.box_tris((-0.4, 0.0, -0.4), (0.4, 0.9, 0.4))
    buv = np.stack([np.stack([t[:, 0] + t[:, 2] + 0.5, t[:, 1] + 0.3 * t[:, 2]], axis=1) for t in box]).astype(np.float32)
    t, n = R.icosphere(3)
    suv = np.stack([np.arctan2(t[..., 0], t[..., 2]) / (2 * math.pi) + 0.5, np.arccos(np.clip(t[..., 1], -1, 1)) / math.pi],
                   axis=-1).astype(np.float32)
    tang = np.concatenate([np.stack([t[..., 2], np.zeros_like(t[..., 0]), -t[..., 0]], axis=-1),
                           np.ones(t.shape[:2] + (1,), np.float32)], axis=-1).astype(np.float32)
    pane = R.box_tris((-0.7, -0.02, -0.45), (0.7, 0.02, 0.45))
    puv = np.stack([np.stack([t_[:, 0] + 0.5, t_[:, 2] + 0.5], axis=1) for t_ in pane]).astype(np.float32)
    g = 12
    panel = np.concatenate([R.quad((x0, 2.2, z0), (x0 + 0.1, 2.2, z0), (x0 + 0.1, 2.2, z0 + 0.1), (x0, 2.2, z0 + 0.1))
                            for x0 in np.linspace(-0.6, 0.5, g) for z0 in np.linspace(-0.6, 0.5, g)])
    pnuv = np.stack([np.stack([(q[:, 0] + 0.6) / 1.2, (q[:, 2] + 0.6) / 1.2], axis=1) for q in panel]).astype(np.float32)
    meshes = [R.RMesh(ground, uvs=guv), R.RMesh(box, uvs=buv), R.RMesh(t, normals=n, uvs=suv),
              R.RMesh(t, normals=n, uvs=suv, tangents=tang), R.RMesh(pane, uvs=puv), R.RMesh(panel, uvs=pnuv)]
    mats = [R.material((0.8, 0.8, 0.8), specular=0.5, roughness=0.7, base_map=0, repeat=(2.0, 2.0)),
            R.material((0.9, 0.5, 0.3), metallic=1.0, roughness=1.0, specular=0.5, roughness_map=1, metalness_map=1),
            R.material((0.7, 0.7, 0.75), roughness=0.3, specular=0.5, normal_map=2, normal_scale=1.5),
            R.material((0.3, 0.6, 0.9), roughness=0.15, metallic=0.5, normal_map=2),
            R.material((1.0, 1.0, 1.0), roughness=0.0, transmission=1.0, ior=1.45, base_map=1),
            R.material((0.0, 0.0, 0.0), emissive=(6.0, 5.0, 4.0), specular=0.0, emissive_map=1),
            R.material((1.0, 1.0, 1.0), roughness=0.3, transmission=1.0, ior=1.5, transmission_map=3)]
    inst = [R.RInstance(0, np.eye(4, dtype=np.float32), [0]),
            R.RInstance(1, S.trs((1.3, 0.0, 0.8), (0.0, 0.4, 0.0), (1, 1, 1)), [1]),
            R.RInstance(2, S.trs((-0.5, 0.6, 0.0), (0.0, 0.3, 0.0), (0.6, 0.6, 0.6)), [2]),
            R.RInstance(3, S.trs((0.7, 0.45, -0.6), (0.2, 0.0, 0.0), (0.45, 0.45, 0.45)), [3]),
            R.RInstance(4, S.trs((0.2, 1.3, -0.2), (0.5, 0.3, 0.2), (1, 1, 1)), [4]),
            R.RInstance(5, np.eye(4, dtype=np.float32), [5]),
            R.RInstance(1, S.trs((-1.4, 0.0, -0.9), (0.0, -0.3, 0.0), (0.8, 1.2, 0.1)), [6])]
    rs = R.build(meshes, inst, mats, [R.directional_light((-0.4, -1.0, 0.5), 3.0)], light_tree=light_tree,
                 textures=textures)
    cam = R.Camera((0.2, 1.6, -4.2), (0.0, 0.6, 0.0), 45.0)
    return rs, cam


def gpu_render(device, rs, p, first, spp, reads=False):
    sc = R.upload(device, rs)
    acc = R.Accum(device, int(p[16]), int(p[17]))
    counts = R.render_reads(sc, p, first, spp, acc) if reads else R.render(sc, p, first, spp, acc)
    out = acc.read()
    device.sync()
    acc.close()
    sc.close()
    return out, counts


def rel_rmse(a, b):
    ca = a[..., :3] / np.maximum(a[..., 3:4], 1)
    cb = b[..., :3] / np.maximum(b[..., 3:4], 1)
    den = np.maximum(np.abs(cb), 1e-3)
    return float(np.sqrt(np.mean(((ca - cb) / den) ** 2)))


@pytest.mark.gpu
@pytest.mark.parametrize("light_tree", [False, True])
def test_textured_scenes_match_the_oracle(device, light_tree):
    """The textured garden on the megakernel against the oracle, same seed and samples: relative RMSE <= 1e-3 (the
    float operations are the oracle's in its order, so in practice the bits agree), with and without the light
    tree."""
    print(f"textured garden 64x48 x 4 spp, light tree {light_tree}", flush=True)
    rs, cam = textured_garden(light_tree)
    assert (rs.tree_entries > 0) == light_tree
    p = rs.params(64, 48, cam, seed=5, max_bounces=6, rr_depth=3, sky=(0.3, 0.4, 0.6), energy_compensation=1.0)
    want = oracle_scene(rs).render(p, 0, 4)
    got, _ = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"textured garden: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_textures_render_on_the_megakernel_whatever_the_setting(device):
    """The wavefront does not read textures: under RENDER_WAVEFRONT, RENDER_AUTO and the counting launch a textured
    scene takes the megakernel, the same bits each way; a texture slot past the pages is refused at upload."""
    import copy

    import threert_amd as T

    rs, cam = textured_garden(False)
    p = rs.params(32, 24, cam, seed=3, max_bounces=5, rr_depth=3, sky=(0.3, 0.4, 0.6))
    outs = []
    try:
        for how in (T.RENDER_WAVEFRONT, T.RENDER_AUTO, T.RENDER_MEGAKERNEL):
            print(f"textured garden under render setting {how}", flush=True)
            device.set_render(how)
            outs.append(gpu_render(device, rs, p, 0, 2)[0])
    finally:
        device.set_render(T.RENDER_AUTO)
    print("textured garden: counting launch", flush=True)
    counted, reads = gpu_render(device, rs, p, 0, 2, reads=True)
    assert reads["nodes"] > 0 and reads["shaded"] > 0, reads
    for o in outs[1:] + [counted]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
    bad = copy.deepcopy(rs)
    bad.materials[2, 13] = np.array([4], np.int32).view(np.float32)[0]  # a normal map past the four textures
    with pytest.raises(T.ThreeRTError) as e:
        R.upload(device, bad)
    assert e.value.code == T.TRT_E_INVALID and "normal map names texture 4" in str(e.value), str(e.value)
