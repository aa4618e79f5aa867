"""three.js's unlit materials (MeshBasicMaterial: pathtrace.wgsl:1366-1372, the reference's F_UNLIT), scene
background (sky_mis 1963-1966, F_BACKGROUND) and vertex colours (vertexColors: vcolor 269-275 in surface 1356 and
glass_at 949, F_VCOLOR; scene.rs:2324-2334 tri_colors), in the oracle, pinned by the reference's analytic render
tests, and on the GPU against it.

- material-unlit.ts: a plane of colour 0.5, unlit, black sky, no light: radiance 0.5.
- material-unlit-lights-nothing.ts: a white Lambert floor before an unlit panel of colour 1: the floor stays black.
  A mirror floor (roughness below 0.3: a mirror-like path) shows the panel at its colour times the mirror's.
- environment-background.ts: camera rays that miss see the background's radiance (0.5 / pi).
- environment-background-lights-nothing.ts: a white floor under a white background and a black sky stays black.
- A smooth glass slab before the background: the paths that pass through the glass see it (0.96 of it: what the
  front face does not reflect).
- material-vertex-colors.ts: colours 0 .. 1 across a white quad in a 1-nit sky read 0.5 at its centre (interpolated
  in linear); -mirrored.ts: on an instance mirrored in x, 0.75 over 0.25; a vertex-coloured glass pane tints a sun's
  shadow by its colour at each surface.
- The GPU (megakernel) against the oracle on scenes with each; the wavefront setting hands such batches to the
  megakernel.
"""
import math

import numpy as np
import pytest

from threert_amd import render as R
from test_glass import gpu_render, oracle_scene, rel_rmse


def build(parts):
    """parts: (tris, material) pairs, one instance each."""
    meshes = [R.RMesh(t) for t, _ in parts]
    inst = [R.RInstance(k, np.eye(4, dtype=np.float32), [k]) for k in range(len(parts))]
    return R.build(meshes, inst, [m for _, m in parts], [])


def plane_y(size, y=0.0):
    s = size / 2
    return R.quad((-s, y, -s), (-s, y, s), (s, y, s), (s, y, -s))


def image(rs, cam, spp, w=64, h=36, bounces=4, **kw):
    p = rs.params(w, h, cam, max_bounces=bounces, rr_depth=8, **kw)
    a = oracle_scene(rs).render(p, 0, spp)
    return a[..., :3] / np.maximum(a[..., 3:4], 1)


def region(c, r):
    h, w = c.shape[:2]
    return c[int(r[1] * h):int(math.ceil(r[3] * h)), int(r[0] * w):int(math.ceil(r[2] * w))].reshape(-1, 3).mean(axis=0)


CENTER = (0.45, 0.45, 0.55, 0.55)


def panel_scene(floor_mat):
    """material-unlit-lights-nothing.ts: a 6 x 6 m floor and a 6 x 4 m unlit panel of colour 1 at (0, 2, -1) facing
    +z."""
    panel = R.quad((-3, 0, -1), (3, 0, -1), (3, 4, -1), (-3, 4, -1))
    return build([(plane_y(6), floor_mat), (panel, R.material((1.0, 1.0, 1.0), unlit=True))])


def test_an_unlit_plane_shows_its_colour():
    """material-unlit.ts: 0.5 +- 2 % with no light at all (a lit material would be black)."""
    rs = build([(plane_y(4), R.material((0.5, 0.5, 0.5), unlit=True))])
    c = image(rs, R.Camera((0, 4, 0.001), (0, 0, 0), 30.0), 4)
    m = region(c, CENTER)
    assert np.all(np.abs(m - 0.5) < 0.01), m
    assert R.material(unlit=True).view(np.int32)[18] == 4


def test_an_unlit_panel_lights_nothing_but_shows_in_a_mirror():
    """material-unlit-lights-nothing.ts: floor [0.4, 0.8, 0.6, 0.9] over panel [0.4, 0.2, 0.6, 0.3] is 0 +- 0.02; the
    same floor as a mirror (metal, colour 0.9, roughness 0.05) shows the panel at about its colour x 0.9."""
    cam = R.Camera((0, 1.2, 5), (0, 1.2, 0), 50.0)
    c = image(panel_scene(R.material((1.0, 1.0, 1.0), specular=0.0)), cam, 16)
    floor, panel = region(c, (0.4, 0.8, 0.6, 0.9)), region(c, (0.4, 0.2, 0.6, 0.3))
    assert np.all(panel > 0.99) and np.all(floor / panel <= 0.02), (floor, panel)
    c = image(panel_scene(R.material((0.9, 0.9, 0.9), metallic=1.0, roughness=0.05)), cam, 16)
    floor = region(c, (0.4, 0.8, 0.6, 0.9))
    assert np.all(np.abs(floor - 0.9) < 0.05), floor


def test_the_background_is_what_camera_rays_see_where_nothing_is():
    """environment-background.ts: 0.5 / pi +- 2 % where nothing is, no light."""
    rs = build([(plane_y(1, y=-100.0), R.material((0.5, 0.5, 0.5)))])
    bg = (0.5 / math.pi,) * 3
    c = image(rs, R.Camera((0, 0, 5), (0, 0, 0), 30.0), 2, background=bg)
    assert np.all(np.abs(region(c, CENTER) - bg[0]) < 0.02 * bg[0])


def test_the_background_lights_nothing():
    """environment-background-lights-nothing.ts: a white floor under a white background and a black sky, no
    lights: floor [0.4, 0.8, 0.6, 0.9] over background [0.4, 0.1, 0.6, 0.2] is 0 +- 0.02."""
    rs = build([(plane_y(6), R.material((1.0, 1.0, 1.0), specular=0.0))])
    c = image(rs, R.Camera((0, 1.2, 5), (0, 1.2, 0), 50.0), 16, background=(1.0, 1.0, 1.0))
    floor, bg = region(c, (0.4, 0.8, 0.6, 0.9)), region(c, (0.4, 0.1, 0.6, 0.2))
    assert np.all(np.abs(bg - 1.0) < 1e-6) and np.all(floor / bg <= 0.02), (floor, bg)


def test_a_path_bent_through_clear_glass_sees_the_background():
    """A clear 5 cm slab (roughness 0, ior 1.5) filling the view before the background, a black sky: a path that
    enters the slab leaves it, after any number of inner reflections, by a pass through the glass (pathtrace.wgsl
    4449-4465 `bent`: glass passed since the last scattering vertex) and sees the background; the 4 % the front face
    reflects leaves by reflection (its pdf above 0) and sees the black sky: 0.96 of the background over the image,
    +- 0.5 %; without the background, 0."""
    slab = R.box_tris((-3.0, -3.0, -0.025), (3.0, 3.0, 0.025))
    rs = build([(slab, R.material((1.0, 1.0, 1.0), roughness=0.0, transmission=1.0, ior=1.5))])
    cam = R.Camera((0, 0, 4), (0, 0, 0), 20.0)
    m = image(rs, cam, 16, bounces=16, background=(2.0, 2.0, 2.0)).reshape(-1, 3).mean(axis=0) / 2.0
    assert np.all(np.abs(m - 0.96) < 0.005 * 0.96), m
    assert np.all(image(rs, cam, 4, bounces=16) == 0.0)


def gradient_quad():
    """material-vertex-colors.ts gradientQuad: a 2 x 2 m quad facing +z, vertex colours from its x, 0 at -1, 1 at +1."""
    tris = R.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))
    cols = np.repeat(((tris[..., 0] + 1) / 2)[..., None], 3, axis=2).astype(np.float32)
    return tris, cols


def quad_image(world, flag=True, spp=16):
    tris, cols = gradient_quad()
    rs = R.build([R.RMesh(tris, colors=cols)], [R.RInstance(0, world, [0])],
                 [R.material((1.0, 1.0, 1.0), specular=0.0, vertex_colors=flag)], [])
    return image(rs, R.Camera((0, 0, 3), (0, 0, 0), 30.0), spp, sky=(1.0, 1.0, 1.0))


def test_vertex_colours_are_interpolated_in_linear():
    """material-vertex-colors.ts: the white Lambert quad in a 1-nit sky, colours 0 .. 1 across it: the centre reads
    0.5 +- 2 % (each corner's colour stored as sRGB 8-bit: 0 and 1 exactly); a material without the vertex-colour
    flag ignores them (1 +- 2 %); tri_colors' encoding of known values."""
    c = quad_image(np.eye(4, dtype=np.float32))
    assert np.all(np.abs(region(c, CENTER) - 0.5) < 0.01), region(c, CENTER)
    c = quad_image(np.eye(4, dtype=np.float32), flag=False)
    assert np.all(np.abs(region(c, CENTER) - 1.0) < 0.02), region(c, CENTER)
    w = R.tri_color_words(np.array([[[0.0, 1.0, 0.5], [0.0031, 0.2, 2.0], [-1.0, 0.25, 1.0]]], np.float32))
    # sRGB x 255 (IEC 61966-2-1): 0.5 -> 187.5 (188), 0.0031 -> 10.2 (10, the linear part), 0.2 -> 123.6 (124),
    # 0.25 -> 137.0 (137); below 0 and above 1 clamped
    assert [int(x) for x in w[0]] == [0xff000000 | 188 << 16 | 255 << 8, 0xff000000 | 255 << 16 | 124 << 8 | 10,
                                      0xff000000 | 255 << 16 | 137 << 8]


def test_vertex_colours_follow_a_mirrored_instance():
    """material-vertex-colors-mirrored.ts: the quad placed mirrored in x (an instance of scale -1): 0.75 at x = -0.5
    over 0.25 at x = +0.5: regions [0.3, 0.4, 0.35, 0.6] over [0.65, 0.4, 0.7, 0.6] = 3 +- 3 %."""
    world = np.diag([-1.0, 1.0, 1.0, 1.0]).astype(np.float32)
    c = quad_image(world)
    a, b = region(c, (0.3, 0.4, 0.35, 0.6)), region(c, (0.65, 0.4, 0.7, 0.6))
    assert np.all(np.abs(a / b - 3.0) < 0.09), (a, b)


def test_vertex_colours_tint_a_glass_shadow():
    """shadow-glass-tint.ts with a sun for its point light, the pane white with vertex colours of 0.5 (sRGB code 188:
    0.5029): the plane under it reads 0.5 E / pi x (0.96 x 0.5029)^2 +- 2 %."""
    plane = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    pane = R.box_tris((-0.5, 0.975, -0.5), (0.5, 1.025, 0.5))
    grey = np.full(pane.shape, 0.5, np.float32)
    rs = R.build([R.RMesh(plane), R.RMesh(pane, colors=grey)],
                 [R.RInstance(0, np.eye(4, dtype=np.float32), [0]), R.RInstance(1, np.eye(4, dtype=np.float32), [1])],
                 [R.material((0.5, 0.5, 0.5), specular=0.0),
                  R.material((1.0, 1.0, 1.0), roughness=0.0, transmission=1.0, ior=1.5, vertex_colors=True)],
                 [R.directional_light((0.0, -1.0, 0.0), 1000.0)])
    assert rs.vcolors.shape == (len(plane) + len(pane), 3) and np.all(rs.vcolors[:len(plane)] == 0xffffffff)
    cam = R.Camera((0.0, 0.6, 3.0), (0.0, 0.0, 0.0), 4.0)
    p = rs.params(16, 16, cam, max_bounces=1, rr_depth=8)
    a = oracle_scene(rs).render(p, 0, 16)
    m = (a[..., :3] / a[..., 3:4])[6:10, 6:10].reshape(-1, 3).mean(axis=0)
    want = 0.5 * 1000 / math.pi * (0.96 * 0.50289) ** 2
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


@pytest.mark.gpu
def test_vertex_colours_match_the_oracle(device):
    """Vertex-coloured meshes (a mirrored instance among them), a vertex-coloured glass pane and a sun, on the
    megakernel against the oracle (relative RMSE <= 1e-3), and the same bits under the wavefront setting."""
    import threert_amd as T
    from threert_amd import scene as S

    tris, cols = gradient_quad()
    t, n = R.icosphere(3)
    rng = np.random.default_rng(4)
    sph_cols = rng.random(t.shape).astype(np.float32)
    pane = R.box_tris((-0.6, 0.9, -0.4), (0.6, 0.95, 0.4))
    meshes = [R.RMesh(plane_y(8)), R.RMesh(tris, colors=cols), R.RMesh(t, normals=n, colors=sph_cols),
              R.RMesh(pane, colors=np.full(pane.shape, 0.6, np.float32))]
    mats = [R.material((0.7, 0.7, 0.7), roughness=0.8),
            R.material((1.0, 0.9, 0.8), specular=0.2, vertex_colors=True),
            R.material((1.0, 1.0, 1.0), metallic=0.5, roughness=0.4, vertex_colors=True),
            R.material((1.0, 1.0, 1.0), roughness=0.0, transmission=1.0, ior=1.5, vertex_colors=True)]
    inst = [R.RInstance(0, np.eye(4, dtype=np.float32), [0]),
            R.RInstance(1, S.trs((-1.0, 1.0, -1.0), (0.0, 0.3, 0.0), (-1.0, 1.0, 1.0)), [1]),
            R.RInstance(2, S.trs((0.8, 0.5, 0.0), (0.0, 0.0, 0.0), (0.5, 0.5, 0.5)), [2]),
            R.RInstance(3, np.eye(4, dtype=np.float32), [3])]
    rs = R.build(meshes, inst, mats, [R.directional_light((0.3, -1.0, -0.4), 3.0)])
    cam = R.Camera((0.5, 2.0, 5.0), (0.0, 0.6, 0.0), 50.0)
    p = rs.params(64, 48, cam, seed=13, max_bounces=6, rr_depth=3, sky=(0.3, 0.35, 0.5))
    print("vertex-colour scene 64x48 x 4 spp", flush=True)
    want = oracle_scene(rs).render(p, 0, 4)
    got, _ = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"vertex colours: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3
    try:
        device.set_render(T.RENDER_WAVEFRONT)
        wf, _ = gpu_render(device, rs, p, 0, 4)
    finally:
        device.set_render(T.RENDER_AUTO)
    assert np.array_equal(wf.view(np.uint32), got.view(np.uint32))


@pytest.mark.gpu
def test_unlit_and_background_match_the_oracle(device):
    """A lit scene with an unlit panel, a mirror floor, a glass slab, a sun and a background, on the megakernel
    against the oracle (relative RMSE <= 1e-3), and the same bits under the wavefront setting."""
    import threert_amd as T

    panel = R.quad((-2, 0, -1), (2, 0, -1), (2, 3, -1), (-2, 3, -1))
    slab = R.box_tris((0.5, 0.0, 1.0), (1.5, 1.5, 1.05))
    box = R.box_tris((-1.2, 0.0, 0.2), (-0.6, 0.6, 0.8))
    parts = [(plane_y(8), R.material((0.8, 0.8, 0.8), metallic=1.0, roughness=0.15)),
             (panel, R.material((0.9, 0.5, 0.2), unlit=True)),
             (slab, R.material((0.9, 1.0, 0.9), roughness=0.0, transmission=1.0, ior=1.5)),
             (box, R.material((0.3, 0.6, 0.9), roughness=0.5))]
    meshes = [R.RMesh(t) for t, _ in parts]
    inst = [R.RInstance(k, np.eye(4, dtype=np.float32), [k]) for k in range(len(parts))]
    rs = R.build(meshes, inst, [m for _, m in parts], [R.directional_light((0.3, -1.0, -0.4), 3.0)])
    cam = R.Camera((0.5, 1.5, 5.0), (0.0, 0.8, 0.0), 50.0)
    p = rs.params(64, 48, cam, seed=9, max_bounces=6, rr_depth=3, sky=(0.2, 0.3, 0.5), background=(0.9, 0.8, 0.7))
    print("unlit + background scene 64x48 x 4 spp", flush=True)
    want = oracle_scene(rs).render(p, 0, 4)
    got, _ = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"unlit + background: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3
    try:
        device.set_render(T.RENDER_WAVEFRONT)
        wf, _ = gpu_render(device, rs, p, 0, 4)
    finally:
        device.set_render(T.RENDER_AUTO)
    assert np.array_equal(wf.view(np.uint32), got.view(np.uint32))
