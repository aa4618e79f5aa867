"""Local lights: point, spot and rect lights (pathtrace.wgsl:2039-2262 local_light / local_lights; scene.rs:1309-1367
the GpuLight records), in the oracle, pinned by the reference's analytic render tests, and on the GPU against it.

- light-point-inverse-square.ts: 1000 cd 2 m over an albedo-0.5 Lambert plane reads 0.5 x 1000 / 2^2 / pi below it;
  an off-axis coloured light checks the cosine and the channels (0.5 / pi x I cos / d^2 over the patch).
- light-spot-cone.ts: nothing outside a hard 20-degree cone; a penumbra point reads glTF KHR_lights_punctual's
  angular attenuation (Lagarde & de Rousiers 2014, Listing 4: the linear ramp in the cosine, squared).
- light-rect-irradiance.ts: a 0.2 m square of 10000 nits 2 m up gives the parallel-rectangle view factor's
  irradiance; the rect is one-sided.
- light-many-shadowed.ts: 24 shadow-casting lights (resampled: more than 8), 10 behind a black slab: the other 14's
  sum at the centre.
- shadow-glass-tint.ts with its point light: (0.96 x 0.5)^2 of the clear value through a grey pane.
- The range window: Karis, "Real Shading in Unreal Engine 4" (SIGGRAPH 2013), eq. 9's saturate(1 - (d/r)^4)^2.
- Lights that cast no shadow are not blocked; few shadowed lights are summed exactly, many resampled: both give the
  closed-form sum.
- The GPU (megakernel) against the oracle, same seed and samples, with few (summed) and many (resampled) lights,
  glass and a sun; the wavefront setting hands such batches to the megakernel; the boundary refuses a directional
  light after a local one and a first-shadowed index outside the local lights.
"""
import math

import numpy as np
import pytest

from threert_amd import render as R
from threert_amd import scene as S
from test_glass import gpu_render, oracle_scene, rel_rmse

INV_PI = 1.0 / math.pi


def plane_scene(lights, extra=(), albedo=0.5):
    """A 20 m albedo-`albedo` Lambert plane at y = 0 (lib.ts plane + lambert) with `extra` (tris, material) pairs."""
    plane = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    meshes = [R.RMesh(plane)] + [R.RMesh(t) for t, _ in extra]
    inst = [R.RInstance(k, np.eye(4, dtype=np.float32), [k]) for k in range(len(meshes))]
    mats = [R.material((albedo, albedo, albedo), specular=0.0)] + [m for _, m in extra]
    return R.build(meshes, inst, mats, lights)


def patch_camera(position, target, half=0.02, fov=30.0):
    """lib.ts camera (fov 30, 16:9) narrowed to the region [0.5 - half, 0.5 + half]^2 around its centre: the whole
    image of the returned camera (render it at 16:9) is that region."""
    return R.Camera(position, target, math.degrees(2.0 * math.atan(2.0 * half * math.tan(math.radians(fov) * 0.5))))


def mean_rgb(acc):
    return (acc[..., :3] / np.maximum(acc[..., 3:4], 1)).reshape(-1, 3).mean(axis=0)


def render_patch(rs, cam, spp, bounces=1, w=32, h=18, seed=1):
    p = rs.params(w, h, cam, seed=seed, max_bounces=bounces, rr_depth=8)
    return mean_rgb(oracle_scene(rs).render(p, 0, spp))


def patch_points(position, target, half=0.02, fov=30.0, n=24):
    """Where the rays of the patch camera's image meet the plane y = 0 (an n x n grid over the image, aspect 16:9)."""
    eye, tgt = np.asarray(position, float), np.asarray(target, float)
    fwd = (tgt - eye) / np.linalg.norm(tgt - eye)
    right = np.cross(fwd, (0.0, 1.0, 0.0))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    t = 2.0 * half * math.tan(math.radians(fov) * 0.5)
    pts = []
    for sy in (np.arange(n) + 0.5) / n * 2 - 1:
        for sx in (np.arange(n) + 0.5) / n * 2 - 1:
            d = fwd + right * sx * t * 16 / 9 + up * sy * t
            pts.append(eye + d * (-eye[1] / d[1]))
    return np.array(pts)


# ---- the reference's analytic render tests, on the oracle --------------------------------------------------------
def test_point_light_inverse_square():
    """light-point-inverse-square.ts: 0.5 x 1000 / 4 / pi +- 2 % in the 4 % patch below the light."""
    rs = plane_scene([R.point_light((0, 2, 0), 1000.0, radius=0.001)])
    m = render_patch(rs, patch_camera((0, 1.5, 3), (0, 0, 0)), 16)
    want = 0.5 * 1000 / 4 / math.pi
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


@pytest.mark.parametrize("range_", [0.0, 3.0])
def test_off_axis_point_light_cosine_channels_and_range(range_):
    """An off-axis light of (1000, 600, 300) cd: 0.5 / pi x I cos / d^2 averaged over where the patch's rays land
    (the cosine is 0.78 there, so a dropped or doubled cosine is 20 % off), times Karis's window
    saturate(1 - (d / r)^4)^2 with a range of 3 m (d = 2.55 m: the window is 0.23); +- 1 %."""
    pos, cd = np.array([1.5, 2.0, 0.5]), np.array([1000.0, 600.0, 300.0])
    rs = plane_scene([R.point_light(tuple(pos), tuple(cd), range_=range_)])
    cam_at = ((0, 1.5, 3), (0, 0, 0))
    m = render_patch(rs, patch_camera(*cam_at), 16)
    x = patch_points(*cam_at)
    dl = pos - x
    d2 = (dl * dl).sum(axis=1)
    k = 0.5 * INV_PI * dl[:, 1] / np.sqrt(d2) / d2
    if range_ > 0:
        k *= np.clip(1 - (d2 / range_ ** 2) ** 2, 0, 1) ** 2
    want = cd * k.mean()
    assert np.all(np.abs(m - want) < 0.01 * want), (m, want)


def test_spot_cone_is_dark_outside_and_ramps_in_the_penumbra():
    """light-spot-cone.ts: a hard 20-degree spot 3 m up, looked at from 4 m straight above (fov 60): the region
    [0.05, 0.6, 0.15, 0.7] (outside the cone) over the centre [0.47, 0.53]^2 is 0 +- 0.005. A penumbra point (a 30
    degree cone, penumbra 0.5, its axis 20 degrees off the point) reads ((cos 20 - cos 30) / (cos 15 - cos 30))^2 of the
    on-axis value, +- 2 %."""
    rs = plane_scene([R.spot_light((0, 3, 0), (0, -1, 0), 2000.0, math.radians(20), 0.0, radius=0.001)])
    w, h = 64, 36
    p = rs.params(w, h, R.Camera((0, 4, 0.001), (0, 0, 0), 60.0), max_bounces=1, rr_depth=8)
    c = oracle_scene(rs).render(p, 0, 4)
    c = c[..., :3] / np.maximum(c[..., 3:4], 1)
    reg = lambda r: c[int(r[1] * h):int(math.ceil(r[3] * h)), int(r[0] * w):int(math.ceil(r[2] * w))].mean()  # noqa: E731
    centre, side = reg((0.47, 0.47, 0.53, 0.53)), reg((0.05, 0.6, 0.15, 0.7))
    assert centre > 1.0 and side / centre <= 0.005, (side, centre)

    tilt = math.radians(20.0)
    axis = (math.sin(tilt), -math.cos(tilt), 0.0)
    cam_at = ((0, 1.5, 3), (0, 0, 0))
    on = render_patch(plane_scene([R.point_light((0, 2, 0), 1000.0)]), patch_camera(*cam_at), 4)
    rs = plane_scene([R.spot_light((0, 2, 0), axis, 1000.0, math.radians(30), 0.5)])
    m = render_patch(rs, patch_camera(*cam_at), 4)
    x = patch_points(*cam_at)
    cos_a = ((np.array([0, 2, 0]) - x) / np.linalg.norm(np.array([0, 2, 0]) - x, axis=1)[:, None]) @ -np.array(axis)
    ramp = np.clip((cos_a - math.cos(math.radians(30))) / (math.cos(math.radians(15)) - math.cos(math.radians(30))),
                   0, 1) ** 2
    assert 0.4 < ramp.mean() < 0.7
    want = on * ramp.mean()
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


def rect_irradiance(half_x: float, half_y: float, h: float, lum: float) -> float:
    """The irradiance under the centre of a parallel uniform rectangle (half sides over height h): four corner
    rectangles' view factor (Howell's catalogue, C-11), times pi L."""
    X, Y = half_x / h, half_y / h
    f = (X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X)) +
         Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))) / (2 * math.pi)
    return 4 * math.pi * lum * f


def test_rect_light_irradiance_and_one_side():
    """light-rect-irradiance.ts: a 0.2 m square of 10000 nits facing down 2 m up: 0.5 E / pi +- 2 % with E from the
    parallel-rectangle view factor (the test's own closed form, and Howell's for a 1 x 0.5 m rect at 1 %); turned to
    face up it lights nothing below."""
    cam = patch_camera((0, 1.5, 3), (0, 0, 0))
    s, h, lum = 0.2, 2.0, 10000.0
    X = s / 2 / h
    E = 4 * lum * (X / math.sqrt(1 + X * X)) * math.atan(X / math.sqrt(1 + X * X))
    assert abs(E - rect_irradiance(s / 2, s / 2, h, lum)) < 1e-9 * E
    m = render_patch(plane_scene([R.rect_light((0, h, 0), (0, -1, 0), (1, 0, 0), s, s, lum)]), cam, 64)
    want = 0.5 * E / math.pi
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)
    m = render_patch(plane_scene([R.rect_light((0, 1.0, 0), (0, -1, 0), (1, 0, 0), 1.0, 0.5, 100.0)]), cam, 256)
    want = 0.5 / math.pi * rect_irradiance(0.5, 0.25, 1.0, 100.0)
    assert np.all(np.abs(m - want) < 0.01 * want), (m, want)
    m = render_patch(plane_scene([R.rect_light((0, h, 0), (0, 1, 0), (1, 0, 0), s, s, lum)]), cam, 4)
    assert np.all(m == 0.0), m


def many_shadowed():
    """light-many-shadowed.ts: 24 point lights of 300 cd at y = 2, 2 m from the axis (14 at 100..260 degrees, 10 at
    -54..54), a black 6 x 0.05 x 12 m slab centred at (3.3, 1, 0) hiding the 10."""
    angles = [100 + k * 160 / 13 for k in range(14)] + [-54 + k * 12 for k in range(10)]
    lights = [R.point_light((2 * math.cos(math.radians(a)), 2.0, 2 * math.sin(math.radians(a))), 300.0, radius=0.001)
              for a in angles]
    slab = (R.box_tris((0.3, 0.975, -6.0), (6.3, 1.025, 6.0)), R.material((0.0, 0.0, 0.0), specular=0.0))
    return plane_scene(lights, [slab])


def test_many_shadowed_lights_are_resampled_to_the_visible_sum():
    """light-many-shadowed.ts: 0.5 / pi x 14 x 300 x cos / d^2 (d^2 = 8, cos = 2 / sqrt 8) +- 2 %, 1024 samples: the
    24 are more than are summed exactly (8), so one of 8 candidates is resampled by unshadowed contribution."""
    rs = many_shadowed()
    first_shadowed = rs.params(4, 4, R.Camera((0, 1, 3), (0, 0, 0)))[39]
    assert first_shadowed == 0 and rs.n_directional == 0
    m = render_patch(rs, patch_camera((0, 1.5, 3), (0, 0, 0)), 1024)
    want = 0.5 / math.pi * 14 * 300 * (2 / math.sqrt(8)) / 8
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


def test_shadowless_lights_pass_blockers_and_few_lights_are_summed():
    """Two lights over an opaque 1 m slab at y = 1 above the patch: the shadow-casting one is blocked, the other
    (cast_shadow off) is not, whatever the order given (light_order puts unshadowed lights first); with 6 lights,
    all summed exactly, the patch reads the closed-form sum of the seen ones."""
    slab = (R.box_tris((-0.5, 0.975, -0.5), (0.5, 1.025, 0.5)), R.material((0.5, 0.5, 0.5), specular=0.0))
    one = 0.5 * 1000 / 4 / math.pi
    cam = patch_camera((0, 0.6, 3), (0, 0, 0))
    lit = R.point_light((0, 2, 0), 1000.0, cast_shadow=False)
    rs = plane_scene([R.point_light((0, 2, 0), 5000.0), lit], [slab])
    assert rs.lights[0, 15] == 0.0 and rs.params(4, 4, cam)[39] == 1
    m = render_patch(rs, cam, 16)
    assert np.all(np.abs(m - one) < 0.02 * one), m
    pos = [(0, 2, 0), (0.6, 1.8, 0.2), (-0.7, 2.2, -0.3), (0.3, 2.5, 0.9), (-0.4, 1.6, 0.6), (1.0, 2.0, -1.0)]
    cam_at = ((0, 1.5, 3), (0, 0, 0))
    rs = plane_scene([R.point_light(q, 200.0 + 50 * k, cast_shadow=k % 2 == 0) for k, q in enumerate(pos)])
    m = render_patch(rs, patch_camera(*cam_at), 16)
    x = patch_points(*cam_at)
    want = 0.0
    for k, q in enumerate(pos):
        dl = np.asarray(q) - x
        d2 = (dl * dl).sum(axis=1)
        want += (0.5 * INV_PI * (200.0 + 50 * k) * dl[:, 1] / np.sqrt(d2) / d2).mean()
    assert np.all(np.abs(m - want) < 0.01 * want), (m, want)


def test_a_soft_lights_half_shadow_under_a_blockers_edge():
    """A sphere light of radius 0.3 m 2 m over the patch, an opaque slab at y = 1 over x >= 0: the shadow ray aims at
    a uniform point of the sphere, so under the slab's edge exactly the points with x > 0 are hidden: half the
    unshadowed value (the patch's spread in x averages out), +- 2 %; a hard light (radius 1 mm) lights the patch
    fully 0.3 m to the -x side and not at all 0.3 m to the +x side."""
    slab = (R.box_tris((0.0, 0.975, -6.0), (6.0, 1.025, 6.0)), R.material((0.5, 0.5, 0.5), specular=0.0))
    one = 0.5 * 1000 / 4 / math.pi
    m = render_patch(plane_scene([R.point_light((0, 2, 0), 1000.0, radius=0.3)], [slab]),
                     patch_camera((0, 0.6, 3), (0, 0, 0)), 256)
    assert np.all(np.abs(m - 0.5 * one) < 0.02 * 0.5 * one), (m, 0.5 * one)
    for x, lit in ((-0.3, True), (0.3, False)):
        rs = plane_scene([R.point_light((x, 2, 0), 1000.0, radius=0.001)], [slab])
        m = render_patch(rs, patch_camera((x, 0.6, 3), (x, 0, 0)), 4)
        assert (np.all(np.abs(m - one) < 0.03 * one) if lit else np.all(m == 0.0)), (x, m)


def test_a_pane_tints_a_point_lights_shadow():
    """shadow-glass-tint.ts as written, with its point light: 0.5 I / 4 / pi x (0.96 x 0.5)^2 +- 2 %."""
    pane = (R.box_tris((-0.5, 0.975, -0.5), (0.5, 1.025, 0.5)),
            R.material((0.5, 0.5, 0.5), metallic=0.0, roughness=0.0, transmission=1.0, ior=1.5))
    rs = plane_scene([R.point_light((0, 2, 0), 1000.0, radius=0.001)], [pane])
    m = render_patch(rs, patch_camera((0, 0.6, 3), (0, 0, 0)), 64)
    want = 0.5 * 1000 / 4 / math.pi * (0.96 * 0.5) ** 2
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


def test_light_order_and_params():
    """light_order (lib.rs:367-380): directional, then unshadowed, then shadowed local lights; Params' counts2.x is
    the directional count and counts2.w the first shadowed light (the light count when none casts a shadow)."""
    given = [R.point_light((0, 2, 0), 1.0), R.directional_light((0, -1, 0), 1.0),
             R.spot_light((0, 2, 0), (0, -1, 0), 1.0, 0.5, cast_shadow=False),
             R.rect_light((0, 2, 0), (0, -1, 0), (1, 0, 0), 1, 1, 1.0)]
    assert R.light_order(given) == [1, 2, 0, 3]
    rs = plane_scene(given)
    p = rs.params(4, 4, R.Camera((0, 1, 3), (0, 0, 0)))
    assert list(rs.lights[:, 3]) == [0.0, 2.0, 1.0, 3.0] and rs.n_directional == 1
    assert (p[20], p[36], p[39]) == (4, 1, 2)
    rs = plane_scene([R.point_light((0, 2, 0), 1.0, cast_shadow=False)])
    p = rs.params(4, 4, R.Camera((0, 1, 3), (0, 0, 0)))
    assert (p[20], p[36], p[39]) == (1, 0, 1)
    co, inv = R.spot_cone(math.radians(30), 0.5)
    assert abs(co - math.cos(math.radians(30))) < 1e-6
    assert abs(inv - 1 / (math.cos(math.radians(15)) - co)) < 1e-3 * inv
    assert R.spot_cone(0.3, 0.0)[1] == pytest.approx(1000.0)


# ---- the GPU against the oracle ---------------------------------------------------------------------------------
def lamp_garden(n_shadowed: int):
    """A ground, an opaque box, a rough glass sphere and a grey pane, a sun, three lights that cast no shadow (a point
    light with a range, a spot with a penumbra, a rect with barn doors) and `n_shadowed` shadow-casting ones (soft
    points, spots and rects in turn): with 12, more than are summed exactly, so they are resampled."""
    t, n = R.icosphere(3)
    ground = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    box = R.box_tris((-0.4, 0.0, -0.4), (0.4, 0.9, 0.4))
    pane = R.box_tris((-0.6, -0.02, -0.4), (0.6, 0.02, 0.4))
    meshes = [R.RMesh(ground), R.RMesh(box), R.RMesh(t, normals=n), R.RMesh(pane)]
    mats = [R.material((0.6, 0.6, 0.6), specular=0.5, roughness=0.7),
            R.material((0.7, 0.3, 0.2), specular=0.5, roughness=0.4),
            R.material((0.9, 0.95, 1.0), roughness=0.3, transmission=1.0, ior=1.5),
            R.material((0.6, 0.8, 0.7), roughness=0.0, transmission=1.0, ior=1.45)]
    inst = [R.RInstance(0, np.eye(4, dtype=np.float32), [0]),
            R.RInstance(1, S.trs((1.2, 0.0, 0.6), (0.0, 0.4, 0.0), (1, 1, 1)), [1]),
            R.RInstance(2, S.trs((-0.6, 0.5, 0.1), (0.0, 0.0, 0.0), (0.5, 0.5, 0.5)), [2]),
            R.RInstance(3, S.trs((0.3, 1.4, -0.3), (0.4, 0.2, 0.1), (1, 1, 1)), [3])]
    lights = [R.directional_light((-0.3, -1.0, 0.4), 2.0),
              R.point_light((-1.5, 1.2, 1.0), (40.0, 30.0, 20.0), range_=4.0, cast_shadow=False),
              R.spot_light((1.0, 2.5, -1.0), (-0.3, -1.0, 0.3), 120.0, math.radians(35), 0.4, cast_shadow=False),
              R.rect_light((0.0, 2.4, 0.0), (0.0, -1.0, 0.1), (1, 0, 0), 0.8, 0.4, 60.0, barn_door_deg=60.0,
                           cast_shadow=False)]
    rng = np.random.default_rng(5)
    for k in range(n_shadowed):
        a = 2 * math.pi * k / max(n_shadowed, 1)
        q = (1.8 * math.cos(a), 1.5 + 0.8 * rng.random(), 1.8 * math.sin(a))
        col = tuple(20.0 + 30.0 * rng.random(3))
        if k % 3 == 0:
            lights.append(R.point_light(q, col, radius=0.05 + 0.1 * rng.random()))
        elif k % 3 == 1:
            lights.append(R.spot_light(q, (-q[0], -q[1], -q[2]), col, math.radians(40), 0.3, radius=0.03))
        else:
            lights.append(R.rect_light(q, (-q[0], -q[1], -q[2]), (0, 1, 0), 0.3, 0.2, col[0], range_=5.0))
    rs = R.build(meshes, inst, mats, lights)
    return rs, R.Camera((0.3, 1.8, -4.0), (0.0, 0.6, 0.0), 45.0)


@pytest.mark.gpu
@pytest.mark.parametrize("n_shadowed", [3, 12])
def test_lamp_gardens_match_the_oracle(device, n_shadowed):
    """The lamp garden on the megakernel against the oracle, same seed and samples, 6 bounces: relative RMSE
    <= 1e-3 (the kernel does the oracle's float operations in its order)."""
    print(f"lamp garden 64x48 x 4 spp, {n_shadowed} shadowed lights", flush=True)
    rs, cam = lamp_garden(n_shadowed)
    p = rs.params(64, 48, cam, seed=11, max_bounces=6, rr_depth=3, sky=(0.2, 0.25, 0.35))
    want = oracle_scene(rs).render(p, 0, 4)
    got, _ = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"lamp garden: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_local_lights_render_on_the_megakernel_and_are_checked(device):
    """Under RENDER_WAVEFRONT a batch with local lights takes the megakernel (the same bits as RENDER_MEGAKERNEL and
    as the counting launch); a directional light after a local one is refused at upload, and Params whose counts2.w
    is not among the local lights at render, each with TRT_E_INVALID."""
    import threert_amd as T

    rs, cam = lamp_garden(12)
    p = rs.params(32, 24, cam, seed=3, max_bounces=4, rr_depth=3)
    outs = []
    try:
        for how in (T.RENDER_WAVEFRONT, T.RENDER_MEGAKERNEL):
            print(f"lamp garden under render setting {how}", flush=True)
            device.set_render(how)
            outs.append(gpu_render(device, rs, p, 0, 2)[0])
    finally:
        device.set_render(T.RENDER_AUTO)
    print("lamp garden: counting launch", flush=True)
    counted, reads = gpu_render(device, rs, p, 0, 2, reads=True)
    assert reads["shaded"] > 0, reads
    for o in outs[1:] + [counted]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))

    import copy
    bad = copy.deepcopy(rs)
    bad.lights = bad.lights[[1, 0] + list(range(2, len(bad.lights)))]
    print("refusal: a directional light after a local one", flush=True)
    with pytest.raises(T.ThreeRTError) as e:
        R.upload(device, bad)
    assert e.value.code == T.TRT_E_INVALID and "directional light after a local one" in str(e.value), str(e.value)
    sc = R.upload(device, rs)
    acc = R.Accum(device, 32, 24)
    for w39 in (0, len(rs.lights) + 1):
        print(f"refusal: counts2.w = {w39}", flush=True)
        q = p.copy()
        q[39] = w39
        with pytest.raises(T.ThreeRTError) as e:
            R.render(sc, q, 0, 1, acc)
        assert e.value.code == T.TRT_E_INVALID and "counts2.w" in str(e.value), str(e.value)
    acc.close()
    sc.close()
