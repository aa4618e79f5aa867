"""The environment map: an equirectangular map lighting the scene and sampled by its CDF (pathtrace.wgsl:1889-1941
env_uv / env_radiance / env_pdf / lower_bound / sample_env, lib.rs:916-942 build_env_cdf, gpu.rs:3537-3561 its
half-float texels), in the oracle, pinned by what the mathematics fixes, and on the GPU against it.

- The kernel's atan2 and acos (Cephes atanf / asinf) within 3 and 1.5 ulp of float64.
- build_env_cdf of a uniform map: the rows' CDF is the running share of sin(theta), each row's (x + 1) / w.
- env_radiance against the sampler's definition (u = atan2(x, -z) / 2 pi + 0.5, v = acos(y) / pi; repeat, linear,
  texel centres at (i + 0.5) / n) on random directions; sample_env's draws fall in the texel the CDF picks (a map
  with one lit texel); env_pdf integrates to 1 over the sphere, and E[1 / pdf] over its own draws is 4 pi.
- sky-uniform-irradiance.ts with its 2-nit sky as a map: an albedo-0.5 plane reads 1 +- 2 %.
- A map with a small bright patch over a Lambert plane: the plane reads albedo / pi x the map's irradiance
  (integrated over the hemisphere by quadrature) with the map sampled (MIS) and without (BSDF rays alone); on a
  glossy metal the two estimators agree.
- The GPU (megakernel) against the oracle, same seed and samples, on an open scene with glass, a sun and textures
  under a map; the wavefront setting hands such batches to the megakernel; the boundary refuses Params asking for a
  map the scene lacks or of another size, and a CDF of the wrong length.
"""
import math

import numpy as np
import pytest

import oracle
from threert_amd import render as R
from threert_amd import scene as S
from test_glass import gpu_render, oracle_scene, rel_rmse
from test_local_lights import patch_camera


def env_scene(env, extra=(), albedo=0.5, mat=None, lights=()):
    """A 2 km Lambert plane (albedo `albedo`, or `mat`) at y = 0 under the map `env`, with `extra` (tris, material)."""
    plane = R.quad((-1000, 0, -1000), (-1000, 0, 1000), (1000, 0, 1000), (1000, 0, -1000))
    meshes = [R.RMesh(plane)] + [R.RMesh(t) for t, _ in extra]
    inst = [R.RInstance(k, np.eye(4, dtype=np.float32), [k]) for k in range(len(meshes))]
    mats = [mat if mat is not None else R.material((albedo, albedo, albedo), specular=0.0)] + [m for _, m in extra]
    return R.build(meshes, inst, mats, list(lights), env=env)


def env_lookup(env, d):
    """The sampler's definition in float64: d to (u, v) as env_uv, then bilinear with repeat in u and v."""
    h, w = env.shape[:2]
    u = np.arctan2(d[..., 0], -d[..., 2]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(d[..., 1], -1, 1)) / np.pi
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = (x - x0)[..., None], (y - y0)[..., None]
    xa, ya = x0.astype(int) % w, y0.astype(int) % h
    xb, yb = (xa + 1) % w, (ya + 1) % h
    e = env.astype(np.float64)
    return (e[ya, xa] * (1 - fx) + e[ya, xb] * fx) * (1 - fy) + (e[yb, xa] * (1 - fx) + e[yb, xb] * fx) * fy


def irradiance(env, n=1024):
    """E = the integral of L cos over the upper hemisphere, by the midpoint rule in (theta, phi)."""
    th = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(4 * n) + 0.5) / (4 * n) * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    wgt = (np.cos(T) * np.sin(T))[..., None] * (np.pi / 2 / n) * (2 * np.pi / (4 * n))
    return (env_lookup(env, d) * wgt).sum(axis=(0, 1))


def mean_patch(rs, cam, spp, env_nee=True, bounces=1):
    p = rs.params(32, 18, cam, max_bounces=bounces, rr_depth=8, env_nee=env_nee)
    a = oracle_scene(rs).render(p, 0, spp)
    return (a[..., :3] / np.maximum(a[..., 3:4], 1)).reshape(-1, 3).mean(axis=0)


def sun_patch_map(h=64, w=128):
    env = np.full((h, w, 3), 0.2, np.float32)
    env[10:13, 40:44] = (500.0, 400.0, 300.0)
    return env


# ---- the functions -----------------------------------------------------------------------------------------------
def test_atan2_and_acos_are_within_a_few_ulp():
    rng = np.random.default_rng(1)
    worst = 0.0
    for y, x in (rng.standard_normal((20000, 2)) * rng.choice([1e-3, 1.0, 1e3], (20000, 1))).astype(np.float32):
        r = math.atan2(float(y), float(x))
        worst = max(worst, abs(oracle.atan2(float(y), float(x)) - r) / float(np.spacing(np.float32(abs(r)))))
    assert worst <= 3.0, worst
    assert oracle.atan2(0.0, -1.0) == pytest.approx(math.pi) and oracle.atan2(1.0, 0.0) == pytest.approx(math.pi / 2)
    worst = 0.0
    for x in np.concatenate([np.linspace(-1, 1, 40001, dtype=np.float32), np.float32([1, -1, 0, 0.5, -0.5, 1e-5])]):
        r = math.acos(float(x))
        ulp = float(np.spacing(np.float32(abs(r)))) if r else float(np.spacing(np.float32(0)))
        worst = max(worst, abs(oracle.acos(float(x)) - r) / ulp)
    assert worst <= 1.5, worst
    assert math.isnan(oracle.acos(1.5))


def test_the_cdf_of_a_uniform_map():
    """build_env_cdf of a constant map: rows by the running share of sin(theta at the row's centre), each row
    uniform ((x + 1) / w), the last entries 1."""
    h, w = 16, 24
    cdf = R.build_env_cdf(np.full((h, w, 4), 3.0, np.float32))
    assert cdf.shape == (h + h * w,) and cdf.dtype == np.float32
    s = np.sin((np.arange(h) + 0.5) / h * np.pi)
    assert np.allclose(cdf[:h], np.cumsum(s) / s.sum(), atol=2e-6)
    assert np.allclose(cdf[h:].reshape(h, w), (np.arange(w) + 1) / w, atol=2e-6)
    assert np.all(np.diff(cdf[:h]) > 0) and abs(cdf[h - 1] - 1) < 1e-6


def test_radiance_sampling_and_pdf():
    """env_radiance equals the sampler's definition within float rounding; with one lit texel every draw of
    sample_env falls in it, where u and v are the draw's fractions of the texel; env_pdf integrates to 1 over the
    sphere and E[1 / pdf] over its own draws is 4 pi (the rows' sin(theta) is taken at their centres: within 0.5 %)."""
    rng = np.random.default_rng(3)
    env = (rng.random((24, 48, 3)) * 4).astype(np.float32)
    rs = env_scene(env)
    o = oracle_scene(rs)
    p = rs.params(4, 4, R.Camera((0, 1, 3), (0, 0, 0)))
    d = rng.standard_normal((2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    want = env_lookup(rs.env_half().view(np.float16).astype(np.float32)[..., :3], d)
    got = np.array([o.env(p, di)[0] for di in d.astype(np.float32)])
    assert np.allclose(got, want, rtol=1e-4, atol=1e-5), np.abs(got - want).max()

    one = np.zeros((16, 32, 3), np.float32)
    one[5, 17] = 10.0
    rs1 = env_scene(one)
    o1 = oracle_scene(rs1)
    p1 = rs1.params(4, 4, R.Camera((0, 1, 3), (0, 0, 0)))
    for u in rng.random((500, 4)).astype(np.float32):
        s = o1.env(p1, (0, 1, 0), u)[2]
        uu = math.atan2(s[0], -s[2]) / (2 * math.pi) + 0.5
        vv = math.acos(max(-1.0, min(1.0, float(s[1])))) / math.pi
        assert abs(uu * 32 - (17 + u[2])) < 1e-3 and abs(vv * 16 - (5 + u[3])) < 1e-3, (u, uu * 32, vv * 16)

    n = 256
    th = (np.arange(n) + 0.5) / n * np.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    tot = 0.0
    for t in th:
        dirs = np.stack([np.sin(t) * np.cos(ph), np.full_like(ph, np.cos(t)), np.sin(t) * np.sin(ph)], -1)
        tot += sum(o.env(p, di)[1] for di in dirs.astype(np.float32)) * math.sin(t) * (np.pi / n) * (np.pi / n)
    assert abs(tot - 1.0) < 5e-3, tot
    inv = [1.0 / o.env(p, o.env(p, (0, 1, 0), u)[2])[1] for u in rng.random((20000, 4)).astype(np.float32)]
    assert abs(np.mean(inv) / (4 * math.pi) - 1.0) < 5e-3, np.mean(inv) / (4 * math.pi)


# ---- renders on the oracle ---------------------------------------------------------------------------------------
def test_a_uniform_map_is_the_uniform_sky():
    """sky-uniform-irradiance.ts with its 2-nit sky as a map: an albedo-0.5 Lambert plane reads 1 +- 2 %, the map
    sampled or not."""
    rs = env_scene(np.full((32, 64, 3), 2.0, np.float32))
    cam = patch_camera((0, 2, 2), (0, 0, 0), half=0.05)
    for nee in (True, False):
        m = mean_patch(rs, cam, 64, env_nee=nee)
        assert np.all(np.abs(m - 1.0) < 0.02), (nee, m)


def test_a_bright_patch_lights_a_plane_as_its_irradiance():
    """A 0.2-nit map with a 3 x 4 texel patch of (500, 400, 300) nits 30 degrees from the zenith, over an albedo-0.5
    Lambert plane: 0.5 / pi x E, E from quadrature over the hemisphere; the map sampled with MIS +- 1 %, BSDF rays
    alone +- 3 % (1024 samples)."""
    env = sun_patch_map()
    rs = env_scene(env)
    want = 0.5 / math.pi * irradiance(rs.env_half().view(np.float16).astype(np.float32)[..., :3])
    cam = patch_camera((0, 2, 2), (0, 0, 0), half=0.05)
    m = mean_patch(rs, cam, 1024, env_nee=True)
    assert np.all(np.abs(m - want) < 0.01 * want), (m, want)
    m = mean_patch(rs, cam, 1024, env_nee=False)
    assert np.all(np.abs(m - want) < 0.03 * want), (m, want)


def test_sampling_the_map_agrees_with_bsdf_rays_on_a_glossy_metal():
    """A rough metal (roughness 0.4) under the patch map: the map sampled (MIS with the BSDF) and BSDF rays alone
    estimate the same radiance, within 3 % at 4096 samples."""
    rs = env_scene(sun_patch_map(), mat=R.material((0.8, 0.8, 0.8), metallic=1.0, roughness=0.4))
    cam = patch_camera((0, 2, 2), (0, 0, 0), half=0.05)
    a = mean_patch(rs, cam, 4096, env_nee=True)
    b = mean_patch(rs, cam, 4096, env_nee=False)
    assert np.all(np.abs(a - b) < 0.03 * a), (a, b)


# ---- the GPU against the oracle ---------------------------------------------------------------------------------
def env_garden():
    """An open scene under a map with a bright patch and a coloured horizon: a ground, an opaque box, a rough glass
    sphere, a metal sphere, a textured pane (a checker), a sun."""
    t, n = R.icosphere(3)
    ground = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    box = R.box_tris((-0.4, 0.0, -0.4), (0.4, 0.9, 0.4))
    pane = R.quad((-0.6, 0.0, 0.0), (0.6, 0.0, 0.0), (0.6, 1.0, 0.0), (-0.6, 1.0, 0.0))
    checker = np.kron((np.indices((4, 4)).sum(axis=0) % 2), np.ones((8, 8)))[..., None] * \
        np.array([200, 180, 40], np.float64) + 30
    tex = R.Texture(np.concatenate([checker, np.full(checker.shape[:2] + (1,), 255)], axis=2).astype(np.uint8))
    uv = np.array([[[0, 0], [4, 0], [4, 4]], [[0, 0], [4, 4], [0, 4]]], np.float32)
    meshes = [R.RMesh(ground), R.RMesh(box), R.RMesh(t, normals=n), R.RMesh(pane, uvs=uv)]
    mats = [R.material((0.6, 0.6, 0.6), specular=0.5, roughness=0.8),
            R.material((0.7, 0.3, 0.2), specular=0.5, roughness=0.4),
            R.material((0.9, 0.95, 1.0), roughness=0.3, transmission=1.0, ior=1.5),
            R.material((0.9, 0.8, 0.6), metallic=1.0, roughness=0.25),
            R.material((1.0, 1.0, 1.0), specular=0.3, roughness=0.6, base_map=0)]
    inst = [R.RInstance(0, np.eye(4, dtype=np.float32), [0]),
            R.RInstance(1, S.trs((1.2, 0.0, 0.6), (0.0, 0.4, 0.0), (1, 1, 1)), [1]),
            R.RInstance(2, S.trs((-0.6, 0.5, 0.1), (0.0, 0.0, 0.0), (0.5, 0.5, 0.5)), [2]),
            R.RInstance(2, S.trs((0.5, 0.35, -0.8), (0.0, 0.0, 0.0), (0.35, 0.35, 0.35)), [3]),
            R.RInstance(3, S.trs((-0.2, 0.0, 1.2), (0.0, 0.3, 0.0), (1, 1, 1)), [4])]
    h, w = 64, 128
    v = (np.arange(h) + 0.5) / h
    env = np.empty((h, w, 3), np.float32)
    env[:] = (0.3 + 0.5 * v)[:, None, None] * np.array([0.6, 0.75, 1.0], np.float32)
    env[v > 0.5] = (0.15, 0.12, 0.1)
    env[12:15, 80:84] = (300.0, 260.0, 200.0)
    rs = R.build(meshes, inst, mats, [R.directional_light((0.4, -1.0, 0.3), 2.0)], textures=[tex], env=env)
    return rs, R.Camera((0.3, 1.6, -4.0), (0.0, 0.5, 0.0), 45.0)


@pytest.mark.gpu
@pytest.mark.parametrize("env_nee", [True, False])
def test_env_garden_matches_the_oracle(device, env_nee):
    """The map-lit garden on the megakernel against the oracle, same seed and samples, 6 bounces, the map sampled
    or not: relative RMSE <= 1e-3 (the kernel does the oracle's float operations in its order)."""
    print(f"env garden 64x48 x 4 spp, map sampled {env_nee}", flush=True)
    rs, cam = env_garden()
    p = rs.params(64, 48, cam, seed=5, max_bounces=6, rr_depth=3, env_nee=env_nee)
    want = oracle_scene(rs).render(p, 0, 4)
    got, _ = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"env garden: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_env_maps_render_on_the_megakernel_and_are_checked(device):
    """Under RENDER_WAVEFRONT a batch with a map takes the megakernel (the same bits as RENDER_MEGAKERNEL and the
    counting launch); Params asking for a map the scene has not, or of another size, are refused at launch, and a
    CDF of the wrong length at upload, each with TRT_E_INVALID."""
    import copy

    import threert_amd as T

    rs, cam = env_garden()
    p = rs.params(32, 24, cam, seed=3, max_bounces=4, rr_depth=3)
    outs = []
    try:
        for how in (T.RENDER_WAVEFRONT, T.RENDER_MEGAKERNEL):
            print(f"env garden under render setting {how}", flush=True)
            device.set_render(how)
            outs.append(gpu_render(device, rs, p, 0, 2)[0])
    finally:
        device.set_render(T.RENDER_AUTO)
    print("env garden: counting launch", flush=True)
    counted, reads = gpu_render(device, rs, p, 0, 2, reads=True)
    assert reads["shaded"] > 0, reads
    for o in outs[1:] + [counted]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))

    bad = copy.deepcopy(rs)
    bad.env_cdf = bad.env_cdf[:-1]
    print("refusal: a CDF one short", flush=True)
    with pytest.raises(T.ThreeRTError) as e:
        R.upload(device, bad)
    assert e.value.code == T.TRT_E_INVALID and "env_cdf" in str(e.value), str(e.value)
    plain = copy.deepcopy(rs)
    plain.env, plain.env_cdf = None, None
    for scene, change, what in ((rs, lambda q: q.view(np.float32).__setitem__(25, 64.0), "size"),
                                (plain, None, "has none")):
        print(f"refusal: environment map ({what})", flush=True)
        sc = R.upload(device, scene)
        acc = R.Accum(device, 32, 24)
        q = p.copy()
        if change:
            change(q)
        with pytest.raises(T.ThreeRTError) as e:
            R.render(sc, q, 0, 1, acc)
        assert e.value.code == T.TRT_E_INVALID and what in str(e.value), str(e.value)
        acc.close()
        sc.close()
