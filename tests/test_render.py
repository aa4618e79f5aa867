"""The offline path loop: the oracle pinned by the reference's furnace fixtures and closed forms, and the GPU's
render held to the oracle's.

Pinning the oracle (CPU): white furnaces (tests/render/furnace-diffuse.ts: an albedo-1 Lambert sphere under a
1-nit sky reads 1.0 +- 2 %; furnace-metal.ts: F0 = 1, roughness 0.5, energy compensation 1 reads 1.0 +- 3 %), the
energy table against an independent quadrature (lib.rs:1125-1156), and an emitter-lit Lambert plane against its
closed-form irradiance.
Parity (GPU): the same scene, seed and samples give the oracle's accumulator; the tolerance is per-pixel relative
RMSE <= 1e-3 (the path's float operations are the oracle's in the same order, so in practice the bits agree)."""
import math
import sys

import numpy as np
import pytest

import oracle
from threert_amd import render as R
from threert_amd import scene as S


def oracle_scene(rs):
    p = rs.prep
    return oracle.RenderScene(p.nodes, p.leaves, p.instances, p.top_first, p.units, **rs.tables())


def sphere_scene(base, metallic, roughness, specular, sky=1.0, clearcoat=0.0, clearcoat_roughness=0.0):
    t, n = R.icosphere(4)
    rs = R.build([R.RMesh(t, normals=n)], [R.RInstance(0, S.translation([0.0, 0.0, 0.0]), [0])],
                 [R.material(base, metallic=metallic, roughness=roughness, specular=specular, clearcoat=clearcoat,
                             clearcoat_roughness=clearcoat_roughness)])
    cam = R.Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), 20.0)
    return rs, cam


def sphere_mean(acc):
    c = acc[..., :3] / np.maximum(acc[..., 3:4], 1)
    h, w = c.shape[:2]
    yy, xx = np.mgrid[0:h, 0:w]
    r = np.hypot(xx - (w - 1) / 2, yy - (h - 1) / 2)
    inside = r < 0.25 * w  # well inside the sphere's silhouette
    return c[inside].mean(axis=0)


def test_white_furnace_lambert():
    """furnace-diffuse.ts: albedo 1, no specular, sky 1: 1.0 +- 2 %."""
    rs, cam = sphere_scene((1, 1, 1), 0.0, 1.0, 0.0)
    p = rs.params(48, 48, cam, max_bounces=8, rr_depth=8, sky=(1.0, 1.0, 1.0))
    acc = oracle_scene(rs).render(p, 0, 64)
    m = sphere_mean(acc)
    assert np.all(np.abs(m - 1.0) < 0.02), m


def test_white_furnace_clearcoat():
    """furnace-clearcoat.ts: an albedo-1 Lambert sphere (no specular) under a clearcoat of roughness 0.3: the base
    gets what the coat does not reflect, so the sphere still reads 1.0 +- 2 %."""
    rs, cam = sphere_scene((1, 1, 1), 0.0, 1.0, 0.0, clearcoat=1.0, clearcoat_roughness=0.3)
    p = rs.params(48, 48, cam, max_bounces=8, rr_depth=8, sky=(1.0, 1.0, 1.0))
    acc = oracle_scene(rs).render(p, 0, 64)
    m = sphere_mean(acc)
    assert np.all(np.abs(m - 1.0) < 0.02), m


def test_a_coat_alone_reflects_its_directional_albedo():
    """A black sphere under a clearcoat (roughness 0.3) and a 1-nit sky reflects, at each pixel, the coat lobe's
    directional albedo at that pixel's n.v: 0.04 E0 + E1 from the energy table (F0 0.04, F90 1), which the energy
    table test holds to an independent quadrature. The pixels well inside the silhouette, averaged, within 3 %."""
    rs, cam = sphere_scene((0, 0, 0), 0.0, 1.0, 0.0, clearcoat=1.0, clearcoat_roughness=0.3)
    w = h = 48
    p = rs.params(w, h, cam, max_bounces=4, rr_depth=4, sky=(1.0, 1.0, 1.0))
    acc = oracle_scene(rs).render(p, 0, 256)
    got = sphere_mean(acc)
    # each pixel centre's ray against the unit sphere (the mesh is an icosphere: its shading normals interpolate
    # the sphere's), then the table at (sqrt(n.v), roughness), bilinear as the kernel reads it
    e = R.energy_lut().reshape(-1, 2)  # node i: (E0, E1)
    tan_half = math.tan(math.radians(cam.fov_deg) * 0.5)
    vals = []
    for y in range(h):
        for x in range(w):
            if math.hypot(x - (w - 1) / 2, y - (h - 1) / 2) >= 0.25 * w:
                continue
            ndc = ((x + 0.5) / w * 2 - 1, 1 - (y + 0.5) / h * 2)
            d = np.array([ndc[0] * tan_half, ndc[1] * tan_half, 1.0])
            d /= np.linalg.norm(d)
            o = np.array([0.0, 0.0, -4.0])
            b = o @ d
            t = -b - math.sqrt(b * b - (o @ o - 1.0))
            n = o + t * d
            nov = float(-(n @ d))
            xx = math.sqrt(min(max(nov, 0.0), 1.0)) * 31
            yy = 0.3 * 31
            x0, y0 = min(int(xx), 30), min(int(yy), 30)
            fx, fy = xx - x0, yy - y0

            def node(i):
                return e[i]

            i = y0 * 32 + x0  # noqa: E741
            a0 = node(i) * (1 - fx) + node(i + 1) * fx
            a1 = node(i + 32) * (1 - fx) + node(i + 33) * fx
            ec = a0 * (1 - fy) + a1 * fy
            vals.append(0.04 * ec[0] + ec[1])
    want = float(np.mean(vals))
    assert 0.03 < want < 0.2
    assert np.all(np.abs(got - want) < 0.03 * want), (got, want)


def test_the_emitter_guide_finds_what_the_full_search_finds():
    """lib.rs:1459-1500 on the library's emitter table: 3000 emitter triangles over six decades of area (the
    reference's LCG, seed 11), a ninth of them black (weight 0); the kernel's pick (pathtrace.wgsl:478-487: the guide
    bucket's range, then the binary search) equals the search over the whole list for 200,000 random 24-bit u, the
    bucket edges and their neighbours, and every entry's own cumulative weight. The table: the entries, the 65,537
    guide words, the header (total weight)."""
    seed = [11]

    def rnd():
        seed[0] = (seed[0] * 1664525 + 1013904223) & 0xFFFFFFFF
        return seed[0]

    def unit():
        return np.float32((rnd() >> 8) / float(1 << 24))

    tris = []
    for i in range(3000):
        c = np.array([300.0 * unit(), 5.0, 100.0 * unit()], np.float32)
        s = np.float32(math.sqrt(2.0 * 10.0 ** (-6.0 * float(unit()))))
        tris.append([c, c + np.array([s, 0, 0], np.float32), c + np.array([0, 0, s], np.float32)])
    tris = np.array(tris, np.float32)
    sections = np.array([0 if i % 9 == 0 else 1 + i % 3 for i in range(3000)], np.uint32)
    mats = [R.material((0, 0, 0), (0.0, 0.0, 0.0))] + [R.material((0, 0, 0), (3.0 * k, k, 0.5 * k)) for k in (1, 2, 3)]
    rs = R.build([R.RMesh(tris, sections=sections)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0, 1, 2, 3])],
                 mats)
    t = rs.emitters
    n = len(rs.emitter_tris)
    assert n > 2000 and len(t) == n + (65537 + 3) // 4 + 1
    total = t[-1, 0]
    assert total > 0 and (t[-1, 1:] == 0).all()
    cdf = t[:n, 1]
    assert (np.diff(cdf) >= 0).all() and abs(float(cdf[-1]) - 1.0) < 1e-6
    assert abs(float(np.sum(t[:n, 2].astype(np.float64))) / float(total) - 1.0) < 1e-5  # the pmf sums to 1
    guide = t[n:n + (65537 + 3) // 4].view(np.uint32).reshape(-1)[:65537]

    def search(lo, hi, u):
        while lo < hi:
            mid = (lo + hi) // 2
            if cdf[mid] < u:
                lo = mid + 1
            else:
                hi = mid
        return min(lo, n - 1)

    def pick(u):
        k = min(int(np.float32(u) * np.float32(65536)), 65535)
        return search(int(guide[k]), int(guide[k + 1]), u)

    us = [rnd() >> 8 for _ in range(200_000)]
    for k in range(65537):
        m = k << 12
        us.extend(x for x in (max(m - 1, 0), m, m + 1) if x < 1 << 24)
    us.extend(min(int(np.float32(c) * np.float32(1 << 24)), (1 << 24) - 1) for c in cdf)
    for m in us:
        u = np.float32(m) / np.float32(1 << 24)
        assert pick(u) == search(0, n, u), m


def test_white_furnace_metal():
    """furnace-metal.ts: F0 = 1 (metallic, base 1), roughness 0.5, energy compensation 1: 1.0 +- 3 %."""
    rs, cam = sphere_scene((1, 1, 1), 1.0, 0.5, 0.5)
    p = rs.params(48, 48, cam, max_bounces=16, rr_depth=16, sky=(1.0, 1.0, 1.0), energy_compensation=1.0)
    acc = oracle_scene(rs).render(p, 0, 64)
    m = sphere_mean(acc)
    assert np.all(np.abs(m - 1.0) < 0.03), m


def test_energy_table_is_the_lobes_integral():
    """lib.rs:1125-1156: specular_albedo against an independent quadrature of the same lobe (uniform hemisphere
    sampling in f64, |a + b - q| < 1e-3 at seven points); a mirror exactly; the table 512 vec4 in [0, 1 + 1e-4]."""
    def quadrature(mu, alpha, n):
        # lib.rs:1125-1142: the lobe's value x cosine over the hemisphere at F = 1, midpoint rule, n x 2n points
        a2 = alpha * alpha
        v0, v1 = math.sqrt(1.0 - mu * mu), mu
        th = (math.pi / 2) * (np.arange(n) + 0.5) / n
        st, c = np.sin(th)[:, None], np.cos(th)[:, None]
        phi = math.pi * (np.arange(2 * n) + 0.5) / (2 * n)
        x, z = st * np.cos(phi)[None, :], st * np.sin(phi)[None, :]
        h0, h1, h2 = v0 + x, v1 + c, z
        noh = h1 / np.sqrt(h0 * h0 + h1 * h1 + h2 * h2)
        d = (noh * a2 - noh) * noh + 1.0
        vis = 0.5 / (c * (mu * (1.0 - alpha) + alpha) + mu * (c * (1.0 - alpha) + alpha))
        s = (a2 / (math.pi * d * d) * vis * c * st).sum()
        return s * math.pi * math.pi / (2 * n * n)

    for r, mu in [(0.3, 0.9), (0.5, 0.7071), (0.5, 0.2), (0.8, 0.5), (1.0, 1.0), (1.0, 0.1), (0.25, 0.4)]:
        a, b = R.specular_albedo(mu, r * r)
        q = quadrature(mu, r * r, 800)
        assert abs(a + b - q) < 1e-3, (r, mu, a + b, q)
    a, b = R.specular_albedo(0.5, 0.0)
    assert abs(a + b - 1.0) < 1e-12 and abs(b - 0.5 ** 5) < 1e-12
    a, b = R.specular_albedo(0.5, 0.02 * 0.02)
    assert abs(a + b - 1.0) < 2e-3 and abs(b - 0.5 ** 5) < 2e-3
    lut = R.energy_lut()
    assert lut.shape == (512, 4) and lut.min() >= 0.0 and lut.max() <= 1.0 + 1e-4


def test_an_emitter_lights_a_plane_as_its_closed_form():
    """A 1 x 1 Lambert plane (albedo 1, no specular) under a small emissive square of radiance L at height h: the
    irradiance at the centre below it is L x the square's projected solid angle (closed form), so the plane reads
    L x F / pi x ... = L * form factor. Direct light only (one bounce)."""
    L, h, a = 10.0, 1.0, 0.05
    plane = R.quad((-0.5, 0, -0.5), (-0.5, 0, 0.5), (0.5, 0, 0.5), (0.5, 0, -0.5))
    lamp = R.quad((-a, h, -a), (a, h, -a), (a, h, a), (-a, h, a))
    rs = R.build([R.RMesh(plane), R.RMesh(lamp)],
                 [R.RInstance(0, np.eye(4, dtype=np.float32), [0]), R.RInstance(1, np.eye(4, dtype=np.float32), [1])],
                 [R.material((1, 1, 1), specular=0.0), R.material((0, 0, 0), (L, L, L), specular=0.0)])
    cam = R.Camera((0.0, 0.3, -0.02), (0.0, 0.0, 0.0), 4.0)
    p = rs.params(16, 16, cam, max_bounces=1, rr_depth=8)
    acc = oracle_scene(rs).render(p, 0, 256)
    c = (acc[..., :3] / acc[..., 3:4]).mean(axis=(0, 1))
    # radiance off a Lambert plane = E / pi, E = L * (4 * ...) : the square's form factor from a point below its
    # centre (Howell's catalogue, a parallel square of half side a at distance h)
    x = a / h
    ff = (2 / math.pi) * (x / math.sqrt(1 + x * x)) * math.atan(x / math.sqrt(1 + x * x)) * 2
    want = L * ff
    assert abs(c[0] - want) / want < 0.03, (c, want)


def sunlit_plane(elevation_deg: float, diameter_deg: float, lux: float = 1000.0):
    """A 2 x 2 Lambert plane (albedo 1, no specular) at y = 0 under one directional light (pathtrace.wgsl:2449-2472,
    1963-1988) at `elevation_deg` from the zenith, seen from straight above."""
    plane = R.quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
    th = math.radians(elevation_deg)
    sun = R.directional_light((math.sin(th), -math.cos(th), 0.0), (lux, lux, lux), diameter_deg=diameter_deg)
    rs = R.build([R.RMesh(plane)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0])],
                 [R.material((1, 1, 1), specular=0.0)], [sun])
    return rs, R.Camera((0.0, 3.0, 0.001), (0.0, 0.0, 0.0), 20.0)


@pytest.mark.parametrize("elevation,diameter", [(30.0, 0.5357), (30.0, 40.0), (20.0, 100.0), (55.0, 60.0)])
def test_the_sun_lights_a_plane_as_its_closed_form(elevation, diameter):
    """The sun's light at a Lambert plane: a cone of half-angle a around the direction at angle t0 from the normal,
    of illuminance E across its axis, holds radiance E / (2 pi (1 - cos a)) and lights the plane with
    E x (pi sin^2 a cos t0) / (2 pi (1 - cos a)) = E cos t0 (1 + cos a) / 2 (the cap's projected solid angle, exact
    while the cap is above the horizon); the plane reads that over pi. Light sampling (the cone sample, its pdf, the
    shadow ray) and BSDF rays that reach the disk (sun_disks) share it by the power heuristic: on the 0.54-degree sun
    nearly all of it comes by light sampling, on the wide cones a good part by BSDF rays, and every case must read
    the closed form (float64 mean over 4096 samples a pixel, within 1 %)."""
    rs, cam = sunlit_plane(elevation, diameter)
    p = rs.params(16, 16, cam, max_bounces=1, rr_depth=8)
    acc = oracle_scene(rs).render(p, 0, 4096, threads=8).astype(np.float64)
    got = (acc[..., :3] / acc[..., 3:4]).mean()
    a = math.radians(diameter * 0.5)
    want = 1000.0 * math.cos(math.radians(elevation)) * (1 + math.cos(a)) / 2 / math.pi
    assert abs(got - want) < 0.01 * want, (got, want)


def test_russian_roulette_keeps_the_expectation():
    """Inside a closed box whose walls are Lambert of albedo 1/2 and emit 1/2 (two-sided), the radiance everywhere
    is Le / (1 - albedo) = 1: every bounce adds half of what is left. With Russian roulette from the first bounce
    (rr_depth 1: q = min(max throughput, 0.95), throughput / q on survival, pathtrace.wgsl:4469-4473) and 64 bounces,
    most paths end early and the survivors carry the rest: the image reads 1 +- 2 %."""
    box = R.box_tris((-1, -1, -1), (1, 1, 1))
    rs = R.build([R.RMesh(box)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0])],
                 [R.material((0.5, 0.5, 0.5), (0.5, 0.5, 0.5), specular=0.0)])
    cam = R.Camera((0.1, -0.2, -0.5), (0.3, 0.2, 1.0), 70.0)
    p = rs.params(24, 24, cam, max_bounces=64, rr_depth=1)
    acc = oracle_scene(rs).render(p, 0, 512, threads=8).astype(np.float64)
    got = (acc[..., :3] / acc[..., 3:4]).mean()
    assert abs(got - 1.0) < 0.02, got


def test_the_sampler_holds_its_known_answers():
    """sample.rs:91-125: the direction table's last entries (Cycles'), and the first 2^m points of each 2D pair
    a (0, m, 2)-net, for the reference's seeds."""
    d = oracle.sobol_dirs()
    assert [int(d[0][31]), int(d[1][31]), int(d[2][31])] == [0xFFFFFFFF, 0xAAAA00A3, 0xC900A00A]
    assert d[0][:16].tolist() == [1, 3, 5, 15, 17, 51, 85, 255, 257, 771, 1285, 3855, 4369, 13107, 21845, 65535]
    for seed, dd in [(0, 0), (0x12345678, 7), (oracle.hash_hp(99), 0xDEAD)]:
        for m in (4, 6, 8):
            n = 1 << m
            pts = np.array([oracle.rng4_at(seed, i, dd) for i in range(n)])
            assert ((pts >= 0) & (pts < 1)).all()

            def cell(v, k):
                return np.minimum((v * np.float32(1 << k)).astype(np.int64), (1 << k) - 1)

            for a in range(m + 1):
                cells = (cell(pts[:, 0], a) << (m - a)) | cell(pts[:, 1], m - a)
                assert len(np.unique(cells)) == n, (seed, m, a)
            for k in range(4):
                assert sorted(cell(pts[:, k], m).tolist()) == list(range(n))


def test_sincos_is_within_two_ulps():
    xs = np.linspace(0.0, 2 * math.pi, 20001, dtype=np.float32)
    for x in xs[::97]:
        s, c = oracle.sincos(float(x))
        assert abs(s - math.sin(float(x))) <= 2.5e-7 and abs(c - math.cos(float(x))) <= 2.5e-7


# ---- GPU parity ---------------------------------------------------------------------------------------------------
def rel_rmse(a, b):
    ca = a[..., :3] / np.maximum(a[..., 3:4], 1)
    cb = b[..., :3] / np.maximum(b[..., 3:4], 1)
    den = np.maximum(np.abs(cb), 1e-3)
    return float(np.sqrt(np.mean(((ca - cb) / den) ** 2)))


def gpu_render(device, rs, p, first, spp):
    sc = R.upload(device, rs)
    acc = R.Accum(device, int(p[16]), int(p[17]))
    R.render(sc, p, first, spp, acc)
    out = acc.read()
    device.sync()
    acc.close()
    sc.close()
    return out


@pytest.mark.gpu
def test_cornell_matches_the_oracle(device):
    print("cornell 96x96 x 8 spp", flush=True)
    rs, cam = R.cornell()
    p = rs.params(96, 96, cam, seed=1, max_bounces=4, rr_depth=3)
    want = oracle_scene(rs).render(p, 0, 8)
    got = gpu_render(device, rs, p, 0, 8)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"cornell: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


def oracle_threads() -> int:
    import os

    return min(16, os.cpu_count() or 1)


@pytest.mark.gpu
def test_cornell_at_full_size_matches_the_oracle(device):
    """The plain-method workload at its size: 512 x 512, 16 samples, 4 bounces, RR from depth 3, seed 1."""
    print("cornell 512x512 x 16 spp", flush=True)
    rs, cam = R.cornell()
    p = rs.params(512, 512, cam, seed=1, max_bounces=4, rr_depth=3)
    want = oracle_scene(rs).render(p, 0, 16, threads=oracle_threads())
    got = gpu_render(device, rs, p, 0, 16)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"cornell 512: {same:.6f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_the_full_city_at_720p_matches_the_oracle(device):
    """The instanced-city workload at its size: 64 meshes (1.0M unique triangles) x 4096 instances (66M instanced),
    metal-rough GGX, the 90,000-lux sun and a constant sky, 1280 x 720, 1 spp, 3 bounces, from a camera above the
    skyline (a real image: most of it lit)."""
    print("city: building 64 meshes x 4096 instances", flush=True)
    rs, cam, sky = R.city(seed=1)
    assert rs.extras["unique_triangles"] > 900_000 and rs.extras["instanced_triangles"] > 55_000_000
    p = rs.params(1280, 720, cam, seed=1, max_bounces=3, rr_depth=3, sky=sky)
    print("city: oracle render", flush=True)
    want = oracle_scene(rs).render(p, 0, 1, threads=oracle_threads())
    print("city: GPU render", flush=True)
    got = gpu_render(device, rs, p, 0, 1)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"city 720p: {same:.6f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3
    assert (want[..., :3].sum(-1) == 0).mean() < 0.2  # at one sample, 11 % (shadowed pixels whose bounce found no light)


@pytest.mark.gpu
def test_a_thin_lens_matches_the_oracle(device):
    """camera_ray's thin lens (pathtrace.wgsl:4017-4023: lens radius cam_right.w, focus distance cam_up.w)."""
    print("cornell through a thin lens 96x96 x 8 spp", flush=True)
    rs, cam = R.cornell()
    p = rs.params(96, 96, cam, seed=3, max_bounces=4, rr_depth=3, lens_radius=0.06, focus=1.6)
    want = oracle_scene(rs).render(p, 0, 8)
    got = gpu_render(device, rs, p, 0, 8)
    print(f"lens: rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3
    pin = oracle_scene(rs).render(rs.params(96, 96, cam, seed=3, max_bounces=4, rr_depth=3), 0, 8)
    assert rel_rmse(pin, want) > 1e-2  # the lens changes the image


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["lit floor", "soup"])
def test_light_tree_scenes_match_the_oracle(device, which):
    """Scenes of 256 emitters or more sample them through the light tree (ltree.rs; lt_imp, emitter_draw and
    emitter_pmf, pathtrace.wgsl:504-618; six tree candidates at camera hits): the GPU's accumulator is the oracle's.
    The soup is the precision workload's generator at 300,000 triangles (15,000 emissive)."""
    from threert_amd import serve
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    import test_light_tree as LT

    if which == "soup":
        rs, cam, sky = serve.soup_scene(300_000, 21)
        p = rs.params(160, 96, cam, seed=5, max_bounces=4, rr_depth=3, sky=sky)
    else:
        rs, cam = LT.lit_floor()
        p = rs.params(96, 64, cam, seed=5, max_bounces=3, rr_depth=2)
    assert rs.tree_entries >= 256 and p.view(np.float32)[87] == 6.0
    print(f"{which}: {rs.tree_entries} emitters in the tree", flush=True)
    want = oracle_scene(rs).render(p, 0, 4, threads=8)
    got = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"{which}: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3 and want[..., :3].sum() > 0


@pytest.mark.gpu
def test_sunlit_instances_match_the_oracle(device):
    print("small city 80x48 x 2 spp", flush=True)
    rs, cam, sky = R.city(seed=3, meshes=6, instances=40, extent=60.0)
    p = rs.params(80, 48, cam, seed=7, max_bounces=3, rr_depth=2, sky=sky)
    want = oracle_scene(rs).render(p, 0, 2)
    got = gpu_render(device, rs, p, 0, 2)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"city: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_clearcoats_match_the_oracle(device):
    """The Cornell box with its boxes and back wall under clearcoats of three roughnesses."""
    print("coated cornell 64x64 x 8 spp", flush=True)
    rs, cam = R.cornell()
    mats = rs.materials.copy()
    for k, (amount, rough) in zip(range(len(mats)), ((1.0, 0.05), (0.5, 0.3), (1.0, 0.6), (0.8, 0.0))):
        mats[k, 30], mats[k, 31] = amount, rough
    rs.materials = mats
    p = rs.params(64, 64, cam, seed=4, max_bounces=4, rr_depth=3)
    want = oracle_scene(rs).render(p, 0, 8)
    got = gpu_render(device, rs, p, 0, 8)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"coated cornell: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_render_ray_counts(device):
    """trt_render_ray_counts: a camera that looks away from the scene traces exactly one closest ray a sample and no
    shadow ray; the Cornell box traces at least that, the same counts for the same render, and reading resets."""
    rs, cam = R.cornell()
    w, h, spp = 48, 32, 4
    away = R.Camera((0.5, 0.5, -1.35), (0.5, 0.5, -3.0), 40.0)
    device.render_ray_counts()
    gpu_render(device, rs, rs.params(w, h, away, seed=1, max_bounces=4, rr_depth=3), 0, spp)
    assert device.render_ray_counts() == (w * h * spp, 0)
    assert device.render_ray_counts() == (0, 0)
    p = rs.params(w, h, cam, seed=1, max_bounces=4, rr_depth=3)
    gpu_render(device, rs, p, 0, spp)
    first = device.render_ray_counts()
    gpu_render(device, rs, p, 0, spp)
    assert device.render_ray_counts() == first
    closest, shadow = first
    print(f"cornell {w}x{h}x{spp}: {closest} closest, {shadow} shadow rays", flush=True)
    assert closest > w * h * spp and shadow > 0
    assert closest <= w * h * spp * 5  # a camera ray and at most 4 bounces a sample


@pytest.mark.gpu
def test_the_fp64_reference_is_the_float64_sum_of_the_launches(device):
    """render.reference_fp64 (trt_accum_add_fp64 on the device) equals adding each launch's FP32 sums in float64 on
    the host, bit for bit, non-finite sums dropped the same way."""
    rs, cam = R.cornell()
    w, h = 40, 32
    p = rs.params(w, h, cam, seed=1, max_bounces=4, rr_depth=3)
    sc = R.upload(device, rs)
    got = R.reference_fp64(device, sc, p, w, h, 96, chunk=16)
    total = np.zeros((h, w, 3), np.float64)
    count = np.zeros((h, w, 1), np.float64)
    for s in range(0, 96, 16):
        acc = R.Accum(device, w, h)
        R.render(sc, p, s, 16, acc)
        px = acc.read()
        acc.close()
        fin = np.isfinite(px[..., 0:3]).all(axis=2, keepdims=True)
        total += np.where(fin, px[..., 0:3], 0.0)
        count += px[..., 3:4]
    sc.close()
    want = total / np.maximum(count, 1.0)
    assert got.dtype == np.float64 and (got == want).all()


@pytest.mark.gpu
def test_gpu_white_furnace(device):
    rs, cam = sphere_scene((1, 1, 1), 0.0, 1.0, 0.0)
    p = rs.params(64, 64, cam, max_bounces=8, rr_depth=8, sky=(1.0, 1.0, 1.0))
    acc = gpu_render(device, rs, p, 0, 128)
    m = sphere_mean(acc)
    assert np.all(np.abs(m - 1.0) < 0.02), m


@pytest.mark.gpu
def test_samples_split_across_calls_equal_one_call(device):
    """Samples 0..3 then 4..7 into one accumulator equal samples 0..7 at once (the host loop's slicing)."""
    rs, cam = R.cornell()
    p = rs.params(64, 64, cam, seed=1)
    sc = R.upload(device, rs)
    a = R.Accum(device, 64, 64)
    R.render(sc, p, 0, 4, a)
    R.render(sc, p, 4, 4, a)
    b = R.Accum(device, 64, 64)
    R.render(sc, p, 0, 8, b)
    ra, rb = a.read(), b.read()
    assert np.allclose(ra, rb, rtol=1e-6, atol=1e-6)
    c, al, n, v = b.resolve()
    assert np.allclose(c, rb[..., :3] / rb[..., 3:4]) and (v >= 0).all()
    a.close()
    b.close()
    sc.close()


@pytest.mark.gpu
def test_exact_ties_keep_their_own_section(device):
    """Four instances of one soup at one place: instance 0 places the mesh whose triangles are in a plain section,
    instances 1-3 a copy whose triangles are in an emissive section. Every ray that meets the soup meets an exact
    tie, won by instance 0, and the copies' boxes (stretched by an off-screen triangle toward the camera) are entered
    first, so the winner replaces an earlier hit on an exact tie; the winner's section must come with it. (Written as a branch around the hit update and inlined into the path tracer, the compiler once
    kept the replaced hit's section.)"""
    print("exact ties: 2000 triangles x 4 stacked instances, 64x64 x 4 spp", flush=True)
    tris, _ = S.clustered_soup(2000, 12, clusters=4, extent=10.0, median_area=0.05)
    plain = R.RMesh(tris, sections=np.zeros(len(tris), np.uint32))
    # the copies carry one more, off-screen triangle toward the camera: their boxes are entered first
    far = np.array([[[8.0, 8.0, -12.0], [8.5, 8.0, -12.0], [8.0, 8.5, -12.0]]], np.float32)
    glowing = R.RMesh(np.concatenate([tris, far]), sections=np.ones(len(tris) + 1, np.uint32))
    eye = np.eye(4, dtype=np.float32)
    rs = R.build([plain, glowing], [R.RInstance(0, eye, [0, 1])] + [R.RInstance(1, eye, [0, 1]) for _ in range(3)],
                 [R.material((0.7, 0.7, 0.7)), R.material((0.0, 0.0, 0.0), (40.0, 40.0, 40.0))],
                 [R.directional_light((-0.3, -1.0, -0.2), (3.0, 3.0, 3.0))])
    cam = R.Camera((0.0, 4.0, -16.0), (0.0, 0.0, 0.0), 50.0)
    p = rs.params(64, 64, cam, seed=3, max_bounces=3, rr_depth=2, sky=(0.2, 0.3, 0.4))
    want = oracle_scene(rs).render(p, 0, 4)
    got = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"ties: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert same == 1.0 and rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_the_serving_job_that_met_a_tie_matches_the_oracle(device):
    """The serving workload's job 163 (a 2.94M-triangle clustered scene, 1788 x 1026, 4 bounces, seed 164) is where
    a miscompiled hit update first showed: at samples 20, 32 and 48 a bounce met an exact tie and kept the replaced
    hit's section, and the path read its 'own' emitter past the end of the emitter list. Whole rows through those
    pixels must equal the oracle's, sample by sample."""
    from threert_amd import serve

    print("serving job 163: scene (2940646, 1003)", flush=True)
    rs, cam, sky = serve.soup_scene(2940646, 1003)
    job = [j for j in serve.workload() if j.id == 163][0]
    p = serve.params_of(serve._Resident(None, rs, cam, sky, 0), job)
    sc = R.upload(device, rs)
    orc = oracle_scene(rs)
    for s, rows in ((20, (194,)), (32, (223, 311)), (48, (131,))):
        acc = R.Accum(device, job.width, job.height)
        R.render(sc, p, s, 1, acc)
        got = acc.read()
        acc.close()
        for y in rows:
            want = orc.render(p, s, 1, None, rows=(y, y + 1), threads=8)[y]
            bad = int((got[y].view(np.uint32) != want.view(np.uint32)).any(axis=1).sum())
            print(f"sample {s} row {y}: {bad} pixels differ", flush=True)
            assert bad == 0
    sc.close()


@pytest.mark.gpu
def test_what_the_library_does_not_render_is_refused(device):
    """Records with cutout sections or with glass among shadow-casting sections (scene.rs INST_CUTOUT 16, INST_GLASS 32
    in the bits word; the Cornell box's records cast shadows); materials with textures, normal maps, function
    flags, transmission textures, alpha cutoffs, or glass without the glass energy table;
    and Params asking for an environment map, more than 64 bounces, more lights than the scene has, fog, volumetric
    scattering, the live light LOD, procedural geometry, aerial perspective or light-tree candidates
    without a tree: each refused with TRT_E_INVALID and a message that names it. The device still renders the unchanged scene afterwards."""
    import copy

    import threert_amd as T

    rs, cam = R.cornell()
    for bit, name in ((16, "cutout sections"), (32, "glass among shadow-casting sections")):
        print(f"refusal: record bit {bit} ({name})", flush=True)
        bad = copy.deepcopy(rs)
        words = bad.prep.instances["n"][:, 1, 3].view(np.uint32)
        words[0] |= np.uint32(bit)
        with pytest.raises(T.ThreeRTError) as e:
            R.upload(device, bad)
        assert e.value.code == T.TRT_E_INVALID and name in str(e.value), str(e.value)
    # materials that need what the kernel does not read (GpuMaterial words, scene.rs:480-491)
    as_int = lambda v: np.array([v], np.int32).view(np.float32)[0]  # noqa: E731
    for word, value, name in ((12, as_int(0), "base-colour texture"), (13, as_int(2), "normal map"),
                              (16, as_int(1), "emissive texture"),
                              (19, as_int(1), "material function"), (22, 1.0, "transmission map"),
                              (27, 0.5, "alpha cutoff"),
                              (28, 1.0, "no glass energy table")):
        print(f"refusal: material word {word} ({name})", flush=True)
        bad = copy.deepcopy(rs)
        bad.materials[1, word] = value
        with pytest.raises(T.ThreeRTError) as e:
            R.upload(device, bad)
        assert e.value.code == T.TRT_E_INVALID and name in str(e.value), str(e.value)
    sc = R.upload(device, rs)
    acc = R.Accum(device, 16, 16)
    f32 = lambda k, v: (lambda p: p.view(np.float32).__setitem__(k, v))  # noqa: E731
    for what, change in (("environment map", f32(24, 1.0)), ("more than 64 bounces", lambda p: p.__setitem__(22, 65)),
                         ("P.counts names more lights", lambda p: p.__setitem__(20, 99)), ("fog", f32(44, 0.01)),
                         ("volumetric scattering", f32(52, 0.5)), ("live light LOD", f32(75, 10.0)),
                         ("procedural geometry", lambda p: p.__setitem__(89, 1)),
                         ("aerial perspective", f32(92, 1.0)),
                         ("light-tree candidates", f32(87, 6.0))):
        print(f"refusal: {what}", flush=True)
        p = rs.params(16, 16, cam)
        change(p)
        with pytest.raises(T.ThreeRTError) as e:
            R.render(sc, p, 0, 1, acc)
        assert e.value.code == T.TRT_E_INVALID and what.split()[0] in str(e.value), str(e.value)
    R.render(sc, rs.params(16, 16, cam, max_bounces=64), 0, 2, acc)
    device.sync()
    c = acc.resolve()[0]
    assert np.isfinite(c).all() and c.mean() > 0
    acc.close()
    sc.close()
