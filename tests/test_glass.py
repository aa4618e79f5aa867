"""Glass: the rough dielectric interface (pathtrace.wgsl:1542-1640, Walter et al. 2007 over GGX) and glass along
shadow rays (pathtrace.wgsl:846-977), in the oracle, pinned by what the reference's own tests fix.

- The glass energy table's estimator (lib.rs:1208 glass_energy, restated in threert_amd/render.py) against an
  independent integral of the interface's BSDF over both sides' directions, the cases of lib.rs:1328-1334, and the
  numbers the reference's render tests expect (lib.rs:1336-1340: R + T of 0.9573, 0.7390 and 0.6474).
- The kernel's table lookup (the oracle's glass_energy) equals the table at its nodes and trilinear interpolation
  between them.
- Renders against the reference's analytic render tests (tests/render/*.ts): glass-rough-enter / -leave /
  -leave-low (one rough interface in a white furnace, by single scattering), their -compensated forms and
  glass-rough-sphere-furnace (1 with energy compensation), glass-internal-reflection (total internal reflection is
  untinted), material-glass-transmission (a clear slab passes (1 - 0.04)^2 / (1 - 0.04^2) of a wall behind it), and
  shadow-glass-tint with a sun in place of its point light (a shadow ray keeps (1 - 0.04) x the colour at each of a
  pane's two surfaces).
- Scenes without glass render as before: the glass code is off without the table.
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


# ---- the interface's R + T, integrated over directions --------------------------------------------------------
def _lam(a2, c):
    c = np.minimum(np.abs(c), 1.0)
    return 0.5 * (np.sqrt(1.0 + a2 * (1.0 - c * c) / (c * c)) - 1.0)


def interface_integral(mu: float, alpha: float, eta: float, n: int = 600) -> tuple[float, float]:
    """(R, T) of one rough interface toward a view at cosine mu: the BSDF x |cos| summed over a midpoint grid in
    (cos theta, phi) on each side. R: GGX D x the kernel's approximate joint visibility x the exact Fresnel at the
    half vector; T: Walter eq. 21 with the height-correlated Smith G2, the half vector v + eta l."""
    a2 = alpha * alpha
    v = np.array([math.sqrt(1.0 - mu * mu), 0.0, mu])
    c = (np.arange(n) + 0.5) / n
    phi = (np.arange(2 * n) + 0.5) * math.pi / n
    C, PH = np.meshgrid(c, phi, indexing="ij")
    st = np.sqrt(1.0 - C * C)
    lx, ly = st * np.cos(PH), st * np.sin(PH)
    dw = (1.0 / n) * (math.pi / n)
    # reflection: l above
    hx, hy, hz = v[0] + lx, v[1] + ly, v[2] + C
    k = np.sqrt(hx * hx + hy * hy + hz * hz)
    noh, voh = hz / k, (v[0] * hx + v[2] * hz) / k
    d = a2 / (math.pi * ((noh * a2 - noh) * noh + 1.0) ** 2)
    vis = 0.5 / np.maximum(C * (mu * (1.0 - alpha) + alpha) + mu * (C * (1.0 - alpha) + alpha), 1e-7)
    r = float((d * vis * R.fresnel_dielectric(voh, eta) * C).sum() * dw)
    # transmission: l below (z = -C)
    hx, hy, hz = v[0] + eta * lx, v[1] + eta * ly, v[2] - eta * C
    k = np.sqrt(hx * hx + hy * hy + hz * hz) * np.sign(hz)
    hx, hy, hz = hx / k, hy / k, hz / k
    voh = v[0] * hx + v[2] * hz
    loh = lx * hx + ly * hy - C * hz
    ok = (voh > 0) & (loh < 0)
    d = a2 / (math.pi * ((hz * a2 - hz) * hz + 1.0) ** 2)
    den = voh + eta * loh
    g2 = 1.0 / (1.0 + _lam(a2, mu) + _lam(a2, C))
    with np.errstate(divide="ignore", invalid="ignore"):
        f = (1.0 - R.fresnel_dielectric(np.where(ok, voh, 1.0), eta)) * d * g2 * voh * eta * eta * -loh / (mu * den * den)
    t = float(np.where(ok, f, 0.0).sum() * dw)
    return r, t


@pytest.mark.parametrize("rough,eta,mu", [(0.3, 1.5, 0.7071), (0.3, 1 / 1.5, 0.7071), (0.6, 1.5, 0.4),
                                          (0.6, 1 / 1.5, 0.9), (1.0, 1.5, 0.9), (1.0, 1 / 1.5, 0.9),
                                          (0.5, 1.33, 0.2), (0.8, 1 / 1.33, 0.5)])
def test_glass_energy_estimator_is_the_interface_integral(rough, eta, mu):
    """lib.rs:1328-1334: the table's estimator (visible normals, 24 x 24) within 1.2 % of the direction integral."""
    e = R.glass_energy(mu, rough * rough, eta, 24)
    qr, qt = interface_integral(mu, rough * rough, eta)
    assert abs(e - qr - qt) < 0.012 * (qr + qt), (e, qr, qt)


@pytest.mark.parametrize("eta,mu,want", [(1.5, math.sqrt(0.5), 0.9573), (1 / 1.5, math.sqrt(0.5), 0.7390),
                                         (1 / 1.5, 0.4, 0.6474)])
def test_glass_energy_matches_the_render_tests_numbers(eta, mu, want):
    """lib.rs:1336-1340 (glass-rough-enter / -leave / -leave-low, roughness 0.6, ior 1.5): the integral within 1e-3
    and the estimator at 192 x 192 within 5e-4 of the reference's numbers."""
    qr, qt = interface_integral(mu, 0.36, eta, 1000)
    assert abs(qr + qt - want) < 1e-3, (qr, qt)
    assert abs(R.glass_energy(mu, 0.36, eta, 192) - want) < 5e-4


def test_fresnel_dielectric_closed_forms():
    """Normal incidence ((eta - 1) / (eta + 1))^2; total internal reflection past the critical angle; the reference
    test's 1 - 0.04 of ior 1.5."""
    assert abs(float(R.fresnel_dielectric(1.0, 1.5)) - 0.04) < 1e-12
    assert abs(float(R.fresnel_dielectric(1.0, 1 / 1.5)) - 0.04) < 1e-12
    crit = math.sqrt(1.0 - (1 / 1.5) ** 2)  # cos of the critical angle leaving ior 1.5
    assert float(R.fresnel_dielectric(crit - 1e-6, 1 / 1.5)) == 1.0
    assert float(R.fresnel_dielectric(crit + 1e-3, 1 / 1.5)) < 1.0
    # s and p polarisations averaged, from Snell's law directly
    ci, eta = 0.6, 1.5
    ct = math.sqrt(1 - (1 - ci * ci) / eta ** 2)
    rs = ((ci - eta * ct) / (ci + eta * ct)) ** 2
    rp = ((eta * ci - ct) / (eta * ci + ct)) ** 2
    assert abs(float(R.fresnel_dielectric(ci, eta)) - 0.5 * (rs + rp)) < 1e-12


def test_glass_energy_table_layout_and_lookup():
    """lib.rs:1341-1344: 2 x 16 x 16 x 8 values in [0.05, 1]. The kernel's lookup (pathtrace.wgsl:1582-1603, the
    oracle's glass_energy) returns the table's value at its nodes (sqrt(cos) = i / 15, roughness = j / 15,
    sqrt((ior - 1) / (ior + 1)) = k / 7 x sqrt(1/2), entering then leaving) and trilinear interpolation between."""
    lut = R.glass_energy_lut()
    assert lut.shape == (1024, 4) and lut.min() >= 0.05 and lut.max() <= 1.0
    flat = lut.reshape(-1).astype(np.float64)
    rs = sphere_glass(0.5, 1.5)[0]
    sc = oracle_scene(rs)

    def node(side, iz, ir, im):
        return flat[((side * 8 + iz) * 16 + ir) * 16 + im]

    def coords(iz):
        z = iz / 7 * R.GLASS_E_ZMAX
        return (1 + z * z) / (1 - z * z)

    for side in (0, 1):
        for iz in (1, 3, 6):  # (iz 0 is ior 1: the kernel's lookup floors it with the clamp)
            for ir in (0, 5, 15):
                for im in (2, 9, 15):
                    got = sc.glass_energy((im / 15) ** 2, ir / 15, coords(iz), bool(side))
                    assert abs(got - node(side, iz, ir, im)) < 2e-5, (side, iz, ir, im, got)
    # between nodes: trilinear in (sqrt(cos), roughness, z)
    rng = np.random.default_rng(3)
    for _ in range(200):
        side = int(rng.integers(0, 2))
        x, y, z = rng.uniform(0, 15), rng.uniform(0, 15), rng.uniform(0.05, 7)
        zz = z / 7 * R.GLASS_E_ZMAX
        ior = (1 + zz * zz) / (1 - zz * zz)
        x0, y0, z0 = int(x), int(y), int(z)
        x1, y1, z1 = min(x0 + 1, 15), min(y0 + 1, 15), min(z0 + 1, 7)
        fx, fy, fz = x - x0, y - y0, z - z0
        want = 0.0
        for iz, wz in ((z0, 1 - fz), (z1, fz)):
            for iy, wy in ((y0, 1 - fy), (y1, fy)):
                for ix, wx in ((x0, 1 - fx), (x1, fx)):
                    want += wz * wy * wx * node(side, iz, iy, ix)
        got = sc.glass_energy((x / 15) ** 2, y / 15, ior, bool(side))
        assert abs(got - want) < 2e-4, (side, x, y, z, got, want)


# ---- renders ----------------------------------------------------------------------------------------------------
def sphere_glass(rough, ior, base=(1.0, 1.0, 1.0)):
    t, n = R.icosphere(4)
    rs = R.build([R.RMesh(t, normals=n)], [R.RInstance(0, S.translation([0.0, 0.0, 0.0]), [0])],
                 [R.material(base, metallic=0.0, roughness=rough, transmission=1.0, ior=ior)])
    cam = R.Camera((0.0, 0.0, -4.0), (0.0, 0.0, 0.0), 20.0)
    return rs, cam


def sheet(rough, leaving=False, cos=math.sqrt(0.5), color=1.0):
    """glass-rough-enter.ts `sheet`: a 200 m glass sheet (ior 1.5) facing up (leaving: down, so the camera sees it
    from inside the glass), seen from above at `cos` through a 5 degree camera."""
    tris = R.quad((-100, 0, -100), (-100, 0, 100), (100, 0, 100), (100, 0, -100))  # normal +y
    if leaving:
        tris = tris[:, ::-1].copy()
    rs = R.build([R.RMesh(tris)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0])],
                 [R.material((color,) * 3, metallic=0.0, roughness=rough, transmission=1.0, ior=1.5)])
    cam = R.Camera((0.0, 3 * cos, 3 * math.sqrt(1 - cos * cos)), (0.0, 0.0, 0.0), 5.0)
    return rs, cam


def image_mean(acc):
    return (acc[..., :3] / np.maximum(acc[..., 3:4], 1)).reshape(-1, 3).mean(axis=0)


def centre_mean(acc, frac):
    c = acc[..., :3] / np.maximum(acc[..., 3:4], 1)
    h, w = c.shape[:2]
    yy, xx = np.mgrid[0:h, 0:w]
    inside = np.hypot(xx - (w - 1) / 2, yy - (h - 1) / 2) < frac * w
    return c[inside].mean(axis=0)


@pytest.mark.parametrize("leaving,cos,comp,want,tol", [
    (False, math.sqrt(0.5), 0.0, 0.9573, 0.015),   # glass-rough-enter.ts
    (True, math.sqrt(0.5), 0.0, 0.7390, 0.015),    # glass-rough-leave.ts
    (True, 0.4, 0.0, 0.6474, 0.015),               # glass-rough-leave-low.ts
    (False, math.sqrt(0.5), 1.0, 1.0, 0.02),       # glass-rough-enter-compensated.ts
    (True, math.sqrt(0.5), 1.0, 1.0, 0.02),        # glass-rough-leave-compensated.ts
])
def test_rough_interface_in_a_white_furnace(leaving, cos, comp, want, tol):
    """One rough interface (roughness 0.6) under a 1-nit sky: R + T by single scattering, or 1 with the energy
    compensation; the sky is sampled on both sides of the sheet (rough glass is two-sided), so the BSDF's value,
    density and sampling are held to each other."""
    rs, cam = sheet(0.6, leaving, cos)
    p = rs.params(16, 16, cam, max_bounces=4, rr_depth=8, sky=(1.0, 1.0, 1.0), energy_compensation=comp)
    m = image_mean(oracle_scene(rs).render(p, 0, 256))
    assert np.all(np.abs(m - want) < tol * want), (m, want)


def test_rough_sphere_white_furnace():
    """glass-rough-sphere-furnace.ts: a closed glass sphere (ior 1.5, roughness 0.5) in a 1-nit sky with energy
    compensation, 64 bounces: every interface returns what it receives, 1 +- 2 %."""
    rs, cam = sphere_glass(0.5, 1.5)
    p = rs.params(24, 24, cam, max_bounces=64, rr_depth=64, sky=(1.0, 1.0, 1.0), energy_compensation=1.0)
    m = centre_mean(oracle_scene(rs).render(p, 0, 128), 0.25)
    assert np.all(np.abs(m - 1.0) < 0.02), m


def test_total_internal_reflection_is_untinted():
    """glass-internal-reflection.ts: smooth grey glass (0.5) seen from inside at 45 degrees (past the critical
    angle, 41.8) reflects the whole sky: 1 +- 2 %."""
    rs, cam = sheet(0.0, leaving=True, color=0.5)
    p = rs.params(16, 16, cam, max_bounces=4, rr_depth=8, sky=(1.0, 1.0, 1.0))
    m = image_mean(oracle_scene(rs).render(p, 0, 64))
    assert np.all(np.abs(m - 1.0) < 0.02), m


def test_clear_slab_transmits_its_two_interfaces():
    """material-glass-transmission.ts: a clear slab (ior 1.5, 5 cm) in front of a 100-nit wall, at normal incidence:
    100 x (1 - 0.04)^2 / (1 - 0.04^2) +- 2 % (the reflections between the two interfaces counted)."""
    want = 100 * (1 - 0.04) ** 2 / (1 - 0.04 ** 2)
    wall = R.quad((-5, -5, -2), (5, -5, -2), (5, 5, -2), (-5, 5, -2))  # normal +z, toward the camera
    slab = R.box_tris((-0.5, -0.5, -0.025), (0.5, 0.5, 0.025))
    rs = R.build([R.RMesh(wall), R.RMesh(slab)],
                 [R.RInstance(0, np.eye(4, dtype=np.float32), [0]), R.RInstance(1, np.eye(4, dtype=np.float32), [1])],
                 [R.material((0, 0, 0), emissive=(100.0, 100.0, 100.0)),
                  R.material((1, 1, 1), metallic=0.0, roughness=0.0, transmission=1.0, ior=1.5)],
                 light_tree=False)
    cam = R.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 10.0)
    p = rs.params(16, 16, cam, max_bounces=16, rr_depth=16)
    m = centre_mean(oracle_scene(rs).render(p, 0, 64), 0.3)
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


def test_a_pane_tints_a_sun_shadow():
    """shadow-glass-tint.ts with a sun (1000 lux, straight down) for its point light: an albedo-0.5 Lambert plane
    under a grey pane (0.5, ior 1.5, smooth, 5 cm, at y = 1) reads 0.5 E / pi x (0.96 x 0.5)^2, each of the pane's
    two surfaces keeping the transmission lobe's share (1 - 0.04 at normal incidence) x its colour. The camera sees
    the patch from beside the pane; direct light only (one bounce)."""
    plane = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    pane = R.box_tris((-0.5, 0.975, -0.5), (0.5, 1.025, 0.5))
    rs = R.build([R.RMesh(plane), R.RMesh(pane)],
                 [R.RInstance(0, np.eye(4, dtype=np.float32), [0]), R.RInstance(1, np.eye(4, dtype=np.float32), [1])],
                 [R.material((0.5, 0.5, 0.5), specular=0.0),
                  R.material((0.5, 0.5, 0.5), metallic=0.0, roughness=0.0, transmission=1.0, ior=1.5)],
                 [R.directional_light((0.0, -1.0, 0.0), 1000.0)])
    cam = R.Camera((0.0, 0.6, 3.0), (0.0, 0.0, 0.0), 4.0)
    p = rs.params(16, 16, cam, max_bounces=1, rr_depth=8)
    m = centre_mean(oracle_scene(rs).render(p, 0, 16), 0.2)
    clear = 0.5 * 1000 / math.pi
    want = clear * (0.96 * 0.5) ** 2
    assert np.all(np.abs(m - want) < 0.02 * want), (m, want)


def test_the_instance_rows_say_glass():
    """scene.rs:127-135: an instance whose rows are all glass casts no shadow and carries INST_GLASS; glass beside
    an opaque section of a shadow-casting instance (the reference's F_IGLASS) is refused; the table is there only
    when a material has glass (gpu.rs features_of: layer.x > 0)."""
    rs, _ = sphere_glass(0.5, 1.5)
    bits = rs.prep.instances["n"][0, 1, 3].view(np.uint32)
    assert bits & R.INST_GLASS and not (bits & 1), hex(int(bits))  # UNIT_SHADOW off
    assert rs.glass_energy is not None and rs.glass_energy.shape == (1024, 4)
    t, n = R.icosphere(1)
    plain = R.build([R.RMesh(t)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0])], [R.material()])
    assert plain.glass_energy is None
    assert plain.prep.instances["n"][0, 1, 3].view(np.uint32) & 1
    sec = np.zeros(len(t), np.uint32)
    sec[::2] = 1
    with pytest.raises(ValueError, match="F_IGLASS"):
        R.build([R.RMesh(t, sections=sec)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0, 1])],
                [R.material(), R.material(transmission=1.0)])
    # the same rows on an instance that casts no shadow anyway are fine
    R.build([R.RMesh(t, sections=sec)], [R.RInstance(0, np.eye(4, dtype=np.float32), [0, 1], cast_shadow=False)],
            [R.material(), R.material(transmission=1.0)])


# ---- GPU parity ---------------------------------------------------------------------------------------------------
def glass_garden(light_tree: bool = False):
    """An open scene with every glass path: a ground plane, an opaque box, a rough tinted glass sphere (roughness
    0.35, ior 1.5), a partly transmissive sphere (transmission 0.6, roughness 0.2, ior 1.33: glass beside the diffuse
    and specular lobes), a tilted clear pane (roughness 0: the delta lobe, ior 1.45, tinted), an emissive panel of
    288 triangles (a light tree over them with light_tree), a sun and a sky."""
    t, n = R.icosphere(3)
    ground = R.quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10))
    box = R.box_tris((-0.4, 0.0, -0.4), (0.4, 0.9, 0.4))
    pane = R.box_tris((-0.7, -0.02, -0.45), (0.7, 0.02, 0.45))
    g = 12
    panel = np.concatenate([R.quad((x0, 2.2, z0), (x0 + 0.1, 2.2, z0), (x0 + 0.1, 2.2, z0 + 0.1), (x0, 2.2, z0 + 0.1))
                            for x0 in np.linspace(-0.6, 0.5, g) for z0 in np.linspace(-0.6, 0.5, g)])
    meshes = [R.RMesh(ground), R.RMesh(box), R.RMesh(t, normals=n), R.RMesh(pane), R.RMesh(panel)]
    mats = [R.material((0.6, 0.6, 0.6), specular=0.5, roughness=0.8),
            R.material((0.7, 0.25, 0.2), specular=0.5, roughness=0.4),
            R.material((0.9, 0.95, 1.0), roughness=0.35, transmission=1.0, ior=1.5),
            R.material((0.3, 0.6, 0.9), roughness=0.2, transmission=0.6, ior=1.33),
            R.material((0.8, 0.9, 0.8), roughness=0.0, transmission=1.0, ior=1.45),
            R.material((0.0, 0.0, 0.0), emissive=(6.0, 5.0, 4.0), specular=0.0)]
    inst = [R.RInstance(0, np.eye(4, dtype=np.float32), [0]),
            R.RInstance(1, S.trs((1.3, 0.0, 0.8), (0.0, 0.4, 0.0), (1, 1, 1)), [1]),
            R.RInstance(2, S.trs((-0.5, 0.6, 0.0), (0.0, 0.0, 0.0), (0.6, 0.6, 0.6)), [2]),
            R.RInstance(2, S.trs((0.7, 0.45, -0.6), (0.0, 0.0, 0.0), (0.45, 0.45, 0.45)), [3]),
            R.RInstance(3, S.trs((0.2, 1.3, -0.2), (0.5, 0.3, 0.2), (1, 1, 1)), [4]),
            R.RInstance(4, np.eye(4, dtype=np.float32), [5])]
    rs = R.build(meshes, inst, mats, [R.directional_light((-0.4, -1.0, 0.5), 3.0)], light_tree=light_tree)
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
@pytest.mark.parametrize("light_tree,comp", [(False, 0.0), (False, 1.0), (True, 1.0)])
def test_glass_scenes_match_the_oracle(device, light_tree, comp):
    """The glass garden on the megakernel against the oracle, same seed and samples: relative RMSE <= 1e-3 (the
    float operations are the oracle's in its order, so in practice the bits agree); with and without the glass
    energy compensation, and with the light tree."""
    print(f"glass garden 64x48 x 4 spp, light tree {light_tree}, compensation {comp}", flush=True)
    rs, cam = glass_garden(light_tree)
    assert (rs.tree_entries > 0) == light_tree
    p = rs.params(64, 48, cam, seed=7, max_bounces=6, rr_depth=3, sky=(0.3, 0.4, 0.6), energy_compensation=comp)
    want = oracle_scene(rs).render(p, 0, 4)
    got, _ = gpu_render(device, rs, p, 0, 4)
    same = (got.view(np.uint32) == want.view(np.uint32)).all(axis=2).mean()
    print(f"glass garden: {same:.4f} of the pixels bit-equal, rel RMSE {rel_rmse(got, want):.2e}", flush=True)
    assert rel_rmse(got, want) <= 1e-3


@pytest.mark.gpu
def test_glass_renders_on_the_megakernel_whatever_the_setting(device):
    """The wavefront does not render glass: under RENDER_WAVEFRONT and RENDER_AUTO a glass scene's launch takes the
    megakernel, and the counting launch (also the megakernel) counts its reads: the same bits each way."""
    import threert_amd as T

    rs, cam = glass_garden(False)
    p = rs.params(32, 24, cam, seed=3, max_bounces=5, rr_depth=3, sky=(0.3, 0.4, 0.6))
    outs = []
    try:
        for how in (T.RENDER_WAVEFRONT, T.RENDER_AUTO, T.RENDER_MEGAKERNEL):
            print(f"glass garden under render setting {how}", flush=True)
            device.set_render(how)
            outs.append(gpu_render(device, rs, p, 0, 2)[0])
    finally:
        device.set_render(T.RENDER_AUTO)
    print("glass garden: counting launch", flush=True)
    counted, reads = gpu_render(device, rs, p, 0, 2, reads=True)
    assert reads["nodes"] > 0 and reads["shaded"] > 0, reads
    for o in outs[1:] + [counted]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))


@pytest.mark.gpu
def test_glass_furnaces_on_the_gpu(device):
    """The rough interface and the rough sphere in a white furnace with energy compensation (glass-rough-enter-
    compensated.ts, glass-rough-sphere-furnace.ts) on the GPU: 1 +- 2 %."""
    rs, cam = sheet(0.6, False)
    p = rs.params(32, 32, cam, max_bounces=4, rr_depth=8, sky=(1.0, 1.0, 1.0), energy_compensation=1.0)
    m = image_mean(gpu_render(device, rs, p, 0, 256)[0])
    assert np.all(np.abs(m - 1.0) < 0.02), m
    rs, cam = sphere_glass(0.5, 1.5)
    p = rs.params(48, 48, cam, max_bounces=64, rr_depth=64, sky=(1.0, 1.0, 1.0), energy_compensation=1.0)
    m = centre_mean(gpu_render(device, rs, p, 0, 256)[0], 0.25)
    assert np.all(np.abs(m - 1.0) < 0.02), m
