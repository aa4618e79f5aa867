"""The wavefront path tracer (trt_device_set_render(TRT_RENDER_WAVEFRONT)) against the megakernel and the oracle.

The wavefront runs the same path loop cut at its traces: every bounce one walk launch traces the queued rays (each
path's next closest ray and the previous bounce's shadow rays), and a shade launch adds the visible direct light, in
the order the megakernel adds it, before it shades the hit. So its accumulator must be the megakernel's bit for bit,
and the oracle's where the megakernel's is: the same scenes as the megakernel's parity tests (suns, emitters through
the guide table and through the light tree, clearcoats, a thin lens, exact ties, the 66M-triangle city at 1280 x 720),
a frame larger than one chunk of paths (4,194,304), batches of jobs, and the ray counts."""
import sys

import numpy as np
import pytest

import threert_amd as T
from threert_amd import render as R
from threert_amd import scene as S

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import test_render as TR  # noqa: E402


def render_as(device, how, sc, p, first, spp, acc=None):
    """One render of `p` with the device's render set to `how`, into a new accumulator (or `acc`); the setting is put
    back to the default afterwards."""
    own = acc is None
    if own:
        acc = R.Accum(device, int(p[16]), int(p[17]))
    device.set_render(how)
    try:
        R.render(sc, p, first, spp, acc)
        out = acc.read()
    finally:
        device.set_render(T.RENDER_AUTO)
    if own:
        acc.close()
    return out


def both(device, rs, p, first, spp):
    sc = R.upload(device, rs)
    device.render_ray_counts()
    mega = render_as(device, T.RENDER_MEGAKERNEL, sc, p, first, spp)
    mega_rays = device.render_ray_counts()
    wave = render_as(device, T.RENDER_WAVEFRONT, sc, p, first, spp)
    wave_rays = device.render_ray_counts()
    sc.close()
    return mega, wave, mega_rays, wave_rays


def equal(a, b) -> float:
    return float((a.view(np.uint32) == b.view(np.uint32)).all(axis=2).mean())


def lit_floor():
    import test_light_tree as LT

    return LT.lit_floor()


def tie_scene():
    tris, _ = S.clustered_soup(2000, 12, clusters=4, extent=10.0, median_area=0.05)
    plain = R.RMesh(tris, sections=np.zeros(len(tris), np.uint32))
    far = np.array([[[8.0, 8.0, -12.0], [8.5, 8.0, -12.0], [8.0, 8.5, -12.0]]], np.float32)
    glowing = R.RMesh(np.concatenate([tris, far]), sections=np.ones(len(tris) + 1, np.uint32))
    eye = np.eye(4, dtype=np.float32)
    rs = R.build([plain, glowing], [R.RInstance(0, eye, [0, 1])] + [R.RInstance(1, eye, [0, 1]) for _ in range(3)],
                 [R.material((0.7, 0.7, 0.7)), R.material((0.0, 0.0, 0.0), (40.0, 40.0, 40.0))],
                 [R.directional_light((-0.3, -1.0, -0.2), (3.0, 3.0, 3.0))])
    cam = R.Camera((0.0, 4.0, -16.0), (0.0, 0.0, 0.0), 50.0)
    return rs, rs.params(64, 64, cam, seed=3, max_bounces=3, rr_depth=2, sky=(0.2, 0.3, 0.4))


def case(which):
    if which == "cornell":
        rs, cam = R.cornell()
        return rs, rs.params(96, 96, cam, seed=1, max_bounces=4, rr_depth=3), 8
    if which == "lens":
        rs, cam = R.cornell()
        return rs, rs.params(96, 96, cam, seed=3, max_bounces=4, rr_depth=3, lens_radius=0.06, focus=1.6), 8
    if which == "coat":
        rs, cam = R.cornell()
        mats = rs.materials.copy()
        for k, (amount, rough) in zip(range(len(mats)), ((1.0, 0.05), (0.5, 0.3), (1.0, 0.6), (0.8, 0.0))):
            mats[k, 30], mats[k, 31] = amount, rough
        rs.materials = mats
        return rs, rs.params(64, 64, cam, seed=4, max_bounces=4, rr_depth=3), 8
    if which == "suns":
        rs, cam, sky = R.city(seed=3, meshes=6, instances=40, extent=60.0)
        return rs, rs.params(80, 48, cam, seed=7, max_bounces=3, rr_depth=2, sky=sky), 2
    if which == "tree":
        rs, cam = lit_floor()
        return rs, rs.params(96, 64, cam, seed=5, max_bounces=3, rr_depth=2), 4
    if which == "ties":
        rs, p = tie_scene()
        return rs, p, 4
    raise ValueError(which)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["cornell", "lens", "coat", "suns", "tree", "ties"])
def test_the_wavefront_is_the_megakernel_and_the_oracle(device, which):
    print(f"wavefront: {which}", flush=True)
    rs, p, spp = case(which)
    mega, wave, mega_rays, wave_rays = both(device, rs, p, 0, spp)
    print(f"{which}: {equal(wave, mega):.4f} of the pixels equal the megakernel's; rays {wave_rays} vs {mega_rays}",
          flush=True)
    assert equal(wave, mega) == 1.0
    assert wave_rays == mega_rays
    want = TR.oracle_scene(rs).render(p, 0, spp, threads=8)
    print(f"{which}: {equal(wave, want):.4f} of the pixels equal the oracle's, rel RMSE {TR.rel_rmse(wave, want):.2e}",
          flush=True)
    assert TR.rel_rmse(wave, want) <= 1e-3
    assert wave[..., :3].sum() > 0


@pytest.mark.gpu
def test_the_full_city_at_720p(device):
    """The instanced-city workload at its size (64 meshes x 4096 instances, 66M instanced triangles, 1280 x 720,
    1 spp, 3 bounces): the wavefront's frame is the megakernel's, bit for bit."""
    print("wavefront: city 1280x720", flush=True)
    rs, cam, sky = R.city(seed=1)
    p = rs.params(1280, 720, cam, seed=1, max_bounces=3, rr_depth=3, sky=sky)
    mega, wave, mega_rays, wave_rays = both(device, rs, p, 0, 1)
    print(f"city: {equal(wave, mega):.6f} equal; rays {wave_rays} vs {mega_rays}", flush=True)
    assert equal(wave, mega) == 1.0 and wave_rays == mega_rays


@pytest.mark.gpu
def test_a_frame_larger_than_a_chunk(device):
    """2400 x 1800 pixels (4,320,000, more than a chunk's 4,194,304 paths) at 3 samples, after 2 samples already in
    the accumulator: the chunks split each sample into pixel ranges, and a pixel's samples still add in order."""
    print("wavefront: cornell 2400x1800 x 3 spp", flush=True)
    rs, cam = R.cornell()
    p = rs.params(2400, 1800, cam, seed=2, max_bounces=2, rr_depth=1)
    sc = R.upload(device, rs)
    out = []
    for how in (T.RENDER_MEGAKERNEL, T.RENDER_WAVEFRONT):
        acc = R.Accum(device, 2400, 1800)
        render_as(device, T.RENDER_MEGAKERNEL, sc, p, 0, 2, acc)
        out.append(render_as(device, how, sc, p, 2, 3, acc))
        acc.close()
    sc.close()
    print(f"2400x1800: {equal(out[0], out[1]):.6f} equal", flush=True)
    assert equal(out[0], out[1]) == 1.0
    assert (out[1][..., 3] == 5).all()


@pytest.mark.gpu
def test_batches(device):
    """trt_render_batch under the wavefront: three jobs of one scene (two sizes, sample counts and seeds), then the
    same jobs with one of a second scene between them (the chunks take one scene at a time), each equal to its job
    rendered alone by the megakernel."""
    from threert_amd import serve

    print("wavefront: batches", flush=True)
    rs, cam = R.cornell()
    rs2, cam2 = lit_floor()
    sc, sc2 = R.upload(device, rs), R.upload(device, rs2)
    specs = [(sc, rs.params(64, 48, cam, seed=1), 0, 3), (sc, rs.params(130, 70, cam, seed=9), 5, 2),
             (sc, rs.params(33, 17, cam, seed=4, max_bounces=6, rr_depth=2), 1, 7),
             (sc2, rs2.params(40, 30, cam2, seed=2, max_bounces=3), 0, 2)]
    alone = [render_as(device, T.RENDER_MEGAKERNEL, s, p, f, n) for s, p, f, n in specs]
    for order in ((0, 1, 2), (0, 3, 1, 2)):
        jobs = [specs[k] for k in order]
        accs = [R.Accum(device, int(p[16]), int(p[17])) for _, p, _, _ in jobs]
        device.set_render(T.RENDER_WAVEFRONT)
        try:
            serve.render_batch(device, [(s, p, f, n, a) for (s, p, f, n), a in zip(jobs, accs)])
            got = [a.read() for a in accs]
        finally:
            device.set_render(T.RENDER_AUTO)
        for a in accs:
            a.close()
        for k, g in zip(order, got):
            assert equal(g, alone[k]) == 1.0, (order, k)
    sc.close()
    sc2.close()


@pytest.mark.gpu
def test_the_render_setting_is_checked(device):
    with pytest.raises(T.ThreeRTError):
        device.set_render(7)
    device.set_render(T.RENDER_AUTO)
