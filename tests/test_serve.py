"""Continuous batching: a job's image is the same alone and in a mix (bit for bit), and the serving workload's
shape (512 Poisson jobs, log-uniform scenes, resolutions and samples).

The reference's `serve` loop (main.rs:345-398) renders one job at a time, so it has no mixed result to compare
with; the bar is the lone render of the same job through trt_render, which the render tests hold to the oracle."""
import numpy as np
import pytest

from threert_amd import serve


def test_the_workload_is_seeded_and_in_range():
    a = serve.workload(n_jobs=512, rate=40.0, seed=1)
    b = serve.workload(n_jobs=512, rate=40.0, seed=1)
    assert [(j.scene, j.width, j.height, j.spp, j.bounces, j.arrival) for j in a] == \
           [(j.scene, j.width, j.height, j.spp, j.bounces, j.arrival) for j in b]
    t = np.array([j.arrival for j in a])
    assert len(a) == 512 and t[0] == 0.0 and (np.diff(t) >= 0).all()
    # Poisson at 40/s: 511 gaps of mean 1/40 s, so the span is 12.8 s within a few standard deviations
    assert abs(t[-1] - 511 / 40) < 4 * np.sqrt(511) / 40
    assert all(256 <= j.width <= 1920 and 256 <= j.height <= 1080 for j in a)
    assert all(1 <= j.spp <= 256 and 2 <= j.bounces <= 6 for j in a)
    sizes = {j.scene[0] for j in a}
    assert min(sizes) >= 10_000 and max(sizes) <= 4_000_000 and len({j.scene for j in a}) <= 24


def test_equal_shares_and_the_last_slice():
    class A:
        def __init__(self, w, h, spp, done, arrival=0.0):
            self.job = serve.Job(0, (1, 1), w, h, spp, 2, arrival)
            self.done = done

    s = serve.Server.__new__(serve.Server)
    s.launch_samples = 1000.0
    s.policy = "equal"
    # two jobs share 1000 samples: 500 each; a 10x10 job takes 5 samples, a 100x100 job at least one
    assert s._slices([A(10, 10, 64, 0), A(100, 100, 64, 0)], 1.0) == [5, 1]
    # never more than what is left
    assert s._slices([A(10, 10, 64, 62)], 1.0) == [2]
    # by age: a job waiting 3 s takes three times the share of one waiting 1 s
    s.policy = "age"
    assert s._slices([A(10, 10, 64, 0, 9.0), A(10, 10, 64, 0, 7.0)], 10.0) == [2, 7]
    with pytest.raises(ValueError):
        serve.Server(None, None, policy="shortest")


def test_percentiles_and_summary():
    r = [serve.Result(k, 0.0, 0.0, float(k + 1), 10) for k in range(100)]
    s = serve.summary(r, 10.0)
    assert s["p50_s"] == pytest.approx(50.5) and s["p99_s"] == pytest.approx(99.01)
    assert s["samples"] == 1000 and s["samples_per_s"] == 100.0


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.gpu
def test_a_job_renders_the_same_alone_and_in_a_batch(device):
    from threert_amd import render as R

    print("batch of 4 jobs over 3 scenes (one with clearcoats) vs each alone", flush=True)
    rs1, cam1 = R.cornell()
    rs2, cam2, sky2 = serve.soup_scene(4000, 5)
    rs3, _ = R.cornell()
    rs3.materials = rs3.materials.copy()
    rs3.materials[:, 30], rs3.materials[:, 31] = 0.7, 0.2
    s1, s2, s3 = R.upload(device, rs1), R.upload(device, rs2), R.upload(device, rs3)
    # a scene with a coat puts the whole launch on the kernel with the coat's terms: the others keep their bits
    jobs = [(s1, rs1.params(40, 24, cam1, seed=3, max_bounces=4, rr_depth=3), 0, 3),
            (s2, rs2.params(33, 17, cam2, seed=9, max_bounces=2, rr_depth=3, sky=sky2), 0, 2),
            (s1, rs1.params(16, 16, cam1, seed=5, max_bounces=3, rr_depth=2), 4, 5),
            (s3, rs3.params(24, 24, cam1, seed=2, max_bounces=3, rr_depth=2), 0, 2)]
    alone = []
    for sc, p, first, spp in jobs:
        acc = R.Accum(device, int(p[16]), int(p[17]))
        R.render(sc, p, first, spp, acc)
        device.sync()
        alone.append(acc.read())
        acc.close()
    accs = [R.Accum(device, int(p[16]), int(p[17])) for _, p, _, _ in jobs]
    serve.render_batch(device, [(sc, p, first, spp, a) for (sc, p, first, spp), a in zip(jobs, accs)])
    device.sync()
    for k, (a, want) in enumerate(zip(accs, alone)):
        got = a.read()
        assert (_bits(got) == _bits(want)).all(), f"job {k} differs in the batch"
        a.close()
    s1.close()
    s2.close()
    s3.close()


@pytest.mark.gpu
def test_the_server_gives_each_job_its_lone_image(device):
    from threert_amd import render as R

    jobs = serve.workload(n_jobs=8, rate=400.0, seed=4, n_scenes=3, min_tris=2000, max_tris=20000, scale=0.06)
    print("server: 8 jobs", [(j.width, j.height, j.spp, j.bounces) for j in jobs], flush=True)
    srv = serve.Server(device, serve.soup_scene, launch_samples=3000, keep_images=True, workers=2)
    res = srv.run(jobs)
    srv.close()
    assert [r.id for r in res] == [j.id for j in jobs] and srv.launches > 1
    for j, r in zip(jobs, res):
        rs, cam, sky = serve.soup_scene(*j.scene)
        e = serve._Resident(None, rs, cam, sky, 0)
        sc = R.upload(device, rs)
        acc = R.Accum(device, j.width, j.height)
        R.render(sc, serve.params_of(e, j), 0, j.spp, acc)
        device.sync()
        want = acc.resolve()[0]
        assert (_bits(r.image) == _bits(want)).all(), f"job {j.id} differs from its lone render"
        assert r.finished >= r.admitted >= r.arrival
        acc.close()
        sc.close()
    s = serve.summary(res, max(r.finished for r in res))
    assert s["jobs"] == 8 and s["samples"] == sum(j.width * j.height * j.spp for j in jobs)


@pytest.mark.gpu
def test_a_full_scale_slice_of_the_workload_gives_each_job_its_lone_image(device):
    """24 jobs of the serving workload at full scale (seed 1): every job on the largest scene (3.56M triangles), the
    largest image, one more job set to 1920 x 1080, and the first jobs on four more scenes; samples capped at 4 (the
    check is the mix, not convergence). Arriving together, they share launches; each job's image is its lone
    render's, bit for bit."""
    import dataclasses

    from threert_amd import render as R

    jobs = serve.workload(n_jobs=512, rate=40.0, seed=1)
    biggest = max(j.scene[0] for j in jobs)
    assert biggest >= 3_000_000
    pick = [j for j in jobs if j.scene[0] == biggest][:6]
    pick.append(max(jobs, key=lambda j: j.width * j.height))
    pick.append(dataclasses.replace(jobs[0], width=1920, height=1080))
    others = sorted({j.scene for j in jobs if j.scene[0] != biggest}, key=lambda sc: sc[0])[:4]
    for sc in others:
        pick += [j for j in jobs if j.scene == sc][:4]
    pick = pick[:24]
    assert len(pick) == 24
    pick = [dataclasses.replace(j, id=k, spp=min(j.spp, 4), arrival=0.001 * k) for k, j in enumerate(pick)]
    print("slice:", [(j.scene[0], j.width, j.height, j.spp, j.bounces) for j in pick], flush=True)
    srv = serve.Server(device, serve.soup_scene, keep_images=True, workers=4)
    res = srv.run(pick)
    srv.close()
    assert [r.id for r in res] == [j.id for j in pick] and srv.launches > 1
    print(f"slice: {srv.launches} launches", flush=True)
    scenes = {}
    for j, r in zip(pick, res):
        if j.scene not in scenes:
            rs, cam, sky = serve.soup_scene(*j.scene)
            scenes[j.scene] = (serve._Resident(None, rs, cam, sky, 0), R.upload(device, rs))
        e, sc = scenes[j.scene]
        acc = R.Accum(device, j.width, j.height)
        R.render(sc, serve.params_of(e, j), 0, j.spp, acc)
        device.sync()
        want = acc.resolve()[0]
        acc.close()
        assert (_bits(r.image) == _bits(want)).all(), f"job {j.id} ({j.scene}, {j.width}x{j.height}) differs"
    for _, sc in scenes.values():
        sc.close()


def test_equal_error_plan_never_raises_a_jobs_error():
    """The denoising server's plan: the fewest calibrated samples whose denoised error is at most the job's plain
    error, else the job as asked; no job's calibrated error goes up."""
    from threert_amd import equal_error as E

    plain = {k: 1.0 / np.sqrt(k) for k in E.CHECKPOINTS}
    den = {k: 0.3 / k ** 0.4 for k in E.CHECKPOINTS}
    cal = {"plain": plain, "denoised": den}
    jobs = serve.workload(n_jobs=64, rate=40.0, seed=3)
    c = {j.scene: cal for j in jobs}
    planned = E.plan(jobs, c)
    for j, d in zip(jobs, planned):
        target = E.plain_error(cal, j.spp)
        assert E.job_error(cal, d) <= target + 1e-12
        if d.denoise:
            assert d.spp < j.spp and (d.spp == 1 or den[d.spp // 2] > target)
        else:
            assert d == j and all(den[k] > target for k in E.CHECKPOINTS if k < j.spp)
    assert any(d.denoise for d in planned) and E.plain_error(cal, 4) == pytest.approx(0.5)
    assert E.plain_error(cal, 1024) == pytest.approx(plain[256] / 2)


@pytest.mark.gpu
def test_the_device_resolve_and_denoise_match_the_host_path(device):
    """trt_accum_resolve_device gives trt_accum_resolve's means, edge-padded; and the server's denoising step (all on
    the device) gives the same image as resolving to the host and denoising there."""
    from threert_amd import render as R
    from threert_amd.unet import UNet, shipped_params

    rs, cam = R.cornell()
    w, h, pw, ph = 72, 40, 80, 48
    sc = R.upload(device, rs)
    acc = R.Accum(device, w, h)
    R.render(sc, rs.params(w, h, cam, seed=1), 0, 2, acc)
    c, a, n, _ = acc.resolve()
    bufs = [device.alloc(12 * pw * ph) for _ in range(3)]
    acc.resolve_device(pw, ph, *bufs)
    device.sync()
    for b, want in zip(bufs, (c, a, n)):
        got = b.download(np.float32, 3 * pw * ph).reshape(ph, pw, 3)
        assert (got == np.pad(want, ((0, ph - h), (0, pw - w), (0, 0)), mode="edge")).all()
        b.free()
    net = UNet(device, pw * ph, shipped_params())
    host = serve.denoise_image(net, c, a, n)
    srv = serve.Server(device, lambda *k: None, unet=net, keep_images=True)

    class Active:
        job = serve.Job(0, (1, 1), w, h, 2, 2, denoise=True)

    Active.acc = acc
    on_device = srv._denoise(Active)
    assert on_device.shape == host.shape and (on_device == host).all()
    srv.close()
    net.close()
    acc.close()
    sc.close()
