"""The `serve` wire format (main.rs:345-398): job lines in, a ready line and one reply a job out; refused jobs get
`ok: false` with the reason (and "unsupported" for what the library does not render, as tools/test-render.ts reads
it); on the GPU, each job's HDR file holds the image a lone render of it gives, bit for bit."""
import io
import json

import numpy as np
import pytest

from threert_amd import wire as W


def test_a_job_line_and_its_defaults():
    jid, job, outs = W.parse_job('{"id": "a7", "scene": {"kind": "cornell"}, "hdr": "x.raw"}', 3)
    assert jid == "a7" and job.id == 3 and job.scene == ("cornell",)
    assert (job.width, job.height, job.spp, job.bounces, job.seed) == (1280, 720, 64, 4, 1)
    assert outs == {"hdr": "x.raw"} and job.keep and not job.denoise and job.camera is None
    _, job, outs = W.parse_job(json.dumps({"id": 1, "scene": {"kind": "soup", "triangles": 5000, "seed": 9},
                                           "spp": 8, "width": 64, "height": 32.0, "bounces": 2, "seed": 77,
                                           "camera": {"position": [0, 1, -5], "target": [0, 0, 0], "fov": 30}}), 0)
    assert job.scene == ("soup", 5000, 9) and (job.width, job.height, job.spp, job.bounces, job.seed) == (64, 32, 8, 2, 77)
    assert job.camera.position == (0.0, 1.0, -5.0) and job.camera.fov_deg == 30.0 and outs == {} and not job.keep
    _, job, _ = W.parse_job('{"scene": {"kind": "city", "seed": 4, "meshes": 8, "instances": 100}}', 0)
    assert job.scene == ("city", 4, 8, 100)


@pytest.mark.parametrize("line, says", [
    ("not json", "not JSON"),
    ("[1, 2]", "JSON object"),
    ('{"id": 1}', "scene"),
    ('{"id": 1, "game": "g.js", "spp": 4}', "unsupported"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "live": 30}', "unsupported"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "overrides": {"a": 1}}', "unsupported"),
    ('{"id": 1, "scene": {"kind": "teapot"}}', "unknown scene"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "spp": 0}', "spp must be in"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "spp": 2.5}', "whole number"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "width": true}', "whole number"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "height": 20000}', "height must be in"),
    ('{"id": 1, "scene": {"kind": "soup", "triangles": 0}}', "triangles must be in"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "camera": {"position": [0, 0], "target": [0, 0, 1]}}', "position"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "camera": {"position": [1, 1, 1], "target": [1, 1, 1]}}', "position =="),
    ('{"id": 1, "scene": {"kind": "cornell"}, "camera": {"position": [0, 0, 0], "target": [0, 0, 1], "fov": 180}}', "fov"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "denoise": "yes"}', "denoise"),
    ('{"id": 1, "scene": {"kind": "cornell"}, "hdr": 5}', "hdr must be"),
])
def test_refused_jobs_say_why(line, says):
    with pytest.raises(W.JobError) as e:
        W.parse_job(line, 0)
    assert says in str(e.value)


def test_the_reader_replies_to_refused_lines_and_queues_the_rest():
    lines = "\n".join([
        '{"id": 1, "scene": {"kind": "cornell"}, "spp": 2, "width": 8, "height": 8}',
        "",
        '{"id": 2, "game": "bundle.js"}',
        '{"id": 3, "scene": {"kind": "cornell"}, "denoise": true}',
        '{"id": 4, "scene": {"kind": "soup", "triangles": 100}, "spp": 1}',
        "{oops",
    ]) + "\n"
    out = io.StringIO()
    ls = W.LineServer(None, io.StringIO(lines), out, workers=1)
    ls._read()
    replies = [json.loads(x) for x in out.getvalue().splitlines()]
    assert [(r["id"], r["ok"]) for r in replies] == [(2, False), (3, False), (None, False)]
    assert replies[0]["error"].startswith("unsupported") and replies[1]["error"].startswith("unsupported")
    jobs = ls._arrivals(0.0)
    assert [j.id for j in jobs] == [0, 1] and [ls.jobs[j.id][0] for j in jobs] == [1, 4]
    assert all(j.arrival >= 0.0 for j in jobs) and ls._closed()
    ls.close()


@pytest.mark.gpu
def test_jobs_over_the_wire_are_their_lone_renders(device, tmp_path):
    from threert_amd import render as R
    from threert_amd import serve as V

    print("wire: 3 jobs (Cornell with hdr + png, a refused line, a soup with a camera)", flush=True)
    jobs = [
        {"id": "c", "scene": {"kind": "cornell"}, "spp": 5, "width": 48, "height": 40, "bounces": 3, "seed": 7,
         "hdr": str(tmp_path / "c.raw"), "out": str(tmp_path / "c.png")},
        {"id": "bad", "scene": {"kind": "cornell"}, "spp": -1},
        {"id": 9, "scene": {"kind": "soup", "triangles": 3000, "seed": 2}, "spp": 3, "width": 33, "height": 17,
         "bounces": 2, "camera": {"position": [10.0, 60.0, -120.0], "target": [0.0, 0.0, 0.0], "fov": 35},
         "hdr": str(tmp_path / "s.raw")},
    ]
    out = io.StringIO()
    ls = W.LineServer(device, io.StringIO("\n".join(json.dumps(j) for j in jobs) + "\n"), out, workers=2,
                      launch_samples=2000)
    ls.run(0.0)
    ls.close()
    replies = [json.loads(x) for x in out.getvalue().splitlines()]
    print(replies, flush=True)
    assert replies[0]["ready"] is True and len(replies) == 4
    by = {r["id"]: r for r in replies[1:]}
    assert by["bad"]["ok"] is False and "spp" in by["bad"]["error"]
    for j in (jobs[0], jobs[2]):
        r = by[j["id"]]
        assert r["ok"], r
        st = r["stats"]
        assert (st["width"], st["height"], st["spp"]) == (j["width"], j["height"], j["spp"]) and st["samples_per_sec"] > 0
        _, job, _ = W.parse_job(json.dumps(j), 0)
        rs, cam, sky = W.make_scene(*job.scene)
        sc = R.upload(device, rs)
        acc = R.Accum(device, job.width, job.height)
        R.render(sc, V.params_of(V._Resident(None, rs, cam, sky, 0), job), 0, job.spp, acc)
        device.sync()
        want = acc.resolve()[0]
        got = np.fromfile(j["hdr"], "<f4").reshape(job.height, job.width, 3)
        assert (got.view(np.uint32) == want.view(np.uint32)).all(), f"job {j['id']} differs from its lone render"
        acc.close()
        sc.close()
    from threert_amd import output as O

    assert (O.read_rgb8(jobs[0]["out"]) == O.srgb8(np.fromfile(jobs[0]["hdr"], "<f4").reshape(40, 48, 3))).all()


@pytest.mark.gpu
def test_the_command_line_server_answers_on_stdout(tmp_path):
    """`python -m threert_amd.wire` as a client starts it: the ready line, then one reply a job in order for a client
    that sends one job at a time, then a clean exit when its input ends."""
    import subprocess
    import sys
    import threading

    print("wire: command-line server, 2 jobs and a refused one", flush=True)
    p = subprocess.Popen([sys.executable, "-m", "threert_amd.wire", "--budget-gb", "4"], stdin=subprocess.PIPE,
                         stdout=subprocess.PIPE, text=True, bufsize=1)
    watchdog = threading.Timer(120.0, p.kill)  # a server that stops answering ends the test, not the run
    watchdog.start()
    try:
        ready = json.loads(p.stdout.readline())
        assert ready["ready"] is True and "gfx" in ready["adapter"]
        replies = []
        for job in ({"id": 1, "scene": {"kind": "cornell"}, "spp": 2, "width": 32, "height": 24,
                     "hdr": str(tmp_path / "a.raw")},
                    {"id": 2, "game": "bundle.js"},
                    {"id": 3, "scene": {"kind": "soup", "triangles": 2000, "seed": 4}, "spp": 1, "width": 16,
                     "height": 16}):
            p.stdin.write(json.dumps(job) + "\n")
            p.stdin.flush()
            replies.append(json.loads(p.stdout.readline()))
        p.stdin.close()
        assert p.wait(timeout=60) == 0
    finally:
        watchdog.cancel()
        if p.poll() is None:
            p.kill()
    assert [(r["id"], r["ok"]) for r in replies] == [(1, True), (2, False), (3, True)]
    assert replies[1]["error"].startswith("unsupported")
    assert replies[0]["stats"]["triangles"] == 36 and (tmp_path / "a.raw").stat().st_size == 32 * 24 * 12


@pytest.mark.gpu
def test_a_denoise_job_over_the_wire_is_the_servers_denoised_image(device, tmp_path):
    """`denoise: true` sends the job's image through the server's U-Net: the HDR file holds what the serving
    denoising step gives for the same accumulator (device resolve into edge-padded buffers, the network), cropped."""
    from threert_amd import render as R
    from threert_amd import serve as V
    from threert_amd.unet import UNet, shipped_params

    print("wire: a denoise job", flush=True)
    job = {"id": 5, "scene": {"kind": "cornell"}, "spp": 2, "width": 40, "height": 24, "seed": 3, "denoise": True,
           "hdr": str(tmp_path / "d.raw")}
    net = UNet(device, 1920 * 1088, shipped_params())
    out = io.StringIO()
    ls = W.LineServer(device, io.StringIO(json.dumps(job) + "\n"), out, unet=net, workers=1)
    ls.run(0.0)
    ls.close()
    replies = [json.loads(x) for x in out.getvalue().splitlines()]
    assert replies[1]["ok"] and replies[1]["stats"]["denoise"] == "unet", replies
    got = np.fromfile(job["hdr"], "<f4").reshape(24, 40, 3)
    _, j, _ = W.parse_job(json.dumps(job), 0)
    rs, cam, sky = W.make_scene(*j.scene)
    sc = R.upload(device, rs)
    acc = R.Accum(device, 40, 24)
    R.render(sc, V.params_of(V._Resident(None, rs, cam, sky, 0), j), 0, 2, acc)
    c, a, n, _ = acc.resolve()
    want = V.denoise_image(net, c, a, n)
    assert got.shape == want.shape and (got == want).all()
    acc.close()
    sc.close()
    net.close()
