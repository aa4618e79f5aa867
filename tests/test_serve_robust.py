"""The serving loop's budget and failure handling, on the CPU with the device calls replaced by fakes: the workers'
uploads never pass the budget (the rest upload at admission); an upload or a launch that fails fails its jobs and
the server goes on serving."""
import threading

import pytest

from threert_amd import render as R
from threert_amd import serve


class FakeScene:
    live = 0
    lock = threading.Lock()

    def __init__(self, rs):
        with FakeScene.lock:
            FakeScene.live += 1
        self.open = True

    def close(self):
        if self.open:
            with FakeScene.lock:
                FakeScene.live -= 1
            self.open = False


class FakeAccum:
    def __init__(self, dev, w, h):
        self.w, self.h = w, h

    def close(self):
        pass


class FakeDev:
    def sync(self):
        pass


@pytest.fixture
def fakes(monkeypatch):
    FakeScene.live = 0
    failing = set()

    def upload(dev, rs):
        if getattr(rs, "tag", None) in failing:
            raise RuntimeError(f"upload of {rs.tag} failed")
        return FakeScene(rs)

    monkeypatch.setattr(R, "upload", upload)
    monkeypatch.setattr(R, "Accum", FakeAccum)
    return failing


_cornell = None


def make_scene(tag):
    global _cornell
    if _cornell is None:
        _cornell = R.cornell()
    rs, cam = _cornell
    rs = R.RenderScene(**{k: getattr(rs, k) for k in rs.__dataclass_fields__})
    rs.tag = tag
    return rs, cam, (0.2, 0.3, 0.4)


def test_the_workers_upload_within_the_budget(fakes):
    one = serve.scene_bytes(make_scene("a")[0])
    store = serve.SceneStore(FakeDev(), make_scene, workers=4, headroom=lambda: int(2.5 * one))
    for k in "abcd":
        store.prefetch((k,))
    for k in "abcd":
        store.host[(k,)].result()
    assert FakeScene.live == 2 and store.pending == 2 * one  # two fit; two wait for their admission
    uploaded = [k for k in "abcd" if store.host[(k,)].result()[3] is not None]
    late = [k for k in "abcd" if k not in uploaded][0]
    e = store.acquire((late,))  # uploads now
    assert FakeScene.live == 3 and store.pending == 2 * one and store.resident
    assert store.pending_except((uploaded[0],)) == one
    assert store.drop_uploads(1, keep=(uploaded[0],)) == one  # frees the other worker upload, not the kept one
    assert FakeScene.live == 2 and store.pending == one
    e2 = store.acquire((uploaded[1],))  # dropped: uploaded again
    assert e2.scene.open and FakeScene.live == 3
    store.release(e)
    store.release(e2)
    store.close()
    assert FakeScene.live == 0


def run_jobs(srv, jobs):
    results = []
    pending = list(jobs)

    def arrivals(t):
        out = list(pending)
        pending.clear()
        return out

    srv.serve(arrivals, lambda: not pending, results.append, lambda t: None)
    return {r.id: r for r in results}


def test_failed_uploads_and_launches_fail_their_jobs_only(fakes, monkeypatch):
    calls = []

    def render_batch(dev, jobs):
        calls.append(len(jobs))
        if len(calls) == 1:
            raise RuntimeError("walk stopped and reported")

    monkeypatch.setattr(serve, "render_batch", render_batch)
    fakes.add("bad")
    srv = serve.Server(FakeDev(), make_scene, budget_bytes=1 << 40, workers=2)
    srv.store.headroom = lambda: 0  # no worker uploads: every scene uploads at admission
    first = [serve.Job(0, ("good",), 16, 16, 2, 2), serve.Job(1, ("bad",), 16, 16, 2, 2)]
    got = run_jobs(srv, first)
    assert got[1].error and "upload of bad failed" in got[1].error  # its upload failed at admission
    assert got[0].error and "walk stopped" in got[0].error        # the first launch failed
    more = [serve.Job(2, ("good",), 16, 16, 3, 2), serve.Job(3, ("other",), 8, 8, 1, 2)]
    got = run_jobs(srv, more)
    assert got[2].error is None and got[3].error is None and got[2].samples == 3 * 16 * 16
    assert srv.used == sum(e.bytes for e in srv.store.resident.values())  # only the resident scenes remain counted
    srv.close()
    assert FakeScene.live == 0
