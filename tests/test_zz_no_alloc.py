"""The trace entry that allocates nothing (trt_trace_rays_device_noalloc): the same hits as trt_trace_rays_device,
through the C ABI, on two calls in a row, the second with a smaller batch; sizes past the limit refused."""
import sys

import numpy as np
import pytest

import oracle
import threert_amd as T
from threert_amd import scene as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a: np.ndarray, b: np.ndarray) -> bool:
    return bool((a.view(np.uint32).reshape(len(a), 6) == b.view(np.uint32).reshape(len(b), 6)).all())


def test_noalloc_trace_gives_the_allocating_entrys_hits(device):
    prep, lo, hi = S.precision_scene(200_000, seed=2)
    rays = S.incoherent_rays(300_000, lo, hi, seed=3)
    sc = T.Scene(device, *prep.desc())
    d_rays = device.alloc(rays.nbytes)
    d_rays.upload(rays)
    d_want, d_got = device.alloc(len(rays) * 24), device.alloc(len(rays) * 24)
    for walk in (T.WALK_WIDE, T.WALK_BINARY):
        for n in (len(rays), 70_001):  # the second call smaller than the first
            say(f"noalloc: walk {walk}, {n} rays")
            sc.trace_device(d_rays, n, d_want, walk)
            d_got.upload(np.full(len(rays) * 6, 0xdeadbeef, np.uint32))  # what the call must overwrite
            sc.trace_device_noalloc(d_rays, n, d_got, walk)
            device.sync()
            want = d_want.download(T.HIT, n)
            got = d_got.download(T.HIT, len(rays))
            assert same(got[:n], want), (walk, n)
            # nothing past the batch is written
            assert (got[n:].view(np.uint32) == 0xdeadbeef).all()
    # and a sample against the oracle's walk
    pick = np.arange(0, 70_001, 7)
    want, _ = oracle.Scene(*prep.desc()).trace_batch(rays[pick].view(oracle.RAY), oracle.WALK_EXACT)
    assert same(got[pick], want)
    with pytest.raises(T.ThreeRTError) as e:
        sc.trace_device_noalloc(d_rays, 1 << 31, d_got)
    assert e.value.code == T.TRT_E_INVALID and "2^31" in str(e.value)
    for b in (d_rays, d_want, d_got):
        b.free()
    sc.close()
