"""The C ABI: libthreert.so loads without a GPU and exports every entry point include/threert.h declares; the
oracle library loads too. No compute call is made here, except in the GPU-marked copy test."""
import ctypes
import os
import re

import pytest

import oracle
import threert_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "threert.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(trt_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(T.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(T.EXPORTS) == names


def test_the_structs_match_the_header_sizes():
    assert T.NODE.itemsize == 32 and T.INSTANCE.itemsize == 160
    assert T.RAY.itemsize == 32 and T.HIT.itemsize == 24
    from threert_amd import render as R

    # ... light_tree, light_tree_vec4s, glass_energy, tex_pages, tex_side + tex_layers, tex_info, tex_info_vec4s,
    # env, env_width + env_height, env_cdf, env_cdf_count, vcolors, vcolor_words
    assert ctypes.sizeof(R._RenderDesc) == 8 * 6 + 8 + 8 * 17 + 16 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8
    assert ctypes.sizeof(T._Tree) == 8 * 4 + 8 + 24


def test_the_oracle_loads():
    L = oracle.lib()
    for name in ("or_build_tris", "or_trace", "or_brute", "or_trace_batch"):
        assert hasattr(L, name)


def test_errors_come_back_as_codes_and_messages():
    import numpy as np
    import pytest

    with pytest.raises(T.ThreeRTError) as e:
        T.mesh_tree(np.zeros((0, 3, 3), np.float32))
    assert e.value.code == T.TRT_E_INVALID and "without triangles" in str(e.value)


@pytest.mark.gpu
def test_device_copy_moves_the_bytes(device):
    import numpy as np

    a = np.arange(1 << 20, dtype=np.uint32) * np.uint32(2654435761)
    src, dst = device.alloc(a.nbytes), device.alloc(a.nbytes)
    src.upload(a)
    dst.copy_from(src, a.nbytes - 64)
    device.sync()
    got = dst.download(np.uint32, len(a) - 16)
    assert (got == a[:-16]).all()
    with pytest.raises(T.ThreeRTError):
        dst.copy_from(src, 24)  # not a multiple of 16
    src.free()
    dst.free()
