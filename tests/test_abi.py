"""The C ABI: libthreert.so loads without a GPU and exports every entry point include/threert.h declares; the
oracle library loads too. No compute call is made here."""
import ctypes
import os
import re

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

    assert ctypes.sizeof(R._RenderDesc) == 8 * 6 + 8 + 8 * 17
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
