"""CPU-side checks of the C ABI library: it loads, exports every function declared in
include/rt_abi.h, keeps the reference's struct layouts and null-session contract
(Native/renderer.cu:309-318).  No GPU work is started."""
import ctypes as c
import re
from pathlib import Path

import numpy as np

import native_api
import oracle

HEADER = Path(__file__).resolve().parents[1] / "include" / "rt_abi.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(Rt[A-Za-z]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 13 + 6
    for name in names:
        assert hasattr(lib, name), name


def test_reference_exports_present(lib):
    for name in ("RtCreate", "RtDestroy", "RtShutdown", "RtReset", "RtAccumulationVersion", "RtSampleCount",
                 "RtEmbreeBuildMilliseconds", "RtCpuThreadCount", "RtBvhInfo", "RtLastError", "RtInitialize",
                 "RtRender", "RtDenoise"):
        assert hasattr(lib, name)


def test_struct_layouts_match_reference():
    assert (c.sizeof(native_api.Triangle), c.sizeof(native_api.Camera), c.sizeof(native_api.Settings)) == (116, 64, 28)
    assert native_api.Triangle.specularColor.offset == 96 and native_api.Triangle.mesh.offset == 92
    assert native_api.Camera.tanHalfFov.offset == 48 and native_api.Settings.seed.offset == 20
    assert native_api.TRIANGLE_DTYPE == oracle.TRIANGLE.newbyteorder("=") or native_api.TRIANGLE_DTYPE.itemsize == 116
    for name in native_api.TRIANGLE_DTYPE.names:
        assert native_api.TRIANGLE_DTYPE.fields[name][1] == oracle.TRIANGLE.fields[name][1]


def test_null_session_contract(lib):
    assert lib.RtSampleCount(None) == 0 and lib.RtAccumulationVersion(None) == 0
    assert lib.RtCpuThreadCount(None) == 0 and lib.RtEmbreeBuildMilliseconds(None) == 0
    n, d, ms = c.c_int(7), c.c_int(7), c.c_float(7)
    lib.RtBvhInfo(None, c.byref(n), c.byref(d), c.byref(ms))
    assert (n.value, d.value, ms.value) == (0, 0, 0)
    assert lib.RtLastError(None)
    lib.RtReset(None); lib.RtShutdown(None); lib.RtDestroy(None)
    out = (c.c_ubyte * 16)()
    assert lib.RtRender(None, None, None, out, 16, 2, None) == 0
    assert lib.RtInitialize(None, None, 0, 1, 1) == 0
    assert lib.RtDenoise(None, out, 16, None) == 0


def test_session_lifecycle_without_initialize(lib):
    s = lib.RtCreate()
    try:
        assert lib.RtSampleCount(s) == 0
        lib.RtReset(s)
        assert lib.RtAccumulationVersion(s) == 1
        assert lib.RtSetOption(s, b"bvh_width", 8) == 1 and lib.RtGetOption(s, b"bvh_width") == 8
        assert lib.RtSetOption(s, b"bvh_width", 3) == 0 and lib.RtSetOption(s, b"nonsense", 1) == 0
        assert lib.RtGetOption(s, b"nonsense") == -1
        assert lib.RtGetOption(s, b"xcd_segments") == 1 and lib.RtSetOption(s, b"xcd_segments", 0) == 1
        assert lib.RtGetOption(s, b"xcd_segments") == 0 and lib.RtSetOption(s, b"xcd_segments", 2) == 0
        # invalid arguments are rejected before any device work
        tris = np.zeros(1, dtype=native_api.TRIANGLE_DTYPE)
        assert lib.RtInitialize(s, native_api.triangle_pointer(tris), 0, 4, 4) == 0
        assert lib.RtInitialize(s, native_api.triangle_pointer(tris), 1, 4097, 4) == 0
        out = (c.c_ubyte * 16)()
        assert lib.RtRender(s, None, None, out, 16, 2, None) == 0
        assert lib.RtDenoise(s, out, 16, c.byref(c.c_float())) == 0
    finally:
        lib.RtDestroy(s)
