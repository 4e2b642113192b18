"""RtTraceRaysNoAlloc (device-resident rays and hits, library-static scratch, no allocation)
returns the same hit records as RtTraceRays, bit for bit, for closest-hit and any-hit
queries, on two calls in a row with the second one smaller; invalid sizes are refused."""
import ctypes as c

import numpy as np
import pytest

import oracle
from native_api import HIT_DTYPE, RAY_DTYPE, Session

pytestmark = pytest.mark.gpu


def _rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, dtype=oracle.RAY)
    rays["origin"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    rays["direction"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["limit"] = np.where(rng.uniform(size=n) < 0.5, 1e30, rng.uniform(0.01, 3, n))
    return rays


class DeviceBuffer:
    """Device memory from the HIP runtime the library links (one runtime per process)."""

    def __init__(self, nbytes):
        self.hip = c.CDLL("libamdhip64.so.7")
        self.ptr, self.nbytes = c.c_void_p(), nbytes
        assert self.hip.hipMalloc(c.byref(self.ptr), c.c_size_t(nbytes)) == 0

    def upload(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(c.c_void_p), c.c_size_t(a.nbytes), 1) == 0   # H2D

    def download(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(c.c_void_p), self.ptr, c.c_size_t(out.nbytes), 2) == 0  # D2H
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None


def _no_alloc(lib, s, rays, any_hit):
    host = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    d_rays, d_hits = DeviceBuffer(host.nbytes), DeviceBuffer(len(host) * HIT_DTYPE.itemsize)
    try:
        d_rays.upload(host)   # synchronous: the rays are in place before the library's stream reads them
        d_hits.upload(np.full(len(host) * HIT_DTYPE.itemsize, 0xAB, np.uint8))
        ok = lib.RtTraceRaysNoAlloc(s.handle, d_rays.ptr, len(host), int(any_hit), d_hits.ptr)
        assert ok, s.error()
        return d_hits.download(HIT_DTYPE, len(host))
    finally:
        d_rays.free()
        d_hits.free()


@pytest.mark.parametrize("scene", ["icosphere", "soup"])
def test_no_alloc_trace_equals_rt_trace_rays(lib, scene):
    from scenes import displaced_icosphere, triangle_soup
    if scene == "icosphere":
        tris, lo, hi = displaced_icosphere(subdiv=6, seed=4), (-3, 0.2, -3), (3, 3.5, 3)
    else:
        tris, lo, hi = triangle_soup(60000, seed=6), (-0.3, -0.3, -0.3), (1.3, 1.3, 1.3)
    with Session(lib) as s:
        s.initialize(tris, 1, 1)
        for any_hit in (False, True):
            for n, seed in ((300000, 1), (40000, 2)):   # second call smaller than the first
                rays = _rays(n, seed + 10 * any_hit, lo, hi)
                ref = s.trace(rays, any_hit)
                got = _no_alloc(lib, s, rays, any_hit)
                assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (any_hit, n)
                assert (ref["index"] >= 0).sum() > n // 20


def test_no_alloc_refuses_invalid_sizes(lib):
    from scenes import cornell
    with Session(lib) as s:
        s.initialize(cornell(), 1, 1)
        buf = c.c_void_p(16)   # never dereferenced: every call below fails validation first
        assert lib.RtTraceRaysNoAlloc(s.handle, buf, (1 << 30) + 1, 0, buf) == 0
        assert lib.RtTraceRaysNoAlloc(s.handle, None, 4, 0, buf) == 0
        assert lib.RtTraceRaysNoAlloc(s.handle, buf, -1, 0, buf) == 0
        assert lib.RtTraceRaysNoAlloc(s.handle, buf, 0, 0, buf) == 1   # nothing to trace
    assert lib.RtTraceRaysNoAlloc(None, None, 1, 0, None) == 0
