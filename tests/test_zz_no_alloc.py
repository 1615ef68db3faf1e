"""lx_trace_closest_noalloc (the allocation-free trace) against
lx_trace_closest on the same rays: bit-identical hits on two calls in a
row, the second with fewer rays, so the shared scratch array's contents from
the first call must not leak into the second.  Device buffers for the rays
and hits come from hipMalloc through ctypes (the caller's memory, not the
library's)."""
import ctypes as C

import numpy as np
import pytest

import luxion_amd as lx
from luxion_amd import _ffi as F
from luxion_amd import scenes

pytestmark = pytest.mark.gpu

HIT_DT = np.dtype([("t", "<f4"), ("instance", "<u4"), ("prim", "<u4"), ("u", "<f4"), ("v", "<f4")])


class _Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(n)) == 0
        return p.value

    def free(self, p):
        assert self.h.hipFree(C.c_void_p(p)) == 0

    def h2d(self, dst, src):
        assert self.h.hipMemcpy(C.c_void_p(dst), C.c_void_p(src.ctypes.data), src.nbytes, 1) == 0

    def d2h(self, dst, src):
        assert self.h.hipMemcpy(C.c_void_p(dst.ctypes.data), C.c_void_p(src), dst.nbytes, 2) == 0


def _err():
    return F.load().lx_last_error().decode()


def _rays(data, n, seed):
    """n rays from random points of the scene's box towards random instance
    origins (most of them hit something)."""
    rng = np.random.default_rng(seed)
    cent = np.array([np.asarray(inst[2], np.float32).reshape(3, 4)[:, 3] for inst in data.instances])
    o = rng.uniform([-30, -3, -30], [30, 14, 30], (n, 3)).astype(np.float32)
    d = cent[rng.integers(0, len(cent), n)] + rng.normal(0, 0.3, (n, 3)).astype(np.float32) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 4:7], r[:, 7] = o, d, 1e16
    return r


def _noalloc(hip, sc, rays, alpha, iteration):
    n = len(rays)
    dr, dh = hip.malloc(max(1, rays.nbytes)), hip.malloc(max(1, n * HIT_DT.itemsize))
    try:
        hip.h2d(dr, np.ascontiguousarray(rays, np.float32))
        lx.check(sc.lib.lx_trace_closest_noalloc(sc.h, dr, n, int(alpha), iteration, dh))
        out = np.zeros(n, HIT_DT)
        hip.d2h(out, dh)
        return out
    finally:
        hip.free(dr)
        hip.free(dh)


def test_noalloc_trace_matches_trace_closest_twice(ctx):
    hip = _Hip()
    data = scenes.instanced_interior(64, 36, n_meshes=6, n_instances=60, tri_range=(80, 1300), n_emissive=2,
                                     tex_size=64, trace_depth=6)
    sc = lx.Scene(ctx, data)
    for k, (n, alpha, it) in enumerate(((131072, True, 3), (20011, False, 5))):
        rays = _rays(data, n, 40 + k)
        print(f"noalloc call {k}: {n} rays, alpha {alpha}, iteration {it}", flush=True)
        got = _noalloc(hip, sc, rays, alpha, it)
        want = lx.trace_closest(sc, rays, alpha=alpha, iteration=it)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), k
        assert (got["t"] > 0).mean() > 0.5


def test_noalloc_trace_rejects_bad_sizes_and_host_pointers(ctx):
    hip = _Hip()
    data = scenes.cornell(16, 16)
    sc = lx.Scene(ctx, data)
    rays = _rays(data, 64, 1)
    out = np.zeros(64, HIT_DT)
    # host pointers are refused before any launch
    st = sc.lib.lx_trace_closest_noalloc(sc.h, rays.ctypes.data, 64, 1, 1, out.ctypes.data)
    assert st == 1 and "device allocation" in _err()
    # more rays than the scratch array holds, and an output buffer too short
    dr, dh = hip.malloc(rays.nbytes), hip.malloc(32 * HIT_DT.itemsize)
    try:
        hip.h2d(dr, rays)
        assert sc.lib.lx_trace_closest_noalloc(sc.h, dr, 131073, 1, 1, dh) == 1
        assert "at most 131072" in _err()
        assert sc.lib.lx_trace_closest_noalloc(sc.h, dr, 64, 1, 1, dh) == 1
        assert "past the end" in _err()
        assert sc.lib.lx_trace_closest_noalloc(sc.h, dr, 0, 1, 1, dh) == 0
    finally:
        hip.free(dr)
        hip.free(dh)
