"""prt_query_rays_device_noalloc (the query that allocates nothing: its traversal-stack overflow is a
static device array) against prt_query_rays_device on the same device buffers: identical closest
hits on two calls in a row, the second with fewer rays, on a scene whose traversal stacks overflow
the LDS part of the stack."""
import ctypes

import numpy as np
import pytest

import protoray_hip as P
import scenegen as G

pytestmark = pytest.mark.gpu


def _rays(seed, n, targets):
    rng = np.random.default_rng(seed)
    o = rng.uniform([-12, -2, -4], [12, 10, 2], size=(n, 3))
    d = targets[rng.integers(0, len(targets), n)] + rng.normal(scale=0.5, size=(n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    far = np.full(n, 3.4028234663852886e38)
    far[: n // 5] = rng.uniform(1, 40, n // 5)
    return np.concatenate([o, d, np.zeros((n, 1)), far[:, None]], axis=1).astype(np.float32)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself links (the test's
    device buffers; hipMemcpy kind 1 = host to device, 2 = device to host)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p.value

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def _query(dev, hip, rays, noalloc):
    n = rays.shape[0]
    r = hip.upload(rays)
    outs = [np.full(n, -7, np.int32)] + [np.zeros(n, np.float32) for _ in range(3)]
    ptrs = [hip.upload(o) for o in outs]
    try:
        dev.queryRaysDevice(n, r, *ptrs, noalloc=noalloc)
        dev.synchronize()
        return [hip.download(p, o) for p, o in zip(ptrs, outs)]
    finally:
        for p in [r] + ptrs:
            hip.free(p)


def test_noalloc_query_matches_the_allocating_query():
    sc = G.spheres_ambient()  # 327,682 triangles
    dev = P.DeviceHip(0)  # loads the library (and its HIP runtime) first
    hip = _Hip()
    try:
        dev.initScene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
        levels = dev.bvhStats()["max_depth"]
        assert 3 * levels > 10  # deep enough that stacks can spill past the 10 LDS entries
        targets = sc["positions"][sc["indices"]].mean(axis=1).astype(np.float64)
        for seed, n in ((1, 400_000), (2, 7_001)):  # the second call is smaller
            rays = _rays(seed, n, targets)
            ref = _query(dev, hip, rays, noalloc=False)
            got = _query(dev, hip, rays, noalloc=True)
            assert (ref[0] >= 0).sum() > n // 4
            assert np.array_equal(got[0], ref[0])
            hit = ref[0] >= 0
            for a, b in zip(got[1:], ref[1:]):
                assert np.array_equal(a[hit].view(np.uint32), b[hit].view(np.uint32))
    finally:
        dev.close()
