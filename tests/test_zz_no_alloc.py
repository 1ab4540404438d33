"""vh_nrc_gradient_noalloc (the radiance cache's gradient on device buffers, allocating nothing)
against vh_nrc_gradient on the same samples: bit for bit, on two calls in a row (the second smaller,
so a stale staging array would show), and the size limit rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from verion_amd import scenes
from verion_amd.backend import INVALID_ARGUMENT, Nrc, Renderer, library

pytestmark = pytest.mark.gpu
MAX_SAMPLES = 65536


def hip_runtime():
    """The HIP runtime libverion_hip.so runs on (the one mapped into this process by it)."""
    library()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert len(paths) == 1, paths
    hip = C.CDLL(paths.pop())
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class DeviceArray:
    H2D, D2H = 1, 2

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes = hip, nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, self.H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def batch(n, seed):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 12), np.float32)
    raw[:, 0:3] = rng.random((n, 3))
    raw[:, 3] = rng.random(n)
    d = rng.normal(size=(n, 3))
    raw[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    raw[:, 8:11] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    target = (0.5 + 0.4 * np.sin(3 * raw[:, 0:3]) * raw[:, 4:7]).astype(np.float32)
    return raw, target


def test_no_alloc_gradient_equals_the_allocating_one(ctx):
    L = library()
    L.vh_nrc_gradient_noalloc.restype = C.c_int
    L.vh_nrc_gradient_noalloc.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    hip = hip_runtime()
    r = Renderer(ctx, scenes.cornell())
    cache = Nrc(r, train_paths=1024)
    cache.init(3)
    first = 8192
    d_raw, d_tgt = DeviceArray(hip, 48 * first), DeviceArray(hip, 12 * first)
    d_grad, d_loss = DeviceArray(hip, 4 * Nrc.PARAMS), DeviceArray(hip, 4)
    try:
        for n, seed in ((first, 21), (1000, 22)):
            raw, target = batch(n, seed)
            want, want_loss = cache.gradient(raw, target)
            d_raw.put(raw)
            d_tgt.put(target)
            assert L.vh_nrc_gradient_noalloc(cache.h, n, d_raw.p, d_tgt.p, d_grad.p, d_loss.p) == 0
            got = d_grad.get(np.float32, Nrc.PARAMS)
            got_loss = d_loss.get(np.float32, 1)[0]
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
            assert np.float32(got_loss) == np.float32(want_loss)
            assert np.linalg.norm(got) > 0
        # beyond the staging array: refused before any device work
        assert L.vh_nrc_gradient_noalloc(cache.h, MAX_SAMPLES + 1, d_raw.p, d_tgt.p, d_grad.p, d_loss.p) == INVALID_ARGUMENT
        assert b"65536" in L.vh_last_error()
    finally:
        for a in (d_raw, d_tgt, d_grad, d_loss):
            a.free()
