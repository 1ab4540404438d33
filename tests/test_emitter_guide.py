"""The emitter CDF's guide table (vh_cdf_guide, a host-only export of libverion_hip.so).

The renderer selects an emitter for a draw u with the reference's binary search
(pt_core.slang:1343-1351: lo = 0, hi = n - 1; mid = (lo + hi) / 2; cdf[mid] >= u ? hi = mid :
lo = mid + 1), restricted to the range between two cutpoints of the guide table. Pinned here on the
CPU: for every u the device can draw (the 2^24 multiples of 2^-24; for the edge cases a million
random draws and every draw next to an interval boundary), the reference search's answer lies
between the cutpoints of u's interval, so the restricted search (emulated step for step) returns
the same emitter; on power CDFs of the packer and on edge cases (plateaus of zero-power emitters,
a total that rounds below 1, two emitters, more emitters than 2^16)."""
import ctypes as C
import os

import numpy as np
import pytest

from verion_amd import pack, scenes

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "verion_amd", "libverion_hip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libverion_hip.so not built")
STEP = np.float32(2.0 ** -24)  # draws are (state >> 8) * 2^-24


def draws(bits, exhaustive):
    if exhaustive:
        return np.arange(1 << 24, dtype=np.int64).astype(np.float32) * STEP
    m = np.random.default_rng(bits).integers(0, 1 << 24, 1 << 20)
    edge = np.arange(1 << bits, dtype=np.int64) << (24 - bits)
    m = np.concatenate([m, edge, np.maximum(edge - 1, 0), np.minimum(edge + 1, (1 << 24) - 1)])
    return m.astype(np.float32) * STEP


def guide(cdf, bits):
    lib = C.CDLL(LIB)
    cdf = np.ascontiguousarray(cdf, np.float32)
    out = np.zeros((1 << bits) + 1, np.uint32)
    rc = lib.vh_cdf_guide(cdf.ctypes.data_as(C.c_void_p), C.c_uint32(len(cdf)), C.c_uint32(bits),
                          out.ctypes.data_as(C.c_void_p))
    return rc, out


def search(cdf, u, lo, hi):
    """The reference's binary search, vectorised over draws, from per-draw [lo, hi]."""
    lo, hi = lo.astype(np.int64), hi.astype(np.int64)
    while True:
        go = lo < hi
        if not go.any():
            return lo
        mid = (lo + hi) // 2
        take = go & (cdf[np.minimum(mid, len(cdf) - 1)] >= u)
        hi = np.where(take, mid, hi)
        lo = np.where(go & ~take, mid + 1, lo)


def bits_for(n):
    b = 0
    while b < 16 and (1 << b) < n:
        b += 1
    return b


def power_cdf(power):
    c = np.cumsum(np.asarray(power, np.float64))
    return (c / c[-1]).astype(np.float32)


def cases():
    rng = np.random.default_rng(3)
    p = pack.pack(scenes.many_lights(emitters=600, clutter_boxes=50))
    yield "many_lights packer", p.emit_cdf, True
    yield "uniform 10000", power_cdf(np.ones(10000)), False
    yield "lognormal sigma 3", power_cdf(rng.lognormal(0.0, 3.0, 10000)), False
    w = rng.lognormal(0.0, 1.0, 5000)
    w[rng.random(5000) < 0.5] = 0.0  # zero-power emitters: plateaus
    w[-300:] = 0.0
    yield "plateaus", power_cdf(w), True
    c = power_cdf(rng.random(777))
    c[-40:] = np.float32(0.9999)  # a total that rounds below 1
    yield "short total", c, False
    yield "two", power_cdf([1.0, 3.0]), False
    yield "one dominant", power_cdf(np.r_[np.full(2000, 1e-6), 1.0, np.full(2000, 1e-6)]), False
    yield "more than 2^16", power_cdf(rng.random(70000)), False


@pytest.mark.parametrize("name,cdf,exhaustive", list(cases()), ids=[c[0] for c in cases()])
def test_restricted_search_selects_the_reference_emitter(name, cdf, exhaustive):
    n = len(cdf)
    bits = bits_for(n)
    U = draws(bits, exhaustive)
    rc, g = guide(cdf, bits)
    assert rc == 0
    full = search(cdf, U, np.zeros(len(U), np.int64), np.full(len(U), n - 1, np.int64))
    first = np.minimum(np.searchsorted(cdf, U, side="left"), n - 1)  # first index reaching u, else n - 1
    assert np.array_equal(full, first)
    k = (U * np.float32(1 << bits)).astype(np.int64)
    assert k.max() == (1 << bits) - 1
    assert (g[k] <= full).all() and (full <= g[k + 1]).all()
    assert np.array_equal(search(cdf, U, g[k], g[k + 1]), full)
    span = (g[k + 1].astype(np.int64) - g[k]).mean()
    assert span <= 8, f"{name}: {span} entries per interval on average"


def test_guide_rejects_a_decreasing_cdf_and_bad_arguments():
    assert guide(np.array([0.5, 0.4, 1.0], np.float32), 2)[0] != 0
    assert guide(np.array([0.5, 1.0], np.float32), 17)[0] != 0
