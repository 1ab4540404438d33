"""The library's own seeded initialisation of the radiance cache (vh_nrc_init_params, host only).

Pinned three ways: the layout and the He-uniform bounds the oracle's `init` states
(oracle/nrc.py: weights uniform in +-sqrt(6 / fan_in), zero biases, the output layer scaled by 0.1
with its 13 unused outputs zero); the moments and the distribution of a uniform law; and exact
values from an independent numpy statement of the documented generator (the PCG hash of
shaders/include/pt_core.slang:69-74, as tests/golden/make_pcg.py computes it).
"""
import math
import os
import sys

import numpy as np
import pytest

from oracle import nrc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "verion_amd", "libverion_hip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libverion_hip.so not built")

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from make_pcg import pcg  # noqa: E402


def lib_init(seed):
    from verion_amd.backend import Nrc

    return Nrc.init_params(seed)


def numpy_init(seed):
    """The generator as the header documents it, in numpy uint32 / float32 arithmetic."""
    theta = np.zeros(nrc.param_count(), np.float32)
    with np.errstate(over="ignore"):
        key = pcg(np.uint32(seed) ^ np.uint32(0xA511E9B3))
    o = 0
    for li, (k, n) in enumerate(nrc.LAYERS):
        last = li == len(nrc.LAYERS) - 1
        lim = np.float32(np.sqrt(np.float32(6.0) / np.float32(k))) * np.float32(0.1 if last else 1.0)
        p = np.arange(o, o + k * n, dtype=np.uint32)
        with np.errstate(over="ignore"):
            u = (pcg(key + p) >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
        w = lim * (np.float32(2.0) * u - np.float32(1.0))
        if last:
            w = w.reshape(k, n)
            w[:, 3:] = 0.0
            w = w.reshape(-1)
        theta[o:o + k * n] = w
        o += k * n + n
    return theta


def layers(theta):
    return nrc.unpack(theta.astype(np.float64))


def test_layout_and_bounds_match_the_oracle_initialisation():
    theta = lib_init(0)
    assert theta.shape == (nrc.param_count(),) and theta.dtype == np.float32
    ours, theirs = layers(theta), layers(nrc.init(0))
    for li, ((W, b), (Wo, bo)) in enumerate(zip(ours, theirs)):
        k = W.shape[0]
        lim = math.sqrt(6.0 / k) * (0.1 if li == len(ours) - 1 else 1.0)
        assert np.all(b == 0.0) and np.all(bo == 0.0)
        live = W[:, :3] if li == len(ours) - 1 else W
        assert np.all(np.abs(live) <= lim * (1 + 1e-6))
        # both fill the interval: the largest draws come within 2 % of the bound
        assert np.abs(live).max() >= 0.98 * lim
        assert np.abs(Wo[:, :3] if li == len(ours) - 1 else Wo).max() >= 0.98 * lim
        if li == len(ours) - 1:
            assert np.all(W[:, 3:] == 0.0) and np.all(Wo[:, 3:] == 0.0)


def test_weights_are_uniform_in_the_he_interval():
    from scipy import stats

    theta = lib_init(3)
    for li, (W, _) in enumerate(layers(theta)):
        last = li == len(nrc.LAYERS) - 1
        k = W.shape[0]
        lim = math.sqrt(6.0 / k) * (0.1 if last else 1.0)
        w = (W[:, :3] if last else W).reshape(-1)
        n = w.size
        # mean 0 and E[w^2] = lim^2 / 3 within 4 standard errors
        assert abs(w.mean()) <= 4 * lim / math.sqrt(3 * n)
        m2 = (w * w).mean()
        assert abs(m2 - lim * lim / 3) <= 4 * math.sqrt(4 * lim ** 4 / 45 / n)
        # the whole law (Kolmogorov-Smirnov against U(-lim, lim))
        assert stats.kstest(w, stats.uniform(loc=-lim, scale=2 * lim).cdf).pvalue > 1e-3


@pytest.mark.parametrize("seed", [0, 1, 7, 0xFFFFFFFF])
def test_values_follow_the_documented_generator(seed):
    np.testing.assert_array_equal(lib_init(seed).view(np.uint32), numpy_init(seed).view(np.uint32))


def test_seeds_are_reproducible_and_distinct():
    a, b, c = lib_init(5), lib_init(5), lib_init(6)
    np.testing.assert_array_equal(a, b)
    assert np.count_nonzero(a != c) > 0.99 * np.count_nonzero(a)


def test_wrong_size_is_rejected():
    import ctypes as C

    from verion_amd.backend import library

    out = np.zeros(10, np.float32)
    assert library().vh_nrc_init_params(0, out.ctypes.data_as(C.POINTER(C.c_float)), out.size) != 0
