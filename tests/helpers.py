"""Shared helpers of the test suite: oracle renders and image statistics."""
import numpy as np

import oracle
from verion_amd import pack


def params(packed, w, h, lighting=False, multiscatter=True, adaptive=None, restir=None):
    return pack.base_params(packed.info, packed.view, w, h, lighting, multiscatter, adaptive, restir)


def oracle_render(scene, w, h, spp, seed=0, multiscatter=True, lighting=False, adaptive=None, restir=None, brute=False):
    """Mean image (pixels, 3) and the raw state of an oracle render of `scene`."""
    p = pack.pack(scene)
    o = oracle.OracleScene(p)
    par = params(p, w, h, lighting, multiscatter, adaptive, restir)
    st = o.render(par, spp, seed=seed, brute=brute)
    return oracle.mean(st, spp), st


def center_mean(img, w, size, channel=0):
    """Mean of one channel over the central size x size block (tests/pathtrace.rs center_mean)."""
    img = np.asarray(img).reshape(-1, w, 3)
    lo = (w - size) // 2
    return float(img[lo:lo + size, lo:lo + size, channel].mean())


def center_pixel(img, w):
    img = np.asarray(img).reshape(-1, w, 3)
    return float(img[w // 2, w // 2, 0])
