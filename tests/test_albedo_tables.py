"""The three directional-albedo tables (crates/verion-gpu/tables/*.f32, copied as data) against
their published quadratures (scripts/gen_{ggx,sheen,dielectric}_albedo.py docstrings), recomputed
here with vectorised numpy. This pins the data the BSDF lookups read."""
import math
import os

import numpy as np
import pytest

TABLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "verion_amd", "tables")


def load(name, shape):
    return np.fromfile(os.path.join(TABLES, name), "<f4").reshape(shape)


def g1(c, a):
    return 2.0 / (1.0 + np.sqrt(1.0 + a * a * (1.0 - c * c) / (c * c)))


def test_ggx_table_matches_its_quadrature():
    # E(mu, alpha) for F = 1: mean of G1(wo) woh / (mu cos_h) G1(wi) over GGX half vectors, 256^2 cells
    tab = load("ggx_albedo.f32", (32, 32))
    n = 256
    u = (np.arange(n) + 0.5) / n
    phi = (np.arange(n) + 0.5) / n * 2 * math.pi
    for r in (0, 5, 13, 22, 31):
        a = (r / 31) ** 2
        for m in (0, 7, 16, 31):
            mu = max(m / 31, 0.01)
            if a == 0.0:
                want = 1.0
            else:
                tan2 = a * a * u / (1 - u)
                ch = 1 / np.sqrt(1 + tan2)
                sh = np.sqrt(1 - ch * ch)
                hx = sh[:, None] * np.cos(phi)[None, :]
                hz = np.broadcast_to(ch[:, None], hx.shape)
                woh = math.sqrt(1 - mu * mu) * hx + mu * hz
                wiz = 2 * woh * hz - mu
                ok = (woh > 0) & (wiz > 0)
                val = np.where(ok, g1(mu, a) * woh / (mu * hz) * g1(np.where(ok, wiz, 1.0), a), 0.0)
                want = min(val.sum() / n / n, 1.0)
            assert abs(tab[r, m] - want) < 2e-6, (r, m, tab[r, m], want)


def test_sheen_table_matches_its_quadrature():
    tab = load("sheen_albedo.f32", (16, 16))
    nu, npp = 1024, 64
    th = (np.arange(nu) + 0.5) / nu * (math.pi / 2)
    ph = (np.arange(npp) + 0.5) / npp * 2 * math.pi
    ci, si = np.cos(th)[:, None], np.sin(th)[:, None]
    for r in (1, 8, 15):
        rr = max(r / 15, 0.07)
        a = rr * rr
        for m in (0, 6, 15):
            mu = max(m / 15, 0.01)
            wo = (math.sqrt(1 - mu * mu), 0.0, mu)
            wx, wy, wz = si * np.cos(ph)[None, :], si * np.sin(ph)[None, :], np.broadcast_to(ci, (nu, npp))
            hx, hy, hz = wo[0] + wx, wo[1] + wy, wo[2] + wz
            hz = hz / np.sqrt(hx * hx + hy * hy + hz * hz)
            sin_h = np.sqrt(np.maximum(1 - hz * hz, 0))
            d = (2 + 1 / a) * sin_h ** (1 / a) / (2 * math.pi)
            v = 1 / (4 * (wz + mu - wz * mu))
            want = min(float((d * v * wz * si).sum()) * (math.pi / 2 / nu) * (2 * math.pi / npp), 1.0)
            assert abs(tab[r, m] - want) < 2e-6, (r, m, tab[r, m], want)


@pytest.mark.parametrize("side,i_ior", [(0, 2), (1, 2), (0, 8), (1, 5)])
def test_dielectric_table_matches_its_quadrature(side, i_ior):
    tab = load("dielectric_albedo.f32", (2, 9, 16, 16))
    ns = 96
    u1 = np.repeat((np.arange(ns) + 0.5) / ns, ns)
    u2 = np.tile((np.arange(ns) + 0.5) / ns, ns)
    ior = 1 + 0.25 * i_ior
    eta = ior if side == 0 else 1 / ior
    for ir in (0, 7, 15):
        alpha = max((ir / 15) ** 2, 1e-3)
        for im in (0, 8, 15):
            mu = max(im / 15, 0.01)
            wo = np.array([math.sqrt(1 - mu * mu), 0.0, mu])
            vh = np.array([alpha * wo[0], alpha * wo[1], wo[2]])
            vh /= np.linalg.norm(vh)
            lensq = vh[0] ** 2 + vh[1] ** 2
            t1 = np.array([-vh[1], vh[0], 0.0]) / math.sqrt(lensq) if lensq > 0 else np.array([1.0, 0, 0])
            t2 = np.cross(vh, t1)
            r = np.sqrt(u1)
            a, b = r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)
            s = 0.5 * (1 + vh[2])
            b = (1 - s) * np.sqrt(1 - a * a) + s * b
            c = np.sqrt(np.maximum(0, 1 - a * a - b * b))
            nh = a[:, None] * t1 + b[:, None] * t2 + c[:, None] * vh
            h = np.stack([alpha * nh[:, 0], alpha * nh[:, 1], np.maximum(0, nh[:, 2])], 1)
            h /= np.linalg.norm(h, axis=1, keepdims=True)
            woh = h @ wo
            s2 = (1 - woh * woh) / (eta * eta)
            cost = np.sqrt(np.maximum(1 - s2, 0))
            rs = (woh - eta * cost) / (woh + eta * cost)
            rp = (eta * woh - cost) / (eta * woh + cost)
            f = np.where(s2 >= 1, 1.0, 0.5 * (rs * rs + rp * rp))
            wiz = 2 * woh * h[:, 2] - wo[2]
            tot = np.where(wiz > 0, f * g1(np.where(wiz > 0, wiz, 1.0), alpha), 0.0)
            rr = 1 / eta
            k = 1 - rr * rr * (1 - woh * woh)
            wtz = -wo[2] * rr + (rr * woh - np.sqrt(np.maximum(k, 0))) * h[:, 2]
            ok = (k > 0) & (wtz < 0)
            tot = tot + np.where(ok, (1 - f) * g1(np.where(ok, -wtz, 1.0), alpha), 0.0)
            want = min(tot.sum() / ns / ns, 1.0)
            assert abs(tab[side, i_ior, ir, im] - want) < 1e-5, (side, i_ior, ir, im, tab[side, i_ior, ir, im], want)


def test_first_ggx_word_is_one():
    # pack.rs:907: GGX albedo at cos = 0, alpha = 0 is 1.0 (a mirror)
    assert load("ggx_albedo.f32", (32, 32))[0, 0] == 1.0
