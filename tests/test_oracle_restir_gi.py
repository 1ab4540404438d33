"""The oracle's ReSTIR GI (pt_core.slang:901-980, 1566-1632), pinned without the reference's GPU:

- the reconnection Jacobian against the ratio of exact solid angles (Van Oosterom & Strackee's
  triangle formula) of a small area element at the reused point, seen from the two shading points;
- the first frame, which has nothing to reuse, against the plain path tracer (its own candidate
  with W = 1 / pdf gives the path's own radiance back, to rounding);
- what the first frame writes: every reservoir's point lies on a surface of the scene and its
  normal faces the pixel's shading point, M = 1, and no point lies on glass, dispersive or not;
- convergence of many frames to the plain path tracer on the closed room of the reference's GI
  tests (crates/verion-gpu-vulkan/tests/restir.rs:152-271, at a size the CPU finishes in seconds),
  and to the closed form of the furnace room;
- sharp surfaces are left to the plain path tracer, bit for bit (restir.rs:263-271)."""
import numpy as np
import pytest

import oracle
from helpers import oracle_render
from verion_amd import scenes

GI = (False, True, 4, 4)
DI_GI = (True, True, 4, 4)
DI = (True, False, 4, 4)


def solid_angle(o, a, b, c):
    """Solid angle of triangle abc seen from o (Van Oosterom & Strackee 1983), in float64."""
    r1, r2, r3 = a - o, b - o, c - o
    l1, l2, l3 = np.linalg.norm(r1), np.linalg.norm(r2), np.linalg.norm(r3)
    num = abs(np.dot(r1, np.cross(r2, r3)))
    den = l1 * l2 * l3 + np.dot(r1, r2) * l3 + np.dot(r1, r3) * l2 + np.dot(r2, r3) * l1
    return 2.0 * np.arctan2(num, den)


def element(x2, n2, eps):
    """A small triangle of size eps around x2 in the plane with normal n2."""
    t = np.cross(n2, [1.0, 0.0, 0.0] if abs(n2[0]) < 0.9 else [0.0, 1.0, 0.0])
    t /= np.linalg.norm(t)
    b = np.cross(n2, t)
    return x2 + eps * t, x2 + eps * (-0.5 * t + 0.866 * b), x2 + eps * (-0.5 * t - 0.866 * b)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_jacobian_is_the_ratio_of_solid_angles():
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(400):
        p = rng.uniform(-1, 1, 3)
        n = unit(rng.normal(size=3))
        pq = p + rng.uniform(-0.3, 0.3, 3)
        x2 = p + unit(n + rng.normal(size=3) * 0.6) * rng.uniform(0.2, 2.0)
        n2 = unit(rng.normal(size=3))
        if np.dot(n2, p - x2) < 0:
            n2 = -n2
        ok, jac = oracle.gi_reconnect(p, n, pq, x2, n2)
        a, b, c = element(x2, n2, 1e-4)
        want = solid_angle(p, a, b, c) / solid_angle(pq, a, b, c)
        cosT = np.dot(n, unit(x2 - p))
        cosR, cosQ = abs(np.dot(n2, unit(x2 - p))), abs(np.dot(n2, unit(pq - x2)))
        if cosT <= 0.0:
            assert not ok and jac == 0.0
            continue
        if cosT < 1e-3 or cosR < 2e-3 or cosQ < 2e-3:
            continue  # at the cut-offs rounding decides
        assert abs(jac / want - 1.0) < 2e-3, (p, pq, x2, n2, jac, want)
        if 0.101 < want < 9.9:
            assert ok
        elif want > 10.1 or want < 0.099:
            assert not ok
        checked += 1
    assert checked > 250


def test_reconnection_rejects_what_cannot_be_moved():
    z = [0.0, 0.0, 0.0]
    up = [0.0, 0.0, 1.0]
    # behind the shading point's surface
    assert oracle.gi_reconnect(z, up, [0.01, 0, 0], [0, 0, -1], up) == (False, 0.0)
    # closer than 1e-3
    assert not oracle.gi_reconnect(z, up, [0.01, 0, 0], [0, 0, 5e-4], [0, 0, -1])[0]
    # the reused point seen edge-on from the new shading point
    assert not oracle.gi_reconnect(z, up, [0.01, 0, 0], [1, 0, 1], [0.70710678, 0, -0.70710678])[0]
    # a shift that grows the solid angle more than tenfold (dQ / dR > sqrt(10))
    ok, jac = oracle.gi_reconnect(z, up, [0.0, 0.0, -3.0], [0, 0, 0.5], [0, 0, -1])
    assert not ok and jac > 10.0
    ok, jac = oracle.gi_reconnect(z, up, [0.0, 0.0, -0.5], [0, 0, 0.5], [0, 0, -1])
    assert ok and abs(jac - 4.0) < 1e-5


@pytest.fixture(scope="module")
def room():
    return scenes.closed_room(0.9)


def test_first_frame_is_the_plain_path_tracer(room):
    plain, _ = oracle_render(room, 32, 32, 1)
    gi, st = oracle_render(room, 32, 32, 1, restir=GI)
    assert np.abs(gi - plain).max() <= 1e-5 * np.abs(plain).max()
    assert (np.abs(gi - plain) <= 1e-5 * np.maximum(np.abs(plain), 1e-3)).all()
    di, _ = oracle_render(room, 32, 32, 1, restir=DI)
    both, _ = oracle_render(room, 32, 32, 1, restir=DI_GI)
    assert (np.abs(both - di) <= 1e-5 * np.maximum(np.abs(di), 1e-3)).all()


def on_surface(x, tol=3e-5):
    """Whether x lies on the room (half 1.5), the block (half 0.4) or the lamp plane y = 1.45."""
    a = np.abs(x)
    room = (np.abs(a.max(axis=1) - 1.5) < tol)
    block = (np.abs(a.max(axis=1) - 0.4) < tol)
    lamps = np.abs(x[:, 1] - 1.45) < tol
    return room | block | lamps


def test_first_frame_reservoirs_are_scene_points(room):
    _, st = oracle_render(room, 32, 32, 1, restir=GI)
    R = st["restir"].reshape(-1, 80)
    valid = R[:, 27].view(np.uint32) != 0
    assert valid.mean() > 0.8, "most pixels of the closed room keep a second-hit sample"
    x2, n2, pq = R[valid, 16:19], R[valid, 19:22], R[valid, 8:11]
    assert on_surface(x2).all()
    assert np.allclose(np.abs(n2).max(axis=1), 1.0, atol=1e-5), "cube faces have axis normals"
    assert (np.einsum("ij,ij->i", n2, pq - x2) > 0).all(), "the normal faces the shading point"
    assert (R[valid, 26] == 1.0).all() and (R[valid, 25] > 0).all()
    assert (R[valid, 22:25] >= 0).all() and (R[valid, 22:25].sum(axis=1) > 0).all()


@pytest.mark.parametrize("dispersion", [8.0, 0.0])
def test_no_reservoir_point_lies_on_glass(dispersion):
    s = scenes.closed_room(0.9, glass=True, dispersion=dispersion)
    _, st = oracle_render(s, 32, 32, 6, restir=GI)
    R = st["restir"].reshape(-1, 80)
    for half in (0, 28):
        valid = R[:, half + 27].view(np.uint32) != 0
        x = R[valid, half + 16:half + 19] - np.array([-0.7, -0.9, 0.0], np.float32)
        on_glass = (np.abs(np.abs(x).max(axis=1) - 0.35) < 3e-5)
        assert valid.sum() > 500 and not on_glass.any()


def worst_block(a, b, w, h):
    """crates/verion-gpu-vulkan/tests/restir.rs:273-293: the largest relative difference of the 8x8
    block sums of two images, per channel."""
    A = a.reshape(h // 8, 8, w // 8, 8, 3).sum(axis=(1, 3))
    B = b.reshape(h // 8, 8, w // 8, 8, 3).sum(axis=(1, 3))
    return float((np.abs(A - B) / np.maximum(B, 1e-3)).max())


def test_gi_converges_to_the_path_traced_image(room):
    # restir.rs:240-250 at 32 x 32 x 512 (the oracle's 64 x 64 x 1024 agrees to 0.4 %)
    plain, _ = oracle_render(room, 32, 32, 512)
    gi, _ = oracle_render(room, 32, 32, 512, restir=GI)
    err = abs(gi[:, 1].mean() / plain[:, 1].mean() - 1.0)
    assert err < 0.01, err
    assert worst_block(gi, plain, 32, 32) < 0.15


def test_gi_keeps_the_furnace_room_closed_form():
    # walls emitting 1 and reflecting 0.5, paths of 8 vertices: sum of 0.5^k, k < 8 = 1.9921875
    s = scenes.furnace_room(0.5, 1.0)
    gi, _ = oracle_render(s, 32, 32, 256, restir=GI)
    assert abs(gi.mean() / 1.9921875 - 1.0) < 0.005


def test_gi_leaves_sharp_surfaces_alone():
    # restir.rs:263-271
    s = scenes.closed_room(0.05)
    _, a = oracle_render(s, 32, 32, 8)
    _, b = oracle_render(s, 32, 32, 8, restir=DI_GI)
    assert np.array_equal(a["film"], b["film"])
