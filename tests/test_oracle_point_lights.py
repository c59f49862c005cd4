"""Point and spot lights in the oracle (point_light_eval, lights.hlsl:15-25, applied by
deferred_lighting_secondary.hlsl:61-70) pinned by closed forms: on a Lambertian plane
(f0 = f90 = 0) seen with one segment the pixel is rho/pi * E * att * cos(theta), with the
range falloff saturate(1 - (d^2/r^2)^2) / max(d^2, 0.001) and, for a spot, the linear
ramp between the outer and inner cone; an occluder between the plane and the light
shadows exactly the points whose segment to the light crosses it, scaled by
1 - shadow_strength, while an occluder beyond the light shadows nothing (the shadow ray
ends at the light) though it blocks a directional light from the same side."""
import math

import numpy as np
import pytest

import oracle
import scenes

RHO = np.array([0.8, 0.6, 0.4], np.float32)
E = np.array([5.0, 4.0, 3.0], np.float32)


def lambert_plane(extra_meshes=(), dir_lights=None, point_lights=None):
    P, N, I = scenes.grid_quad((-15, 0, -15), (30, 0, 0), (0, 0, 30), 4, 4, (0, 1, 0))
    meshes = [(P, N, I)] + list(extra_meshes)
    inst = np.array([(scenes.translate_scale((0, 0, 0)), k, 0) for k in range(len(meshes))], scenes.INSTANCE_DTYPE)
    m = scenes.material(RHO, 0.5, 0.0)
    m[4:7] = 0.0
    m[8:11] = 0.0
    sky, _ = scenes.constant_sky()
    dl = np.zeros((0, 16), np.float32) if dir_lights is None else np.asarray(dir_lights, np.float32).reshape(-1, 16)
    sc = scenes.Scene("point_plane", meshes, inst, m[None], dl, sky, np.zeros(3, np.float32),
                      np.eye(3, dtype=np.float32).reshape(-1))
    sc.point_lights = None if point_lights is None else np.asarray(point_lights, np.float32).reshape(-1, 16)
    return sc


def plane_points(cam):
    """World position (float64) of each pixel centre's primary ray on y = 0, and the ray."""
    P = np.zeros((cam.width * cam.height, 3))
    O = np.zeros_like(P)
    D = np.zeros_like(P)
    for py in range(cam.height):
        for px in range(cam.width):
            o, d = oracle.camera_ray32(cam, px, py)
            o, d = o.astype(np.float64), d.astype(np.float64)
            i = py * cam.width + px
            O[i], D[i] = o, d
            P[i] = o + d * (-o[1] / d[1])
    return P, O, D


def closed_form(P, pos, radius, spot_axis=None, inner=0.0, outer=0.0):
    lv = np.asarray(pos, np.float64) - P
    d2 = (lv * lv).sum(axis=1)
    ldir = lv / np.sqrt(d2)[:, None]
    att = np.clip(1.0 - (d2 / radius ** 2) ** 2, 0.0, 1.0) / np.maximum(d2, 0.001)
    if spot_axis is not None:
        a = np.asarray(spot_axis, np.float64) / np.linalg.norm(spot_axis)
        ci, co = math.cos(math.radians(inner)), math.cos(math.radians(outer))
        att = att * (np.clip(ldir @ a, co, ci) - co) / (ci - co)
    cos = np.maximum(ldir[:, 1], 0.0)
    return (RHO.astype(np.float64) / math.pi)[None, :] * E.astype(np.float64)[None, :] * (att * cos)[:, None], d2


CAM = dict(eye=(0.3, 7.0, 0.4), target=(0.0, 0.0, 0.0), w=64, h=48, fov=90.0)


def camera():
    return scenes.make_camera(CAM["eye"], CAM["target"], CAM["w"], CAM["h"], CAM["fov"])


@pytest.mark.parametrize("spot", [False, True])
def test_point_and_spot_light_on_lambert_plane(oracle_lib, spot):
    pos, radius = (0.7, 2.5, -0.4), 5.0
    kw = dict(spot_axis=(0.2, 1.0, 0.1), inner_deg=25.0, outer_deg=50.0) if spot else {}
    rec = scenes.point_light(E, pos, radius, **kw)
    sc = lambert_plane(point_lights=rec)
    cam = camera()
    sums = oracle_lib.OracleScene(sc).render32(cam, 2, 7, spp=1)["sums"].astype(np.float64)
    P, _, _ = plane_points(cam)
    want, d2 = closed_form(P, pos, radius, kw.get("spot_axis"), kw.get("inner_deg", 0.0), kw.get("outer_deg", 0.0))
    # every pixel sees the plane; the range cut-off (d^2 = r^2) is a zero on both sides
    lit = want[:, 1] > 1e-4 * want[:, 1].max()
    assert 0.05 < lit.mean() < 0.95
    np.testing.assert_allclose(sums[lit], want[lit], rtol=2e-4)
    far = d2 > radius ** 2 * 1.001
    assert far.any() and np.all(sums[far] == 0.0)
    if spot:   # outside the outer cone: exactly dark though within range
        a = np.asarray(kw["spot_axis"]) / np.linalg.norm(kw["spot_axis"])
        lv = np.asarray(pos) - P
        cos_a = (lv @ a) / np.linalg.norm(lv, axis=1)
        outside = (cos_a < math.cos(math.radians(50.0)) - 1e-4) & (d2 < radius ** 2 * 0.999)
        assert outside.any() and np.all(sums[outside] == 0.0)


def quad(y, half, normal=(0, 1, 0)):
    return scenes.grid_quad((-half, y, -half), (2 * half, 0, 0), (0, 0, 2 * half), 1, 1, normal)


@pytest.mark.parametrize("strength,cast", [(1.0, True), (0.25, True), (1.0, False)])
def test_point_light_shadow_ends_at_the_light(oracle_lib, strength, cast):
    """Occluder [-0.5, 0.5]^2 at y = 1 under a light at (0, 2, 0): plane points with |x|, |z|
    < 1 are shadowed (scaled by 1 - strength; unshadowed when the light casts none). A wide
    quad at y = 3, above the light, never shadows the point light but fully shadows a
    directional light straight overhead, whose contribution is therefore absent."""
    pos, radius = (0.0, 2.0, 0.0), 6.0
    rec = scenes.point_light(E, pos, radius, cast_shadow=cast, shadow_strength=strength)
    sun = scenes.dir_light((9.0, 9.0, 9.0), (0, 1, 0))
    sc = lambert_plane(extra_meshes=[quad(1.0, 0.5), quad(3.0, 20.0, (0, -1, 0))], dir_lights=sun[None],
                       point_lights=rec)
    cam = scenes.make_camera((0.2, 2.9, 0.3), (0.0, 0.0, 0.0), 64, 48, 100.0)
    sums = oracle_lib.OracleScene(sc).render32(cam, 2, 7, spp=1)["sums"].astype(np.float64)
    P, O, D = plane_points(cam)
    want, _ = closed_form(P, pos, radius)
    # pixels whose primary ray meets the occluder itself are not plane pixels
    t1 = (1.0 - O[:, 1]) / D[:, 1]
    X1 = O + D * t1[:, None]
    on_occluder = (np.abs(X1[:, 0]) <= 0.5 + 1e-3) & (np.abs(X1[:, 2]) <= 0.5 + 1e-3)
    # the segment to the light crosses y = 1 at P/2 (the light is at height 2 above x = z = 0)
    m = np.maximum(np.abs(P[:, 0]), np.abs(P[:, 2]))
    inside, outside = (m < 0.99) & ~on_occluder, (m > 1.01) & ~on_occluder
    assert inside.sum() > 50 and outside.sum() > 500
    keep = 1.0 if not cast else 1.0 - strength
    np.testing.assert_allclose(sums[outside], want[outside], rtol=2e-4)
    np.testing.assert_allclose(sums[inside], want[inside] * keep, rtol=2e-4, atol=1e-12)
    out = oracle_lib.OracleScene(sc).render32(cam, 2, 7, spp=1, record=True)
    # shadow rays: one per plane pixel towards the sun and one towards the casting point light
    n_plane = int((~on_occluder).sum())
    assert out["counts"][2] >= n_plane * (2 if cast else 1)


def test_directional_and_point_lights_add(oracle_lib):
    """Three directional lights (one from below: no contribution) and a point light: the
    pixel is the sum of each light's closed form."""
    pos, radius = (0.7, 2.5, -0.4), 5.0
    rec = scenes.point_light(E, pos, radius)
    sun_a = scenes.dir_light((1.5, 1.0, 0.5), (0.3, 1.0, 0.2))
    sun_b = scenes.dir_light((0.2, 0.4, 0.9), (-0.5, 0.7, 0.1))
    below = scenes.dir_light((3.0, 3.0, 3.0), (0, -1, 0))
    cam = camera()
    P, _, _ = plane_points(cam)
    want, _ = closed_form(P, pos, radius)
    for s in (sun_a, sun_b):
        d = s[4:7].astype(np.float64)
        want = want + (RHO.astype(np.float64) / math.pi * s[0:3].astype(np.float64) * d[1])[None, :]
    sc = lambert_plane(dir_lights=np.stack([sun_a, below, sun_b]), point_lights=rec)
    sums = oracle_lib.OracleScene(sc).render32(cam, 2, 7, spp=1)["sums"].astype(np.float64)
    np.testing.assert_allclose(sums, want, rtol=2e-4)
