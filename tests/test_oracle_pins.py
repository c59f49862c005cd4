"""Independent pins of three oracle functions that the rest of the suite only uses:

* instance flattening (``ora_get_world``): world positions and normals of every
  triangle against a float64 numpy evaluation of the 3x4 row-major instance
  transform (``rhi::AccelerationStructureInstanceDesc::transform``,
  include/bisemutum/rhi/accel.hpp:48-55) and of the inverse transpose that
  hit.hlsl:155 applies to normals (WorldToObject4x3 * n), including rotated,
  non-uniformly scaled, sheared and mirrored (det < 0) instances;
* the radiance-cache image mode (``ora_render32_cached``): closed forms -- a cache at
  segment 0 replaces every hit with exactly its constant, and inside a closed
  Lambertian box with no light reaching it the cached image's mean is rho^b * k for a
  cache at segment b (the path throughput after b Lambertian bounces has mean rho^b);
* the training-sample mode (``ora_train_samples``): on a Lambertian plane under a
  constant sky and a sun, every record is the plane hit (position on the plane, unit
  direction to the camera, normal, roughness, base colour) and the targets' mean is
  rho * L_sky + rho / pi * cos(theta_sun) * E_sun; inside the closed box every slot
  records a wall hit, at the segment asked for.

A transposed transform, a flipped cofactor sign, an off-by-one cache segment or a
cache contribution not weighted by the throughput each fails one of these.
"""
import json
import math
import os

import numpy as np
import pytest

import scenes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scene_basic")

RHO = np.array([0.5, 0.4, 0.3])
LS = np.array([0.5, 0.6, 0.8])
E_SUN = np.array([2.0, 1.9, 1.7])
SUN = np.array([0.3, 0.8, 0.2]) / np.linalg.norm([0.3, 0.8, 0.2])


# ---------------------------------------------------------------- flattening

def world_numpy(scene):
    """Float64 world triangles of a scenes.Scene: positions A p + t, normals |det A| A^-T n."""
    vs, ns, ms = [], [], []
    for inst in scene.instances:
        m = inst["xform"].astype(np.float64).reshape(3, 4)
        a, t = m[:, :3], m[:, 3]
        p, n, idx = scene.meshes[int(inst["mesh"])]
        p = p.astype(np.float64)
        n = n.astype(np.float64)
        idx = np.asarray(idx, np.int64).reshape(-1, 3)
        wp = p @ a.T + t
        cof_t = abs(np.linalg.det(a)) * np.linalg.inv(a).T   # = sign(det) * cofactor matrix
        wn = n @ cof_t.T
        vs.append(wp[idx].reshape(-1, 9))
        ns.append(wn[idx].reshape(-1, 9))
        ms.append(np.full(len(idx), int(inst["material"]), np.uint32))
    return np.concatenate(vs), np.concatenate(ns), np.concatenate(ms)


def rot(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def xform(a, t):
    m = np.zeros((3, 4))
    m[:, :3] = a
    m[:, 3] = t
    return m.reshape(-1).astype(np.float32)


def instanced_scene():
    """One random-triangle mesh and one icosphere placed by five instances: identity, an
    oblique rotation with non-uniform scale, a shear, a mirror (det < 0) and a mirror
    composed with rotation and non-uniform scale."""
    rng = np.random.default_rng(11)
    tv = rng.uniform(-1, 1, size=(40, 3)).astype(np.float32)
    tn = rng.normal(size=(40, 3))
    tn = (tn / np.linalg.norm(tn, axis=1, keepdims=True)).astype(np.float32)
    tidx = rng.integers(0, 40, size=(30, 3)).astype(np.uint32)
    sphere = scenes.icosphere(1)
    a_rot = rot((0.3, 1.0, -0.5), 0.9) @ np.diag([2.0, 0.5, 1.3])
    a_shear = np.array([[1.0, 0.7, 0.0], [0.0, 1.0, -0.4], [0.2, 0.0, 1.0]])
    a_mirror = np.diag([-1.0, 1.0, 1.0])
    a_mix = rot((1.0, 0.2, 0.4), -1.7) @ np.diag([0.7, -2.2, 1.1])
    assert np.linalg.det(a_mirror) < 0 and np.linalg.det(a_mix) < 0
    inst = np.array([
        (xform(np.eye(3), (0, 0, 0)), 0, 0),
        (xform(a_rot, (3.0, -1.0, 2.5)), 0, 1),
        (xform(a_shear, (-2.0, 4.0, 0.5)), 1, 2),
        (xform(a_mirror, (0.0, 1.0, -3.0)), 1, 3),
        (xform(a_mix, (1.5, 2.0, 7.0)), 0, 4),
    ], scenes.INSTANCE_DTYPE)
    mats = np.stack([scenes.material((0.5, 0.5, 0.5), 0.5, 0.0) for _ in range(5)])
    sky, col = scenes.constant_sky()
    return scenes.Scene("instanced", [(tv, tn, tidx), sphere], inst, mats, scenes.dir_light((1, 1, 1), (0, 1, 0))[None],
                        sky, col, np.eye(3, dtype=np.float32).reshape(-1))


def assert_world_matches(oracle_lib, scene):
    got_v, got_n, got_m = oracle_lib.OracleScene(scene).world()
    want_v, want_n, want_m = world_numpy(scene)
    assert got_v.shape == want_v.shape
    assert np.array_equal(got_m, want_m)
    # float32 arithmetic of a 3x4 product: a few ulps of the largest term
    scale_v = np.maximum(np.abs(want_v), 1.0)
    assert np.max(np.abs(got_v - want_v) / scale_v) < 4e-6
    scale_n = np.maximum(np.linalg.norm(want_n.reshape(-1, 3), axis=1), 1e-3).repeat(3).reshape(-1, 9)
    assert np.max(np.abs(got_n - want_n) / scale_n) < 4e-6


def test_flattening_matches_numpy_on_oblique_and_mirrored_instances(oracle_lib):
    assert_world_matches(oracle_lib, instanced_scene())


def test_flattening_matches_numpy_on_box_room(oracle_lib, c0_scene):
    assert_world_matches(oracle_lib, c0_scene)


def test_flattening_matches_numpy_on_example_project(oracle_lib):
    import test_assets
    assert_world_matches(oracle_lib, test_assets.fixture_scene(64, 32).scene)


def test_flattening_pin_rejects_plausible_mistakes(oracle_lib):
    """The numpy restatement separates the right transform from the mistakes it is meant
    to catch: a transposed linear part and an unflipped cofactor of a mirror."""
    scene = instanced_scene()
    got_v, got_n, _ = oracle_lib.OracleScene(scene).world()
    bad = dataclass_copy(scene)
    x = bad.instances["xform"].copy()
    for i in range(len(x)):
        m = x[i].reshape(3, 4).astype(np.float64)
        m[:, :3] = m[:, :3].T
        x[i] = m.reshape(-1)
    bad.instances["xform"] = x
    tv, _, _ = world_numpy(bad)
    assert np.max(np.abs(tv - got_v)) > 0.1
    # a normal transformed by the cofactor without the determinant's sign points inwards
    _, want_n, _ = world_numpy(scene)
    mirrored = slice(30 + 30 + 80, 30 + 30 + 80 + 80)   # triangles of instance 3 (the mirror)
    cof_unsigned = -want_n[mirrored]
    assert np.all((cof_unsigned.reshape(-1, 3) * got_n[mirrored].reshape(-1, 3)).sum(1) < 0)


def dataclass_copy(scene):
    import dataclasses
    return dataclasses.replace(scene, instances=scene.instances.copy())


# ---------------------------------------------------------------- cache modes

def plane_scene():
    p, n, i = scenes.grid_quad((-100, 0, -100), (200, 0, 0), (0, 0, 200), 1, 1, (0, 1, 0))
    mat = scenes.material(RHO, 1.0, 0.0)
    mat[4:7] = 0.0   # f0 = f90 = 0: the lit BSDF reduces to Lambert (lit.hlsl:24)
    mat[8:11] = 0.0
    inst = np.array([(scenes.translate_scale((0, 0, 0)), 0, 0)], scenes.INSTANCE_DTYPE)
    sky, col = scenes.constant_sky(LS)
    return scenes.Scene("plane", [(p, n, i)], inst, mat[None], scenes.dir_light(E_SUN, SUN)[None], sky, col,
                        np.eye(3, dtype=np.float32).reshape(-1))


def closed_box_scene():
    """The box room with a ceiling (every ray from inside hits a wall), one Lambertian
    material of albedo RHO, a sky that no ray reaches and no directional light."""
    half, height = 6.0, 4.0
    room = scenes.merge([scenes.room_c0(half, height),
                         scenes.grid_quad((-half, height, -half), (0, 0, 2 * half), (2 * half, 0, 0), 2, 2,
                                          (0, -1, 0))])
    mat = scenes.material(RHO, 1.0, 0.0)
    mat[4:7] = 0.0
    mat[8:11] = 0.0
    inst = np.array([(scenes.translate_scale((0, 0, 0)), 0, 0)], scenes.INSTANCE_DTYPE)
    sky, col = scenes.constant_sky(LS)
    return scenes.Scene("closed_box", [room], inst, mat[None], np.zeros((0, 16), np.float32), sky, col,
                        np.eye(3, dtype=np.float32).reshape(-1))


def plane_camera(w=32, h=24):
    return scenes.make_camera((0.0, 2.0, 0.0), (0.0, 0.0, 0.3), w, h, 40.0)


def box_camera(w=24, h=16):
    return scenes.make_camera((0.5, 2.0, -1.0), (2.0, 1.2, 3.0), w, h, 70.0)


K = np.array([0.7, 1.9, 3.1], np.float32)


def test_cache_at_segment_zero_is_the_constant_exactly(oracle_lib):
    sc = oracle_lib.OracleScene(plane_scene())
    # half the image looks at the plane, half at the sky
    cam = scenes.make_camera((0.0, 2.0, 0.0), (0.0, 2.0, 5.0), 16, 12, 60.0)
    spp = 3
    sums = sc.render32_cached(cam, max_bounces=4, seed=9, spp=spp, cache_bounce=0, cache_rgb=K)
    o, d = zip(*[oracle_lib.camera_ray32(cam, px, py) for py in range(cam.height) for px in range(cam.width)])
    hits = np.array(d)[:, 1] < 0.0          # rays going down meet the plane
    assert 0 < hits.sum() < len(hits)
    np.testing.assert_array_equal(sums[hits], np.tile(spp * K, (int(hits.sum()), 1)))
    np.testing.assert_allclose(sums[~hits], np.tile(spp * LS, (int((~hits).sum()), 1)), rtol=1e-6)
    # negative cache outputs are clamped to zero
    neg = sc.render32_cached(cam, max_bounces=4, seed=9, spp=1, cache_bounce=0, cache_rgb=-K)
    assert np.all(neg[hits] == 0.0)


@pytest.mark.parametrize("bounce", [1, 2])
def test_cache_image_mean_in_a_closed_box_is_rho_power_b_times_k(oracle_lib, bounce):
    """No light reaches the inside of a closed box, so the plain image is black and a path
    contributes only what the cache returns: throughput x k at its hit of segment `bounce`,
    whose mean is rho^bounce * k (each Lambertian bounce's sample weight f cos / pdf has mean
    rho; VNDF sampling at roughness 1 covers the whole hemisphere)."""
    sc = oracle_lib.OracleScene(closed_box_scene())
    cam = box_camera()
    spp = 64
    plain = sc.render32(cam, max_bounces=5, seed=3, spp=2)["sums"]
    assert np.all(plain == 0.0)
    sums = sc.render32_cached(cam, max_bounces=5, seed=3, spp=spp, cache_bounce=bounce, cache_rgb=K)
    per = sums / spp                      # per-pixel mean over spp samples
    mean = per.mean(axis=0)
    se = per.std(axis=0, ddof=1) / math.sqrt(len(per))
    want = RHO ** bounce * K
    z = (mean - want) / se
    assert np.all(np.abs(z) < 4.5), (mean, want, z)
    # the other segment's value is many standard errors away (off-by-one detection)
    other = RHO ** (bounce + 1) * K
    assert np.all(np.abs(mean - other) / se > 8)


def test_training_records_and_targets_on_the_plane(oracle_lib):
    """Training slots record the plane hit of segment 0 and gather the suffix radiance:
    mean target = rho * L_sky + rho / pi * cos(theta_sun) * E_sun (sky by NEE + MIS)."""
    sc = oracle_lib.OracleScene(plane_scene())
    cam = plane_camera(40, 30)
    n = 4096
    rec, tgt = sc.train_samples(cam, max_bounces=3, seed=21, cache_bounce=0, n=n)
    tag = rec[:, 13].copy().view(np.uint32)
    assert np.all(tag == 1)
    eye = np.array([0.0, 2.0, 0.0])
    np.testing.assert_allclose(rec[:, 1], np.zeros(n, np.float32), atol=1e-6)    # on the plane y = 0 (o + t d)
    wo = eye - rec[:, 0:3].astype(np.float64)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    np.testing.assert_allclose(rec[:, 3:6], wo, atol=2e-6)                         # towards the camera
    np.testing.assert_allclose(rec[:, 6:9], np.tile([0, 1, 0], (n, 1)), atol=1e-7)  # normal
    np.testing.assert_array_equal(rec[:, 9], np.ones(n, np.float32))               # roughness
    np.testing.assert_array_equal(rec[:, 10:13], np.tile(RHO.astype(np.float32), (n, 1)))
    # jittered random pixels: records spread over the camera's footprint on the plane
    assert len(np.unique(rec[:, 0])) > n // 2
    want = RHO * LS + RHO / math.pi * SUN[1] * E_SUN
    mean = tgt.mean(axis=0)
    se = tgt.std(axis=0, ddof=1) / math.sqrt(n)
    assert np.all(np.abs(mean - want) / se < 4.5), (mean, want)
    # segment 1 never hits (extension rays leave the plane): nothing is recorded there,
    # and a slot whose tag is 0 is not a training sample (its target is not used)
    rec1, _ = sc.train_samples(cam, max_bounces=3, seed=21, cache_bounce=1, n=256)
    assert np.all(rec1 == 0.0)


def test_training_records_in_the_closed_box(oracle_lib):
    """Inside a closed box a training slot reaches its record segment unless its path
    ended before (a VNDF sample below the horizon weighs zero); a recorded point lies on a
    wall, its direction points back along a segment inside the box, and with no light
    reaching the inside every recorded target is zero.  Slots that never reach the
    segment record nothing."""
    sc = oracle_lib.OracleScene(closed_box_scene())
    cam = box_camera()
    eye = np.array([0.5, 2.0, -1.0])
    prev = None
    for bounce in (0, 1, 2):
        rec, tgt = sc.train_samples(cam, max_bounces=5, seed=4, cache_bounce=bounce, n=512)
        tag = rec[:, 13].copy().view(np.uint32)
        ok = tag == 1
        assert np.all((tag == 0) | ok)
        assert np.all(rec[~ok] == 0.0)
        if bounce == 0:
            assert np.all(ok)
        else:
            assert 0.2 < ok.mean() <= prev.mean()   # a slot that reaches segment b reached b - 1
            assert np.all(prev[ok])
        prev = ok
        assert np.all(tgt[ok] == 0.0)
        p = rec[ok, 0:3].astype(np.float64)
        on_wall = ((np.abs(np.abs(p[:, 0]) - 6.0) < 1e-5) | (np.abs(np.abs(p[:, 2]) - 6.0) < 1e-5) |
                   (np.abs(p[:, 1]) < 1e-5) | (np.abs(p[:, 1] - 4.0) < 1e-5))
        assert np.all(on_wall)
        # the normal faces the inside and the incoming direction lies on its side
        assert np.all((rec[ok, 3:6] * rec[ok, 6:9]).sum(1) > 0)
        wo = eye - p
        wo /= np.linalg.norm(wo, axis=1, keepdims=True)
        if bounce == 0:
            np.testing.assert_allclose(rec[ok, 3:6], wo, atol=3e-6)
        else:
            # the previous vertex is a wall point, not the camera
            assert np.mean(np.abs((wo * rec[ok, 3:6]).sum(1) - 1.0) > 1e-3) > 0.9
