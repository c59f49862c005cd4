"""Pins of the whole per-pixel estimator of the CPU restatement.

Closed forms (a Lambertian plane under a constant sky and a directional light),
sample-order invariances, and convergence of the FP32 transcription to the FP64
oracle on the box-room scene (room + 64 icospheres).
"""
import math

import numpy as np
import pytest

import scenes

RHO = np.array([0.5, 0.4, 0.3])
LS = np.array([0.5, 0.6, 0.8])
E_SUN = np.array([2.0, 1.9, 1.7])
SUN = np.array([0.3, 0.8, 0.2]) / np.linalg.norm([0.3, 0.8, 0.2])


def plane_scene(sun=True):
    p, n, i = scenes.grid_quad((-100, 0, -100), (200, 0, 0), (0, 0, 200), 1, 1, (0, 1, 0))
    mat = scenes.material(RHO, 1.0, 0.0)
    mat[4:7] = 0.0   # f0 = f90 = 0: the lit BSDF reduces to Lambert (lit.hlsl:24)
    mat[8:11] = 0.0
    inst = np.array([(scenes.translate_scale((0, 0, 0)), 0, 0)], scenes.INSTANCE_DTYPE)
    sky, col = scenes.constant_sky(LS)
    light = scenes.dir_light(E_SUN if sun else (0, 0, 0), SUN)
    return scenes.Scene("plane", [(p, n, i)], inst, mat[None], light[None], sky, col,
                        np.eye(3, dtype=np.float32).reshape(-1))


def plane_camera(w=32, h=24):
    return scenes.make_camera((0.0, 2.0, 0.0), (0.0, 0.0, 0.3), w, h, 40.0)


def test_plane_direct_light_is_exact(oracle_lib):
    """One segment: only the sun lights the plane (no sky at the last vertex), no randomness."""
    sc = oracle_lib.OracleScene(plane_scene())
    cam = plane_camera()
    out = sc.render32(cam, max_bounces=2, seed=1, spp=1)
    want = RHO / math.pi * SUN[1] * E_SUN
    np.testing.assert_allclose(out["sums"], np.tile(want, (cam.width * cam.height, 1)), rtol=2e-6)
    assert out["counts"][1] == 0 and out["counts"][3] == 0


@pytest.mark.parametrize("flags", [0, 2])
def test_plane_sky_and_sun_closed_form(oracle_lib, flags):
    """Two segments: E[L] = rho * Ls (sky, by NEE + BSDF sampling with MIS, or BSDF
    sampling alone) + rho/pi * cos(theta_sun) * E_sun."""
    sc = oracle_lib.OracleScene(plane_scene())
    cam = plane_camera()
    spp = 256
    out = sc.render32(cam, max_bounces=3, seed=5, spp=spp)
    mean = out["sums"].mean(axis=0) / spp
    want = RHO * LS + RHO / math.pi * SUN[1] * E_SUN
    np.testing.assert_allclose(mean, want, rtol=0.01)
    out = sc.render32(cam, max_bounces=3, seed=5, spp=spp, flags=flags)
    np.testing.assert_allclose(out["sums"].mean(axis=0) / spp, want, rtol=0.015)


def test_plane_fp64_oracle_closed_form(oracle_lib):
    sc = oracle_lib.OracleScene(plane_scene())
    cam = plane_camera()
    mean, var = sc.render64(cam, 3, np.arange(0, cam.width * cam.height, 37, dtype=np.uint32), 4096)
    want = RHO * LS + RHO / math.pi * SUN[1] * E_SUN
    np.testing.assert_allclose(mean.mean(axis=0), want, rtol=3e-3)


def test_all_misses_see_the_sky_exactly(oracle_lib):
    sc = oracle_lib.OracleScene(plane_scene())
    cam = scenes.make_camera((0.0, 2.0, 0.0), (0.0, 5.0, 0.2), 16, 16, 30.0)  # looking up
    out = sc.render32(cam, max_bounces=5, seed=1, spp=3)
    np.testing.assert_allclose(out["sums"], np.tile(3 * LS, (256, 1)), rtol=1e-6)
    assert out["counts"][1] == 0 and out["counts"][2] == 0


def test_sample_order_and_determinism(c0_oracle):
    cam = scenes.camera_c0(64, 48)
    a = c0_oracle.render32(cam, 5, 7, spp=3, threads=3)["sums"]
    b = c0_oracle.render32(cam, 5, 7, spp=3, threads=8)["sums"]
    assert np.array_equal(a, b)
    c = c0_oracle.render32(cam, 5, 7, spp=1)["sums"]
    c = c0_oracle.render32(cam, 5, 7, spp=2, sample_base=1, sums=c)["sums"]
    assert np.array_equal(a, c)
    d = c0_oracle.render32(cam, 5, 8, spp=3)["sums"]
    assert not np.array_equal(a, d)


def test_recorded_path_structure(c0_oracle):
    cam = scenes.camera_c0(64, 48)
    out = c0_oracle.render32(cam, 5, 7, spp=2, record=True)
    segs = out["segments"]
    bits = out["shadow_bits"]
    assert segs.min() >= 1 and segs.max() <= 4
    assert out["counts"][0] == 2 * 64 * 48
    assert out["counts"][1] == int(segs.sum()) - len(segs)
    n_light = sum(int(((bits[:, b] & 1) != 0).sum()) for b in range(16))
    n_sky = sum(int(((bits[:, b] & 0x80) != 0).sum()) for b in range(16))
    assert out["counts"][2] == n_light and out["counts"][3] == n_sky
    # no sky shadow ray at the last possible vertex (segment max_bounces - 2)
    assert np.all((bits[:, 3] & 0x80) == 0)
    assert np.any((bits[:, 2] & 0x80) != 0)


def test_fp32_converges_to_fp64_c0(c0_oracle):
    """FP32 transcription at 512 spp vs FP64 oracle at 4096 spp on a pixel subset."""
    cam = scenes.camera_c0(80, 60)
    spp32 = 512
    s32 = c0_oracle.render32(cam, 5, 0x5EED0000, spp=spp32)["sums"] / spp32
    rng = np.random.default_rng(0)
    pix = np.sort(rng.choice(cam.width * cam.height, 120, replace=False)).astype(np.uint32)
    m64, v64 = c0_oracle.render64(cam, 5, pix, 4096)
    # standard error of the difference; the floor covers FP32 rounding of a 512-term running sum
    se = np.sqrt(v64 / spp32 + v64 / 4096) + 1e-4 * np.abs(m64) + 1e-6
    z = (s32[pix] - m64) / se
    assert np.median(np.abs(z)) < 1.0
    assert np.mean(np.abs(z) > 4) < 0.02
    # aggregate over the subset: the summed difference is within 4 combined standard errors
    zsum = (s32[pix] - m64).sum(axis=0) / np.sqrt((se ** 2).sum(axis=0))
    assert np.all(np.abs(zsum) < 4), zsum


RHO_ODD = np.array([0.8, 0.7, 0.6])
RHO_EVEN = np.array([0.1, 0.2, 0.3])


def checker_scene():
    """A Lambertian floor whose material is the world checkerboard (odd cells RHO_ODD,
    even cells RHO_EVEN) under a vertical sun; black sky."""
    p, n, i = scenes.grid_quad((-100, 0, -100), (200, 0, 0), (0, 0, 200), 1, 1, (0, 1, 0))
    mats = []
    for rho in (RHO_ODD, RHO_EVEN):
        m = scenes.material(rho, 1.0, 0.0)
        m[4:7] = 0.0
        m[8:11] = 0.0
        mats.append(m)
    mats[0].view(np.uint32)[13] = 1   # checkerboard, even cells take material 1
    mats[0].view(np.uint32)[14] = 1
    inst = np.array([(scenes.translate_scale((0, 0, 0)), 0, 0)], scenes.INSTANCE_DTYPE)
    sky, col = scenes.constant_sky((0.0, 0.0, 0.0))
    light = scenes.dir_light(E_SUN, (0.0, 1.0, 0.0), cast_shadow=False)
    return scenes.Scene("checker", [(p, n, i)], inst, np.stack(mats), light[None], sky, col,
                        np.eye(3, dtype=np.float32).reshape(-1))


def pixel_floor_points(cam):
    """FP64 hit points of the pixel-centre camera rays on the plane y = 0."""
    w, h = cam.width, cam.height
    x, y = np.meshgrid((np.arange(w) + 0.5) / w, (np.arange(h) + 0.5) / h)
    ndc = np.stack([2 * x - 1, 1 - 2 * y, np.ones_like(x), np.ones_like(x)], -1).reshape(-1, 4)
    v = ndc @ cam.inv_proj.astype(np.float64).T
    v = v[:, :3] / np.linalg.norm(v[:, :3], axis=1, keepdims=True)
    d = v @ cam.inv_view[:3, :3].astype(np.float64).T
    o = cam.inv_view[:3, 3].astype(np.float64)
    t = -o[1] / d[:, 1]
    return o + d * t[:, None]


def test_checker_floor_closed_form(oracle_lib):
    """Each pixel shows rho/pi * E of its cell: odd (int(floor(x)) ^ int(floor(z))) the
    material itself, even its pattern material; pixels near a cell edge are skipped."""
    sc = oracle_lib.OracleScene(checker_scene())
    cam = scenes.make_camera((0.3, 6.0, 0.2), (0.0, 0.0, -1.5), 96, 64, 60.0)
    out = sc.render32(cam, max_bounces=2, seed=1, spp=1)["sums"]
    P = pixel_floor_points(cam)
    fx, fz = np.floor(P[:, 0]), np.floor(P[:, 2])
    odd = ((fx.astype(np.int64) ^ fz.astype(np.int64)) & 1) == 1
    clear = (np.minimum(P[:, 0] - fx, fx + 1 - P[:, 0]) > 1e-3) & (np.minimum(P[:, 2] - fz, fz + 1 - P[:, 2]) > 1e-3)
    want = np.where(odd[:, None], RHO_ODD, RHO_EVEN) / math.pi * E_SUN
    assert clear.mean() > 0.95 and 0.3 < odd[clear].mean() < 0.7      # both kinds of cell in view
    np.testing.assert_allclose(out[clear], want[clear], rtol=2e-6)
    # the FP64 oracle shades the same cells
    pix = np.flatnonzero(clear)[::17].astype(np.uint32)
    mean, _ = sc.render64(cam, 2, pix, 1)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)   # the scene stores float32
    want64 = np.where(odd[:, None], f32(RHO_ODD), f32(RHO_EVEN)) / math.pi * f32(E_SUN)
    np.testing.assert_allclose(mean, want64[pix], rtol=1e-12)
