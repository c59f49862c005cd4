"""Textured rect lights in the restatement (rect_light_eval_ltc with a texture,
shaders/renderer/lights.hlsl:429-445,495-511) and the mip chain the engine generates for
texture assets (shaders/core/mipmap.hlsl, src/graphics/command_helpers.cpp:66-199).

Pinned by closed forms:
* log2 against numpy (exact at powers of two);
* SampleLevel's level choice on textures whose levels are constants (nearest halves to the
  lower level, linear blends, the LOD bias and clamps, magnification vs minification);
* the lookup's (u, v, level) against the ray / plane intersection in float64 on a
  non-square light (a swap of width and height, or of the corners, fails);
* the integrals' directions: straight up under a centred light, the vector irradiance
  sum(theta_i n_i) of the polygon off-centre, towards a small distant light for the
  specular lobe of any invertible transform (a transposed or inverted transform fails);
* renders: a constant 1 texture leaves the image unchanged, a constant 1/2 halves it
  exactly;
* the mip helper: block averages, the odd-size footprints with zero reads beyond the
  level, 8-bit rounding and the sRGB round trip.
"""
import dataclasses

import numpy as np
import pytest

import bi_pt
import oracle
import scenes
from test_oracle_point_lights import lambert_plane

UP = np.array([0.0, 1.0, 0.0])


def const_levels(values, w=8, h=4, fmt=scenes.TEXEL_RGBA32F):
    """Levels of a w x h texture, level l filled with values[l]."""
    out, lw, lh = [], w, h
    for v in values:
        out.append(np.full((lh, lw, 4), v, np.float32))
        lw, lh = max(lw // 2, 1), max(lh // 2, 1)
    return out


def tex_scene(textures, rect=None):
    sc = lambert_plane()
    if rect is None:
        rect = scenes.rect_light((5, 5, 5), (0.0, 2.0, 0.0), 4.0, 2.0, (1, 0, 0), (0, 0, 1))   # faces down
    sc.rect_lights = np.asarray(rect, np.float32).reshape(-1, 28)
    sc.ltc = scenes.smooth_ltc(4)
    sc.rect_textures = textures
    return sc


def test_log2_matches_numpy():
    x = np.concatenate([np.float32(2.0) ** np.arange(0, 100, dtype=np.float32),
                        np.random.default_rng(1).uniform(0, 40, 20000).astype(np.float32) ** 2 + 1,
                        np.array([1.0, 1.0000001, 1.4142135, 1.4142137, 3.0e38], np.float32)]).astype(np.float32)
    got = oracle.log2_32(x).astype(np.float64)
    want = np.log2(x.astype(np.float64))
    p2 = np.arange(100)
    assert np.array_equal(got[:100], p2)          # exact at powers of two
    assert np.all(np.abs(got - want) <= 2.5e-7 * np.maximum(1.0, np.abs(want)))
    assert oracle.log2_32(np.array([np.inf], np.float32))[0] == 128.0


@pytest.mark.parametrize("mode", [0, 1])
def test_level_choice_on_constant_levels(mode):
    t = scenes.RectTexture(const_levels([1, 2, 3, 4]), mipmap_mode=mode)
    o = oracle.OracleScene(tex_scene([t]))
    s = lambda lv: float(o.rect_tex_sample32(0, 0.3, 0.6, lv)[0])
    assert s(-3.0) == 1 and s(0.0) == 1 and s(7.5) == 4
    if mode == 0:   # nearest: ceil(d + 1/2) - 1, halves to the lower level
        assert (s(0.5), s(0.5001), s(1.49), s(1.5), s(1.51), s(2.9)) == (1, 2, 2, 2, 3, 4)
    else:
        assert s(1.25) == 2.25 and s(2.5) == 3.5 and s(0.75) == 1.75 and s(3.0) == 4
    # the sampler's bias and clamps
    tb = dataclasses.replace(t, lod_bias=1.0, lod_max=2.0, lod_min=0.0)
    ob = oracle.OracleScene(tex_scene([tb]))
    assert float(ob.rect_tex_sample32(0, 0.3, 0.6, 0.0)[0]) == 2     # 0 + 1
    assert float(ob.rect_tex_sample32(0, 0.3, 0.6, 2.0)[0]) == 3     # 3 clamped to lod_max 2
    tm = dataclasses.replace(t, lod_min=1.0)
    assert float(oracle.OracleScene(tex_scene([tm])).rect_tex_sample32(0, 0.3, 0.6, 0.0)[0]) == 2


def test_magnification_and_minification_filters():
    lv0 = np.zeros((2, 2, 4), np.float32)
    lv0[..., 0] = [[0.0, 1.0], [2.0, 3.0]]
    t = scenes.RectTexture([lv0], mag_filter=0, min_filter=1, mipmap_mode=1)
    o = oracle.OracleScene(tex_scene([t]))
    # lambda <= 0: the magnification filter (nearest texel); lambda > 0 on one level: the
    # minification filter (bilinear: the centre averages the four texels)
    assert float(o.rect_tex_sample32(0, 0.5, 0.5, 0.0)[0]) == 3.0
    assert float(o.rect_tex_sample32(0, 0.5, 0.5, 0.5)[0]) == 1.5
    assert float(o.rect_tex_sample32(0, 0.2, 0.2, 0.0)[0]) == 0.0


def plane_hit(rect, P, d):
    """(u, v, t) of the ray P + t d on the light's plane in float64: u along position2 ->
    position3 over the width, v along position2 -> position1 over the height, clamped."""
    r = np.asarray(rect, np.float64)
    p1, p2, p3, n = r[12:15], r[16:19], r[20:23], r[24:27]
    t = abs(np.dot(P - p2, n)) / abs(np.dot(d, n))
    x = P + d * t - p2
    ex, ey = p3 - p2, p1 - p2
    return (np.clip(np.dot(x, ex) / np.dot(ex, ex), 0, 1), np.clip(np.dot(x, ey) / np.dot(ey, ey), 0, 1), t)


def test_lookup_coordinates_and_level():
    rect = scenes.rect_light((5, 5, 5), (0.4, 2.0, -0.3), 4.0, 1.5, (1, 0, 0.2), (0, 0.1, -1))
    rect = scenes.textured_rect_light(rect, 0, 16, 4, 4.0, 1.5)   # 16 x 4 texels on 4 x 1.5
    t = scenes.RectTexture(const_levels([1, 2, 3, 4, 5], 16, 4))
    o = oracle.OracleScene(tex_scene([t], rect))
    rng = np.random.default_rng(5)
    for _ in range(200):
        P = np.array([rng.uniform(-2, 2), rng.uniform(0, 1.0), rng.uniform(-2, 2)])
        d = rng.normal(size=3); d[1] = abs(d[1]) + 0.2; d /= np.linalg.norm(d)
        rough = rng.uniform(0.05, 1.0)
        _, uvl = o.rect_tex_lookup32(0, d.astype(np.float32), P.astype(np.float32), rough)
        u, v, tt = plane_hit(rect, P.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64))
        assert abs(uvl[0] - u) < 2e-5 and abs(uvl[1] - v) < 2e-5
        lvl = np.log2(max(1.0, tt * rough * 4.0))
        assert abs(uvl[2] - lvl) < 2e-5
    # a direction along the plane sees nothing
    along = np.cross(rect[24:27].astype(np.float64), (0.0, 0.0, 1.0))
    rgb, _ = o.rect_tex_lookup32(0, (along / np.linalg.norm(along)).astype(np.float32), np.zeros(3, np.float32), 1.0)
    assert not rgb.any()


def polygon_vector(rect, P):
    """Vector irradiance direction of the quad (position3, 2, 1, 0) seen from P: sum over
    edges of the angle times the unit normal of the edge's great circle (float64)."""
    r = np.asarray(rect, np.float64)
    q = [r[20:23], r[16:19], r[12:15], r[8:11]]
    v = [(x - P) / np.linalg.norm(x - P) for x in q]
    s = np.zeros(3)
    for k in range(4):
        a, b = v[k], v[(k + 1) % 4]
        c = np.cross(a, b)
        s += np.arccos(np.clip(np.dot(a, b), -1, 1)) * c / np.linalg.norm(c)
    return s / np.linalg.norm(s)


def angle(a, b):
    return float(np.arccos(np.clip(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b), -1, 1)))


def test_integral_directions():
    rect = scenes.textured_rect_light(
        scenes.rect_light((5, 5, 5), (0.0, 2.0, 0.0), 2.0, 2.0, (1, 0, 0), (0, 0, 1)), 0, 8, 8, 2.0, 2.0)
    t = scenes.RectTexture(const_levels([1, 1, 1, 1], 8, 8))
    o = oracle.OracleScene(tex_scene([t], rect))
    V = np.array([0.3, 0.9, 0.1]); V /= np.linalg.norm(V)
    args = (0.4, 0.4, (0.5, 0.5, 0.5), (0.04, 0.04, 0.04), (1, 1, 1))
    # centred under the light: straight up
    _, terms, dirs = o.rect_light_scene32(0, (0, 0, 0), UP, V, *args)
    assert terms[0] > 0 and angle(dirs[0], UP) < 1e-5
    # off-centre: the polygon's vector irradiance (the fit of theta / sin(theta) within 2e-3)
    rng = np.random.default_rng(8)
    for _ in range(40):
        P = np.array([rng.uniform(-3, 3), 0.0, rng.uniform(-3, 3)])
        _, terms, dirs = o.rect_light_scene32(0, P, UP, V, *args)
        assert terms[0] > 0 and angle(dirs[0], polygon_vector(rect, P)) < 3e-3
    # a small distant light: both lobes look up a point of the light (within its angular
    # radius of the centre), whatever the transform
    small = scenes.textured_rect_light(
        scenes.rect_light((5, 5, 5), (1.5, 40.0, -2.0), 1.0, 0.8, (1, 0, 0.1), (0, 0.1, 1)), 0, 8, 8, 1.0, 0.8)
    radius = np.arctan(0.5 * np.hypot(1.0, 0.8) / 39.9)
    os_ = oracle.OracleScene(tex_scene([t], small))
    for rx, ry in ((0.3, 0.3), (0.6, 0.2), (0.15, 0.7)):
        for vz in (0.2, 0.6, 0.95):
            V = np.array([np.sqrt(1 - vz * vz) * 0.6, vz, np.sqrt(1 - vz * vz) * -0.8])
            P = np.array([0.3, 0.0, -0.2])
            _, terms, dirs = os_.rect_light_scene32(0, P, UP, V, rx, ry, *args[2:])
            want = np.array([1.5, 40.0, -2.0]) - P
            assert terms[0] > 0 and angle(dirs[0], want) < radius
            if terms[1] > 0:
                assert angle(dirs[1], want) < radius


@pytest.mark.parametrize("fmt", [scenes.TEXEL_RGBA32F, scenes.TEXEL_RGBA8_UNORM])
def test_constant_textures_scale_the_image(oracle_lib, fmt):
    cam = scenes.make_camera((0.3, 4.0, 0.4), (0.0, 0.0, 0.0), 24, 18, 90.0)
    base = tex_scene([])
    plain = oracle_lib.OracleScene(base).render32(cam, 3, 11, spp=1)["sums"]
    rect = scenes.textured_rect_light(base.rect_lights[0], 0, 8, 4, 4.0, 2.0)
    one = 1.0 if fmt == scenes.TEXEL_RGBA32F else 255
    dt = np.float32 if fmt == scenes.TEXEL_RGBA32F else np.uint8
    lv = [l.astype(dt) for l in const_levels([one] * 4)]
    white = oracle_lib.OracleScene(tex_scene([scenes.RectTexture(lv, format=fmt)], rect)).render32(cam, 3, 11, spp=1)
    assert np.array_equal(white["sums"], plain)
    assert plain.sum() > 0
    if fmt == scenes.TEXEL_RGBA32F:
        half = oracle_lib.OracleScene(tex_scene([scenes.RectTexture(const_levels([0.5] * 4))], rect))
        assert np.array_equal(half.render32(cam, 3, 11, spp=1)["sums"], plain * np.float32(0.5))


# ------------------------------------------------------------ the mip chain helper
def footprint_mips(a, levels):
    """mipmap.hlsl in float64: destination (x, y) averages source (2x + i, 2y + j), i, j < 2,
    or < 3 along an axis whose destination size is odd; reads beyond the source are zero."""
    out = [a.astype(np.float64)]
    for _ in range(levels - 1):
        s = out[-1]
        h, w = s.shape[:2]
        dh, dw = max(h // 2, 1), max(w // 2, 1)
        pad = np.zeros((2 * dh + 3, 2 * dw + 3, 4))
        pad[:h, :w] = s
        nx, ny = (3 if dw % 2 else 2), (3 if dh % 2 else 2)
        d = np.zeros((dh, dw, 4))
        for j in range(ny):
            for i in range(nx):
                d += pad[j:j + 2 * dh:2, i:i + 2 * dw:2]
        out.append(d / (nx * ny))
    return out


@pytest.mark.parametrize("w,h", [(4, 4), (8, 2), (6, 6), (3, 3), (7, 5), (16, 1)])
def test_mip_chain_float(w, h):
    a = np.random.default_rng(w * 31 + h).uniform(0, 4, (h, w, 4)).astype(np.float32)
    n = int(np.floor(np.log2(max(w, h)))) + 1
    got = bi_pt.rect_texture_mips(a, n, scenes.TEXEL_RGBA32F)
    want = footprint_mips(a, n)
    assert np.array_equal(got[0], a)
    for g, e in zip(got, want):
        assert g.shape == e.shape
        np.testing.assert_allclose(g, e, rtol=2e-6, atol=1e-6)
    if (w, h) == (4, 4):
        # 2 x 2 block averages, then the 1 x 1 level reading a 3 x 3 footprint of which the
        # four texels of level 1 exist: their sum over nine
        np.testing.assert_allclose(got[1][0, 0], a[:2, :2].reshape(-1, 4).mean(0), rtol=1e-6)
        np.testing.assert_allclose(got[2][0, 0], got[1].reshape(-1, 4).sum(0) / 9, rtol=1e-6)
    if (w, h) == (3, 3):
        np.testing.assert_allclose(got[1][0, 0], a.reshape(-1, 4).mean(0), rtol=1e-6)


def test_mip_chain_8bit():
    a = np.full((4, 4, 4), 128, np.uint8)
    got = bi_pt.rect_texture_mips(a, 3, scenes.TEXEL_RGBA8_UNORM)
    assert (got[1] == 128).all()
    assert (got[2] == 57).all()           # round(128 * 4 / 9)
    # sRGB: decoded, averaged, encoded: constant texels keep their value
    for c in range(256):
        s = bi_pt.rect_texture_mips(np.full((4, 4, 4), c, np.uint8), 2, scenes.TEXEL_RGBA8_SRGB)
        assert (s[1] == c).all()
    # colour averaged in linear light, alpha as stored
    b = np.zeros((2, 2, 4), np.uint8)
    b[0, :, :] = 255
    s = bi_pt.rect_texture_mips(b, 2, scenes.TEXEL_RGBA8_SRGB)[1][0, 0]
    lin = (2 * 1.0 + 0) / 9        # two white texels of four, over the nine-texel footprint
    enc = 1.055 * lin ** (1 / 2.4) - 0.055
    assert s[0] == round(enc * 255) and s[3] == round(2 * 255 / 9)


def test_mip_chain_rejects_bad_sizes():
    a = np.zeros((4, 4, 4), np.float32)
    with pytest.raises(bi_pt.PtError):
        bi_pt.rect_texture_mips(a, 4, scenes.TEXEL_RGBA32F)
