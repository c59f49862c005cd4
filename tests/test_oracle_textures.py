"""Pins of the oracle's texture-driven materials by closed forms and textbook rules.

* level-0 sampling: bilinear interpolation reproduces a linear ramp exactly, nearest
  filtering picks floor(u W), Vulkan's repeat / mirrored-repeat / clamp-to-edge
  addressing at and beyond the edges, RGBA8 UNORM as c / 255 and sRGB through the
  IEC 61966-2-1 decode (the alpha channel stays linear);
* base-colour textures: a Lambertian plane under a vertical sun, one segment, shows
  rho(texel) / pi * E in every pixel whose hit lies inside a texel (nearest filter);
* normal maps: a constant tilted tangent-space normal on that plane turns the shading
  normal to normalize(n.x T + n.y B + n.z N) with B = normalize(cross(N, T)) * w, so the
  sun's cosine is the closed form for either bitangent sign;
* alpha test: a cut-out plane over a lit floor shows the floor through its holes (the
  shadow rays pass the holes too) and itself elsewhere;
* stochastic translucency: each ray's surface is present with probability opacity, so a
  translucent plane of opacity a over a floor has mean a rho_t/pi E + (1-a)^2 rho_f/pi E
  (the floor's shadow ray crosses the translucent plane once more).
"""
import math

import numpy as np
import pytest

import scenes

E = np.array([2.0, 1.5, 1.0])


def lambert(rho, two_sided=True):
    m = scenes.material(rho, 1.0, 0.0, two_sided)
    m[4:7] = 0.0    # f0 = f90 = 0: the lit BSDF reduces to Lambert
    m[8:11] = 0.0
    return m


def plane(y, half, uv=True):
    """Quad at height y over [-half, half]^2 facing +y, texture coordinates (x, z) -> ((x + half) / 2half, (z + half) / 2half)."""
    p = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], np.float32)
    n = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    i = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    t = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return p, n, i, t


def scene_of(meshes, uvs, tans, mats, inst_mats, textures, sun_dir=(0.0, 1.0, 0.0)):
    inst = np.array([(scenes.translate_scale((0, 0, 0)), k, m) for k, m in enumerate(inst_mats)],
                    scenes.INSTANCE_DTYPE)
    sky, col = scenes.constant_sky((0.0, 0.0, 0.0))
    return scenes.Scene("tex", meshes, inst, np.stack(mats), scenes.dir_light(E, sun_dir)[None], sky, col,
                        np.eye(3, dtype=np.float32).reshape(-1), mesh_uv=uvs, mesh_tan=tans, textures=textures)


def down_camera(w=40, h=40, y=3.0):
    return scenes.make_camera((0.0, y, 0.0), (0.0, 0.0, 0.0), w, h, 40.0, up=(0, 0, -1))


def primary_hits(oracle_lib, cam, y):
    """Where each pixel's camera ray meets the plane at height y (float64)."""
    pts = []
    for py in range(cam.height):
        for px in range(cam.width):
            o, d = oracle_lib.camera_ray32(cam, px, py)
            o, d = o.astype(np.float64), d.astype(np.float64)
            pts.append(o + d * ((y - o[1]) / d[1]))
    return np.array(pts)


# ---------------------------------------------------------------- sampling

def tex_scene(textures):
    p, n, i, t = plane(0.0, 1.0)
    return scene_of([(p, n, i)], [t], None, [lambert((0.5, 0.5, 0.5))], [0], textures)


@pytest.mark.parametrize("address", [0, 1, 2])
def test_bilinear_ramp_and_addressing(oracle_lib, address):
    W = 8
    ramp = np.zeros((1, W, 4), np.float32)
    ramp[0, :, 0] = np.arange(W)                 # r = texel index
    ramp[0, :, 1] = 2.0 * np.arange(W) + 1.0     # g: another linear function
    sc = oracle_lib.OracleScene(tex_scene([scenes.Texture(ramp, scenes.TEXEL_RGBA32F, 1, address, 0)]))
    u = np.linspace(0.0625 + 1e-3, 1.0 - 0.0625 - 1e-3, 97)     # inside the texel centres
    got = sc.tex_sample(0, np.stack([u, np.full_like(u, 0.5)], 1))
    x = u * W - 0.5
    np.testing.assert_allclose(got[:, 0], x, atol=2e-6)
    np.testing.assert_allclose(got[:, 1], 2.0 * x + 1.0, atol=4e-6)
    # beyond the edges: u = 0 is half way between texel -1 and texel 0
    edge = sc.tex_sample(0, [[0.0, 0.5], [1.0, 0.5], [-0.25, 0.5], [1.25, 0.5]])[:, 0]
    want = {0: [0.5 * (W - 1), 0.5 * (W - 1), 5.5, 1.5],     # repeat: texel -1 is texel W - 1
            1: [0.0, W - 1.0, 1.5, 5.5],                     # mirrored repeat: texel -1 is texel 0
            2: [0.0, W - 1.0, 0.0, W - 1.0]}[address]        # clamp to edge
    np.testing.assert_allclose(edge, want, atol=2e-6)
    # repeat and mirrored repeat have periods 1 and 2
    if address == 0:
        np.testing.assert_allclose(sc.tex_sample(0, [[0.3, 0.5]]), sc.tex_sample(0, [[2.3, 0.5]]), atol=2e-6)
    if address == 1:
        np.testing.assert_allclose(sc.tex_sample(0, [[0.3, 0.5]]), sc.tex_sample(0, [[-0.3, 0.5]]), atol=1e-6)


def test_nearest_and_rgba8_decoding(oracle_lib):
    rng = np.random.default_rng(9)
    t8 = rng.integers(0, 256, (4, 6, 4), dtype=np.uint8)
    sc = oracle_lib.OracleScene(tex_scene([scenes.Texture(t8, scenes.TEXEL_RGBA8_UNORM, 0, 0, 0),
                                           scenes.Texture(t8, scenes.TEXEL_RGBA8_SRGB, 0, 0, 0)]))
    ii, jj = np.meshgrid(np.arange(6), np.arange(4), indexing="xy")
    uv = np.stack([(ii.ravel() + 0.3) / 6.0, (jj.ravel() + 0.8) / 4.0], 1)
    texel = t8[jj.ravel(), ii.ravel()].astype(np.float64)
    unorm = sc.tex_sample(0, uv)
    np.testing.assert_array_equal(unorm, (texel / 255.0).astype(np.float32))
    srgb = sc.tex_sample(1, uv).astype(np.float64)
    c = texel[:, :3] / 255.0
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    np.testing.assert_allclose(srgb[:, :3], lin, rtol=1e-6, atol=1e-9)
    np.testing.assert_array_equal(srgb[:, 3].astype(np.float32), (texel[:, 3] / 255.0).astype(np.float32))  # linear alpha
    assert srgb[:, :3].max() <= 1.0 and np.all(np.diff(np.sort(lin.ravel())) >= 0)


# ---------------------------------------------------------------- shading

def test_base_color_texture_on_a_lit_plane(oracle_lib):
    rng = np.random.default_rng(4)
    cells = rng.integers(20, 250, (2, 2, 4), dtype=np.uint8)
    p, n, i, t = plane(0.0, 1.0)
    m = lambert((0.5, 0.5, 0.5))
    m.view(np.uint32)[scenes.M_BASE_TEX] = 1
    sc = oracle_lib.OracleScene(scene_of([(p, n, i)], [t], None, [m], [0],
                                         [scenes.Texture(cells, scenes.TEXEL_RGBA8_UNORM, 0, 2, 2)]))
    cam = down_camera()
    out = sc.render32(cam, max_bounces=2, seed=1, spp=1)["sums"]
    hit = primary_hits(oracle_lib, cam, 0.0)
    u, v = (hit[:, 0] + 1.0) / 2.0, (hit[:, 2] + 1.0) / 2.0
    inside = (np.abs(u - 0.5) > 0.02) & (np.abs(v - 0.5) > 0.02) & (u > 0.01) & (u < 0.99) & (v > 0.01) & (v < 0.99)
    ci, cj = np.floor(u * 2).astype(int).clip(0, 1), np.floor(v * 2).astype(int).clip(0, 1)
    rho = cells[cj, ci, :3].astype(np.float64) / 255.0
    want = rho / math.pi * E
    assert inside.sum() > 500
    np.testing.assert_allclose(out[inside], want[inside], rtol=2e-6)


@pytest.mark.parametrize("w", [1.0, -1.0])
def test_normal_map_tilts_the_shading_normal(oracle_lib, w):
    a = 0.5
    nt = np.array([0.3 * math.sin(a), math.sin(a), math.cos(a)])
    nt /= np.linalg.norm(nt)
    texel = np.round((nt * 0.5 + 0.5) * 255.0).astype(np.uint8)
    nmap = np.tile(np.append(texel, 255).astype(np.uint8), (2, 2, 1))
    p, n, i, t = plane(0.0, 1.0)
    tan = np.tile(np.array([1.0, 0.0, 0.0, w], np.float32), (4, 1))
    m = lambert((0.6, 0.5, 0.4))
    m.view(np.uint32)[scenes.M_NORMAL_TEX] = 1
    b = 0.4
    sun = (0.0, math.cos(b), math.sin(b))
    sc = oracle_lib.OracleScene(scene_of([(p, n, i)], [t], [tan], [m], [0],
                                         [scenes.Texture(nmap, scenes.TEXEL_RGBA8_UNORM, 1, 0, 0)], sun_dir=sun))
    cam = down_camera(24, 24)
    out = sc.render32(cam, max_bounces=2, seed=1, spp=1)["sums"]
    # tangent frame of the plane: T = +x, N = +y, B = normalize(cross(N, T)) * w = (0, 0, -w)
    q = texel.astype(np.float64) / 255.0 * 2.0 - 1.0
    T, N, B = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, -w])
    nn = q[0] * T + q[1] * B + q[2] * N
    nn /= np.linalg.norm(nn)
    cos = max(float(nn @ np.array(sun)), 0.0)
    want = np.array([0.6, 0.5, 0.4]) / math.pi * E * cos
    hit = primary_hits(oracle_lib, cam, 0.0)
    on = (np.abs(hit[:, 0]) < 0.99) & (np.abs(hit[:, 2]) < 0.99)
    assert on.sum() > 300
    np.testing.assert_allclose(out[on], np.tile(want, (int(on.sum()), 1)), rtol=3e-6)
    # the other bitangent sign gives a measurably different cosine
    nn2 = q[0] * T - q[1] * B + q[2] * N
    assert abs(float(nn2 @ np.array(sun)) / np.linalg.norm(nn2) - cos) > 0.1


def two_planes(top_mat, textures, top_uv=True):
    p1, n1, i1, t1 = plane(0.0, 1.0)
    p2, n2, i2, t2 = plane(-1.0, 50.0)
    floor = lambert((0.2, 0.7, 0.3))
    return scene_of([(p1, n1, i1), (p2, n2, i2)], [t1, t2] if top_uv else None, None, [top_mat, floor], [0, 1],
                    textures)


def test_alpha_test_cut_out(oracle_lib):
    mask = np.zeros((1, 2, 4), np.uint8)
    mask[0, 1, 3] = 255          # right half (u > 0.5) opaque, left half cut out
    mask[..., :3] = 200
    top = lambert((0.9, 0.1, 0.1))
    w = top.view(np.uint32)
    w[scenes.M_BLEND] = scenes.BLEND_ALPHA_TEST
    w[scenes.M_OPACITY_TEX], w[scenes.M_OPACITY_CHANNEL], w[scenes.M_OPACITY_STEP] = 1, 3, 1
    top[scenes.M_STEP_AT], top[scenes.M_STEP_BELOW], top[scenes.M_STEP_ABOVE] = 0.5, 0.0, 1.0
    sc = oracle_lib.OracleScene(two_planes(top, [scenes.Texture(mask, scenes.TEXEL_RGBA8_UNORM, 0, 2, 2)]))
    cam = down_camera(40, 40)
    out = sc.render32(cam, max_bounces=2, seed=3, spp=1)["sums"]
    hit = primary_hits(oracle_lib, cam, 0.0)
    u, v = (hit[:, 0] + 1.0) / 2.0, (hit[:, 2] + 1.0) / 2.0
    on_top = (u > 0.0) & (u < 1.0) & (v > 0.0) & (v < 1.0)
    solid = on_top & (u > 0.52)
    hole = on_top & (u < 0.48)
    assert solid.sum() > 100 and hole.sum() > 100
    np.testing.assert_allclose(out[solid], np.tile(np.array([0.9, 0.1, 0.1]) / math.pi * E, (solid.sum(), 1)),
                               rtol=2e-6)
    # through the hole: the floor, lit through the hole (vertical sun)
    np.testing.assert_allclose(out[hole], np.tile(np.array([0.2, 0.7, 0.3]) / math.pi * E, (hole.sum(), 1)),
                               rtol=2e-6)
    # the same material made opaque hides the floor everywhere on the plane
    w[scenes.M_BLEND] = scenes.BLEND_OPAQUE
    sc2 = oracle_lib.OracleScene(two_planes(top, [scenes.Texture(mask, scenes.TEXEL_RGBA8_UNORM, 0, 2, 2)]))
    out2 = sc2.render32(cam, max_bounces=2, seed=3, spp=1)["sums"]
    np.testing.assert_allclose(out2[hole], out[solid][:1].repeat(hole.sum(), 0), rtol=2e-6)


def test_stochastic_translucency_mean(oracle_lib):
    a = 0.3
    top = lambert((0.9, 0.1, 0.1))
    top.view(np.uint32)[scenes.M_BLEND] = scenes.BLEND_TRANSLUCENT
    top[scenes.M_OPACITY] = a
    sc = oracle_lib.OracleScene(two_planes(top, [], top_uv=False))
    cam = down_camera(96, 96, y=1.5)
    hit = primary_hits(oracle_lib, cam, 0.0)
    on_top = (np.abs(hit[:, 0]) < 0.98) & (np.abs(hit[:, 2]) < 0.98)
    spp = 8
    out = sc.render32(cam, max_bounces=2, seed=5, spp=spp, flags=1)["sums"] / spp   # jittered samples
    rt, rf = np.array([0.9, 0.1, 0.1]), np.array([0.2, 0.7, 0.3])
    want = a * rt / math.pi * E + (1 - a) * (1 - a) * rf / math.pi * E
    vals = out[on_top]
    mean = vals.mean(axis=0)
    se = vals.std(axis=0, ddof=1) / math.sqrt(len(vals))
    assert np.all(np.abs(mean - want) / se < 4.5), (mean, want)
    # the draw is per ray: a ray sees the plane or the floor, never a blend
    one = sc.render32(cam, max_bounces=2, seed=5, spp=1)["sums"][on_top]
    lit_top = np.isclose(one, rt / math.pi * E, rtol=1e-5).all(1)
    lit_floor = np.isclose(one, rf / math.pi * E, rtol=1e-5).all(1)
    dark = (one == 0.0).all(1)
    assert np.all(lit_top | lit_floor | dark)
    assert abs(lit_top.mean() - a) < 0.05 and abs(lit_floor.mean() - (1 - a) ** 2) < 0.05
