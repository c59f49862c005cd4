"""The oracle's primary-hit GBuffer (bi_pt_render_gbuffer) pinned by what the engine's
packing fixes: its float -> half conversion equals numpy's; on a plane every field has
a closed form (half of the base colour and roughness, the octahedral code of +-y, the
tangent code of the default tangent, the rg11b10 words of f0 and f90 as unorm16 halves,
anisotropy and 1/ior as unorm8, zeros on misses, the two-sided flip); decoding with the
engine's own unpacking (pack.hlsl, frame.hlsl) recovers arbitrary normals and tangents."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
import scenes


def test_half_conversion_matches_numpy(oracle_lib):
    rng = np.random.default_rng(3)
    xs = np.concatenate([
        rng.normal(size=20000).astype(np.float32) * np.float32(3.0),
        (rng.uniform(-1, 1, 20000) * 2.0 ** rng.integers(-30, 17, 20000)).astype(np.float32),
        np.array([0.0, -0.0, 1.0, 65504.0, 65519.99, 65520.0, 1e6, -1e6, np.inf, -np.inf, 2.0 ** -24, 2.0 ** -25,
                  3 * 2.0 ** -26, 2.0 ** -14, 2.0 ** -14 - 2.0 ** -26, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11],
                 np.float32)])
    with np.errstate(over="ignore"):
        want = xs.astype(np.float16).view(np.uint16)
    got = np.array([oracle.half32(x) for x in xs], np.uint16)
    assert np.array_equal(got, want)


def plane_scene(two_sided, base=(0.3, 0.5, 0.7), rough=0.6, f0=(0.04, 0.5, 0.9), f90=(1.0, 0.8, 0.25), aniso=0.2,
                ior=1.5):
    P, N, I = scenes.grid_quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), 2, 2, (0, 1, 0))
    inst = np.array([(scenes.translate_scale((0, 0, 0)), 0, 0)], scenes.INSTANCE_DTYPE)
    m = scenes.material(base, rough, 0.0, two_sided=two_sided, anisotropy=aniso)
    m[4:7] = f0
    m[8:11] = f90
    m[11] = ior
    sky, col = scenes.constant_sky()
    return scenes.Scene("gbuf_plane", [(P, N, I)], inst, m[None], scenes.dir_light((1, 1, 1), (0, 1, 0))[None],
                        sky, col, np.eye(3, dtype=np.float32).reshape(-1))


def h(x):
    return int(np.float32(x).astype(np.float16).view(np.uint16))


def rg11b10(c):
    c = np.asarray(c, np.float32)
    def u(x):   # the shaders' float -> uint (saturating)
        return int(x) if x > 0 else 0
    return u(c[0] * np.float32(2047)) | (u(c[1] * np.float32(2047)) << 11) | (u(c[2] * np.float32(1023)) << 22)


def unorm16(halfword):
    return int(Fraction(halfword * 65535, 65536) + Fraction(1, 2))   # round half up (no exact tie but 32768)


@pytest.mark.parametrize("two_sided,from_below", [(True, False), (True, True), (False, True)])
def test_plane_closed_form(oracle_lib, two_sided, from_below):
    sc = plane_scene(two_sided)
    eye, target = ((0.0, 3.0, 0.5), (0.0, 0.0, 0.0)) if not from_below else ((0.0, -3.0, 0.5), (0.0, 0.0, 0.0))
    cam = scenes.make_camera(eye, target, 48, 36, 120.0)
    o = oracle_lib.OracleScene(sc)
    pix = np.arange(cam.width * cam.height, dtype=np.uint32)
    g = o.gbuffer32(cam, pix)
    hitmask = g["base_color"][:, 3] != 0
    assert 100 < hitmask.sum() < len(pix)
    for k in g:
        assert not g[k][~hitmask].any(), k          # misses hold zeros
    b = g["base_color"][hitmask]
    assert np.all(b == np.array([h(0.3), h(0.5), h(0.7), h(1.0)], np.uint16))
    flipped = two_sided and from_below
    # oct code of (0, 1, 0) is (0, 1); the flipped normal -N = (-0, -1, -0) keeps its
    # signed zeros: (-0, -1).  The default tangent (1, 0, 0) in the frame of +-y: frame.x =
    # (1, 0, 0), frame.y = (0, 0, -+1), projected (1, 0) -> third component +1
    nr = g["normal_roughness"][hitmask]
    want = [h(-0.0), h(-1.0)] if flipped else [h(0.0), h(1.0)]
    assert np.all(nr == np.array(want + [h(1.0), h(0.6)], np.uint16)), np.unique(nr, axis=0)
    p0, p1 = rg11b10((0.04, 0.5, 0.9)), rg11b10((1.0, 0.8, 0.25))
    want_f = [unorm16(p0 & 0xFFFF), unorm16(p0 >> 16), unorm16(p1 & 0xFFFF), unorm16(p1 >> 16)]
    assert np.all(g["fresnel"][hitmask] == np.array(want_f, np.uint16))
    want_m = [round(0.2 * 255), round(255 / 1.5), 0, 1]
    assert np.all(g["material_0"][hitmask] == np.array(want_m, np.uint8))


def oct_decode(f):
    n = np.array([f[0], f[1], 1.0 - abs(f[0]) - abs(f[1])])
    t = min(max(-n[2], 0.0), 1.0)
    n[0] += -t if n[0] >= 0 else t
    n[1] += -t if n[1] >= 0 else t
    return n / np.linalg.norm(n)


def frame(n):
    sign = 1.0 if n[2] > 0 else -1.0
    a = -1.0 / (sign + n[2])
    b = n[0] * n[1] * a
    return np.array([1 + sign * n[0] * n[0] * a, sign * b, -sign * n[0]]), np.array([b, sign + n[1] * n[1] * a, -n[1]])


def unpack(packed):
    N = oct_decode(packed[:2])
    sign = -1.0 if packed[2] < 0 else 1.0
    px = sign * packed[2] * 2.0 - 1.0
    py = sign * (1.0 - abs(px))
    fx, fy = frame(N)
    T = fx * px + fy * py
    return N, T / np.linalg.norm(T)


def test_decoded_normals_and_tangents(oracle_lib):
    """Flat triangles of random orientation (one instance each, its rotation carrying the
    default tangent): the engine's unpacking of the half-precision code gives back the face
    normal and the tangent within half precision."""
    rng = np.random.default_rng(11)
    worst_n = worst_t = 0.0
    for trial in range(24):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        P = np.array([[-5, 0, -5], [5, 0, -5], [0, 0, 5]], np.float32)
        N = np.tile(np.array([0, 1, 0], np.float32), (3, 1))
        I = np.array([[0, 2, 1]], np.uint32)
        xf = np.zeros(12, np.float32)
        xf.reshape(3, 4)[:, :3] = R
        inst = np.array([(xf, 0, 0)], scenes.INSTANCE_DTYPE)
        mat = scenes.material((0.5, 0.5, 0.5), 0.5, 0.0, two_sided=False)
        sky, col = scenes.constant_sky()
        sc = scenes.Scene("tri", [(P, N, I)], inst, mat[None], scenes.dir_light((1, 1, 1), (0, 1, 0))[None], sky,
                          col, np.eye(3, dtype=np.float32).reshape(-1))
        n_w = R @ np.array([0.0, 1.0, 0.0])
        t_w = R @ np.array([1.0, 0.0, 0.0])
        cam = scenes.make_camera(tuple(n_w * 4.0 + 1e-3), (0.0, 0.0, 0.0), 8, 8, 30.0,
                                 up=tuple(t_w))
        g = oracle_lib.OracleScene(sc).gbuffer32(cam, np.arange(64, dtype=np.uint32))
        hit = g["base_color"][:, 3] != 0
        assert hit.sum() > 30
        for code in g["normal_roughness"][hit]:
            vals = code.view(np.float16).astype(np.float64)
            Nd, Td = unpack(vals[:3])
            worst_n = max(worst_n, np.linalg.norm(Nd - n_w))
            worst_t = max(worst_t, np.linalg.norm(Td - t_w))
    assert worst_n < 2e-3 and worst_t < 4e-3, (worst_n, worst_t)


# ----------------------------------------------------- drift mode of the restatement
QUANTIZE = oracle.GBUFFER_QUANTIZE


def test_drift_mode_keeps_exactly_representable_surfaces(oracle_lib):
    """On a plane whose surface survives the GBuffer exactly (half-exact base colour and
    roughness, zero f0/f90 and anisotropy, normal +Y), the drift mode renders the same image;
    a base colour that is not a half moves it by half precision only."""
    cam = scenes.make_camera((0.3, 3.0, 0.4), (0.0, 0.0, 0.0), 24, 18, 70.0)
    exact = oracle_lib.OracleScene(plane_scene(True, base=(0.5, 0.25, 0.75), rough=0.5, f0=(0, 0, 0), f90=(0, 0, 0),
                                               aniso=0.0))
    a = exact.render32(cam, 3, 5, spp=2)["sums"]
    b = exact.render32(cam, 3, 5, spp=2, flags=QUANTIZE)["sums"]
    assert a.sum() > 0 and np.array_equal(a, b)
    off = oracle_lib.OracleScene(plane_scene(True, base=(0.3, 0.3, 0.3), rough=0.5, f0=(0, 0, 0), f90=(0, 0, 0),
                                             aniso=0.0))
    a = off.render32(cam, 2, 5, spp=1)["sums"].astype(np.float64)
    b = off.render32(cam, 2, 5, spp=1, flags=QUANTIZE)["sums"].astype(np.float64)
    rel = np.abs(b - a) / np.maximum(a, 1e-12)
    lit_px = a[:, 0] > 1e-3
    # one segment: the image scales with the base colour, 0.3 -> half(0.3) = 0.300048828125
    want = abs(float(np.float32(0.3).astype(np.float16)) / float(np.float32(0.3)) - 1)
    assert lit_px.all()
    np.testing.assert_allclose(rel[lit_px], want, rtol=1e-3)


def test_drift_of_the_reference_gbuffer_on_c0(oracle_lib, c0_scene):
    """The reference's GBuffer quantisation against the FP32 path on the C0 room: a small,
    non-zero difference of the mean (f90 = 1 does not survive the unorm16 split: it reads
    back as 2046/2047)."""
    cam = scenes.camera_c0(48, 36)
    o = oracle_lib.OracleScene(c0_scene)
    a = o.render32(cam, 5, 9, spp=4)["sums"].astype(np.float64)
    b = o.render32(cam, 5, 9, spp=4, flags=QUANTIZE)["sums"].astype(np.float64)
    assert np.isfinite(b).all()
    drift = abs(b.sum() / a.sum() - 1)
    print(f"c0 drift of the mean radiance: {drift:.2e}", flush=True)
    assert 0 < drift < 2e-2
