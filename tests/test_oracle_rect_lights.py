"""Rect lights in the oracle (rect_light_eval_ltc, lights.hlsl:161-513, and
surface_eval_lut_lit, material/lit.hlsl:37-58) pinned by what the method fixes:

* the diffuse term integrates the clipped quad edge by edge, sum_e theta_e n_e.z =
  2 * integral of cos(theta) over the light's solid angle (Lambert's formula), with
  theta / sin(theta) by the reference's rational fit: against the closed form of a
  parallel rectangle and a quadrature of a light straddling the horizon;
* one-sided lights are dark from behind, two-sided ones are not;
* with identity transforms the specular integral is the cosine integral in every azimuth
  quadrant and for either roughness order (the quadrant folding, the row flips and the
  winding reversals undo each other);
* the table lookup: tables that store their own texel coordinates return the coordinate
  get_ltc_tex3d_coord maps to, trilinearly;
* the lit combination (1 - F) base / pi * E * I_d + (f0 norm + f90 fresnel) * E * I_s;
* a render adds the rect lighting of every shaded vertex."""
import math

import numpy as np
import pytest

import oracle
import scenes


def corner_ff(a, b, h):
    """Form factor from a point to a rectangle [0, a] x [0, b] in a parallel plane at height h
    above it, with one corner straight above the point."""
    x, y = a / h, b / h
    return (x / math.sqrt(1 + x * x) * math.atan(y / math.sqrt(1 + x * x)) +
            y / math.sqrt(1 + y * y) * math.atan(x / math.sqrt(1 + y * y))) / (2 * math.pi)


def parallel_ff(x0, x1, z0, z1, h):
    """Form factor to the rectangle [x0, x1] x [z0, z1] at height h (signed corner sums)."""
    f = 0.0
    for xs, sx in ((x1, 1), (x0, -1)):
        for zs, sz in ((z1, 1), (z0, -1)):
            f += sx * sz * math.copysign(1, xs) * math.copysign(1, zs) * corner_ff(abs(xs), abs(zs), h)
    return f


TABLES_ID = oracle.ltc_tables(*scenes.identity_ltc())
UP = np.array([0.0, 1.0, 0.0], np.float32)


def lit(rect, P, N, V, rx=0.3, ry=0.3, base=(0.8, 0.6, 0.4), f0=(0, 0, 0), f90=(0, 0, 0), tables=TABLES_ID):
    return oracle.rect_light32(rect, tables, P, N, V, rx, ry, base, f0, f90)


@pytest.mark.parametrize("cx,cz,w,h,height", [(0.0, 0.0, 2.0, 2.0, 2.0), (0.7, -0.4, 1.5, 3.0, 1.2),
                                               (2.5, 1.0, 1.0, 1.0, 0.8), (-0.2, 0.3, 6.0, 4.0, 0.5)])
def test_diffuse_integral_of_a_parallel_rectangle(cx, cz, w, h, height):
    rect = scenes.rect_light((5, 5, 5), (cx, height, cz), w, h, (1, 0, 0), (0, 0, 1))   # normal -y: faces down
    _, terms = lit(rect, [0, 0, 0], UP, UP)
    ff = parallel_ff(cx - w / 2, cx + w / 2, cz - h / 2, cz + h / 2, height)
    # sum of theta_e n_e.z = 2 * integral of cos = 2 pi ff, up to the theta / sin(theta) fit
    assert terms[0] == pytest.approx(2 * math.pi * ff, rel=2e-3)


def test_one_and_two_sided():
    down = scenes.rect_light((5, 5, 5), (0.3, 1.5, 0.2), 2.0, 1.0, (1, 0, 0), (0, 0, 1))
    # the same rectangle with normal +y: facing away
    up = scenes.rect_light((5, 5, 5), (0.3, 1.5, 0.2), 1.0, 2.0, (0, 0, 1), (1, 0, 0))
    up2 = scenes.rect_light((5, 5, 5), (0.3, 1.5, 0.2), 1.0, 2.0, (0, 0, 1), (1, 0, 0), two_sided=True)
    rgb_d, t_d = lit(down, [0, 0, 0], UP, UP)
    rgb_u, t_u = lit(up, [0, 0, 0], UP, UP)
    rgb_2, t_2 = lit(up2, [0, 0, 0], UP, UP)
    assert t_d[0] > 0.1 and t_u[0] == 0.0 and np.all(rgb_u == 0.0)
    assert t_2[0] == pytest.approx(t_d[0], rel=1e-6)


def quad_cos_integral(rect, P, N, n=600):
    """2 * integral of cos(theta) over the part of the light above the tangent plane at P
    (midpoint quadrature over the light's surface, float64)."""
    p0, p1, p2, p3 = (rect[8 + 4 * k:11 + 4 * k].astype(np.float64) for k in range(4))
    ex, ey = p3 - p2, p1 - p2   # light's width and height directions from corner p2
    nl = np.cross(ex, ey); area = np.linalg.norm(nl); nl /= area
    s = (np.arange(n) + 0.5) / n
    u, v = np.meshgrid(s, s, indexing="ij")
    X = p2 + u[..., None] * ex + v[..., None] * ey
    r = X - np.asarray(P, np.float64)
    d2 = (r * r).sum(-1); d = np.sqrt(d2)
    cos_p = (r @ np.asarray(N, np.float64)) / d
    cos_l = np.abs(r @ nl) / d
    f = np.where(cos_p > 0, cos_p * cos_l / d2, 0.0)
    return 2.0 * f.mean() * area


@pytest.mark.parametrize("tilt", [0.0, 0.6])
def test_horizon_clipping(tilt):
    """A light standing on the tangent plane and reaching below it: only its part above counts."""
    y_axis = np.array([0.0, math.cos(tilt), math.sin(tilt)])
    rect = scenes.rect_light((3, 3, 3), (0.4, 0.3, 1.2), 2.0, 2.0, (1, 0, 0), y_axis, two_sided=True)
    _, terms = lit(rect, [0, 0, 0], UP, UP)
    want = quad_cos_integral(rect, [0, 0, 0], UP)
    assert want > 0.05
    assert terms[0] == pytest.approx(want, rel=3e-3)


@pytest.mark.parametrize("rx,ry", [(0.3, 0.3), (0.5, 0.1), (0.1, 0.5)])
@pytest.mark.parametrize("az", [0.3, 1.9, 3.6, 5.4])   # one view azimuth in each quadrant
def test_identity_transform_specular_is_the_cosine_integral(rx, ry, az):
    rect = scenes.rect_light((5, 5, 5), (0.6, 1.4, -0.3), 1.6, 1.1, (0.8, 0.0, 0.6), (0.0, -0.3, 1.0))
    rect2 = scenes.rect_light((5, 5, 5), (0.6, 1.4, -0.3), 1.6, 1.1, (0.8, 0.0, 0.6), (0.0, -0.3, 1.0), True)
    N = np.array([0.1, 0.97, 0.2]); N /= np.linalg.norm(N)
    fr = frame(N)
    V = math.cos(0.8) * N + math.sin(0.8) * (math.cos(az) * fr[0] + math.sin(az) * fr[1])
    for r in (rect, rect2):
        _, terms = lit(r, [0.05, 0.0, -0.1], N, V, rx, ry)
        assert terms[0] > 0.01
        assert terms[1] == pytest.approx(terms[0], rel=1e-5)


def frame(n):
    """The shading frame the restatement builds from n (Duff et al. 2017), float64."""
    sign = 1.0 if n[2] > 0 else -1.0
    a = -1.0 / (sign + n[2]); b = n[0] * n[1] * a
    return (np.array([1 + sign * n[0] * n[0] * a, sign * b, -sign * n[0]]),
            np.array([b, sign + n[1] * n[1] * a, -n[1]]))


def test_table_lookup_coordinates():
    """Tables storing their texel's (x, y, z) coordinates: the fetch returns the texel-space
    position of get_ltc_tex3d_coord's two depth slices, blended by the roughness fraction."""
    i = np.arange(4096)
    coords = np.stack([i % 8, (i // 8) % 8, i // 64, np.zeros(4096)], 1).astype(np.float32)
    norm = np.stack([i % 8 + 100 * ((i // 8) % 8), i // 64], 1).astype(np.float32)
    t = oracle.ltc_tables(coords, coords[:, [1, 2, 0, 3]], coords[:, [2, 0, 1, 3]], norm)
    rng = np.random.default_rng(4)
    for u3, u2, u1, u0 in rng.uniform(0, 1, (200, 4)):
        m, b = oracle.ltc_fetch32(t, u3, u2, u1, u0)
        ws = u0 * 7; wf = math.floor(ws); wc = min(math.floor(ws + 1), 7); w = ws - wf
        x, y = u3 * 7, u2 * 7
        z = (u1 * 7 + 8 * wf) * (1 - w) + (u1 * 7 + 8 * wc) * w
        np.testing.assert_allclose(m[0], [x, y, z], atol=2e-4)
        np.testing.assert_allclose(m[1], [y, z, x], atol=2e-4)
        np.testing.assert_allclose(m[2], [z, x, y], atol=2e-4)
        np.testing.assert_allclose(b, [x + 100 * y, z], atol=2e-3)


def test_lit_combination():
    rect = scenes.rect_light((4.0, 3.0, 2.0), (0.2, 1.3, 0.1), 1.4, 0.9, (1, 0, 0), (0, 0, 1))
    t = oracle.ltc_tables(*scenes.identity_ltc(norm=(0.7, 0.2)))
    N = UP.astype(np.float64)
    V = np.array([0.3, 0.8, -0.2]); V /= np.linalg.norm(V)
    base, f0, f90 = np.array([0.8, 0.5, 0.3]), np.array([0.04, 0.3, 0.6]), np.array([1.0, 0.9, 0.5])
    rgb, terms = oracle.rect_light32(rect, t, [0, 0, 0], N, V, 0.2, 0.2, base, f0, f90)
    ndotv = float(N @ V)
    F = f0 + (f90 - f0) * (1 - ndotv) ** 5
    E = np.array([4.0, 3.0, 2.0])
    want = (1 - F) * base / math.pi * E * terms[0] + (f0 * 0.7 + f90 * 0.2) * E * terms[1]
    np.testing.assert_allclose(rgb, want, rtol=1e-5)
    assert terms[2] == pytest.approx(0.7) and terms[3] == pytest.approx(0.2)


def test_render_adds_rect_lighting(oracle_lib):
    """One segment on a plane under a rect light (black sky, no other light): each pixel is
    the rect lighting at its point."""
    from test_oracle_point_lights import RHO, lambert_plane, plane_points
    sc = lambert_plane()
    sc.rect_lights = scenes.rect_light((6, 5, 4), (0.5, 1.5, -0.2), 2.0, 1.2, (1, 0, 0.3), (0, 0, 1))[None]
    sc.ltc = scenes.smooth_ltc(2)
    cam = scenes.make_camera((0.3, 4.0, 0.4), (0.0, 0.0, 0.0), 40, 30, 90.0)
    sums = oracle_lib.OracleScene(sc).render32(cam, 2, 7, spp=1)["sums"].astype(np.float64)
    P, O, _ = plane_points(cam)
    tables = oracle.ltc_tables(*sc.ltc)
    want = np.zeros_like(sums)
    for k in range(len(P)):
        V = O[k] - P[k]; V /= np.linalg.norm(V)
        want[k] = lit(sc.rect_lights[0], P[k], UP, V, 1.0, 1.0, RHO, (0, 0, 0), (0, 0, 0), tables)[0]
    lit_px = want[:, 0] > 1e-3
    assert lit_px.mean() > 0.5
    np.testing.assert_allclose(sums[lit_px], want[lit_px], rtol=1e-3)


def local_to_world(fr_t, fr_b, n, v):
    return v[0] * fr_t + v[1] * fr_b + v[2] * n


@pytest.mark.parametrize("mirror", ["y", "x", "xy", "diag"])
def test_specular_term_is_mirror_symmetric(mirror):
    """A two-sided light and the view mirrored together about a plane through the normal (or
    the roughnesses swapped with the mirror about x = y): the table lookup folds to the same
    coordinate and the unfolding flips must undo the mirror, whatever the transform."""
    tables = oracle.ltc_tables(*scenes.smooth_ltc(5))
    N = np.array([0.0, 0.0, 1.0])
    t, b = frame(N)
    corners = np.array([[0.3, 0.5, 1.1], [-0.6, 0.9, 1.3], [-0.8, 0.1, 0.9], [0.2, -0.3, 0.7]])   # local, planar below
    # make the quad planar: project on the plane through the first three corners
    n_q = np.cross(corners[1] - corners[0], corners[2] - corners[0]); n_q /= np.linalg.norm(n_q)
    corners[3] -= n_q * ((corners[3] - corners[0]) @ n_q)
    v_local = np.array([0.45, 0.25, 0.86]); v_local /= np.linalg.norm(v_local)
    rx, ry = 0.35, 0.2
    F = {"y": np.diag([1.0, -1.0, 1.0]), "x": np.diag([-1.0, 1.0, 1.0]), "xy": np.diag([-1.0, -1.0, 1.0]),
         "diag": np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])}[mirror]

    def rect_of(cs):
        r = np.zeros(28, np.float32)
        r[0:3] = 1.0
        r[3:4].view(np.int32)[0] = -1
        r[7:8].view(np.uint32)[0] = 1   # two-sided
        for k in range(4):
            r[8 + 4 * k:11 + 4 * k] = local_to_world(t, b, N, cs[k])
        return r

    def spec(cs, v, rx_, ry_):
        return lit(rect_of(cs), [0, 0, 0], N, local_to_world(t, b, N, v), rx_, ry_, tables=tables)[1][1]

    a = spec(corners, v_local, rx, ry)
    rxm, rym = (ry, rx) if mirror == "diag" else (rx, ry)
    m = spec(corners @ F.T, F @ v_local, rxm, rym)
    assert a > 0.01
    assert m == pytest.approx(a, rel=2e-4)


REF_TABLES = "/root/reference/bisemutum/assets/textures"


@pytest.mark.skipif(not __import__("os").path.isdir(REF_TABLES), reason="the engine's LTC table assets are not here")
def test_engine_tables_load_and_fit_ggx_roughly():
    """The engine's own LTC tables (3D texture assets) read as bi_pt_set_ltc_luts takes them:
    finite, invertible transforms everywhere, and a specular integral over a light around the
    mirror direction within 40% of a quadrature of the GGX lobe (F = 1) — the tables are the
    engine's fit, not restated here."""
    import assets
    luts = assets.read_ltc_tables(REF_TABLES)
    assert all(np.all(np.isfinite(t)) for t in luts)
    m = np.stack([luts[0][:, :3], luts[1][:, :3], luts[2][:, :3]], axis=1).astype(np.float64)
    assert np.all(np.linalg.det(m) != 0.0)   # near-singular (tiny roughness, grazing view) but invertible
    T = oracle.ltc_tables(*luts)
    N = np.array([0.0, 0.0, 1.0])
    for alpha in (0.1, 0.3, 0.6):
        for vth in (0.3, 0.8):
            V = np.array([math.sin(vth), 0.0, math.cos(vth)])
            d = np.array([-math.sin(vth), 0.0, math.cos(vth)])
            x = np.cross([0.0, 1.0, 0.0], d); x /= np.linalg.norm(x); y = np.cross(d, x)
            size = 0.6
            r = scenes.rect_light((1, 1, 1), d, size, size, x, y, two_sided=True)
            _, terms = oracle.rect_light32(r, T, [0, 0, 0], N, V, alpha, alpha, [0, 0, 0], [1, 1, 1], [1, 1, 1])
            n = 160
            s = (np.arange(n) + 0.5) / n - 0.5
            u, v = np.meshgrid(s * size, s * size, indexing="ij")
            X = d + u[..., None] * x + v[..., None] * y
            dist = np.linalg.norm(X, axis=-1)
            w = X / dist[..., None]
            dw = (size / n) ** 2 * np.abs(w @ d) / dist ** 2
            mat = scenes.material((0, 0, 0), math.sqrt(alpha), 1.0)
            mat[4:7] = 1.0
            mat[8:11] = 1.0
            f = oracle.bsdf64_batch(mat, N, V, w.reshape(-1, 3)).reshape(n, n, 3)[..., 0]
            ggx = (f * dw).sum()
            ltc = terms[1] / (2 * math.pi) * (terms[2] + terms[3])
            assert 0.6 < ltc / ggx < 1.4, (alpha, vth, ltc, ggx)
