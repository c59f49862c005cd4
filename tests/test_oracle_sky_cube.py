"""The oracle's cube-map sky (bi_pt_set_sky_cube) pinned by what its definition fixes:
a cube filled from an equirect through the engine's face mapping reproduces the
equirect at every face-texel centre; seamless filtering is continuous across the
twelve edges and gives closed forms on a cube of constant faces (edge: mean of two
faces, corner: mean of three); the sampling density integrates to one and agrees with
the pdf lookup; the importance estimator of the sphere's area is 4 pi."""
import math

import numpy as np
import pytest

import oracle
import scenes
from test_oracle_shading import sphere_grid, tiny_sky_scene


def cube_scene(faces, col=None, rot=None):
    sc = tiny_sky_scene(faces, np.ones(3, np.float32) if col is None else np.asarray(col, np.float32))
    sc.sky_layout = scenes.SKY_CUBE
    if rot is not None:
        sc.sky_rot = np.asarray(rot, np.float32).reshape(-1)
    return sc


def smooth_equirect(w=96, h=48):
    """Asymmetric smooth radiance: distinct along +-x, +-y, +-z and in every channel."""
    th = (np.arange(h) + 0.5) / h * math.pi
    ph = (np.arange(w) + 0.5) / w * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    x, y, z = np.sin(T) * np.sin(P), np.cos(T), np.sin(T) * np.cos(P)
    r = 1.0 + 0.6 * x + 0.2 * y * y + 0.1 * z
    g = 1.0 - 0.3 * y + 0.25 * x * z
    b = 1.2 + 0.5 * z - 0.2 * x * y
    return np.stack([r, g, b, np.ones_like(r)], -1).astype(np.float32)


def test_face_texel_centres_reproduce_the_equirect(oracle_lib):
    """Cube made by equitangular_to_cubemap's mapping (numpy, float64): at each face-texel
    centre the oracle's cube lookup gives that texel, and the equirect sky gives the
    same radiance there.  A transposed or mirrored face reads a different texel."""
    eq = smooth_equirect()
    S = 12
    faces = scenes.equirect_to_cube(eq, S)
    cube = oracle_lib.OracleScene(cube_scene(faces))
    ere = oracle_lib.OracleScene(tiny_sky_scene(eq, np.ones(3, np.float32)))
    d = scenes.cube_face_directions(S).reshape(-1, 3)
    rgb_c, _ = oracle.sky_batch32(cube, d)
    rgb_e, _ = oracle.sky_batch32(ere, d)
    np.testing.assert_allclose(rgb_c, faces.reshape(-1, 4)[:, :3], rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(rgb_c, rgb_e, rtol=2e-4, atol=2e-5)
    # the same check with each face swapped for its opposite fails
    swapped = faces[[1, 0, 3, 2, 5, 4]]
    rgb_s, _ = oracle.sky_batch32(oracle_lib.OracleScene(cube_scene(swapped)), d)
    assert np.abs(rgb_s - rgb_e).max() > 0.1


def test_layer_orientation_matches_cubemap_hlsl(oracle_lib):
    """One texel lit alone on each face is seen in the direction cubemap.hlsl gives for
    its (u, v): +X (1, -b, -a), -X (-1, -b, a), +Y (a, 1, b), -Y (a, -1, -b),
    +Z (a, -b, 1), -Z (-a, -b, -1) with a = 2u - 1, b = 2v - 1."""
    S = 8
    i, j = 2, 5   # row (v), column (u)
    a, b = (2 * j + 1) / S - 1, (2 * i + 1) / S - 1
    expect = [(1, -b, -a), (-1, -b, a), (a, 1, b), (a, -1, -b), (a, -b, 1), (-a, -b, -1)]
    for f in range(6):
        faces = np.zeros((6, S, S, 4), np.float32)
        faces[f, i, j, :3] = 1.0
        s = oracle_lib.OracleScene(cube_scene(faces))
        d = np.array(expect[f], np.float32)
        d /= np.linalg.norm(d)
        np.testing.assert_allclose(s.sky_eval(d), [1, 1, 1], atol=1e-5)
        # and the texel with mirrored column is dark there
        faces[f, i, j, :3] = 0.0
        faces[f, i, S - 1 - j, :3] = 1.0
        s2 = oracle_lib.OracleScene(cube_scene(faces))
        assert float(s2.sky_eval(d).max()) < 1e-6


def test_constant_faces_edges_and_corners(oracle_lib):
    """Face f filled with f + 1: inside a face the value is exact; on an edge the mean of
    its two faces; at a corner the mean of its three (seamless filtering)."""
    S = 8
    faces = np.zeros((6, S, S, 4), np.float32)
    for f in range(6):
        faces[f] = f + 1
    s = oracle_lib.OracleScene(cube_scene(faces))
    val = {(1, 0, 0): 1, (-1, 0, 0): 2, (0, 1, 0): 3, (0, -1, 0): 4, (0, 0, 1): 5, (0, 0, -1): 6}
    for axis, v in val.items():
        d = np.array(axis, np.float32)
        np.testing.assert_allclose(s.sky_eval(d), [v] * 3, rtol=1e-6)
    axes = list(val)
    for p in range(6):
        for q in range(p + 1, 6):
            if np.dot(axes[p], axes[q]) != 0:
                continue
            d = np.array(axes[p], np.float32) + np.array(axes[q], np.float32)
            d /= np.linalg.norm(d)
            np.testing.assert_allclose(s.sky_eval(d), [(val[axes[p]] + val[axes[q]]) / 2] * 3, rtol=1e-5)
    for sx in (1, -1):
        for sy in (1, -1):
            for sz in (1, -1):
                d = np.array([sx, sy, sz], np.float32) / np.float32(math.sqrt(3))
                m = (val[(sx, 0, 0)] + val[(0, sy, 0)] + val[(0, 0, sz)]) / 3
                np.testing.assert_allclose(s.sky_eval(d), [m] * 3, rtol=1e-5)


def test_lookup_is_continuous_across_every_edge(oracle_lib):
    """Random texels: points just either side of each of the twelve edges read nearly the
    same value (a wrongly oriented neighbour face would jump by the texel contrast)."""
    S = 8
    rng = np.random.default_rng(7)
    faces = rng.uniform(0.0, 1.0, (6, S, S, 4)).astype(np.float32)
    s = oracle_lib.OracleScene(cube_scene(faces))
    axes = np.eye(3)
    worst = 0.0
    for p in range(3):
        for q in range(3):
            if p == q:
                continue
            r = 3 - p - q
            for sp in (1, -1):
                for sq in (1, -1):
                    for t in np.linspace(-0.95, 0.95, 23):
                        edge = sp * axes[p] + sq * axes[q] + t * axes[r]
                        across = sp * axes[p] - sq * axes[q]   # leaves face p for face q
                        for eps in (1e-4,):
                            d0 = (edge + eps * across).astype(np.float32)
                            d1 = (edge - eps * across).astype(np.float32)
                            v0, v1 = s.sky_eval(d0 / np.linalg.norm(d0)), s.sky_eval(d1 / np.linalg.norm(d1))
                            worst = max(worst, float(np.abs(v0 - v1).max()))
    assert worst < 5e-3


def test_cube_pdf_normalised_and_sampling_consistent(oracle_lib):
    eq, _ = scenes.preetham_like_sky((0.45, 0.8, 0.35), 128, 64)
    faces = scenes.equirect_to_cube(eq, 24)
    rot = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]], np.float32)
    s = oracle_lib.OracleScene(cube_scene(faces, (1.0, 0.9, 0.8), rot))
    d, dw = sphere_grid(512, 1024)
    rgb, pdf = oracle.sky_batch32(s, d)
    assert abs((pdf * dw).sum() - 1.0) < 0.01
    lum_grid = 0.2126 * rgb[:, 0] + 0.7152 * rgb[:, 1] + 0.0722 * rgb[:, 2]
    integral = float((lum_grid * dw).sum())
    rng = np.random.default_rng(3)
    wis, ps = [], []
    for u0, u1 in rng.uniform(size=(20000, 2)).astype(np.float32):
        wi, p = s.sky_sample(float(u0), float(u1))
        wis.append(wi)
        ps.append(p)
    wis, ps = np.array(wis), np.array(ps)
    assert np.all(np.abs(np.linalg.norm(wis, axis=1) - 1) < 1e-5)
    rgb_s, pdf_s = oracle.sky_batch32(s, wis)
    assert np.mean(np.abs(ps / np.maximum(pdf_s, 1e-30) - 1) < 1e-3) > 0.99
    # importance estimates: the sphere's area and the luminance integral
    area = 1.0 / ps
    assert abs(area.mean() - 4 * math.pi) < 4 * area.std() / math.sqrt(len(area)) + 1e-3
    lum = (0.2126 * rgb_s[:, 0] + 0.7152 * rgb_s[:, 1] + 0.0722 * rgb_s[:, 2]) / ps
    assert abs(lum.mean() - integral) / integral < 0.03


def test_cube_sky_renders_like_its_equirect(oracle_lib):
    """The C0 room lit by a cube made from its equirect sky renders like the equirect
    (the two skies agree up to the cube's resampling)."""
    import dataclasses
    eq, col = scenes.preetham_like_sky((0.3, 0.85, 0.4), 128, 64)
    sc = dataclasses.replace(scenes.scene_c0(), sky_rgba=eq, sky_color=col)
    cube = dataclasses.replace(sc, sky_rgba=scenes.equirect_to_cube(eq, 48), sky_layout=scenes.SKY_CUBE)
    cam = scenes.camera_c0(80, 60)
    pix = np.arange(cam.width * cam.height, dtype=np.uint32)
    means = []
    for s_ in (sc, cube):
        o = oracle_lib.OracleScene(s_)
        means.append(o.render32_pixels(cam, 4, 11, 16, pix).mean(0))
    np.testing.assert_allclose(means[0], means[1], rtol=0.03)
