"""GPU parity of the cube-map sky (bi_pt_set_sky_cube): radiance bit for bit against the
oracle on the C0 room under a rotated, coloured cube sky with sky next-event
estimation, and on a view of the open sky whose rays cross every face, edge and corner
(the seamless taps)."""
import dataclasses

import numpy as np
import pytest

import bi_pt
import scenes
from test_gpu_parity import radiance_check

pytestmark = pytest.mark.gpu

ROT = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]], np.float32).reshape(-1)


def cube_c0(size):
    eq, _ = scenes.preetham_like_sky((0.3, 0.85, 0.4), 128, 64)
    return dataclasses.replace(scenes.scene_c0(), sky_rgba=scenes.equirect_to_cube(eq, size),
                               sky_color=np.array([1.0, 0.9, 0.8], np.float32), sky_rot=ROT,
                               sky_layout=scenes.SKY_CUBE)


@pytest.mark.parametrize("size,flags", [(24, 0), (7, 0), (24, bi_pt.FLAG_NO_SKY_NEE)])
def test_cube_sky_radiance_matches_oracle(oracle_lib, size, flags):
    print(f"case: c0 under a {size}^2 cube sky, flags {flags}, 96x72 2spp depth 5", flush=True)
    sc = cube_c0(size)
    cam = scenes.camera_c0(96, 72)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    g.render(cam, spp=2, max_bounces=5, seed=0x5EED0C0B, flags=flags)
    gs = g.sums(cam.width * cam.height)
    os_ = oracle_lib.OracleScene(sc).render32(cam, 5, 0x5EED0C0B, spp=2, flags=flags)["sums"]
    assert radiance_check(gs, os_)
    assert g.stats()["rays_sky_shadow"] > 0 or flags
    g.close()


def test_cube_sky_seen_directly_matches_oracle(oracle_lib):
    """A wide camera inside a single far-away triangle's scene: nearly every ray is a sky
    lookup, over all six faces (jittered primary rays), with the face texels random."""
    print("case: open sky through a 170-degree camera, 8^2 random cube, 128x96 4spp", flush=True)
    rng = np.random.default_rng(5)
    faces = rng.uniform(0.0, 2.0, (6, 8, 8, 4)).astype(np.float32)
    sc = scenes.scene_random_triangles(20, 9)
    sc = dataclasses.replace(sc, sky_rgba=faces, sky_color=np.ones(3, np.float32), sky_rot=ROT,
                             sky_layout=scenes.SKY_CUBE)
    results = []
    for eye, target in (((0.0, 0.0, 50.0), (0.0, 0.0, 100.0)), ((0.0, 0.0, 50.0), (0.0, 0.0, 0.0)),
                        ((0.0, 50.0, 0.0), (0.0, 100.0, 0.1))):
        cam = scenes.make_camera(eye, target, 128, 96, 170.0)
        g = bi_pt.PathTracer()
        g.set_scene(sc)
        g.render(cam, spp=4, max_bounces=3, seed=0x5EED0C0C, flags=bi_pt.FLAG_JITTER)
        gs = g.sums(cam.width * cam.height)
        os_ = oracle_lib.OracleScene(sc).render32(cam, 3, 0x5EED0C0C, spp=4, flags=bi_pt.FLAG_JITTER)["sums"]
        results.append(radiance_check(gs, os_))
        g.close()
    assert all(results)


def test_cube_sky_replaces_equirect_and_clears(oracle_lib):
    """Setting a cube after an equirect (and clearing it with face size 0) restarts the
    running image and renders as the oracle with that sky."""
    print("case: equirect -> cube -> none on one context, 48x36 1spp", flush=True)
    sc = cube_c0(16)
    cam = scenes.camera_c0(48, 36)
    g = bi_pt.PathTracer()
    g.set_scene(dataclasses.replace(sc, sky_rgba=scenes.preetham_like_sky((0.3, 0.85, 0.4), 64, 32)[0],
                                    sky_layout=scenes.SKY_EQUIRECT))
    g.render(cam, spp=1, max_bounces=4, seed=5)
    g.set_sky_cube(sc.sky_rgba, sc.sky_color, sc.sky_rot)
    g.render(cam, spp=1, max_bounces=4, seed=5, accumulate=True)
    assert g.stats()["samples_accumulated"] == 1
    os_ = oracle_lib.OracleScene(sc).render32(cam, 4, 5, spp=1)["sums"]
    assert radiance_check(g.sums(cam.width * cam.height), os_)
    g.set_sky_cube(np.zeros((0,), np.float32), sc.sky_color, sc.sky_rot)
    g.render(cam, spp=1, max_bounces=4, seed=5, accumulate=True)
    dark = dataclasses.replace(sc, sky_rgba=np.zeros((0, 0, 4), np.float32), sky_layout=scenes.SKY_EQUIRECT)
    os_ = oracle_lib.OracleScene(dark).render32(cam, 4, 5, spp=1)["sums"]
    assert radiance_check(g.sums(cam.width * cam.height), os_)
    g.close()
