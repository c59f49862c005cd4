"""GPU parity of point and spot lights (bi_pt_set_point_lights, PointLightData semantics of
lights.hlsl:15-25): radiance bit for bit against the oracle on the C0 room lit by the sun,
a point light and a spot; the shadow ray to a point light ends at the light (an occluder
beyond it shadows nothing) with shadow strength and lights that cast none; the light
list's limit and its validation; clearing point lights returns to the directional-only
image; a batch mixing contexts with and without point lights."""
import dataclasses

import numpy as np
import pytest

import bi_pt
import scenes
from test_gpu_parity import radiance_check
from test_oracle_point_lights import E, lambert_plane, quad

pytestmark = pytest.mark.gpu


def lit_c0():
    pl = np.stack([scenes.point_light((6.0, 5.0, 3.5), (1.5, 3.2, -0.8), 7.0),
                   scenes.point_light((2.0, 6.0, 9.0), (-2.5, 3.6, 2.0), 9.0, spot_axis=(0.3, 1.0, -0.2),
                                      inner_deg=20.0, outer_deg=45.0, shadow_strength=0.8)])
    sc = scenes.scene_c0()
    sc.point_lights = pl
    return sc


def test_point_lights_c0_match_oracle(oracle_lib):
    print("case: c0 + point light + spot, 96x72 2spp depth 5", flush=True)
    sc = lit_c0()
    cam = scenes.camera_c0(96, 72)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    g.render(cam, spp=2, max_bounces=5, seed=0x5EED0D01)
    gs = g.sums(cam.width * cam.height)
    out = oracle_lib.OracleScene(sc).render32(cam, 5, 0x5EED0D01, spp=2)
    assert radiance_check(gs, out["sums"])
    st = g.stats()
    assert st["rays_light_shadow"] == int(out["counts"][2])
    assert st["rays_sky_shadow"] == int(out["counts"][3])
    # without point lights: the directional-only image
    g.set_point_lights(None)
    g.render(cam, spp=2, max_bounces=5, seed=0x5EED0D01)
    plain = dataclasses.replace(sc)
    plain.point_lights = None
    assert radiance_check(g.sums(cam.width * cam.height),
                          oracle_lib.OracleScene(plain).render32(cam, 5, 0x5EED0D01, spp=2)["sums"])
    g.close()


@pytest.mark.parametrize("strength,cast", [(1.0, True), (0.25, True), (1.0, False)])
def test_point_light_shadow_extent_matches_oracle(oracle_lib, strength, cast):
    print(f"case: plane + occluders, point light strength {strength} cast {cast}, 64x48 1spp depth 2 and 4", flush=True)
    rec = scenes.point_light(E, (0.0, 2.0, 0.0), 6.0, cast_shadow=cast, shadow_strength=strength)
    sun = scenes.dir_light((9.0, 9.0, 9.0), (0, 1, 0))
    sc = lambert_plane(extra_meshes=[quad(1.0, 0.5), quad(3.0, 20.0, (0, -1, 0))], dir_lights=sun[None],
                       point_lights=rec)
    cam = scenes.make_camera((0.2, 2.9, 0.3), (0.0, 0.0, 0.0), 64, 48, 100.0)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    ok = []
    for depth in (2, 4):
        g.render(cam, spp=1, max_bounces=depth, seed=11)
        ok.append(radiance_check(g.sums(cam.width * cam.height),
                                 oracle_lib.OracleScene(sc).render32(cam, depth, 11, spp=1)["sums"]))
    g.close()
    assert all(ok)


def test_point_light_limits_and_validation():
    print("case: light list limits", flush=True)
    g = bi_pt.PathTracer()
    g.set_scene(scenes.scene_c0())   # one directional light
    rec = scenes.point_light(E, (0.0, 2.0, 0.0), 6.0)
    g.set_point_lights(np.stack([rec] * 3))
    with pytest.raises(bi_pt.PtError):
        g.set_point_lights(np.stack([rec] * 4))   # 1 directional + 4 point
    with pytest.raises(bi_pt.PtError):
        g.set_lights(np.stack([scenes.dir_light((1, 1, 1), (0, 1, 0))] * 2))   # 2 + 3
    bad = rec.copy()
    bad[4] = np.inf
    with pytest.raises(bi_pt.PtError):
        g.set_point_lights(bad[None])
    # a scene with four directional lights replaces the three point lights
    sc = dataclasses.replace(scenes.scene_c0(), dir_lights=np.stack([scenes.dir_light((1, 1, 1), (0, 1, 0))] * 4))
    g.set_scene(sc)
    g.close()


def test_batch_with_point_lights_matches_oracle(oracle_lib):
    print("case: batch of c0 with and without point lights, 48x36 1spp depth 4", flush=True)
    specs = [(lit_c0(), scenes.camera_c0(48, 36)), (scenes.scene_c0(), scenes.camera_c0(40, 30))]
    b = bi_pt.Batch()
    ts, jobs = [], []
    for sc, cam in specs:
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        ts.append(t)
        jobs.append(bi_pt.Job(t, cam, spp=1, max_bounces=4, seed=21))
    b.render(jobs)
    ok = []
    for (sc, cam), t in zip(specs, ts):
        want = oracle_lib.OracleScene(sc).render32(cam, 4, 21, spp=1)["sums"]
        ok.append(radiance_check(t.sums(cam.width * cam.height), want))
        t.close()
    b.close()
    assert all(ok)
