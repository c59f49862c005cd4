"""GPU parity of rect lights (bi_pt_set_rect_lights, bi_pt_set_ltc_luts; RectLightData and
rect_light_eval_ltc, lights.hlsl:161-513): radiance bit for bit against the oracle on the C0
room lit by the sun, a ceiling panel and a two-sided light standing through the floor,
with synthetic LTC tables, with point lights as well, batched with a job without rect
lights; and the boundary's checks (tables required, untextured lights only, at most 16)."""
import dataclasses

import numpy as np
import pytest

import bi_pt
import scenes
from test_gpu_parity import radiance_check

pytestmark = pytest.mark.gpu


def rect_c0(seed=3):
    sc = scenes.scene_c0()
    sc.rect_lights = np.stack([
        scenes.rect_light((3.0, 2.6, 2.2), (0.5, 3.9, -0.8), 3.0, 1.5, (1, 0, 0), (0, 0, 1)),              # faces down
        scenes.rect_light((0.8, 1.6, 2.4), (-2.0, 0.6, 1.5), 1.6, 2.4, (0.6, 0, 0.8), (0, 1, 0), True),   # crosses the floor
    ])
    sc.ltc = scenes.smooth_ltc(seed)
    return sc


def test_rect_lights_c0_match_oracle(oracle_lib):
    print("case: c0 + two rect lights, smooth LTC tables, 96x72 2spp depth 5", flush=True)
    sc = rect_c0()
    cam = scenes.camera_c0(96, 72)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    g.render(cam, spp=2, max_bounces=5, seed=0x5EED0E01)
    out = oracle_lib.OracleScene(sc).render32(cam, 5, 0x5EED0E01, spp=2)
    assert radiance_check(g.sums(cam.width * cam.height), out["sums"])
    # the rect lights add light: more than the same scene without them
    plain = dataclasses.replace(sc)
    plain.rect_lights = None
    g.set_rect_lights(None)
    g.render(cam, spp=2, max_bounces=5, seed=0x5EED0E01)
    base = g.sums(cam.width * cam.height)
    assert radiance_check(base, oracle_lib.OracleScene(plain).render32(cam, 5, 0x5EED0E01, spp=2)["sums"])
    assert out["sums"].sum() > base.sum() * 1.05
    g.close()


def test_rect_and_point_lights_with_identity_tables(oracle_lib):
    print("case: c0 + rect lights + point light + sun, identity LTC tables, jittered 64x48 2spp depth 6", flush=True)
    sc = rect_c0()
    sc.ltc = scenes.identity_ltc(norm=(0.8, 0.1))
    sc.point_lights = scenes.point_light((4.0, 3.0, 2.0), (1.0, 2.5, 1.0), 8.0)[None]
    cam = scenes.camera_c0(64, 48)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    g.render(cam, spp=2, max_bounces=6, seed=12, flags=bi_pt.FLAG_JITTER)
    want = oracle_lib.OracleScene(sc).render32(cam, 6, 12, spp=2, flags=bi_pt.FLAG_JITTER)["sums"]
    assert radiance_check(g.sums(cam.width * cam.height), want)
    g.close()


def test_batch_with_rect_lights_matches_oracle(oracle_lib):
    print("case: batch of c0 with and without rect lights, 48x36 1spp depth 4", flush=True)
    specs = [(rect_c0(), scenes.camera_c0(48, 36)), (scenes.scene_c0(), scenes.camera_c0(40, 30))]
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
        ok.append(radiance_check(t.sums(cam.width * cam.height),
                                 oracle_lib.OracleScene(sc).render32(cam, 4, 21, spp=1)["sums"]))
        t.close()
    b.close()
    assert all(ok)


def test_rect_light_checks():
    print("case: rect light boundary checks", flush=True)
    g = bi_pt.PathTracer()
    g.set_scene(scenes.scene_c0())
    r = scenes.rect_light((1, 1, 1), (0, 2, 0), 1.0, 1.0, (1, 0, 0), (0, 0, 1))
    g.set_rect_lights(r[None])
    with pytest.raises(bi_pt.PtError):   # no LTC tables yet
        g.render(scenes.camera_c0(16, 12), spp=1, max_bounces=3)
    tex = r.copy()
    tex[3:4].view(np.int32)[0] = 0
    with pytest.raises(bi_pt.PtError):
        g.set_rect_lights(tex[None])
    with pytest.raises(bi_pt.PtError):
        g.set_rect_lights(np.stack([r] * 17))
    bad = r.copy()
    bad[9] = np.nan
    with pytest.raises(bi_pt.PtError):
        g.set_rect_lights(bad[None])
    g.set_ltc_luts(*scenes.identity_ltc())
    g.render(scenes.camera_c0(16, 12), spp=1, max_bounces=3)
    g.close()
