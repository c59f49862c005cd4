"""GPU parity of rect lights (bi_pt_set_rect_lights, bi_pt_set_ltc_luts; RectLightData and
rect_light_eval_ltc, lights.hlsl:161-513): radiance bit for bit against the oracle on the C0
room lit by the sun, a ceiling panel and a two-sided light standing through the floor,
with synthetic LTC tables, with point lights as well, batched with a job without rect
lights; textured rect lights (bi_pt_set_rect_light_textures: RGBA8 sRGB with the engine's mip
chain, RGBA32F levels with a nearest mip filter, a LOD bias and clamp) beside an untextured
one, solo and batched; and the boundary's checks (tables required, a textured light's
texture required at render time, at most 16 lights and 16 textures)."""
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
    tex[3:4].view(np.int32)[0] = 16
    with pytest.raises(bi_pt.PtError):   # beyond the 16 textures
        g.set_rect_lights(tex[None])
    with pytest.raises(bi_pt.PtError):
        g.set_rect_lights(np.stack([r] * 17))
    with pytest.raises(bi_pt.PtError):   # more levels than the size allows
        g.set_rect_light_textures([scenes.RectTexture([np.zeros((2, 2, 4), np.float32)] * 3)])
    bad = r.copy()
    bad[9] = np.nan
    with pytest.raises(bi_pt.PtError):
        g.set_rect_lights(bad[None])
    g.set_ltc_luts(*scenes.identity_ltc())
    g.render(scenes.camera_c0(16, 12), spp=1, max_bounces=3)
    tex[3:4].view(np.int32)[0] = 0
    g.set_rect_lights(tex[None])
    with pytest.raises(bi_pt.PtError):   # texture 0 is not there
        g.render(scenes.camera_c0(16, 12), spp=1, max_bounces=3)
    g.set_rect_light_textures([scenes.RectTexture([np.ones((1, 1, 4), np.float32)])])
    g.render(scenes.camera_c0(16, 12), spp=1, max_bounces=3)
    g.close()


def textured_c0(seed=5):
    """C0 with three rect lights: an sRGB texture with the engine's mip chain (linear mips,
    repeat), RGBA32F levels of their own (nearest mips, magnification nearest, clamp, bias
    and clamp of the LOD), and an untextured one."""
    rng = np.random.default_rng(seed)
    sc = rect_c0()
    srgb0 = rng.integers(0, 256, (16, 32, 4), dtype=np.uint8)
    t0 = scenes.RectTexture(bi_pt.rect_texture_mips(srgb0, 6, scenes.TEXEL_RGBA8_SRGB), format=scenes.TEXEL_RGBA8_SRGB,
                            address_u=scenes.ADDRESS_REPEAT, address_v=scenes.ADDRESS_REPEAT)
    lv, w, h = [], 8, 8
    for _ in range(4):
        lv.append(rng.uniform(0.2, 3.0, (h, w, 4)).astype(np.float32))
        w, h = max(w // 2, 1), max(h // 2, 1)
    t1 = scenes.RectTexture(lv, mag_filter=0, min_filter=1, mipmap_mode=0, lod_bias=0.5, lod_max=2.0)
    sc.rect_textures = [t0, t1]
    r0 = scenes.textured_rect_light(sc.rect_lights[0], 0, 32, 16, 3.0, 1.5)
    r1 = scenes.textured_rect_light(sc.rect_lights[1], 1, 8, 8, 1.6, 2.4)
    r2 = scenes.rect_light((1.5, 1.2, 0.9), (2.2, 2.0, 2.5), 1.0, 0.7, (0, 0, 1), (0, 1, 0), True)
    sc.rect_lights = np.stack([r0, r1, r2])
    return sc


def test_textured_rect_lights_match_oracle(oracle_lib):
    print("case: c0 + textured rect lights (sRGB mips, RGBA32F nearest mips) + untextured, 80x60 2spp depth 5",
          flush=True)
    sc = textured_c0()
    cam = scenes.camera_c0(80, 60)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    g.render(cam, spp=2, max_bounces=5, seed=0x7E57, flags=bi_pt.FLAG_JITTER)
    got = g.sums(cam.width * cam.height)
    want = oracle_lib.OracleScene(sc).render32(cam, 5, 0x7E57, spp=2, flags=bi_pt.FLAG_JITTER)["sums"]
    assert radiance_check(got, want)
    # setting the same textures again (as a renderer does every frame) keeps the running image
    g.set_rect_light_textures(sc.rect_textures)
    g.render(cam, spp=1, max_bounces=5, seed=0x7E57, flags=bi_pt.FLAG_JITTER, accumulate=True)
    cont = oracle_lib.OracleScene(sc).render32(cam, 5, 0x7E57, spp=1, sample_base=2, sums=want.copy(),
                                               flags=bi_pt.FLAG_JITTER)["sums"]
    assert radiance_check(g.sums(cam.width * cam.height), cont)
    # the textures change the image
    plain = oracle_lib.OracleScene(dataclasses.replace(sc, rect_textures=[], rect_lights=np.stack(
        [rect_c0().rect_lights[0], rect_c0().rect_lights[1], sc.rect_lights[2]]))).render32(
        cam, 5, 0x7E57, spp=2, flags=bi_pt.FLAG_JITTER)["sums"]
    assert not np.array_equal(plain, want)
    g.close()


def test_batch_with_textured_rect_lights_matches_oracle(oracle_lib):
    print("case: batch of textured-rect c0 and plain c0, 48x36 1spp depth 4", flush=True)
    specs = [(textured_c0(7), scenes.camera_c0(48, 36)), (scenes.scene_c0(), scenes.camera_c0(40, 30))]
    b = bi_pt.Batch()
    ts, jobs = [], []
    for sc, cam in specs:
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        ts.append(t)
        jobs.append(bi_pt.Job(t, cam, spp=1, max_bounces=4, seed=33))
    b.render(jobs)
    ok = []
    for (sc, cam), t in zip(specs, ts):
        ok.append(radiance_check(t.sums(cam.width * cam.height),
                                 oracle_lib.OracleScene(sc).render32(cam, 4, 33, spp=1)["sums"]))
        t.close()
    b.close()
    assert all(ok)
