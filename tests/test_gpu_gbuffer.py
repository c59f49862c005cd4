"""GPU parity of the primary-hit GBuffer (bi_pt_render_gbuffer): the four packed images
(rgba16f base colour, rgba16f normal / tangent / roughness, rgba16 unorm fresnel, rgba8
unorm material_0) bit for bit against the oracle, on the C0 room with jittered camera
rays (misses through the open top included), on the example project (textures, normal
maps, non-opaque surfaces, the checkerboard), and on a continued accumulation (the
GBuffer of the call's first sample).  The colour image equals bi_pt_render's."""
import numpy as np
import pytest

import bi_pt
import scenes
from test_assets import fixture_scene

pytestmark = pytest.mark.gpu

PARTS = ("base_color", "normal_roughness", "fresnel", "material_0")


def check(gb, ob, tag):
    for k in PARTS:
        assert gb[k].shape == ob[k].shape, (tag, k)
        bad = np.nonzero((gb[k] != ob[k]).any(1))[0]
        assert len(bad) == 0, (tag, k, len(bad), bad[:5], gb[k][bad[:3]], ob[k][bad[:3]])


def test_gbuffer_c0_matches_oracle(oracle_lib, c0_scene):
    print("case: c0 GBuffer 160x120, jittered, 2 spp", flush=True)
    cam = scenes.make_camera((0.0, 1.6, 5.5), (0.0, 2.6, -2.0), 160, 120, 75.0)
    g = bi_pt.PathTracer()
    g.set_scene(c0_scene)
    gb = g.render_gbuffer(cam, spp=2, max_bounces=5, seed=0x6B0F, flags=bi_pt.FLAG_JITTER)
    img = g.image().copy()
    pix = np.arange(cam.width * cam.height, dtype=np.uint32)
    ob = oracle_lib.OracleScene(c0_scene).gbuffer32(cam, pix, sample=0, seed=0x6B0F, flags=bi_pt.FLAG_JITTER)
    miss = ob["base_color"][:, 3] == 0
    assert 0.02 < miss.mean() < 0.9, miss.mean()
    check(gb, ob, "c0")
    r = bi_pt.PathTracer()
    r.set_scene(c0_scene)
    r.render(cam, spp=2, max_bounces=5, seed=0x6B0F, flags=bi_pt.FLAG_JITTER)
    assert np.array_equal(r.image().view(np.uint32), img.view(np.uint32))
    # continued accumulation: the GBuffer of the call's first sample (2)
    gb2 = g.render_gbuffer(cam, spp=1, max_bounces=5, seed=0x6B0F, flags=bi_pt.FLAG_JITTER, accumulate=True)
    assert g.stats()["samples_accumulated"] == 3
    ob2 = oracle_lib.OracleScene(c0_scene).gbuffer32(cam, pix, sample=2, seed=0x6B0F, flags=bi_pt.FLAG_JITTER)
    check(gb2, ob2, "c0 sample 2")
    assert not all(np.array_equal(gb2[k], gb[k]) for k in PARTS)   # jitter moved the rays
    g.close()
    r.close()


def test_gbuffer_example_project_matches_oracle(oracle_lib):
    print("case: scene_basic GBuffer 260x128", flush=True)
    ps = fixture_scene(260, 128, on_unsupported="raise")
    st = ps.settings
    g = bi_pt.PathTracer()
    g.set_scene(ps.scene)
    gb = g.render_gbuffer(ps.camera, spp=1, max_bounces=st["max_bounces"], seed=0x5EED5B,
                          ray_length=st["ray_length"])
    pix = np.arange(260 * 128, dtype=np.uint32)
    ob = oracle_lib.OracleScene(ps.scene).gbuffer32(ps.camera, pix, sample=0, seed=0x5EED5B,
                                                    ray_length=st["ray_length"])
    check(gb, ob, "scene_basic")
    # several materials and normal-mapped surfaces are on screen
    assert len(np.unique(gb["base_color"], axis=0)) > 50
    assert len(np.unique(gb["material_0"], axis=0)) >= 2
    g.close()


def test_gbuffer_parts_are_optional(oracle_lib, c0_scene):
    print("case: c0 GBuffer, normal_roughness only", flush=True)
    cam = scenes.camera_c0(64, 48)
    g = bi_pt.PathTracer()
    g.set_scene(c0_scene)
    gb = g.render_gbuffer(cam, spp=1, max_bounces=4, seed=3, parts=("normal_roughness",))
    assert set(gb) == {"normal_roughness"}
    ob = oracle_lib.OracleScene(c0_scene).gbuffer32(cam, np.arange(64 * 48, dtype=np.uint32), seed=3)
    assert np.array_equal(gb["normal_roughness"], ob["normal_roughness"])
    g.close()
