"""The engine's example project (scene_basic: camera, directional light, an earth sphere
with a base-colour texture and a normal map, a white cube, a translucent cube, an
alpha-tested cage cube and a checkerboard plane, from StaticMesh and Texture assets)
loaded from its own formats and rendered by the HIP pass through the C ABI, against the
restatement, bit for bit.  The textures are the fixtures' 1/8-size copies."""
import dataclasses

import numpy as np
import pytest

import bi_pt
from test_assets import fixture_scene
from test_gpu_parity import radiance_check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,spp", [(260, 128, 2), (1298, 635, 1)])
def test_scene_basic_radiance_matches_oracle(oracle_lib, w, h, spp):
    print(f"case: scene_basic {w}x{h} {spp}spp", flush=True)
    ps = fixture_scene(w, h, on_unsupported="raise")
    assert ps.unsupported == {} and len(ps.scene.textures) == 3
    st = ps.settings
    g = bi_pt.PathTracer()
    g.set_scene(ps.scene)
    g.render(ps.camera, spp=spp, max_bounces=st["max_bounces"], seed=0x5EED5B, ray_length=st["ray_length"])
    gs = g.sums(w * h)
    os_ = oracle_lib.OracleScene(ps.scene).render32(ps.camera, st["max_bounces"], 0x5EED5B, spp=spp,
                                                     ray_length=st["ray_length"])["sums"]
    assert radiance_check(gs, os_)
    lit = (os_.sum(1) > 0).mean()
    assert 0.3 < lit < 0.9, lit      # the floor and objects lit, the black sky above
    g.close()


def test_checker_floor_matches_oracle(oracle_lib):
    """The checkerboard material on a floor (odd cells the material, even cells its
    pattern material), direct light and two bounces, against the restatement."""
    print("case: checker floor 192x128 2spp", flush=True)
    import scenes
    from test_oracle_estimator import checker_scene
    sc = checker_scene()
    cam = scenes.make_camera((0.3, 6.0, 0.2), (0.0, 0.0, -1.5), 192, 128, 60.0)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    for mb in (2, 4):
        g.render(cam, spp=2, max_bounces=mb, seed=0xC4EC)
        gs = g.sums(cam.width * cam.height)
        os_ = oracle_lib.OracleScene(sc).render32(cam, mb, 0xC4EC, spp=2)["sums"]
        assert radiance_check(gs, os_)
    g.close()


def test_checker_pattern_material_is_validated():
    import scenes
    from test_oracle_estimator import checker_scene
    sc = checker_scene()
    g = bi_pt.PathTracer()
    for word, value in ((14, 2), (13, 2)):     # pattern material out of range; unknown pattern
        bad = sc.materials.copy()
        bad.view(np.uint32)[0, word] = value
        with pytest.raises(Exception, match="pattern"):
            g.set_scene(dataclasses.replace(sc, materials=bad))
    chained = sc.materials.copy()
    chained.view(np.uint32)[1, 13] = 1          # the even cells' record has a pattern itself
    chained.view(np.uint32)[1, 14] = 0
    with pytest.raises(Exception, match="without a pattern"):
        g.set_scene(dataclasses.replace(sc, materials=chained))
    g.set_scene(sc)
    g.close()


def test_textured_closed_form_scenes_match_oracle(oracle_lib):
    """The oracle's texture pins (base colour, normal map, alpha test, translucency)
    rendered by the kernels: sums equal the restatement's bit for bit, and the ray queries
    see the cut-out and translucent surfaces as the restatement's traversal does."""
    import test_oracle_textures as T
    import scenes
    print("case: textured closed-form scenes", flush=True)
    rng = np.random.default_rng(4)
    cells = rng.integers(20, 250, (2, 2, 4), dtype=np.uint8)
    p, n, i, t = T.plane(0.0, 1.0)
    m = T.lambert((0.5, 0.5, 0.5))
    m.view(np.uint32)[scenes.M_BASE_TEX] = 1
    base = T.scene_of([(p, n, i)], [t], None, [m], [0], [scenes.Texture(cells, scenes.TEXEL_RGBA8_SRGB, 1, 1, 0)])
    nm = np.tile(np.array([170, 120, 230, 255], np.uint8), (3, 5, 1))
    tan = np.tile(np.array([0.6, 0.0, 0.8, -1.0], np.float32), (4, 1))
    mn = T.lambert((0.6, 0.5, 0.4))
    mn.view(np.uint32)[scenes.M_NORMAL_TEX] = 1
    mn.view(np.uint32)[scenes.M_BASE_TEX] = 2
    normal = T.scene_of([(p, n, i)], [t], [tan], [mn], [0],
                        [scenes.Texture(nm, scenes.TEXEL_RGBA8_UNORM, 1, 0, 0), scenes.Texture(cells, 0, 0, 2, 2)],
                        sun_dir=(0.0, 0.92106099, 0.38941834))
    mask = np.zeros((1, 2, 4), np.uint8)
    mask[0, 1, 3] = 255
    mask[..., :3] = 200
    top = T.lambert((0.9, 0.1, 0.1))
    w = top.view(np.uint32)
    w[scenes.M_BLEND], w[scenes.M_OPACITY_TEX], w[scenes.M_OPACITY_CHANNEL], w[scenes.M_OPACITY_STEP] = 1, 1, 3, 1
    top[scenes.M_STEP_AT], top[scenes.M_STEP_BELOW], top[scenes.M_STEP_ABOVE] = 0.5, 0.0, 1.0
    cut = T.two_planes(top, [scenes.Texture(mask, scenes.TEXEL_RGBA8_UNORM, 0, 2, 2)])
    tr = T.lambert((0.9, 0.1, 0.1))
    tr.view(np.uint32)[scenes.M_BLEND] = scenes.BLEND_TRANSLUCENT
    tr[scenes.M_OPACITY] = 0.3
    trans = T.two_planes(tr, [], top_uv=False)
    cam = T.down_camera(64, 48, y=1.5)
    for name, sc in (("base", base), ("normal", normal), ("cut", cut), ("translucent", trans)):
        g = bi_pt.PathTracer()
        g.set_scene(sc)
        o = oracle_lib.OracleScene(sc)
        for mb, spp, flags in ((2, 2, 1), (5, 3, 0)):
            g.render(cam, spp=spp, max_bounces=mb, seed=0x7E7, flags=flags)
            gs = g.sums(cam.width * cam.height)
            os_ = o.render32(cam, mb, 0x7E7, spp=spp, flags=flags)["sums"]
            assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32)), (name, mb)
        # ray queries through the non-opaque surfaces: hits of the restatement's traversal
        rng2 = np.random.default_rng(11)
        org = np.column_stack([rng2.uniform(-1.2, 1.2, 3000), np.full(3000, 2.0), rng2.uniform(-1.2, 1.2, 3000)])
        dirs = np.column_stack([rng2.normal(0, 0.3, 3000), -np.ones(3000), rng2.normal(0, 0.3, 3000)])
        tt, tri, _, _ = g.trace_closest(org, dirs, tmax=100.0)
        rays = np.zeros((3000, 8), np.float32)
        rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = org, 0.001, dirs, 100.0
        ot, oleaf, _, _ = o.trace(rays)
        assert np.array_equal(np.isinf(tt), oleaf == 0xFFFFFFFF), name
        hit = oleaf != 0xFFFFFFFF
        assert np.array_equal(tt[hit].view(np.uint32), ot[hit].view(np.uint32)), name
        occ = g.trace_occlusion(org, dirs, tmax=100.0)
        (oocc,) = o.trace(rays, any_hit=True)
        assert np.array_equal(occ.astype(np.uint32), oocc), name
        print(f"{name}: sums and ray queries equal the oracle", flush=True)
        g.close()
