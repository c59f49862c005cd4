"""The engine's example project (scene_basic: camera, directional light, sphere, three
cubes and a checkerboard plane from StaticMesh assets) loaded from its own formats and
rendered by the HIP pass through the C ABI, against the restatement.  Materials
outside the pass's subset (textured, translucent) take the default surface on both
sides, as the loader reports."""
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
    ps = fixture_scene(w, h)
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
