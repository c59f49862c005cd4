"""The committed FP64 reference of the atrium view (tests/golden/c1_fp64_4096spp.npz, which
the equal-RMSE figures compare renders with) was made for the scene and pixels the
generator builds today: its scene digest and pixel list match, so a change to the scene
or its records cannot leave the fixture silently unused."""
import importlib.util
import os

import numpy as np

import scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def load_generator():
    spec = importlib.util.spec_from_file_location("make_c1_fp64", os.path.join(HERE, "golden", "make_c1_fp64.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_c1_fp64_fixture_matches_scene():
    gen = load_generator()
    ref = np.load(os.path.join(HERE, "golden", "c1_fp64_4096spp.npz"))
    cfg = scenes.CONFIGS["c1"]
    scene, cam = cfg["scene"](), cfg["camera"]()
    assert str(ref["scene_sha256"]) == gen.scene_digest(scene)
    rng = np.random.default_rng(gen.PIXEL_SEED)
    pix = np.sort(rng.choice(cam.width * cam.height, gen.PIXELS, replace=False)).astype(np.uint32)
    assert np.array_equal(ref["pixels"], pix)
    assert int(ref["spp"]) == gen.SPP and int(ref["max_bounces"]) == cfg["max_bounces"]
    assert np.all(np.isfinite(ref["mean"])) and np.all(ref["mean"] >= 0.0)
