"""GPU parity on the 262,144-triangle atrium (the benchmark scene)."""
import numpy as np
import pytest

import bi_pt
import oracle
import scenes
from test_gpu_parity import compare_bvh, compare_bvh4, radiance_check, random_rays, expected_queues

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    sc = scenes.scene_c1()
    return sc, oracle.OracleScene(sc)


@pytest.fixture(scope="module")
def c1_gpu(c1):
    p = bi_pt.PathTracer()
    p.set_scene(c1[0])
    yield p
    p.close()


def test_c1_bvh_bit_exact(c1, c1_gpu):
    print("case: c1 bvh 262144 tris", flush=True)
    assert c1_gpu.num_tris() == 262144
    compare_bvh(c1_gpu.bvh(), c1[1].bvh())
    compare_bvh4(c1_gpu, c1[1])


def test_c1_trace_bit_exact(c1, c1_gpu):
    print("case: c1 trace", flush=True)
    rng = np.random.default_rng(7)
    rays = random_rays(rng, 50000, np.array([0.0, 3.0, 0.0]), 9.0)
    g = c1_gpu.trace(rays)
    o = c1[1].trace(rays)
    for a, b in zip(g, o):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(c1_gpu.trace(rays, any_hit=True)[0], c1[1].trace(rays, any_hit=True)[0])


def test_c1_radiance_and_queues(c1, c1_gpu):
    print("case: c1 radiance 240x136 1spp depth 8", flush=True)
    cam = scenes.camera_c1(240, 136)
    c1_gpu.render(cam, spp=1, max_bounces=9, seed=0x5EED0001, flags=bi_pt.FLAG_RECORD_QUEUES)
    ref = c1[1].render32(cam, 9, 0x5EED0001, spp=1, record=True)
    assert radiance_check(c1_gpu.sums(cam.width * cam.height), ref["sums"])
    ext, sh = expected_queues(ref["segments"], ref["shadow_bits"], 8)
    for b in range(8):
        assert np.array_equal(c1_gpu.queue(0, b, 0), ext[b]), b
        assert np.array_equal(c1_gpu.queue(0, b, 1), sh[b]), b


def test_c1_full_frame_properties(c1_gpu):
    """At 1920x1080x16 spp: sample-split and wave-split invariance (bit-exact) and finite output."""
    print("case: c1 full frame 1920x1080 16 spp", flush=True)
    cam = scenes.camera_c1()
    npix = cam.width * cam.height
    c1_gpu.render(cam, spp=16, max_bounces=9, seed=0x5EED0001)
    full = c1_gpu.sums(npix).copy()
    c1_gpu.render(cam, spp=8, max_bounces=9, seed=0x5EED0001, wave_paths=npix)
    c1_gpu.render(cam, spp=8, max_bounces=9, seed=0x5EED0001, accumulate=True, wave_paths=4 * npix)
    split = c1_gpu.sums(npix)
    assert np.array_equal(full, split)
    assert np.all(np.isfinite(full)) and full.min() >= 0.0
    st = c1_gpu.stats()
    assert st["rays_extension"] > 0 and st["rays_sky_shadow"] > 0
