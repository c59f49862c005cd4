"""bi_pt_trace_closest_noalloc: closest-hit queries whose scratch is the library's static
device array.  Called through the C ABI, its results equal bi_pt_trace_closest's bit for
bit, on two calls in a row (the second smaller, so stale scratch from the first would
show), on the box-room scene and on a scene without triangles; sizes beyond the maximum
and a scene whose LBVH is not current are refused."""
import numpy as np
import pytest

import bi_pt
import scenes

pytestmark = pytest.mark.gpu


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform((-5.5, 0.2, -5.5), (5.5, 3.8, 5.5), (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    return o, d


def assert_same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))


def test_noalloc_equals_allocating_query(c0_scene):
    print("case: noalloc closest hits vs bi_pt_trace_closest, c0, 200000 then 70001 rays", flush=True)
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    with pytest.raises(bi_pt.PtError) as e:   # nothing built yet: refused, nothing allocated
        t.trace_closest_noalloc(*random_rays(8, 1))
    assert e.value.status == 6
    t.render(scenes.camera_c0(16, 16), spp=1, max_bounces=3, seed=1)   # builds the LBVH
    for n, seed in ((200000, 2), (70001, 3)):
        o, d = random_rays(n, seed)
        want = t.trace_closest(o, d, tmax=1000.0)
        got = t.trace_closest_noalloc(o, d, tmax=1000.0)
        assert_same(got, want)
        hits = int((want[1] != 0xFFFFFFFF).sum())
        print(f"{n} rays: {hits} hits, equal", flush=True)
        assert 0 < hits < n
    # a shorter extent, then the maximum size
    o, d = random_rays(bi_pt.NOALLOC_MAX_RAYS, 4)
    assert_same(t.trace_closest_noalloc(o, d, tmax=2.5), t.trace_closest(o, d, tmax=2.5))
    with pytest.raises(bi_pt.PtError) as e:
        t.trace_closest_noalloc(*random_rays(bi_pt.NOALLOC_MAX_RAYS + 1, 5))
    assert e.value.status == 1
    # a changed scene is refused until it is built again
    t.update_transforms(c0_scene.instances["xform"])
    with pytest.raises(bi_pt.PtError) as e:
        t.trace_closest_noalloc(*random_rays(8, 6))
    assert e.value.status == 6
    o, d = random_rays(5000, 7)
    want = t.trace_closest(o, d)          # builds again
    assert_same(t.trace_closest_noalloc(o, d), want)
    t.stats()                             # no device error flag
    t.close()


def test_noalloc_scene_without_triangles():
    print("case: noalloc closest hits on a scene without triangles", flush=True)
    sky, col = scenes.constant_sky()
    empty = scenes.Scene("empty", [], np.zeros(0, scenes.INSTANCE_DTYPE),
                         scenes.material((0.5, 0.5, 0.5), 0.5, 0.0)[None],
                         scenes.dir_light((1, 1, 1), (0, 1, 0))[None], sky, col, np.eye(3, dtype=np.float32).reshape(-1))
    t = bi_pt.PathTracer()
    t.set_scene(empty)
    t.render(scenes.camera_c0(8, 8), spp=1, max_bounces=3, seed=1)
    for n in (3000, 17):
        o, d = random_rays(n, n)
        got = t.trace_closest_noalloc(o, d)
        assert np.all(np.isinf(got[0])) and np.all(got[1] == 0xFFFFFFFF)
        assert_same(got, t.trace_closest(o, d))
    t.close()
