"""GPU parity: the HIP pass (through the C ABI) against the CPU restatement.

Bit-exact: Morton codes, sorted order, BVH topology and refit boxes, closest and
any-hit traversal results, queue contents per segment, and per-pixel radiance
sums.  The radiance comparison states its tolerance anyway (relative 1e-6 per
channel on at least 99.9% of pixels), but the expected and observed result is
exact equality, because the kernels and the restatement evaluate the same float
operations in the same order.
"""
import numpy as np
import pytest

import bi_pt
import oracle
import scenes

pytestmark = pytest.mark.gpu

RTOL = 1e-6
MIN_FRACTION_WITHIN = 0.999


@pytest.fixture(scope="module")
def pt():
    p = bi_pt.PathTracer()
    yield p
    p.close()


@pytest.fixture(scope="module")
def c0_gpu(pt, c0_scene):
    pt.set_scene(c0_scene)
    return pt


def compare_bvh(g, o):
    for k in ("morton", "order", "children", "parents"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(g["bounds"].view(np.uint32), o["bounds"].view(np.uint32))
    assert np.array_equal(g["centroid_bounds"], o["centroid_bounds"])


def compare_bvh4(gpu, ora):
    gc, gb = gpu.bvh4()
    oc, ob = ora.bvh4()
    assert np.array_equal(gc, oc)
    assert np.array_equal(gb.view(np.uint32), ob.view(np.uint32))


def test_bvh_bit_exact_c0(c0_gpu, c0_oracle):
    print("case: c0 bvh (LBVH + the default 3 rounds of treelet restructuring)", flush=True)
    compare_bvh(c0_gpu.bvh(), c0_oracle.bvh())
    compare_bvh4(c0_gpu, c0_oracle)


@pytest.mark.parametrize("rounds", [0, 1, 5])
def test_bvh_bit_exact_c0_rounds(oracle_lib, c0_scene, rounds):
    print(f"case: c0 bvh, {rounds} treelet rounds", flush=True)
    g = bi_pt.PathTracer()
    g.set_scene(c0_scene)
    g.set_build_options(treelet_rounds=rounds)
    o = oracle_lib.OracleScene(c0_scene)
    o.set_treelet_rounds(rounds)
    compare_bvh(g.bvh(), o.bvh())
    compare_bvh4(g, o)
    # images do not depend on the tree: compare a small render with the default build
    cam = scenes.camera_c0(48, 32)
    g.render(cam, spp=2, max_bounces=5, seed=3)
    a = g.image().copy()
    d = bi_pt.PathTracer()
    d.set_scene(c0_scene)
    d.render(cam, spp=2, max_bounces=5, seed=3)
    assert np.array_equal(a.view(np.uint32), d.image().view(np.uint32))
    g.close()
    d.close()


def test_small_render_builds_plain_tree_inspection_shows_configured(oracle_lib, c0_scene):
    """A call tracing fewer than max(4 n, 2^20) paths of a scene builds its LBVH without
    treelet rounds; the inspection hooks rebuild the configured tree."""
    print("case: c0 small render, then the BVH as configured", flush=True)
    g = bi_pt.PathTracer()
    g.set_scene(c0_scene)
    cam = scenes.camera_c0(48, 32)   # 1,536 paths << 2^20
    g.render(cam, spp=1, max_bounces=5, seed=5)
    o = oracle_lib.OracleScene(c0_scene)
    compare_bvh(g.bvh(), o.bvh())
    compare_bvh4(g, o)
    g.close()


@pytest.mark.parametrize("n,seed", [(12, 1), (20, 2), (40, 3), (64, 4), (100, 5), (200, 6)])
def test_treelets_small_trees_bit_exact(oracle_lib, n, seed):
    """Trees of 12-200 triangles: the root's treelet (and most others) has 6 or 7 leaves,
    so the split search over the full leaf set decides the result."""
    print(f"case: {n} random triangles, seed {seed}", flush=True)
    sc = scenes.scene_random_triangles(n, seed)
    for rounds in (1, 2):
        g = bi_pt.PathTracer()
        g.set_scene(sc)
        g.set_build_options(treelet_rounds=rounds)
        o = oracle_lib.OracleScene(sc)
        o.set_treelet_rounds(rounds)
        compare_bvh(g.bvh(), o.bvh())
        g.close()


def random_rays(rng, n, centre, radius, tmax=1000.0):
    o = centre + rng.uniform(-radius, radius, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.001
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def test_trace_bit_exact_c0(c0_gpu, c0_oracle):
    print("case: c0 trace", flush=True)
    rng = np.random.default_rng(1)
    rays = random_rays(rng, 20000, np.array([0.0, 2.0, 0.0]), 6.0)
    g = c0_gpu.trace(rays)
    o = c0_oracle.trace(rays)
    for a, b in zip(g, o):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.mean(g[1] != 0xFFFFFFFF) > 0.3   # open-top room: many rays escape
    ga = c0_gpu.trace(rays, any_hit=True)[0]
    oa = c0_oracle.trace(rays, any_hit=True)[0]
    assert np.array_equal(ga, oa)


def test_anyhit_longest_rays_all_octants(c0_gpu, c0_oracle):
    """Longest shadow-ray extent (1.7e38) in every direction octant: empty BVH4 slots
    are rejected and the GPU answers match the oracle bit for bit."""
    print("case: c0 any-hit tmax 1.7e38, 8 octants", flush=True)
    rng = np.random.default_rng(2)
    rays = random_rays(rng, 16000, np.array([0.0, 2.0, 0.0]), 6.0, tmax=1.70141173e38)
    signs = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32)
    rays[:, 4:7] = np.abs(rays[:, 4:7]) * signs[np.arange(len(rays)) % 8]
    ga = c0_gpu.trace(rays, any_hit=True)[0]
    oa = c0_oracle.trace(rays, any_hit=True)[0]
    assert np.array_equal(ga, oa)
    g = c0_gpu.trace(rays)
    o = c0_oracle.trace(rays)
    for a, b in zip(g, o):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    too_long = rays[:4].copy()
    too_long[:, 7] = np.float32(3.0e38)
    with pytest.raises(bi_pt.PtError):
        c0_gpu.trace(too_long, any_hit=True)


def radiance_check(gs, os_):
    assert gs.shape == os_.shape
    same = np.array_equal(gs.view(np.uint32), os_.view(np.uint32))
    rel = np.abs(gs - os_) / np.maximum(np.abs(os_), 1e-30)
    within = np.all(rel <= RTOL, axis=1).mean()
    print(f"radiance: bit-identical={same} within-rtol={within:.6f} max-rel={rel.max():.3e}", flush=True)
    assert within >= MIN_FRACTION_WITHIN
    return same


def test_radiance_c0_matches_oracle(c0_gpu, c0_oracle):
    print("case: c0 radiance 320x240 1spp depth 4", flush=True)
    cam = scenes.camera_c0()
    c0_gpu.render(cam, spp=1, max_bounces=5, seed=0x5EED0000)
    gs = c0_gpu.sums(cam.width * cam.height)
    os_ = c0_oracle.render32(cam, 5, 0x5EED0000, spp=1)["sums"]
    assert radiance_check(gs, os_)
    img = c0_gpu.image()
    np.testing.assert_array_equal(img[:, :3], gs / np.float32(1))
    assert np.all(img[:, 3] == 1.0)


def test_radiance_all_camera_rays_miss(oracle_lib):
    """Every camera ray escapes: the segments after the first have no entries (zero shading
    tiles, zero queue lengths); the image is the sky, as in the restatement."""
    print("case: random triangles, camera facing away, 40x24 2spp", flush=True)
    sc = scenes.scene_random_triangles(50, 3)
    cam = scenes.make_camera((0.0, 0.0, 20.0), (0.0, 0.0, 40.0), 40, 24, 40.0)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    g.render(cam, spp=2, max_bounces=5, seed=0x5EED0077)
    gs = g.sums(cam.width * cam.height)
    os_ = oracle_lib.OracleScene(sc).render32(cam, 5, 0x5EED0077, spp=2)["sums"]
    assert radiance_check(gs, os_)
    st = g.stats()
    assert st["rays_extension"] == 0 and st["rays_light_shadow"] == 0 and st["rays_sky_shadow"] == 0
    g.close()


@pytest.mark.parametrize("nl", [2, 4])
def test_radiance_c0_several_lights(oracle_lib, c0_scene, nl):
    """Shading keeps only as many light contributions in registers as the launch's jobs
    have lights (1, 2 or 4): two and four lights, with and without shadows and with a
    partial shadow strength, against the restatement."""
    print(f"case: c0 radiance 96x72 2spp, {nl} lights", flush=True)
    import dataclasses
    lights = [scenes.dir_light((2.6, 2.5, 2.3), (0.35, 0.85, 0.4)),
              scenes.dir_light((0.8, 0.9, 1.2), (-0.6, 0.5, 0.2), cast_shadow=False),
              scenes.dir_light((1.5, 0.4, 0.3), (0.1, 0.7, -0.7), shadow_strength=0.5),
              scenes.dir_light((0.3, 1.1, 0.5), (-0.2, 0.9, -0.4))][:nl]
    sc = dataclasses.replace(c0_scene, dir_lights=np.stack(lights))
    cam = scenes.camera_c0(96, 72)
    g = bi_pt.PathTracer()
    g.set_scene(sc)
    g.render(cam, spp=2, max_bounces=5, seed=0x5EED0042)
    gs = g.sums(cam.width * cam.height)
    os_ = oracle_lib.OracleScene(sc).render32(cam, 5, 0x5EED0042, spp=2)["sums"]
    assert radiance_check(gs, os_)
    g.close()


def expected_queues(segments, bits, D, nl=1):
    n = len(segments)
    ext, sh = [], []
    paths = np.arange(n, dtype=np.uint32)
    for b in range(D):
        alive = paths[segments > b]
        ext.append(alive)
        entries = []
        bb = bits[alive, b]
        for p, m in zip(alive, bb):
            for l in range(nl):
                if m & (1 << l):
                    entries.append((int(p) << 3) | l)
            if m & 0x80:
                entries.append((int(p) << 3) | 7)
        sh.append(np.array(entries, np.uint32))
    return ext, sh


def test_queue_contents_c0(c0_gpu, c0_oracle):
    print("case: c0 queues", flush=True)
    cam = scenes.camera_c0()
    c0_gpu.render(cam, spp=1, max_bounces=5, seed=0x5EED0000, flags=bi_pt.FLAG_RECORD_QUEUES)
    ref = c0_oracle.render32(cam, 5, 0x5EED0000, spp=1, record=True)
    ext, sh = expected_queues(ref["segments"], ref["shadow_bits"], 4)
    for b in range(4):
        assert np.array_equal(c0_gpu.queue(0, b, 0), ext[b]), f"extension queue {b}"
        assert np.array_equal(c0_gpu.queue(0, b, 1), sh[b]), f"shadow queue {b}"
    st = c0_gpu.stats()
    assert st["frames"] >= 1


def test_accumulation_waves_and_stack_are_invisible(c0_gpu, c0_oracle):
    print("case: c0 split invariance", flush=True)
    cam = scenes.camera_c0(160, 120)
    npix = cam.width * cam.height
    c0_gpu.render(cam, spp=4, max_bounces=5, seed=3)
    one = c0_gpu.sums(npix).copy()
    c0_gpu.render(cam, spp=1, max_bounces=5, seed=3)
    c0_gpu.render(cam, spp=3, max_bounces=5, seed=3, accumulate=True)
    two = c0_gpu.sums(npix).copy()
    c0_gpu.render(cam, spp=4, max_bounces=5, seed=3, wave_paths=npix, stack_lds=8)
    three = c0_gpu.sums(npix).copy()
    c0_gpu.render(cam, spp=4, max_bounces=5, seed=3, wave_paths=2 * npix, stack_lds=32)
    four = c0_gpu.sums(npix).copy()
    assert np.array_equal(one, two) and np.array_equal(one, three) and np.array_equal(one, four)
    ref = c0_oracle.render32(cam, 5, 3, spp=4)["sums"]
    assert radiance_check(one, ref)


def test_depth_output(c0_gpu, c0_oracle):
    print("case: c0 depth", flush=True)
    cam = scenes.camera_c0(64, 48)
    c0_gpu.render(cam, spp=1, max_bounces=2, seed=1, depth=True)
    depth = c0_gpu.depth_image()
    rays = []
    for p in range(cam.width * cam.height):
        o, d = oracle.camera_ray32(cam, p % cam.width, p // cam.width)
        rays.append(np.concatenate([o, [0.001], d, [1000.0]]))
    t, leaf, _, _ = c0_oracle.trace(np.array(rays, np.float32))
    hit = leaf != 0xFFFFFFFF
    P = np.array(rays)[:, 0:3] + np.array(rays)[:, 4:7] * t[:, None]
    clip = (cam.proj_view.astype(np.float64) @ np.concatenate([P, np.ones((len(P), 1))], 1).T).T
    want = np.where(hit, clip[:, 2] / clip[:, 3], 0.0)
    np.testing.assert_allclose(depth, want, rtol=1e-5, atol=1e-7)
    assert np.all(depth[~hit] == 0.0) and np.all(depth[hit] > 0.0)


def convergence_z(pt, ors, cam, spp=1024, spp64=4096, npix=96):
    pt.render(cam, spp=spp, max_bounces=5, seed=0x5EED0000)
    mean = pt.sums(cam.width * cam.height) / spp
    rng = np.random.default_rng(0)
    pix = np.sort(rng.choice(cam.width * cam.height, npix, replace=False)).astype(np.uint32)
    m64, v64 = ors.render64(cam, 5, pix, spp64)
    # standard error of the difference; the floor covers FP32 rounding of the running sums
    se = np.sqrt(v64 / spp + v64 / spp64) + 1e-4 * np.abs(m64) + 1e-6
    z = (mean[pix] - m64) / se
    zsum = (mean[pix] - m64).sum(axis=0) / np.sqrt((se ** 2).sum(axis=0))
    print(f"median|z|={np.median(np.abs(z)):.3f} frac(|z|>4)={np.mean(np.abs(z) > 4):.4f} zsum={zsum}", flush=True)
    return z, zsum


def test_gpu_converges_to_fp64_rough_c0(c0_scene):
    """C0 geometry with roughness >= 0.4: near-Gaussian per-pixel errors, strict bounds."""
    print("case: c0-rough convergence 1024 spp vs fp64 4096 spp", flush=True)
    rough = scenes.Scene(**{**c0_scene.__dict__})
    m = c0_scene.materials.copy()
    m[:, 3] = np.maximum(m[:, 3], 0.4)
    rough.materials = m
    pt = bi_pt.PathTracer()
    pt.set_scene(rough)
    z, zsum = convergence_z(pt, oracle.OracleScene(rough), scenes.camera_c0(80, 60))
    pt.close()
    assert np.median(np.abs(z)) < 1.0
    assert np.mean(np.abs(z) > 4) < 0.02
    assert np.all(np.abs(zsum) < 4), zsum


def test_gpu_converges_to_fp64_c0(c0_gpu, c0_oracle):
    """The box-room materials include near-mirror metals (roughness 0.05): a few pixels
    carry heavy-tailed (firefly) estimates, so the per-pixel bound is looser while the
    aggregate over the subset stays within 4 standard errors."""
    print("case: c0 convergence 1024 spp vs fp64 4096 spp", flush=True)
    z, zsum = convergence_z(c0_gpu, c0_oracle, scenes.camera_c0(80, 60))
    assert np.median(np.abs(z)) < 1.0
    assert np.mean(np.abs(z) > 4) < 0.10
    assert np.all(np.abs(zsum) < 4), zsum


def test_gpu_plane_closed_form():
    """Lambertian plane, constant sky, sun: E[L] = rho*Ls + rho/pi*cos(theta)*E on every plane pixel."""
    print("case: plane closed form", flush=True)
    from test_oracle_estimator import E_SUN, LS, RHO, SUN, plane_camera, plane_scene
    import math
    pt = bi_pt.PathTracer()
    pt.set_scene(plane_scene())
    cam = plane_camera()
    spp = 512
    pt.render(cam, spp=spp, max_bounces=3, seed=11)
    mean = pt.sums(cam.width * cam.height).mean(axis=0) / spp
    pt.close()
    want = RHO * LS + RHO / math.pi * SUN[1] * E_SUN
    np.testing.assert_allclose(mean, want, rtol=5e-3)


def test_instance_update_rebuilds(pt, c0_scene):
    print("case: c0 moved instances", flush=True)
    pt.set_scene(c0_scene)
    moved = scenes.Scene(**{**c0_scene.__dict__})
    inst = c0_scene.instances.copy()
    inst["xform"][1:, 3] += 0.5    # shift every sphere along +x
    inst["xform"][1:, 7] *= 0.9    # and scale heights
    moved.instances = inst
    pt.update_transforms(inst["xform"])
    ref = oracle.OracleScene(moved)
    compare_bvh(pt.bvh(), ref.bvh())
    cam = scenes.camera_c0(96, 72)
    pt.render(cam, spp=1, max_bounces=5, seed=9)
    assert radiance_check(pt.sums(cam.width * cam.height), ref.render32(cam, 5, 9, spp=1)["sums"])
    pt.set_scene(c0_scene)


def test_error_paths(pt, c0_scene):
    print("case: error paths", flush=True)
    pt.set_scene(c0_scene)
    cam = scenes.camera_c0(16, 16)
    with pytest.raises(bi_pt.PtError):
        pt.render(cam, spp=0)
    bad = scenes.Scene(**{**c0_scene.__dict__})
    inst = c0_scene.instances.copy()
    inst["mesh"][3] = 99
    bad.instances = inst
    with pytest.raises(bi_pt.PtError):
        pt.set_scene(bad)
    with pytest.raises(bi_pt.PtError):
        pt.set_lights(np.zeros((5, 16), np.float32))
    small = bi_pt.PathTracer(hbm_budget_bytes=1 << 20)
    with pytest.raises(bi_pt.PtError) as e:
        small.set_scene(c0_scene)
        small.render(scenes.camera_c0(), spp=1)
    assert e.value.status == 2
    small.close()
    pt.render(cam, spp=1)   # the context still works after the failures


def test_rgba16f_output_is_rounded_rgba32f():
    """BI_PT_FLAG_OUTPUT_RGBA16F: the same running mean, each channel rounded to the
    nearest even IEEE half (numpy's float32 -> float16 conversion)."""
    print("case: c0 64x48 2spp, RGBA16F output", flush=True)
    sc = scenes.scene_c0()
    cam = scenes.camera_c0(64, 48)
    a = bi_pt.PathTracer()
    a.set_scene(sc)
    a.render(cam, spp=2, max_bounces=5, seed=21)
    f32 = a.image().copy()
    b = bi_pt.PathTracer()
    b.set_scene(sc)
    b.render(cam, spp=2, max_bounces=5, seed=21, flags=bi_pt.FLAG_OUTPUT_RGBA16F)
    f16 = b.image()
    assert f16.dtype == np.float16
    assert np.array_equal(f16.view(np.uint16), f32.astype(np.float16).view(np.uint16))
    a.close()
    b.close()


def test_ray_queries_match_restatement(c0_oracle, c0_scene):
    """bi_pt_trace_closest / bi_pt_trace_occlusion (rays in device memory, the pass's
    traversal kernels): hits, triangles (in the scene's order), barycentrics and occlusion
    equal the restatement's traversal bit for bit; a far tmax and unnormalised directions."""
    print("case: c0 ray queries, 30000 rays", flush=True)
    rng = np.random.default_rng(9)
    rays = random_rays(rng, 30000, np.array([0.0, 2.0, 0.0]), 6.0)
    g = bi_pt.PathTracer()
    g.set_scene(c0_scene)
    t, tri, u, v = g.trace_closest(rays[:, 0:3], rays[:, 4:7], tmax=1000.0)
    ot, oleaf, ou, ov = c0_oracle.trace(rays)
    hit = oleaf != 0xFFFFFFFF
    order = c0_oracle.bvh()["order"]
    assert np.array_equal(tri[hit], order[oleaf[hit]]) and np.all(tri[~hit] == 0xFFFFFFFF)
    assert np.array_equal(t[hit].view(np.uint32), ot[hit].view(np.uint32)) and np.all(np.isinf(t[~hit]))
    assert np.array_equal(u[hit].view(np.uint32), ou[hit].view(np.uint32))
    assert np.array_equal(v[hit].view(np.uint32), ov[hit].view(np.uint32))
    assert 0.3 < hit.mean() < 1.0
    occ = g.trace_occlusion(rays[:, 0:3], rays[:, 4:7], tmax=1000.0)
    assert np.array_equal(occ, c0_oracle.trace(rays, any_hit=True)[0])
    # directions scaled by 2: t in units of the direction halves (float rounding of the
    # slab and triangle tests, and TMin = 0.001 now in the doubled unit, may change a few
    # grazing or near-origin rays)
    t2, tri2, _, _ = g.trace_closest(rays[:, 0:3], 2.0 * rays[:, 4:7], tmax=500.0)
    same = (tri2 == tri) & hit
    assert same.sum() >= 0.999 * hit.sum()
    assert np.allclose(t2[same], 0.5 * t[same], rtol=1e-5)
    with pytest.raises(bi_pt.PtError):
        g.trace_closest(rays[:4, 0:3], rays[:4, 4:7], tmax=3.0e38)
    g.close()
