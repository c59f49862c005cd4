"""Batched rendering (bi_pt_render_batch): several contexts' jobs share one set of waves.

Each job must come out bit for bit as the same job rendered alone with bi_pt_render
and as the CPU restatement renders it, whatever else is in the batch, with or without
the hipGraph replay, and under any tuning the autotuner picks.
"""
import numpy as np
import pytest

import bi_pt
import scenes
from test_oracle_estimator import plane_camera, plane_scene

pytestmark = pytest.mark.gpu


def empty_scene():
    sky, col = scenes.constant_sky((0.5, 0.6, 0.8))
    return scenes.Scene("empty", [], np.zeros(0, scenes.INSTANCE_DTYPE),
                        scenes.material((0.5, 0.5, 0.5), 0.5, 0.0)[None],
                        scenes.dir_light((1, 1, 1), (0, 1, 0))[None], sky, col,
                        np.eye(3, dtype=np.float32).reshape(-1))


def job_specs(c0_scene):
    """(scene, camera, render kwargs): different sizes, spp, bounces, lights and scenes."""
    return [
        (c0_scene, scenes.camera_c0(64, 48), dict(spp=4, max_bounces=5, seed=11)),
        (plane_scene(), plane_camera(40, 30), dict(spp=3, max_bounces=3, seed=12, flags=bi_pt.FLAG_JITTER)),
        (c0_scene, scenes.camera_c0(33, 17), dict(spp=5, max_bounces=9, seed=13)),
        (empty_scene(), plane_camera(16, 8), dict(spp=2, max_bounces=4, seed=14)),
    ]


def tracers(specs):
    out = []
    for sc, _, _ in specs:
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        out.append(t)
    return out


def solo_sums(specs):
    res = []
    for (sc, cam, kw), t in zip(specs, tracers(specs)):
        t.render(cam, **kw)
        res.append((t.sums(cam.width * cam.height).copy(), t.image().copy()))
        t.close()
    return res


def batch_sums(specs, batch, ts, stream=None, accumulate=False):
    jobs = [bi_pt.Job(t, cam, accumulate=accumulate, **kw) for (sc, cam, kw), t in zip(specs, ts)]
    batch.render(jobs, stream=stream)
    return [(t.sums(cam.width * cam.height).copy(), t.image().copy()) for (sc, cam, kw), t in zip(specs, ts)]


def assert_same(a, b):
    for (sa, ia), (sb, ib) in zip(a, b):
        assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
        assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32))


def assert_oracle(specs, got, oracle_lib):
    """Each job's running sums equal the CPU restatement's (the scene-less job sees the
    constant sky on every sample)."""
    for (sc, cam, kw), (sums, _) in zip(specs, got):
        if sc.num_tris == 0:
            sky = np.asarray(sc.sky_rgba[0, 0, :3], np.float32) * np.asarray(sc.sky_color, np.float32)
            want = np.tile(kw["spp"] * sky, (cam.width * cam.height, 1))
            np.testing.assert_allclose(sums, want, rtol=1e-6)
            continue
        want = oracle_lib.OracleScene(sc).render32(cam, kw["max_bounces"], kw["seed"], spp=kw["spp"],
                                                   flags=kw.get("flags", 0))["sums"]
        assert np.array_equal(sums.view(np.uint32), want.view(np.uint32)), sc.name


def test_batch_equals_solo(c0_scene, oracle_lib):
    print("case: batch of 4 jobs (c0 64x48, plane, c0 33x17 depth 8, empty scene) vs solo and oracle", flush=True)
    specs = job_specs(c0_scene)
    solo = solo_sums(specs)
    assert np.all(np.isfinite(solo[0][0])) and solo[0][0].max() > 0
    b = bi_pt.Batch()
    ts = tracers(specs)
    got = batch_sums(specs, b, ts)
    assert_same(got, solo)
    assert_oracle(specs, got, oracle_lib)
    # small waves: the jobs' samples split over many waves and mix differently
    b.set_tuning(wave_paths=5000)
    ts2 = tracers(specs)
    assert_same(batch_sums(specs, b, ts2), solo)
    st = b.stats()
    assert st["rays_camera"] == 2 * sum(cam.width * cam.height * kw["spp"] for _, cam, kw in specs)
    for t in ts + ts2:
        t.close()
    b.close()


def test_batch_graph_replay_and_accumulation(c0_scene):
    print("case: batch hipGraph capture + 3 replays, accumulating", flush=True)
    import torch
    specs = job_specs(c0_scene)[:3]
    stream = torch.cuda.Stream()
    plain = bi_pt.Batch()
    ts_plain = tracers(specs)
    graph = bi_pt.Batch()
    graph.set_tuning(use_graph=True)
    ts_graph = tracers(specs)
    for k in range(4):
        a = batch_sums(specs, plain, ts_plain, accumulate=True)
        with torch.cuda.stream(stream):
            g = batch_sums(specs, graph, ts_graph, stream=stream.cuda_stream, accumulate=True)
        stream.synchronize()
        # eager work between replays
        torch.cuda.synchronize()
        assert_same(g, a)
    assert graph.tuning()["use_graph"] == 1
    for t in ts_plain + ts_graph:
        t.close()
    plain.close()
    graph.close()


def test_batch_autotune_keeps_results(c0_scene):
    print("case: autotune on 2 jobs, then render", flush=True)
    specs = job_specs(c0_scene)[:2]
    solo = solo_sums(specs)
    b = bi_pt.Batch()
    ts = tracers(specs)
    jobs = [bi_pt.Job(t, cam, **kw) for (sc, cam, kw), t in zip(specs, ts)]
    best = b.autotune(jobs, rounds=1)
    assert best["wave_paths"] in (0, 1 << 23, 1 << 24, 1 << 25)
    assert best["stack_lds"] in (0, 8, 12, 16)
    assert best["block_threads"] in (0, 128, 256, 512)
    assert_same(batch_sums(specs, b, ts), solo)
    for t in ts:
        t.close()
    b.close()


def test_batch_errors(c0_scene):
    print("case: batch error paths", flush=True)
    b = bi_pt.Batch()
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    cam = scenes.camera_c0(16, 16)
    with pytest.raises(bi_pt.PtError):
        b.render([bi_pt.Job(t, cam, spp=1), bi_pt.Job(t, cam, spp=1)])   # one context twice
    with pytest.raises(bi_pt.PtError):
        b.render([bi_pt.Job(t, cam, spp=0)])
    b.set_tuning(use_graph=True)
    with pytest.raises(bi_pt.PtError):
        b.render([bi_pt.Job(t, cam, spp=1)])                            # graph needs a created stream
    b.set_tuning()
    b.render([bi_pt.Job(t, cam, spp=1)])
    assert np.all(np.isfinite(t.image()))
    t.close()
    b.close()


def test_batch_mixed_shadow_casting_lights_match_oracle(oracle_lib):
    """Two jobs in one wave whose light 0 differs: casting shadows (its radiance slot written
    by the shadow launch, which overlaps the next segment) and not (the same slot written by
    the shading): both bit for bit as the oracle renders them, over several calls."""
    print("case: batch of c0 with a shadow-casting and a shadowless sun, 64x48 2spp depth 6, 3 calls", flush=True)
    import dataclasses
    base = scenes.scene_c0()
    lit = dataclasses.replace(base, dir_lights=scenes.dir_light((2.6, 2.5, 2.3), (0.35, 0.85, 0.4))[None])
    bare = dataclasses.replace(base, dir_lights=scenes.dir_light((2.6, 2.5, 2.3), (0.35, 0.85, 0.4),
                                                                 cast_shadow=False)[None])
    cam = scenes.camera_c0(64, 48)
    b = bi_pt.Batch()
    ts = []
    for sc in (lit, bare):
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        ts.append(t)
    ok = []
    for call in range(3):
        b.render([bi_pt.Job(t, cam, spp=2, max_bounces=6, seed=40 + call) for t in ts])
        for sc, t in zip((lit, bare), ts):
            want = oracle_lib.OracleScene(sc).render32(cam, 6, 40 + call, spp=2)["sums"]
            ok.append(np.array_equal(t.sums(cam.width * cam.height).view(np.uint32), want.view(np.uint32)))
    for t in ts:
        t.close()
    b.close()
    assert all(ok), ok
