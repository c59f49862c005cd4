"""The multi-scene, serving and radiance-cache workloads at their real shapes.

* Multi-scene batch: the 64 box-room scenes of 1k-500k triangles (log-uniform), one job
  each at 256x256 / 640x480 / 1280x720 and 1-8 spp, depth 6, in one bi_pt_render_batch
  call captured in a hipGraph and replayed.  Every job equals its solo bi_pt_render bit
  for bit, on the capture and on the replay, and seeded pixel subsets of four jobs equal
  the CPU restatement.
* Serving: the first 100 Poisson arrivals (200 jobs/s) over the 8 resident templates of
  10k-1M triangles, orbit cameras, 256x256 / 512x512 / 1280x720, spp 1/4/16, depth 6,
  96 GB budget, submitted while earlier jobs run.  Every result equals the same job
  rendered alone; sampled jobs' pixels equal the CPU restatement's running mean; the
  budget holds at every round.
* Radiance cache: bi_pt_render_nrc on the 262,144-triangle atrium at 1920x1080, depth 8,
  65,536 training paths per frame: the first 4,096 training records and targets equal
  the restatement's, and the image is finite and non-negative.
"""
import numpy as np
import pytest

import bi_pt
import scenes

pytestmark = pytest.mark.gpu


def subset(rng, cam, n):
    return np.sort(rng.choice(cam.width * cam.height, n, replace=False)).astype(np.uint32)


def test_batch_c2_graph_replay_equals_solo_and_oracle(oracle_lib):
    import torch
    specs = scenes.batch_c2()
    print(f"case: 64-scene batch ({sum(s.num_tris for s, _, _ in specs)} tris), graph capture + replay", flush=True)
    solo = []
    for sc, cam, kw in specs:
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        t.render(cam, **kw)
        solo.append(t.sums(cam.width * cam.height).copy())
        t.close()
    print("solo renders done", flush=True)
    ts = []
    for sc, _, _ in specs:
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        ts.append(t)
    b = bi_pt.Batch()
    b.set_tuning(use_graph=True)
    stream = torch.cuda.Stream()
    jobs = [bi_pt.Job(t, cam, **kw) for (sc, cam, kw), t in zip(specs, ts)]
    for call in ("capture", "replay"):
        with torch.cuda.stream(stream):
            b.render(jobs, stream=stream.cuda_stream)
        stream.synchronize()
        for k, ((sc, cam, kw), t) in enumerate(zip(specs, ts)):
            got = t.sums(cam.width * cam.height)
            assert np.array_equal(got.view(np.uint32), solo[k].view(np.uint32)), (call, k)
        print(f"batched call ({call}) equals the solo renders", flush=True)
    st = b.stats()
    assert st["rays_camera"] == 2 * sum(cam.width * cam.height * kw["spp"] for _, cam, kw in specs)
    rng = np.random.default_rng(0xC2)
    for k in sorted(rng.choice(len(specs), 4, replace=False)):
        sc, cam, kw = specs[k]
        pix = subset(rng, cam, 48)
        want = oracle_lib.OracleScene(sc).render32_pixels(cam, kw["max_bounces"], kw["seed"], kw["spp"], pix)
        got = ts[k].sums(cam.width * cam.height)[pix]
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (k, sc.name)
        print(f"job {k} ({sc.num_tris} tris, {cam.width}x{cam.height}, {kw['spp']} spp): 48 pixels equal the "
              f"oracle", flush=True)
    b.close()
    for t in ts:
        t.close()


def test_batch_c2_trees_equal_oracle(oracle_lib):
    """After one batched call of the 64 scenes (their LBVHs built concurrently on four
    streams, treelet rounds where the call traces enough paths), every scene's BVH4 is the
    restatement's, node for node and bit for bit."""
    import torch
    specs = scenes.batch_c2()
    print("case: 64-scene batch, eager call, every BVH4 against the oracle's", flush=True)
    ts, rounds = [], []
    for sc, cam, kw in specs:
        paths = cam.width * cam.height * kw["spp"]
        r = 3 if paths >= max(4 * sc.num_tris, 1 << 20) else 0   # the call's rule for treelet rounds
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        t.set_build_options(treelet_rounds=r)
        ts.append(t)
        rounds.append(r)
    b = bi_pt.Batch()
    stream = torch.cuda.Stream()
    jobs = [bi_pt.Job(t, cam, **kw) for (sc, cam, kw), t in zip(specs, ts)]
    with torch.cuda.stream(stream):
        b.render(jobs, stream=stream.cuda_stream)
    stream.synchronize()
    b.stats()   # raises on any device flag
    assert sum(1 for r in rounds if r) > 8
    for k, ((sc, _, _), t, r) in enumerate(zip(specs, ts, rounds)):
        o = oracle_lib.OracleScene(sc)
        o.set_treelet_rounds(r)
        oc, ob = o.bvh4()
        o.close()
        gc, gb = t.bvh4()
        assert np.array_equal(gc, oc) and np.array_equal(gb.view(np.uint32), ob.view(np.uint32)), (k, sc.num_tris, r)
    print(f"64 trees equal ({sum(1 for r in rounds if r)} restructured)", flush=True)
    b.close()
    for t in ts:
        t.close()


def test_serving_c3_first_arrivals(oracle_lib):
    templates = scenes.templates_c3()
    arrivals = scenes.arrivals_c3()[:100]
    print(f"case: serving, first {len(arrivals)} arrivals over templates {[s.num_tris for s in templates]}",
          flush=True)
    owners = []
    for sc in templates:
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        owners.append(t)
    budget = 96 << 30
    kw = dict(max_bounces=scenes.C3_MAX_BOUNCES, seed=scenes.C3_SEED)
    srv = bi_pt.Server(hbm_budget_bytes=budget, spp_per_round=4)
    sids = [srv.add_scene(t) for t in owners]
    ids = []
    i = 0
    peak_hbm = 0
    peak_running = 0
    rounds = 0
    while True:
        # three arrivals per round: later jobs join while earlier ones are running
        for _ in range(3):
            if i < len(arrivals):
                _, k, cam, spp = arrivals[i]
                ids.append(srv.submit(sids[k], cam, spp=spp, **kw))
                i += 1
        active = srv.step()
        rounds += 1
        st = srv.stats()
        peak_hbm = max(peak_hbm, st["hbm_used"])
        peak_running = max(peak_running, st["running"])
        assert st["hbm_used"] <= budget
        if active == 0 and i == len(arrivals):
            break
        assert rounds < 2000
    print(f"served in {rounds} rounds, peak {peak_running} jobs running, peak {peak_hbm / 2**20:.0f} MiB", flush=True)
    assert peak_running > 1
    results = []
    for jid, (_, k, cam, spp) in zip(ids, arrivals):
        s = srv.poll(jid)
        assert s["state"] == 2 and s["spp_done"] == spp
        results.append(srv.take_result(jid, cam.width * cam.height))
    srv.close()
    # each result is the job rendered alone (a view of the same resident scene)
    views = [bi_pt.PathTracer(view_of=t) for t in owners]
    for n, (jid, (_, k, cam, spp)) in enumerate(zip(ids, arrivals)):
        views[k].render(cam, spp=spp, flags=bi_pt.FLAG_REUSE_BVH, **kw)
        want = views[k].image()
        assert np.array_equal(results[n].view(np.uint32), want.view(np.uint32)), (n, k, cam.width, spp)
    print("every served job equals its solo render", flush=True)
    rng = np.random.default_rng(0xC3)
    oras = {}
    for n in sorted(rng.choice(len(arrivals), 4, replace=False)):
        _, k, cam, spp = arrivals[n]
        if k not in oras:
            oras[k] = oracle_lib.OracleScene(templates[k])
        pix = subset(rng, cam, 32)
        sums = oras[k].render32_pixels(cam, kw["max_bounces"], kw["seed"], spp, pix)
        mean = sums / np.float32(spp)
        assert np.array_equal(results[n][pix, :3].view(np.uint32), mean.view(np.uint32)), (n, k)
        print(f"job {n} (template {k}, {cam.width}x{cam.height}, {spp} spp): 32 pixels equal the oracle", flush=True)
    for t in views + owners:
        t.close()


def test_nrc_c4_atrium_1080p_training_samples(oracle_lib):
    sc = scenes.scene_c1()
    cam = scenes.camera_c1()
    mb = 9
    print("case: radiance cache on the atrium at 1920x1080, depth 8, 65,536 training paths", flush=True)
    t = bi_pt.PathTracer()
    t.set_scene(sc)
    nrc = bi_pt.Nrc(t, train_samples=65536, train_steps=4, cache_bounce=2, seed=1)
    seed = 0x5EED0004
    for frame in range(3):
        nrc.render(cam, spp=1, max_bounces=mb, seed=seed)
    img = t.image()
    assert np.all(np.isfinite(img)) and np.all(img >= 0.0)
    assert img[:, :3].mean() > 0.0
    rec, tgt = nrc.samples(4096)
    orec, otgt = oracle_lib.OracleScene(sc).train_samples(cam, mb, bi_pt.nrc_train_seed(seed, 2), 2, 4096)
    valid = orec[:, 13].view(np.uint32) == 1
    print(f"training samples: {int(valid.sum())} of 4096 reach segment 2", flush=True)
    assert valid.sum() > 1000
    assert np.array_equal(rec.view(np.uint32), orec.view(np.uint32))
    assert np.array_equal(tgt.view(np.uint32), otgt.view(np.uint32))
    st = nrc.stats()
    assert st["frames"] == 3 and st["steps"] == 12
    assert np.isfinite(st["last_loss"])
    nrc.close()
    t.close()
