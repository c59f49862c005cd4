"""Serving loop (bi_pt_server) and scene views (bi_pt_create_view).

Each job finished by the server must equal, bit for bit, the same camera and settings
rendered alone by bi_pt_render on a context holding its scene, and the CPU
restatement's image -- whatever other jobs it shared its rounds with, however its
samples were split over rounds, and whenever it was admitted.  Admission must respect
the HBM budget.
"""
import numpy as np
import pytest

import bi_pt
import scenes
from test_oracle_estimator import plane_camera, plane_scene

pytestmark = pytest.mark.gpu


def solo_image(scene, cam, **kw):
    t = bi_pt.PathTracer()
    t.set_scene(scene)
    t.render(cam, **kw)
    img = t.image().copy()
    t.close()
    return img


def test_view_renders_like_its_scene(c0_scene):
    print("case: view of the c0 scene vs the scene context", flush=True)
    owner = bi_pt.PathTracer()
    owner.set_scene(c0_scene)
    view = bi_pt.PathTracer(view_of=owner)
    cam = scenes.camera_c0(48, 40)
    owner.render(cam, spp=3, max_bounces=5, seed=21)
    a = owner.image().copy()
    view.render(cam, spp=3, max_bounces=5, seed=21)
    assert np.array_equal(view.image().view(np.uint32), a.view(np.uint32))
    assert view.num_tris() == owner.num_tris()
    with pytest.raises(bi_pt.PtError):
        view.set_scene(c0_scene)
    # accumulation of the view continues across calls like a context's
    view.render(cam, spp=2, max_bounces=5, seed=21, accumulate=True)
    owner.render(cam, spp=5, max_bounces=5, seed=21)
    assert np.array_equal(view.image().view(np.uint32), owner.image().view(np.uint32))
    view.close()
    owner.close()


def test_server_results_equal_solo(c0_scene, oracle_lib):
    print("case: server, 9 jobs over 2 scenes, 3 spp per round, vs solo and oracle", flush=True)
    plane = plane_scene()
    owners = []
    for sc in (c0_scene, plane):
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        owners.append(t)
    srv = bi_pt.Server(spp_per_round=3, max_jobs_per_round=4)
    sids = [srv.add_scene(t) for t in owners]
    oracles = [oracle_lib.OracleScene(sc) for sc in (c0_scene, plane)]
    specs = []
    rng = np.random.default_rng(5)
    for k in range(9):
        which = k % 2
        w, h = [(40, 30), (24, 24), (64, 20)][k % 3]
        cam = scenes.camera_c0(w, h) if which == 0 else plane_camera(w, h)
        kw = dict(spp=int(rng.integers(1, 9)), max_bounces=int(rng.integers(2, 8)), seed=100 + k,
                  flags=bi_pt.FLAG_JITTER if k % 4 == 0 else 0)
        specs.append((which, cam, kw))
    ids = []
    for k, (which, cam, kw) in enumerate(specs):
        ids.append(srv.submit(sids[which], cam, **kw))
        if k % 3 == 2:
            srv.step()    # later jobs arrive while earlier ones are running
    rounds = 0
    while srv.step() and rounds < 100:
        rounds += 1
    st = srv.stats()
    assert st["waiting"] == 0 and st["running"] == 0 and st["rays"] > 0
    for jid, (which, cam, kw) in zip(ids, specs):
        s = srv.poll(jid)
        assert s["state"] == 2 and s["spp_done"] == kw["spp"]
        assert s["submit_s"] <= s["admit_s"] <= s["done_s"]
        got = srv.take_result(jid, cam.width * cam.height)
        want = solo_image((c0_scene, plane)[which], cam, **kw)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (jid, kw)
        # the restatement's running mean: sums / spp, both correctly rounded
        sums = oracles[which].render32(cam, kw["max_bounces"], kw["seed"], spp=kw["spp"], flags=kw["flags"])["sums"]
        mean = sums / np.float32(kw["spp"])
        assert np.array_equal(got[:, :3].view(np.uint32), mean.view(np.uint32)), (jid, kw)
    srv.close()
    for t in owners:
        t.close()


def test_server_admission_by_budget(c0_scene):
    print("case: server admission under a small HBM budget", flush=True)
    owner = bi_pt.PathTracer()
    owner.set_scene(c0_scene)
    cam = scenes.camera_c0(256, 256)
    per_job = 256 * 256 * 28
    # measure the resident scene plus wave buffers at their cap (2 jobs x 65,536 paths)
    probe = bi_pt.Server(spp_per_round=1, wave_paths=1 << 17)
    sid = probe.add_scene(owner)
    for k in range(2):
        probe.submit(sid, cam, spp=1, max_bounces=3, seed=1 + k)
    while probe.step():
        pass
    fixed = probe.stats()["hbm_used"] - 2 * per_job
    probe.close()
    # room for the fixed part and two and a half jobs: at most two run at once
    srv = bi_pt.Server(hbm_budget_bytes=fixed + 2 * per_job + per_job // 2, spp_per_round=1,
                       wave_paths=1 << 17)
    sid = srv.add_scene(owner)
    ids = [srv.submit(sid, cam, spp=2, max_bounces=3, seed=7 + k) for k in range(6)]
    peak = 0
    done_order = []
    for _ in range(200):
        active = srv.step()
        peak = max(peak, srv.stats()["running"])
        for jid in ids:
            if jid not in done_order and srv.poll(jid)["state"] == 2:
                done_order.append(jid)
                got = srv.take_result(jid, 256 * 256)   # frees the job's memory
                k = ids.index(jid)
                want = solo_image(c0_scene, cam, spp=2, max_bounces=3, seed=7 + k)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        if active == 0:
            break
    assert 1 <= peak <= 2
    assert sorted(done_order) == sorted(ids)
    with pytest.raises(bi_pt.PtError):
        srv.submit(sid, scenes.camera_c0(8192, 8192), spp=1)   # can never fit
    srv.close()
    owner.close()


def test_server_rejects_what_a_round_would_reject(c0_scene):
    """Jobs a render would refuse are refused at submission (a zero-initialised ray length,
    an oversized image, the RGBA16F output flag), so they never fail the other jobs'
    rounds; the valid job next to them completes and equals its solo render."""
    print("case: server submission checks next to a valid job", flush=True)
    owner = bi_pt.PathTracer()
    owner.set_scene(c0_scene)
    srv = bi_pt.Server(spp_per_round=2)
    sid = srv.add_scene(owner)
    cam = scenes.camera_c0(32, 24)
    good = srv.submit(sid, cam, spp=3, max_bounces=4, seed=77)
    with pytest.raises(bi_pt.PtError) as e:
        srv.submit(sid, cam, spp=2, max_bounces=4, seed=78, ray_length=0.0)
    assert e.value.status == 1
    with pytest.raises(bi_pt.PtError) as e:
        srv.submit(sid, scenes.camera_c0(8192, 8193), spp=1)
    assert e.value.status == 6
    with pytest.raises(bi_pt.PtError) as e:
        srv.submit(sid, cam, spp=2, flags=bi_pt.FLAG_OUTPUT_RGBA16F)
    assert e.value.status == 6
    while srv.step():
        pass
    assert srv.poll(good)["state"] == 2
    got = srv.take_result(good, cam.width * cam.height)
    want = solo_image(c0_scene, cam, spp=3, max_bounces=4, seed=77)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    srv.close()
    owner.close()


def test_server_failed_round_fails_its_jobs_only(c0_scene):
    """A round whose kernels raise an internal-limit flag (raised here from the host) fails
    the jobs of that round with BI_PT_JOB_FAILED; jobs still waiting and later ones go on."""
    print("case: server round failure (flag raised from the host)", flush=True)
    owner = bi_pt.PathTracer()
    owner.set_scene(c0_scene)
    srv = bi_pt.Server(spp_per_round=1, max_jobs_per_round=2)
    sid = srv.add_scene(owner)
    cam = scenes.camera_c0(24, 16)
    a = srv.submit(sid, cam, spp=3, max_bounces=4, seed=1)
    b = srv.submit(sid, cam, spp=3, max_bounces=4, seed=2)
    c = srv.submit(sid, cam, spp=2, max_bounces=4, seed=3)   # waits: two jobs per round
    assert srv.step() == 3
    srv.inject_error(1)
    with pytest.raises(bi_pt.PtError) as e:
        srv.step()
    assert e.value.status == 5 and "stack overflow" in str(e.value)
    assert srv.poll(a)["state"] == 3 and srv.poll(b)["state"] == 3
    assert srv.poll(c)["state"] == 0
    with pytest.raises(bi_pt.PtError) as e:
        srv.take_result(a, cam.width * cam.height)
    assert e.value.status == 5
    d = srv.submit(sid, cam, spp=2, max_bounces=4, seed=4)
    while srv.step():
        pass
    for jid, seed in ((c, 3), (d, 4)):
        assert srv.poll(jid)["state"] == 2
        got = srv.take_result(jid, cam.width * cam.height)
        want = solo_image(c0_scene, cam, spp=2, max_bounces=4, seed=seed)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    srv.close()
    owner.close()


def test_error_flags_are_sticky_until_checked(c0_scene):
    """A flag raised before or during one render is still reported after later renders,
    by the next statistics call, which clears it."""
    print("case: sticky device error flags", flush=True)
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    cam = scenes.camera_c0(24, 16)
    t.render(cam, spp=1, max_bounces=3, seed=1)
    t.stats()
    t.inject_error(8)
    t.render(cam, spp=1, max_bounces=3, seed=2)
    t.render(cam, spp=1, max_bounces=3, seed=3)
    with pytest.raises(bi_pt.PtError) as e:
        t.stats()
    assert e.value.status == 5 and "step bound" in str(e.value)
    t.render(cam, spp=1, max_bounces=3, seed=4)
    t.stats()   # cleared by the report
    t.close()


def test_server_with_textured_rect_lights_matches_oracle(c0_scene, oracle_lib):
    print("case: server, 4 jobs over textured-rect c0 and c0, 2 spp per round, vs oracle", flush=True)
    from test_gpu_rect_lights import textured_c0
    lit = textured_c0(9)
    owners = []
    for sc in (lit, c0_scene):
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        owners.append(t)
    srv = bi_pt.Server(spp_per_round=2, max_jobs_per_round=3)
    sids = [srv.add_scene(t) for t in owners]
    specs = [(0, scenes.camera_c0(40, 30), dict(spp=3, max_bounces=5, seed=71, flags=bi_pt.FLAG_JITTER)),
             (1, scenes.camera_c0(24, 24), dict(spp=2, max_bounces=4, seed=72, flags=0)),
             (0, scenes.camera_c0(32, 20), dict(spp=5, max_bounces=6, seed=73, flags=0)),
             (1, scenes.camera_c0(16, 16), dict(spp=1, max_bounces=3, seed=74, flags=0))]
    ids = [srv.submit(sids[w], cam, **kw) for w, cam, kw in specs]
    rounds = 0
    while srv.step() and rounds < 100:
        rounds += 1
    for jid, (which, cam, kw) in zip(ids, specs):
        s = srv.poll(jid)
        assert s["state"] == 2 and s["spp_done"] == kw["spp"]
        got = srv.take_result(jid, cam.width * cam.height)
        sums = oracle_lib.OracleScene((lit, c0_scene)[which]).render32(cam, kw["max_bounces"], kw["seed"],
                                                                       spp=kw["spp"], flags=kw["flags"])["sums"]
        mean = sums / np.float32(kw["spp"])
        assert np.array_equal(got[:, :3].view(np.uint32), mean.view(np.uint32)), (jid, kw)
    srv.close()
    for t in owners:
        t.close()
