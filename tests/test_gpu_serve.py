"""Serving loop (bi_pt_server) and scene views (bi_pt_create_view).

Each job finished by the server must equal, bit for bit, the same camera and settings
rendered alone by bi_pt_render on a context holding its scene -- whatever other jobs it
shared its rounds with, however its samples were split over rounds, and whenever it was
admitted.  Admission must respect the HBM budget.
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


def test_server_results_equal_solo(c0_scene):
    print("case: server, 9 jobs over 2 scenes, 3 spp per round", flush=True)
    plane = plane_scene()
    owners = []
    for sc in (c0_scene, plane):
        t = bi_pt.PathTracer()
        t.set_scene(sc)
        owners.append(t)
    srv = bi_pt.Server(spp_per_round=3, max_jobs_per_round=4)
    sids = [srv.add_scene(t) for t in owners]
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
