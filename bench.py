"""Benchmark of the wavefront path-tracing pass on one MI355X.

A step is one bi_pt_render call: the per-frame LBVH rebuild plus 16 spp of the
1920x1080 atrium (262,144 triangles, depth 8), as one render job.  Prints one JSON
line with Mrays/s (all traced rays: camera, extension, light-shadow, sky-shadow),
p50/p99 job latency, the trace kernel's roofline figures and the CPU baseline
(the CPU restatement timed on this host's cores on a bounded sample).

Two more legs report the other forms beside the plain pass: "batched" (64 scenes of
1k-500k triangles in one bi_pt_render_batch call, autotuned, replayed from a hipGraph,
against the same jobs rendered one by one) and "serving" (Poisson arrivals over 8
resident scenes through bi_pt_server, p50/p99 job latency against first-come
first-served single renders).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("bisemutum-engine_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BYTES_PER_TRACED_RAY = 40      # SoA ray read (o, d: 24 B) + hit write (t, leaf, u, v: 16 B)
BYTES_PER_CAMERA_RAY = 16      # camera rays are generated in the kernel: hit write only


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_k_trace_closest_c1.md")


def measured_traffic():
    """HBM bytes per k_trace_closest launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a copy of known size), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def reachable_bvh4_nodes(pt):
    """BVH4 nodes reachable from the root (the nodes a traversal can read)."""
    import numpy as np
    child4, _ = pt.bvh4()
    seen = np.zeros(len(child4), bool)
    stack = [0]
    while stack:
        k = stack.pop()
        if seen[k]:
            continue
        seen[k] = True
        stack.extend(int(c) for c in child4[k] if c > 0)   # 0 = the root: only empty slots name it
    return int(seen.sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1", choices=["c0", "c1"])
    ap.add_argument("--wave-paths", type=int, default=0)
    ap.add_argument("--stack-lds", type=int, default=0)
    ap.add_argument("--refill-lanes", type=int, default=0)
    ap.add_argument("--work-parts", type=int, default=0)
    ap.add_argument("--block-threads", type=int, default=0, help="traversal workgroup size: 128, 256 or 512")
    ap.add_argument("--treelet-rounds", type=int, default=-1,
                    help="rounds of treelet restructuring after the LBVH (-1: the library's default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-rmse", action="store_true", help="skip the FP64 oracle RMSE leg")
    ap.add_argument("--no-nrc", action="store_true", help="skip the radiance-cache leg")
    ap.add_argument("--rmse-frames", type=int, default=16, help="frames averaged in each RMSE figure")
    ap.add_argument("--nrc-warmup", type=int, default=96, help="frames of online training before the cache's RMSE")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched leg")
    ap.add_argument("--batch-calls", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="arrival window of the serving leg (0 skips it; 20 s = about 4,100 jobs)")
    ap.add_argument("--serve-spp-per-round", type=int, default=16,
                    help="samples of each admitted job per scheduling round of the serving leg")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed runs of the CPU baseline (after one warm-up)")
    ap.add_argument("--profile-nrc", action="store_true",
                    help="only run the radiance-cache frames that profiles/kernels_nrc_c1_1080p.md profiles "
                         "(8 frames at 1 spp of online training, then 4 at 16 spp) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device-to-device copy first (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this pass runs on one GPU only; multi-GPU scaling is not part of this build "
              "(image tiles would shard with no exchange but the final gather).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import bi_pt
    import scenes

    if args.profile_nrc:
        profile_nrc()
        return
    cfg = scenes.CONFIGS[args.config]
    t = time.perf_counter()
    scene = cfg["scene"]()
    cam = cfg["camera"]()
    log(f"scene {scene.name}: {scene.num_tris} tris, generated in {time.perf_counter() - t:.1f}s")
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {src.numel() * 4} bytes read, {dst.numel() * 4} bytes written")
        del src, dst
    pt = bi_pt.PathTracer(0)
    pt.set_scene(scene)
    if args.treelet_rounds >= 0:
        pt.set_build_options(treelet_rounds=args.treelet_rounds)
    stream = torch.cuda.current_stream().cuda_stream
    spp, mb, seed = cfg["spp"], cfg["max_bounces"], cfg["seed"]

    def step(k, flags=0):
        pt.render(cam, spp=spp, max_bounces=mb, seed=seed + k, flags=flags, wave_paths=args.wave_paths,
                  stack_lds=args.stack_lds, refill_lanes=args.refill_lanes, work_parts=args.work_parts,
                  block_threads=args.block_threads, stream=stream)

    for k in range(args.warmup):
        step(k)
        torch.cuda.synchronize()
        log(f"warmup step {k} done")
    before = pt.stats()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        step(1000 + k)
        ends[k].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    after = pt.stats()
    lat = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    rays = sum(after[k] - before[k] for k in ("rays_camera", "rays_extension", "rays_light_shadow",
                                                "rays_sky_shadow"))
    secs = t1 - t0
    mrays = rays / secs / 1e6
    log(f"timed {args.steps} steps: {secs:.3f}s, {rays} rays, {mrays:.1f} Mrays/s")

    # roofline leg: one more frame with per-kernel HIP events on the render stream
    s0 = pt.stats()
    step(5000, flags=bi_pt.FLAG_TIMING)
    s1 = pt.stats()
    n = pt.num_tris()
    nodes4 = reachable_bvh4_nodes(pt)
    bvh_bytes = 128 * nodes4 + 48 * n   # reachable BVH4 nodes + leaf intersection records
    cam_rays = s1["rays_camera"] - s0["rays_camera"]
    ext_rays = s1["rays_extension"] - s0["rays_extension"]
    launches = s1["launches_trace_closest"]
    alg = cam_rays * BYTES_PER_CAMERA_RAY + ext_rays * BYTES_PER_TRACED_RAY + launches * bvh_bytes
    ms_trace = s1["ms_trace_closest"]
    avg_launch_ms = ms_trace / max(launches, 1)
    achieved = (alg / launches) / (avg_launch_ms * 1e-3) / 1e9
    timing = {k: round(s1[k], 3) for k in ("ms_build", "ms_trace_closest", "ms_shade", "ms_trace_shadow",
                                           "ms_resolve", "ms_total")}
    rays_by_kind = {k: int(s1[k] - s0[k]) for k in ("rays_camera", "rays_extension", "rays_light_shadow",
                                                    "rays_sky_shadow")}
    log(f"kernel classes (ms, one frame): {timing}")
    log(f"rays by kind (one frame): {rays_by_kind}")

    result = {
        "metric": "Mrays/s and p99 render-job latency at equal RMSE vs FP64 oracle",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, generated in-repo)",
        "config": {
            "workload": f"plain wavefront path tracing, {scene.name} ({n} tris, LBVH rebuilt per frame), "
                        f"{cam.width}x{cam.height}, {spp} spp, depth {mb - 1}",
            "width": cam.width, "height": cam.height, "spp": spp, "max_bounces": mb,
            "wave_paths": args.wave_paths, "stack_lds": args.stack_lds,
            "refill_lanes": args.refill_lanes, "work_parts": args.work_parts, "block_threads": args.block_threads,
            "treelet_rounds": args.treelet_rounds,
        },
        "job_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3)},
        "rays_per_step": int(rays // args.steps),
        "kernel_ms_one_frame": timing,
        "rays_by_kind_one_frame": rays_by_kind,
        "roofline": {
            "kernel": "k_trace_closest (camera + extension rays)",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(),
            "alg_bytes_per_launch": int(alg / launches),
            "avg_launch_ms": round(avg_launch_ms, 4),
            "launches": int(launches),
        },
    }

    ref = fp64_reference(scene) if not args.no_rmse else None
    if ref is None and not args.no_rmse:
        log("rmse and nrc legs skipped: the FP64 reference fixture is missing or was made for another scene "
            "(python tests/golden/make_c1_fp64.py writes it)")
    if ref is not None:
        result["rmse_vs_fp64"] = rmse_leg(pt, cam, cfg, ref, args.rmse_frames)
        if not args.no_nrc:
            result["nrc"] = nrc_leg(pt, cam, cfg, ref, result["rmse_vs_fp64"], args.rmse_frames, args.nrc_warmup)
    if not args.no_batch:
        result["batched"] = batch_leg(args.batch_calls)
    if args.serve_seconds > 0:
        result["serving"] = serve_leg(args.serve_seconds, spp_per_round=args.serve_spp_per_round)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_leg(scene, cfg, args.cpu_runs)
    pt.close()
    print(json.dumps(result), flush=True)


RAY_KINDS = ("rays_camera", "rays_extension", "rays_light_shadow", "rays_sky_shadow")


def batch_leg(steps):
    """Multi-scene form: 64 c0-style scenes (1k-500k tris, log-uniform), one job each
    (256x256 / 640x480 / 1280x720, 1-8 spp, depth 6), LBVHs rebuilt every call.
    The plain form renders the 64 jobs one after another with bi_pt_render; the
    batched form renders them in one bi_pt_render_batch call (shared waves) with the
    autotuned knobs, replayed from a hipGraph."""
    import torch

    import bi_pt
    import scenes
    t = time.perf_counter()
    specs = scenes.batch_c2()
    tracers = []
    for sc, _, _ in specs:
        tr = bi_pt.PathTracer()
        tr.set_scene(sc)
        tracers.append(tr)
    tris = sum(sc.num_tris for sc, _, _ in specs)
    paths = sum(cam.width * cam.height * kw["spp"] for _, cam, kw in specs)
    log(f"batch leg: {len(specs)} scenes, {tris} tris, {paths} paths, set up in {time.perf_counter() - t:.1f}s")
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream

    def rays_of(objs):
        return sum(sum(o.stats()[k] for k in RAY_KINDS) for o in objs)

    # plain: one bi_pt_render per job, back to back on one stream
    def plain():
        for (sc, cam, kw), tr in zip(specs, tracers):
            tr.render(cam, stream=sh, **kw)

    with torch.cuda.stream(stream):
        plain()
        stream.synchronize()
        r0 = rays_of(tracers)
        t0 = time.perf_counter()
        for _ in range(steps):
            plain()
        stream.synchronize()
        plain_s = (time.perf_counter() - t0) / steps
        plain_rays = (rays_of(tracers) - r0) // steps

        b = bi_pt.Batch()
        jobs = [bi_pt.Job(tr, cam, **kw) for (sc, cam, kw), tr in zip(specs, tracers)]
        t = time.perf_counter()
        tuned = b.autotune(jobs, rounds=1, stream=sh)
        tune_s = time.perf_counter() - t
        b.set_tuning(use_graph=True, **{k: tuned[k] for k in ("wave_paths", "stack_lds", "refill_lanes",
                                                              "work_parts", "block_threads")})
        b.render(jobs, stream=sh)          # capture
        stream.synchronize()
        r0 = rays_of([b])
        t0 = time.perf_counter()
        for _ in range(steps):
            b.render(jobs, stream=sh)      # replay
        stream.synchronize()
        batch_s = (time.perf_counter() - t0) / steps
        batch_rays = (rays_of([b]) - r0) // steps
    b.close()
    for tr in tracers:
        tr.close()
    log(f"batch leg: plain {plain_s * 1e3:.1f} ms ({plain_rays / plain_s / 1e6:.0f} Mrays/s), "
        f"batched {batch_s * 1e3:.1f} ms ({batch_rays / batch_s / 1e6:.0f} Mrays/s), autotune {tune_s:.1f}s")
    return {"workload": f"{len(specs)} c0-style scenes, {tris} tris in all (1k-500k each, log-uniform), one job "
                        f"each: 256x256 / 640x480 / 1280x720, 1-8 spp, depth 6, LBVHs rebuilt per call; "
                        f"{paths} paths per call",
            "calls": steps, "rays_per_call": int(batch_rays),
            "plain_ms_per_call": round(plain_s * 1e3, 2), "plain_mrays_s": round(plain_rays / plain_s / 1e6, 1),
            "batched_ms_per_call": round(batch_s * 1e3, 2), "batched_mrays_s": round(batch_rays / batch_s / 1e6, 1),
            "tuning": {"wave_paths": int(tuned["wave_paths"]) or 1 << 25,   # 0 = the engine's default
                       "stack_lds": int(tuned["stack_lds"]) or 16, "refill_lanes": int(tuned["refill_lanes"]) or 32,
                       "block_threads": int(tuned["block_threads"]) or 256},
            "autotune_s": round(tune_s, 1)}


def serve_leg(seconds, rate=200.0, spp_per_round=16):
    """Serving form: Poisson arrivals at `rate` jobs/s for `seconds` over 8 resident
    templates (10k-1M tris); orbit cameras, 256x256 / 512x512 / 1280x720, spp 1/4/16,
    depth 6, 96 GB HBM budget.  The plain form serves the same arrivals first-come
    first-served with one bi_pt_render per job (latencies from measured job times)."""
    import numpy as np
    import torch

    import bi_pt
    import scenes
    templates = scenes.templates_c3()
    tracers = []
    for sc in templates:
        tr = bi_pt.PathTracer()
        tr.set_scene(sc)
        tracers.append(tr)
    jobs = scenes.arrivals_c3(seconds, rate, len(templates))
    kw = dict(max_bounces=scenes.C3_MAX_BOUNCES, seed=scenes.C3_SEED)
    log(f"serve leg: {len(jobs)} jobs over {seconds}s, templates {[s.num_tris for s in templates]}")

    srv = bi_pt.Server(hbm_budget_bytes=96 << 30, spp_per_round=spp_per_round)
    sids = [srv.add_scene(tr) for tr in tracers]
    torch.cuda.synchronize()
    ids = []
    i = 0
    t0 = srv.clock()
    peak = 0
    while True:
        now = srv.clock() - t0
        while i < len(jobs) and jobs[i][0] <= now:
            a, k, cam, spp = jobs[i]
            ids.append(srv.submit(sids[k], cam, spp=spp, arrival_s=t0 + a, **kw))
            i += 1
        active = srv.step()
        peak = max(peak, srv.stats()["running"])
        if active == 0:
            if i == len(jobs):
                break
            time.sleep(max(0.0, jobs[i][0] - (srv.clock() - t0)))
    t_end = srv.clock() - t0
    lat = []
    for jid in ids:
        s = srv.poll(jid)
        lat.append(s["done_s"] - s["submit_s"])
        srv.take_result(jid, None)
    st = srv.stats()
    srv.close()
    lat = np.array(lat) * 1e3

    # plain form: first come first served, one bi_pt_render per job, BVHs kept resident
    dur = []
    rays0 = sum(sum(tr.stats()[k] for k in RAY_KINDS) for tr in tracers)
    torch.cuda.synchronize()
    t_plain = time.perf_counter()
    for a, k, cam, spp in jobs:
        t = time.perf_counter()
        tracers[k].render(cam, spp=spp, flags=bi_pt.FLAG_REUSE_BVH, **kw)
        torch.cuda.synchronize()
        dur.append(time.perf_counter() - t)
    t_plain = time.perf_counter() - t_plain
    plain_rays = sum(sum(tr.stats()[k] for k in RAY_KINDS) for tr in tracers) - rays0
    done = 0.0
    plat = []
    for (a, _, _, _), d in zip(jobs, dur):
        done = max(done, a) + d
        plat.append(done - a)
    plat = np.array(plat) * 1e3
    for tr in tracers:
        tr.close()
    log(f"serve leg: served {len(ids)} jobs in {t_end:.2f}s, p50 {np.percentile(lat, 50):.1f} ms, "
        f"p99 {np.percentile(lat, 99):.1f} ms; plain FCFS p50 {np.percentile(plat, 50):.1f} ms, "
        f"p99 {np.percentile(plat, 99):.1f} ms")
    return {"workload": f"Poisson arrivals {rate:g} jobs/s for {seconds:g}s ({len(jobs)} jobs) over 8 resident "
                        f"c0-style templates ({min(s.num_tris for s in templates)}-"
                        f"{max(s.num_tris for s in templates)} tris), orbit cameras, 256x256 / 512x512 / "
                        f"1280x720, spp 1/4/16, depth 6, 96 GB HBM budget, {spp_per_round} spp per job per round",
            "jobs": len(jobs), "served_s": round(t_end, 3), "rays": int(st["rays"]),
            "mrays_s": round(st["rays"] / t_end / 1e6, 1),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                           "p99": round(float(np.percentile(lat, 99)), 2)},
            "peak_running_jobs": int(peak), "rounds": int(st["rounds"]),
            "plain_fcfs": {"latency_ms": {"p50": round(float(np.percentile(plat, 50)), 2),
                                          "p99": round(float(np.percentile(plat, 99)), 2)},
                           "busy_s": round(t_plain, 3), "mrays_s": round(plain_rays / t_plain / 1e6, 1)}}


FP64_REFERENCE = os.path.join(ROOT, "tests", "golden", "c1_fp64_4096spp.npz")


def fp64_reference(scene):
    """FP64 oracle means at 4096 spp on 2048 seeded pixels of the atrium view (a committed
    fixture, tests/golden/make_c1_fp64.py), or None when the scene differs from it."""
    import hashlib

    import numpy as np
    if not os.path.exists(FP64_REFERENCE):
        return None
    ref = np.load(FP64_REFERENCE)
    h = hashlib.sha256()
    for p, n, i in scene.meshes:
        h.update(p.tobytes()); h.update(n.tobytes()); h.update(i.tobytes())
    h.update(scene.instances.tobytes()); h.update(scene.materials.tobytes()); h.update(scene.dir_lights.tobytes())
    h.update(scene.sky_rgba.tobytes())
    return ref if str(ref["scene_sha256"]) == h.hexdigest() else None


def image_errors(img, ref):
    """(relative RMSE, relative MSE) of an image's reference pixels against the FP64 means."""
    import numpy as np
    m = ref["mean"]
    e = img[ref["pixels"], :3].astype(np.float64) - m
    return float(np.sqrt((e ** 2).mean() / (m ** 2).mean())), float((e ** 2 / (m ** 2 + 1e-2)).mean())


def timed(fn):
    """Milliseconds of fn's work on the current stream (HIP events)."""
    import torch
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(torch.cuda.current_stream())
    fn()
    b.record(torch.cuda.current_stream())
    b.synchronize()
    return a.elapsed_time(b)


def rmse_leg(pt, cam, cfg, ref, frames):
    """Plain frames (16 spp, LBVH rebuilt) against the FP64 reference: RMSE averaged over frames."""
    import numpy as np
    ms, mse, rel = [], [], []
    for k in range(frames):
        t = timed(lambda: pt.render(cam, spp=cfg["spp"], max_bounces=cfg["max_bounces"], seed=cfg["seed"] + 7000 + k),
                  )
        r, q = image_errors(pt.image(), ref)
        ms.append(t)
        mse.append(r * r)
        rel.append(q)
    out = {"relative_rmse": round(float(np.sqrt(np.mean(mse))), 4), "relative_mse": round(float(np.mean(rel)), 4),
           "frames": frames, "ms_per_frame": round(float(np.mean(ms)), 2), "spp": cfg["spp"],
           "pixels": int(len(ref["pixels"])), "oracle_spp": int(ref["spp"]),
           "relative_mse_per_frame": [round(float(x), 4) for x in rel]}
    log(f"rmse leg: plain {out}")
    return out


def bootstrap_ci(values, n=2000, seed=0x5EED):
    """Mean and 95% percentile bootstrap interval of the mean over frames."""
    import numpy as np
    v = np.asarray(values, np.float64)
    rng = np.random.default_rng(seed)
    means = v[rng.integers(0, len(v), (n, len(v)))].mean(1)
    return float(v.mean()), float(np.percentile(means, 2.5)), float(np.percentile(means, 97.5))


def paired_gain(base, other):
    """Mean of base - other over frames and its 95% percentile bootstrap interval."""
    import numpy as np
    m, lo, hi = bootstrap_ci(np.asarray(base, np.float64) - np.asarray(other, np.float64))
    return [round(m, 4), round(lo, 4), round(hi, 4)]


def constant_cache_params(rgb):
    """Network weights whose output is the constant rgb for every input: the constant input
    (59) carried through one hidden unit of each layer, the output row scaled by rgb."""
    import numpy as np
    p = np.zeros(4 * 4096 + 64 * 16, np.float32)
    p[59 * 64 + 0] = 1.0                       # layer 1: input 59 -> unit 0
    for layer in (1, 2, 3):
        p[layer * 4096 + 0] = 1.0              # unit 0 -> unit 0
    p[16384:16384 + 3] = np.asarray(rgb, np.float32)   # output layer row 0
    return p


def nrc_leg(pt, cam, cfg, ref, plain, frames, warmup):
    """Neural radiance cache form at equal time: paths end in the cache after 2 bounces,
    65,536 training paths and 4 Adam steps per frame (online).  After `warmup` frames of
    online training, the cached frame's spp is set so that its time matches the plain
    16-spp frame; per-frame relative MSE against the FP64 reference over `frames` frames
    (at least 16), with 95% bootstrap intervals.  Control: a cache whose output is one
    constant (the mean of the last frame's training targets), frozen, at the same spp."""
    import numpy as np

    import bi_pt
    mb = cfg["max_bounces"]
    frames = max(frames, 16)
    nrc = bi_pt.Nrc(pt, train_samples=65536, train_steps=4, cache_bounce=2, seed=1)
    t0 = time.perf_counter()
    for k in range(warmup):
        nrc.render(cam, spp=1, max_bounces=mb, seed=cfg["seed"] + 9000 + k)
    nrc.stats()
    warm_s = time.perf_counter() - t0
    rec, tgt = nrc.samples(65536)
    valid = rec[:, 13].view(np.uint32) == 1
    const_rgb = tgt[valid].astype(np.float64).mean(0) if valid.any() else np.zeros(3)
    t16 = timed(lambda: nrc.render(cam, spp=16, max_bounces=mb, seed=cfg["seed"] + 9500))
    spp_eq = max(1, int(cfg["spp"] * plain["ms_per_frame"] / t16))

    def run(cache, spp, base):
        ms, rel = [], []
        for k in range(frames):
            ms.append(timed(lambda: cache.render(cam, spp=spp, max_bounces=mb, seed=base + k)))
            rel.append(image_errors(pt.image(), ref)[1])
        m, lo, hi = bootstrap_ci(rel)
        return {"spp": spp, "frames": frames, "ms_per_frame": round(float(np.mean(ms)), 2),
                "relative_mse": round(m, 4), "relative_mse_ci95": [round(lo, 4), round(hi, 4)],
                "relative_mse_per_frame": [round(float(x), 4) for x in rel]}

    eq = run(nrc, spp_eq, cfg["seed"] + 10000)
    same_spp = run(nrc, cfg["spp"], cfg["seed"] + 11000)
    st = nrc.stats()
    nrc.close()
    ctl = bi_pt.Nrc(pt, train_samples=65536, train_steps=4, cache_bounce=2, seed=1, flags=bi_pt.NRC_FREEZE)
    ctl.set_params(constant_cache_params(const_rgb))
    control = run(ctl, spp_eq, cfg["seed"] + 10000)
    ctl.close()
    plain_rel = plain.get("relative_mse_per_frame")
    out = {"workload": f"cache after 2 bounces, 65,536 training paths and 4 Adam steps per frame, "
                       f"{warmup} warm-up frames of online training at 1 spp",
           "equal_time": eq, "equal_spp": same_spp,
           "constant_cache_equal_time": dict(control, constant_rgb=[round(float(x), 5) for x in const_rgb]),
           "plain": {k: plain[k] for k in ("spp", "ms_per_frame", "relative_rmse", "relative_mse")},
           "relmse_ratio_equal_time": round(eq["relative_mse"] / plain["relative_mse"], 4),
           "relmse_ratio_to_constant_cache": round(eq["relative_mse"] / control["relative_mse"], 4),
           "network_beats_constant_beyond_ci": bool(eq["relative_mse_ci95"][1] < control["relative_mse_ci95"][0]),
           # paired over frames (the network and the constant cache render the same paths):
           # the constant's relative MSE minus the network's, mean and 95% bootstrap interval
           "paired_relmse_gain_ci95": paired_gain(control["relative_mse_per_frame"], eq["relative_mse_per_frame"]),
           "frames_network_better": int(sum(c > a for c, a in zip(control["relative_mse_per_frame"],
                                                                    eq["relative_mse_per_frame"]))),
           "warmup_s": round(warm_s, 2), "train_steps": int(st["steps"]), "last_loss": round(st["last_loss"], 4),
           "valid_training_samples": int(st["last_valid_samples"])}
    if plain_rel:
        m, lo, hi = bootstrap_ci(plain_rel)
        out["plain"]["relative_mse_ci95"] = [round(lo, 4), round(hi, 4)]
    log(f"nrc leg: {out}")
    return out


def host_threads():
    """Worker threads for the CPU baseline: the CPUs this process may run on, at most 16
    (the host share that goes with one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(scene, cfg, runs):
    """The FP32 CPU restatement on this host, a bounded sample of the same workload: the
    full frame at 4 of the 16 spp, one warm-up run, then the median of `runs` runs; the
    restatement's scene set-up (flattening and LBVH build) is timed on its own."""
    import numpy as np

    import oracle
    threads = host_threads()
    w, h = cfg["width"], cfg["height"]
    cam = cfg["camera"](w, h)
    t = time.perf_counter()
    ors = oracle.OracleScene(scene)
    build_s = time.perf_counter() - t
    spp = 4
    secs, rays = [], 0
    for k in range(runs + 1):
        t = time.perf_counter()
        out = ors.render32(cam, cfg["max_bounces"], cfg["seed"] + k, spp=spp, threads=threads)
        dt = time.perf_counter() - t
        if k > 0:
            secs.append(dt)
            rays += int(out["counts"].sum())
        log(f"cpu leg: run {k} {'(warm-up) ' if k == 0 else ''}{dt:.2f}s on {threads} threads")
    ors.close()
    med = float(np.median(secs))
    rate = rays / sum(secs) / 1e6   # rays per run vary by a few per mille; the rate of the median run
    per_run = rays / len(secs)
    value = per_run / med / 1e6
    log(f"cpu leg: median {med:.2f}s, {value:.2f} Mrays/s (mean {rate:.2f})")
    return {"value": round(value, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"same scene and camera, {w}x{h}, {spp} of the {cfg['spp']} spp per run, depth "
                      f"{cfg['max_bounces'] - 1}; median of {runs} runs after 1 warm-up; render only (the "
                      f"restatement's flattening + LBVH build, {build_s:.2f}s, timed apart)",
            "cpu_model": cpu_model(), "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(),
            "median_run_s": round(med, 3), "runs_s": [round(x, 3) for x in secs], "build_s": round(build_s, 3)}


def profile_nrc():
    """Radiance-cache frames on the 1920x1080 atrium (262,144 triangles, depth 8):
    8 frames of bi_pt_render_nrc at 1 spp (online training), then 4 at 16 spp."""
    import torch

    import bi_pt
    import scenes
    cfg = scenes.CONFIGS["c1"]
    cam = cfg["camera"]()
    pt = bi_pt.PathTracer()
    pt.set_scene(cfg["scene"]())
    nrc = bi_pt.Nrc(pt, train_samples=65536, train_steps=4, cache_bounce=2, seed=1)
    for k in range(8):
        nrc.render(cam, spp=1, max_bounces=cfg["max_bounces"], seed=100 + k)
    for k in range(4):
        nrc.render(cam, spp=16, max_bounces=cfg["max_bounces"], seed=200 + k)
    torch.cuda.synchronize()
    log(f"cache frames done: {nrc.stats()}")
    nrc.close()
    pt.close()


if __name__ == "__main__":
    main()
