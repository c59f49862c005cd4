// server.hip -- the serving loop: render jobs of uneven scene size, resolution and spp
// arrive at any time; each is admitted when the HBM budget has room for its running
// image and output, renders as a view of a resident scene (BVH built once), and every
// scheduling round renders the next samples of all admitted jobs in one batched call
// (shared ray queues).  Jobs join at round boundaries and leave when their samples are
// done (continuous batching); latencies are measured on the host clock.
//
// No reference counterpart: the reference renders one camera per frame
// (src/renderer/basic.cpp:157-161); this drives the same pass for many requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "pt_host.hpp"

using namespace bipt;

namespace {

using Clock = std::chrono::steady_clock;

struct SrvJob {
    uint32_t scene = 0;
    bi_pt_camera cam{};
    bi_pt_settings st{};          // st.spp = the job's total samples
    uint32_t state = BI_PT_JOB_WAITING;
    uint32_t spp_done = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bi_pt_status failed = BI_PT_OK;   // BI_PT_JOB_FAILED: the status of the round that failed it
    std::string why;
    bi_pt_ctx* view = nullptr;
    DevBuf<float> out;            // RGBA32F running mean
    uint64_t bytes = 0;           // charged to the budget while admitted
};

}  // namespace

struct bi_pt_server : bipt::Owner {
    bi_pt_batch* batch = nullptr;
    std::vector<bi_pt_ctx*> scenes;
    uint64_t scene_bytes = 0;      // resident scenes and their BVHs
    uint64_t job_bytes = 0;        // admitted jobs' images and outputs
    uint32_t max_jobs = 64;        // admitted jobs per round
    uint32_t spp_per_round = 16;   // samples of each job per round
    std::deque<uint64_t> waiting;
    std::vector<uint64_t> running;
    std::unordered_map<uint64_t, SrvJob> jobs;
    uint64_t next_id = 1;
    Clock::time_point t0 = Clock::now();
    uint64_t rounds = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    uint64_t reserve = 0;          // wave buffers at full capacity for the resident scenes
    uint32_t wave_cap = kDefaultWavePaths;   // the wave engine's capacity (what `reserve` covers)
    uint64_t used() const { return scene_bytes + job_bytes + std::max<uint64_t>(reserve, batch->allocated); }
};

namespace {

uint64_t job_cost(const SrvJob& j) {
    const uint64_t npix = (uint64_t)j.cam.width * j.cam.height;
    return npix * 12 + npix * 16;   // running sums + RGBA32F output
}

void release(bi_pt_server* s, SrvJob& j) {
    if (j.view) (void)bi_pt_destroy(j.view);
    j.view = nullptr;
    dfree(s, j.out);
    s->job_bytes -= j.bytes;
    j.bytes = 0;
}

}  // namespace

extern "C" {

bi_pt_status bi_pt_server_create(const bi_pt_server_desc* d, bi_pt_server** out) {
    if (!d || !out) return BI_PT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    bi_pt_server* s = new (std::nothrow) bi_pt_server();
    if (!s) return BI_PT_ERROR_OUT_OF_MEMORY;
    s->device = d->device;
    s->budget = d->hbm_budget_bytes;
    if (d->max_jobs_per_round) s->max_jobs = d->max_jobs_per_round;
    if (d->spp_per_round) s->spp_per_round = d->spp_per_round;
    bi_pt_device_desc dd{};
    dd.device = d->device;
    if (bi_pt_status st = bi_pt_batch_create(&dd, &s->batch); st != BI_PT_OK) {
        delete s;
        return st;
    }
    if (d->wave_paths) s->wave_cap = std::min(d->wave_paths, kMaxWavePaths);
    bi_pt_tuning t{};
    t.wave_paths = s->wave_cap;
    (void)bi_pt_batch_set_tuning(s->batch, &t);
    *out = s;
    return BI_PT_OK;
}

bi_pt_status bi_pt_server_destroy(bi_pt_server* s) {
    if (!s) return BI_PT_ERROR_INVALID_ARGUMENT;
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    for (auto& kv : s->jobs) release(s, kv.second);
    (void)bi_pt_batch_destroy(s->batch);
    delete s;
    return BI_PT_OK;
}

const char* bi_pt_server_last_error(const bi_pt_server* s) { return s ? s->err.c_str() : "null server"; }

double bi_pt_server_clock(const bi_pt_server* s) { return s ? s->now() : 0.0; }

bi_pt_status bi_pt_server_add_scene(bi_pt_server* s, bi_pt_ctx* scene, uint32_t* scene_id) {
    if (!s || !scene || !scene_id) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (scene->scene_src || !scene->has_scene)
        return fail(s, BI_PT_ERROR_INVALID_ARGUMENT, "a scene must be a context with a scene of its own");
    if (scene->device != s->device) return fail(s, BI_PT_ERROR_INVALID_ARGUMENT, "scene on another device");
    BI_HIP_TRY(s, hipSetDevice(s->device));
    // resident: build the BVH now and keep it (jobs render with BI_PT_FLAG_REUSE_BVH)
    if (!scene->bvh_built) {
        BI_TRY(engine_init(scene, scene->eng));
        if (bi_pt_status e = ensure_build_buffers(scene); e != BI_PT_OK) return fail(s, e, scene->err);
        if (bi_pt_status e = build_bvh(scene, nullptr, &scene->eng.counters.p[kCntError], scene->treelet_rounds);
            e != BI_PT_OK)
            return fail(s, e, scene->err);
        BI_HIP_TRY(s, hipDeviceSynchronize());
        scene->bvh_built = true;
    }
    // the wave buffers are reserved at full capacity for the scenes' light counts, so
    // admission never counts on memory a later round's buffers will take
    auto shadow_mult = [](const bi_pt_ctx* c) {
        uint32_t m = 1;
        for (const LightRec& L : c->lights) m += L.sm_index >= 0 ? 1u : 0u;
        return m;
    };
    uint32_t slots = (uint32_t)scene->lights.size() + 2u, sh = shadow_mult(scene);
    for (const bi_pt_ctx* c : s->scenes) {
        slots = std::max<uint32_t>(slots, (uint32_t)c->lights.size() + 2u);
        sh = std::max(sh, shadow_mult(c));
    }
    const uint64_t reserve = std::max(engine_bytes(s->wave_cap, slots, sh), s->reserve);
    if (s->budget && s->scene_bytes + scene->allocated + reserve + s->job_bytes > s->budget)
        return fail(s, BI_PT_ERROR_OUT_OF_MEMORY, "scene does not fit the HBM budget");
    s->reserve = reserve;
    s->scene_bytes += scene->allocated;
    s->scenes.push_back(scene);
    *scene_id = (uint32_t)(s->scenes.size() - 1);
    return BI_PT_OK;
}

bi_pt_status bi_pt_server_submit(bi_pt_server* s, const bi_pt_request* r, uint64_t* job_id) {
    if (!s || !r || !job_id) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (r->scene_id >= s->scenes.size()) return fail(s, BI_PT_ERROR_INVALID_ARGUMENT, "unknown scene id");
    // everything a round would reject is rejected here, so one job cannot fail the others' rounds
    BI_TRY(validate_job(s, &r->camera, &r->settings));
    if (r->settings.flags & BI_PT_FLAG_OUTPUT_RGBA16F)
        return fail(s, BI_PT_ERROR_UNSUPPORTED, "the server returns RGBA32F images (BI_PT_FLAG_OUTPUT_RGBA16F)");
    SrvJob j;
    j.scene = r->scene_id;
    j.cam = r->camera;
    j.st = r->settings;
    j.st.accumulate = 0;
    j.st.flags &= ~(uint32_t)(BI_PT_FLAG_RECORD_QUEUES | BI_PT_FLAG_TIMING);
    j.st.flags |= BI_PT_FLAG_REUSE_BVH;
    j.t_submit = r->arrival_s >= 0.0 ? r->arrival_s : s->now();
    if (s->budget && s->scene_bytes + s->reserve + job_cost(j) > s->budget)
        return fail(s, BI_PT_ERROR_OUT_OF_MEMORY, "job can never fit the HBM budget");
    const uint64_t id = s->next_id++;
    s->jobs.emplace(id, j);
    s->waiting.push_back(id);
    *job_id = id;
    return BI_PT_OK;
}

bi_pt_status bi_pt_server_step(bi_pt_server* s, bi_pt_stream stream, uint32_t* active) {
    if (!s) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(s, hipSetDevice(s->device));
    // admission: first come, first served, while the budget and the round have room
    while (!s->waiting.empty() && s->running.size() < s->max_jobs) {
        SrvJob& j = s->jobs[s->waiting.front()];
        const uint64_t cost = job_cost(j);
        if (s->budget && s->used() + cost > s->budget) break;
        if (bi_pt_status e = bi_pt_create_view(s->scenes[j.scene], &j.view); e != BI_PT_OK)
            return fail(s, e, "view: " + std::string(bi_pt_last_error(s->scenes[j.scene])));
        if (bi_pt_status e = dalloc(s, j.out, (size_t)4 * j.cam.width * j.cam.height); e != BI_PT_OK) {
            (void)bi_pt_destroy(j.view);
            j.view = nullptr;
            if (s->running.empty()) return e;   // nothing will free memory: report
            break;
        }
        j.bytes = cost;
        s->job_bytes += cost;
        j.state = BI_PT_JOB_RUNNING;
        j.t_admit = s->now();
        s->running.push_back(s->waiting.front());
        s->waiting.pop_front();
    }
    if (!s->running.empty()) {
        // one round: the next samples of every admitted job, in one batched call
        std::vector<bi_pt_settings> st(s->running.size());
        std::vector<bi_pt_job> bj(s->running.size());
        std::vector<uint32_t> chunk(s->running.size());
        for (size_t k = 0; k < s->running.size(); ++k) {
            SrvJob& j = s->jobs[s->running[k]];
            chunk[k] = std::min(s->spp_per_round, j.st.spp - j.spp_done);
            st[k] = j.st;
            st[k].spp = chunk[k];
            st[k].accumulate = j.spp_done > 0 ? 1u : 0u;
            bj[k] = bi_pt_job{j.view, &j.cam, &st[k], j.out.p, nullptr};
        }
        // a round that fails (a HIP error, or a kernel's internal-limit flag) fails its jobs:
        // they leave with state BI_PT_JOB_FAILED and their memory is freed; waiting jobs and
        // later rounds go on
        const hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
        bi_pt_status e = bi_pt_render_batch(s->batch, bj.data(), (uint32_t)bj.size(), stream);
        std::string why = e != BI_PT_OK ? std::string("round: ") + bi_pt_batch_last_error(s->batch) : "";
        if (e == BI_PT_OK && hipStreamSynchronize(hs) != hipSuccess) {
            e = BI_PT_ERROR_HIP;
            why = "round: stream synchronisation failed";
        }
        if (e == BI_PT_OK && (e = engine_take_errors(s->batch, s->batch->eng, hs)) != BI_PT_OK)
            why = "round: " + s->batch->err;
        if (e != BI_PT_OK) {
            (void)hipStreamSynchronize(hs);
            const double t = s->now();
            for (uint64_t id : s->running) {
                SrvJob& j = s->jobs[id];
                release(s, j);
                j.state = BI_PT_JOB_FAILED;
                j.t_done = t;
                j.failed = e;
                j.why = why;
            }
            s->running.clear();
            ++s->rounds;
            if (active) *active = (uint32_t)s->waiting.size();
            return fail(s, e, why);
        }
        const double t = s->now();
        ++s->rounds;
        std::vector<uint64_t> still;
        for (size_t k = 0; k < s->running.size(); ++k) {
            SrvJob& j = s->jobs[s->running[k]];
            j.spp_done += chunk[k];
            if (j.spp_done >= j.st.spp) {
                j.state = BI_PT_JOB_DONE;
                j.t_done = t;
            } else {
                still.push_back(s->running[k]);
            }
        }
        s->running.swap(still);
    }
    if (active) *active = (uint32_t)(s->waiting.size() + s->running.size());
    return BI_PT_OK;
}

bi_pt_status bi_pt_server_poll(bi_pt_server* s, uint64_t id, bi_pt_job_status* out) {
    if (!s || !out) return BI_PT_ERROR_INVALID_ARGUMENT;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(s, BI_PT_ERROR_INVALID_ARGUMENT, "unknown job id");
    const SrvJob& j = it->second;
    out->state = j.state;
    out->spp_done = j.spp_done;
    out->submit_s = j.t_submit;
    out->admit_s = j.t_admit;
    out->done_s = j.t_done;
    return BI_PT_OK;
}

bi_pt_status bi_pt_server_take_result(bi_pt_server* s, uint64_t id, float* host_rgba, uint64_t capacity_floats) {
    if (!s) return BI_PT_ERROR_INVALID_ARGUMENT;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(s, BI_PT_ERROR_INVALID_ARGUMENT, "unknown job id");
    SrvJob& j = it->second;
    if (j.state == BI_PT_JOB_FAILED) {   // nothing to copy: report why and forget the job
        const bi_pt_status e = j.failed;
        const std::string why = j.why;
        s->jobs.erase(it);
        return fail(s, e, "job failed: " + why);
    }
    if (j.state != BI_PT_JOB_DONE) return fail(s, BI_PT_ERROR_INVALID_ARGUMENT, "job not finished");
    const uint64_t n = (uint64_t)4 * j.cam.width * j.cam.height;
    if (host_rgba) {
        if (capacity_floats < n) return fail(s, BI_PT_ERROR_INVALID_ARGUMENT, "result buffer too small");
        BI_HIP_TRY(s, hipSetDevice(s->device));
        BI_HIP_TRY(s, hipMemcpy(host_rgba, j.out.p, n * sizeof(float), hipMemcpyDeviceToHost));
    }
    release(s, j);
    s->jobs.erase(it);
    return BI_PT_OK;
}

bi_pt_status bi_pt_server_debug_inject_error(bi_pt_server* s, uint32_t flags) {
    if (!s) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(s, hipSetDevice(s->device));
    if (bi_pt_status e = engine_inject_error(s->batch, s->batch->eng, flags); e != BI_PT_OK)
        return fail(s, e, s->batch->err);
    return BI_PT_OK;
}

bi_pt_status bi_pt_server_get_stats(bi_pt_server* s, bi_pt_server_stats* out) {
    if (!s || !out) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(s, hipSetDevice(s->device));
    bi_pt_stats st{};
    BI_TRY(bi_pt_batch_get_stats(s->batch, &st));
    out->rays = st.rays_camera + st.rays_extension + st.rays_light_shadow + st.rays_sky_shadow;
    out->rounds = s->rounds;
    out->waiting = (uint32_t)s->waiting.size();
    out->running = (uint32_t)s->running.size();
    out->hbm_used = s->used();
    return BI_PT_OK;
}

}  // extern "C"
