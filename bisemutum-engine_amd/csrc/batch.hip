// batch.hip -- batched rendering of several contexts' jobs in shared waves
// (bi_pt_render_batch), the hipGraph replay of repeated calls and the autotuner of the
// wave engine's knobs.  The reference renders one view per PathTracingPass::render
// (src/renderer/pass/path_tracing.cpp:224-488); this object serves many jobs at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "pt_host.hpp"

using namespace bipt;

namespace {

Knobs knobs_of(const bi_pt_tuning& t) {
    return Knobs{t.wave_paths, t.stack_lds, t.refill_lanes, t.work_parts, t.block_threads};
}

bi_pt_tuning tuning_of(const bi_pt_batch* b) {
    bi_pt_tuning t{};
    t.wave_paths = b->knobs.wave_paths;
    t.stack_lds = b->knobs.stack_lds;
    t.refill_lanes = b->knobs.refill_lanes;
    t.work_parts = b->knobs.work_parts;
    t.block_threads = b->knobs.block_threads;
    t.use_graph = b->use_graph ? 1u : 0u;
    return t;
}

void drop_graph(bi_pt_batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph_key.clear();
}

std::vector<JobIn> jobs_in(const bi_pt_job* jobs, uint32_t n) {
    std::vector<JobIn> v(n);
    for (uint32_t j = 0; j < n; ++j)
        v[j] = JobIn{jobs[j].ctx, jobs[j].camera, jobs[j].settings, jobs[j].out_rgba32f, jobs[j].out_depth};
    return v;
}

// A BVH may be kept only if every job asks for it.
bool all_reuse(const bi_pt_job* jobs, uint32_t n) {
    for (uint32_t j = 0; j < n; ++j)
        if (!jobs[j].settings || !(jobs[j].settings->flags & BI_PT_FLAG_REUSE_BVH)) return false;
    return n > 0;
}

}  // namespace

extern "C" {

bi_pt_status bi_pt_batch_create(const bi_pt_device_desc* desc, bi_pt_batch** out) {
    if (!out) return BI_PT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    bi_pt_batch* b = new (std::nothrow) bi_pt_batch();
    if (!b) return BI_PT_ERROR_OUT_OF_MEMORY;
    b->device = desc ? desc->device : 0;
    b->budget = desc ? desc->hbm_budget_bytes : 0;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0 || b->device < 0 || b->device >= ndev) {
        delete b;
        return e != hipSuccess ? BI_PT_ERROR_HIP : BI_PT_ERROR_INVALID_ARGUMENT;
    }
    if (hipSetDevice(b->device) != hipSuccess) { delete b; return BI_PT_ERROR_HIP; }
    if (bi_pt_status st = engine_init(b, b->eng); st != BI_PT_OK) {
        engine_free(b, b->eng);
        delete b;
        return st;
    }
    *out = b;
    return BI_PT_OK;
}

bi_pt_status bi_pt_batch_destroy(bi_pt_batch* b) {
    if (!b) return BI_PT_ERROR_INVALID_ARGUMENT;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    drop_graph(b);
    engine_free(b, b->eng);
    delete b;
    return BI_PT_OK;
}

const char* bi_pt_batch_last_error(const bi_pt_batch* b) { return b ? b->err.c_str() : "null batch"; }

bi_pt_status bi_pt_batch_set_tuning(bi_pt_batch* b, const bi_pt_tuning* t) {
    if (!b || !t) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (t->stack_lds > 32 || t->refill_lanes > 64 || t->work_parts > kWorkPartsMax ||
        (t->block_threads != 0 && t->block_threads != 128 && t->block_threads != 256 && t->block_threads != 512))
        return fail(b, BI_PT_ERROR_INVALID_ARGUMENT, "tuning out of range");
    b->knobs = knobs_of(*t);
    b->use_graph = t->use_graph != 0;
    drop_graph(b);
    return BI_PT_OK;
}

bi_pt_status bi_pt_batch_get_tuning(const bi_pt_batch* b, bi_pt_tuning* t) {
    if (!b || !t) return BI_PT_ERROR_INVALID_ARGUMENT;
    *t = tuning_of(b);
    return BI_PT_OK;
}

bi_pt_status bi_pt_render_batch(bi_pt_batch* b, const bi_pt_job* jobs, uint32_t n, bi_pt_stream stream) {
    if (!b) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (n && !jobs) return fail(b, BI_PT_ERROR_INVALID_ARGUMENT, "null job array");
    BI_HIP_TRY(b, hipSetDevice(b->device));
    for (uint32_t j = 0; j < n; ++j)
        if (jobs[j].ctx && jobs[j].ctx->device != b->device)
            return fail(b, BI_PT_ERROR_INVALID_ARGUMENT, "job context on another device");
    if (b->use_graph && !stream)
        return fail(b, BI_PT_ERROR_INVALID_ARGUMENT, "graph replay needs a created stream (not the default stream)");
    std::vector<JobIn> v = jobs_in(jobs, n);
    RenderOpts opt;
    opt.knobs = b->knobs;
    opt.reuse_bvh = all_reuse(jobs, n);
    return render_jobs(b, b->eng, v.data(), n, reinterpret_cast<hipStream_t>(stream), opt, b);
}

bi_pt_status bi_pt_batch_get_stats(bi_pt_batch* b, bi_pt_stats* out) {
    if (!b || !out) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(b, hipSetDevice(b->device));
    return engine_check(b, b->eng, out);
}

// Coordinate descent over the knobs that matter most (wave capacity 8M / 16M / 32M paths,
// traversal block shape 128 / 256 / 512 threads, LDS stack depth, refill threshold), each
// candidate timed with HIP events over `rounds` calls; the
// jobs render with accumulation off so each call does the same work.
bi_pt_status bi_pt_batch_autotune(bi_pt_batch* b, const bi_pt_job* jobs, uint32_t n, uint32_t rounds,
                                  bi_pt_stream stream, bi_pt_tuning* best_out) {
    if (!b || (n && !jobs) || n == 0) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(b, hipSetDevice(b->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    rounds = std::max(rounds, 1u);
    std::vector<bi_pt_settings> st(n);
    for (uint32_t j = 0; j < n; ++j) {
        if (!jobs[j].settings) return fail(b, BI_PT_ERROR_INVALID_ARGUMENT, "null settings");
        st[j] = *jobs[j].settings;
        st[j].accumulate = 0;
    }
    std::vector<JobIn> v = jobs_in(jobs, n);
    for (uint32_t j = 0; j < n; ++j) v[j].st = &st[j];
    const bool graph = b->use_graph;
    b->use_graph = false;
    drop_graph(b);
    hipEvent_t e0, e1;
    BI_HIP_TRY(b, hipEventCreate(&e0));
    BI_HIP_TRY(b, hipEventCreate(&e1));
    auto time_knobs = [&](const Knobs& k, float* ms) -> bi_pt_status {
        RenderOpts opt;
        opt.knobs = k;
        opt.reuse_bvh = all_reuse(jobs, n);
        BI_TRY(render_jobs(b, b->eng, v.data(), n, s, opt, nullptr));   // warm: sizes the buffers
        BI_HIP_TRY(b, hipEventRecord(e0, s));
        for (uint32_t r = 0; r < rounds; ++r) BI_TRY(render_jobs(b, b->eng, v.data(), n, s, opt, nullptr));
        BI_HIP_TRY(b, hipEventRecord(e1, s));
        BI_HIP_TRY(b, hipEventSynchronize(e1));
        BI_HIP_TRY(b, hipEventElapsedTime(ms, e0, e1));
        return BI_PT_OK;
    };
    Knobs best = b->knobs;
    float best_ms = 0.0f;
    bi_pt_status status = time_knobs(best, &best_ms);
    const uint32_t waves[] = {1u << 23, 1u << 24, 1u << 25};
    const uint32_t stacks[] = {8u, 12u, 16u};
    const uint32_t refills[] = {32u, 44u, 52u};
    const uint32_t blocks[] = {128u, 256u, 512u};
    for (int axis = 0; axis < 4 && status == BI_PT_OK; ++axis) {
        for (int c = 0; c < 3 && status == BI_PT_OK; ++c) {
            Knobs k = best;
            if (axis == 0) k.wave_paths = waves[c];
            if (axis == 1) k.stack_lds = stacks[c];
            if (axis == 2) k.refill_lanes = refills[c];
            if (axis == 3) k.block_threads = blocks[c];
            float ms = 0.0f;
            status = time_knobs(k, &ms);
            if (status == BI_PT_OK && ms < best_ms) {
                best_ms = ms;
                best = k;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->use_graph = graph;
    if (status != BI_PT_OK) return status;
    b->knobs = best;
    if (best_out) *best_out = tuning_of(b);
    return BI_PT_OK;
}

}  // extern "C"
