// engine.hip -- the wave engine: plans the waves of a call (one or several render
// jobs), sizes the queues, and issues the per-segment kernels on one stream; the batch
// object (bi_pt_render_batch) with its hipGraph capture of a whole call and the
// autotuner of the engine's knobs.
//
// Frame structure (PathTracingPass::render, src/renderer/pass/path_tracing.cpp:224-488,
// re-designed as compacted waves instead of full-screen passes):
//   LBVH build of every job's scene (every call unless the caller keeps a built one:
//   the reference rebuilds the TLAS every frame, accel.cpp:156-158)
//   for each wave (whole samples of whole images, possibly of several jobs):
//     segment 0:   trace camera rays (generated in the kernel) -> shade -> shadow rays
//     segment b:   trace extension queue -> shade -> shadow rays
//     resolve:     per-pixel running sums += radiance of each sample, in sample order
//   output:        running mean -> RGBA32F (pt_accumulate.hlsl:9 as sum / count)
// A job's samples are visited in increasing order across waves and each path's
// arithmetic depends only on (seed, pixel, sample, bounce) and its own scene, so a job
// renders bit for bit the same alone, in any wave split, or batched with other jobs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "pt_host.hpp"
#include "pt_nrc.hpp"

namespace bipt {

bi_pt_status engine_init(Owner* own, Engine& eng) {
    if (eng.ready) return BI_PT_OK;
    BI_TRY(dalloc(own, eng.counters, kCntSlots));
    BI_HIP_TRY(own, hipMemset(eng.counters.p, 0, kCntSlots * sizeof(uint32_t)));
    BI_TRY(dalloc(own, eng.empty_tree, 1));
    Bvh4Node e{};
    for (int j = 0; j < 4; ++j) {
        e.lox[j] = e.loy[j] = e.loz[j] = e.hix[j] = e.hiy[j] = e.hiz[j] = INFINITY;
        e.child[j] = 0;
        e.pad[j] = 0;
    }
    BI_HIP_TRY(own, hipMemcpy(eng.empty_tree.p, &e, sizeof e, hipMemcpyHostToDevice));
    for (auto& ev : eng.ev) BI_HIP_TRY(own, hipEventCreate(&ev));
    BI_HIP_TRY(own, hipEventCreate(&eng.plan_ev));
    BI_HIP_TRY(own, hipStreamCreateWithFlags(&eng.aux_s, hipStreamNonBlocking));
    BI_HIP_TRY(own, hipEventCreateWithFlags(&eng.ev_shade, hipEventDisableTiming));
    for (auto& ev : eng.ev_shadow) BI_HIP_TRY(own, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    eng.ready = true;
    return BI_PT_OK;
}

void engine_free(Owner* own, Engine& eng) {
    for (int k = 0; k < 6; ++k) { dfree(own, eng.qa[k]); dfree(own, eng.qb[k]); }
    for (int k = 0; k < 4; ++k) dfree(own, eng.ps_f[k]);
    for (int t = 0; t < 2; ++t) {
        for (int k = 0; k < 9; ++k) dfree(own, eng.sq_f[t][k]);
        dfree(own, eng.sq_pk[t]);
    }
    dfree(own, eng.acc);
    dfree(own, eng.qa_path); dfree(own, eng.qb_path);
    dfree(own, eng.hit_rec); dfree(own, eng.tile_state);
    dfree(own, eng.counters); dfree(own, eng.plan); dfree(own, eng.empty_tree);
    if (eng.ready) {
        for (auto& ev : eng.ev) (void)hipEventDestroy(ev);
        (void)hipEventDestroy(eng.plan_ev);
    }
    if (eng.aux_s) {
        (void)hipStreamDestroy(eng.aux_s);
        eng.aux_s = nullptr;
    }
    if (eng.ev_shade) (void)hipEventDestroy(eng.ev_shade);
    for (auto& ev : eng.ev_shadow)
        if (ev) (void)hipEventDestroy(ev);
    eng.ev_shade = nullptr;
    eng.ev_shadow[0] = eng.ev_shadow[1] = nullptr;
    if (eng.build_fork) {
        (void)hipEventDestroy(eng.build_fork);
        for (int k = 0; k < Engine::kBuildStreams; ++k) {
            (void)hipEventDestroy(eng.build_join[k]);
            (void)hipStreamDestroy(eng.build_s[k]);
        }
        eng.build_fork = nullptr;
    }
    if (eng.plan_host) (void)hipHostFree(eng.plan_host);
    eng.plan_host = nullptr;
    eng.plan_host_n = 0;
    eng.ready = false;
}

std::string device_error_text(uint32_t flags) {
    std::string m = "device check failed:";
    if (flags & 1u) m += " traversal stack overflow;";
    if (flags & 4u) m += " refit wait limit;";
    if (flags & 8u) m += " traversal step bound exceeded;";
    if (flags & 16u) m += " treelet start list overflow;";
    if (flags & 32u) m += " shade look-back wait limit;";
    if (flags & ~(1u | 4u | 8u | 16u | 32u)) m += " flags " + std::to_string(flags) + ";";
    return m;
}

// The error word is sticky: render calls never clear it, so a flag raised by any call is
// reported by the next check, which clears it.
bi_pt_status engine_check(Owner* own, Engine& eng, bi_pt_stats* out) {
    BI_HIP_TRY(own, hipDeviceSynchronize());
    uint32_t hc[kCntSlots];
    BI_HIP_TRY(own, hipMemcpy(hc, eng.counters.p, sizeof hc, hipMemcpyDeviceToHost));
    uint64_t r[4];
    std::memcpy(r, &hc[kCntRays], sizeof r);
    eng.stats.rays_extension = r[1];
    eng.stats.rays_light_shadow = r[2];
    eng.stats.rays_sky_shadow = r[3];
    if (out) *out = eng.stats;
    if (hc[kCntError]) {
        BI_HIP_TRY(own, hipMemset(eng.counters.p + kCntError, 0, sizeof(uint32_t)));
        return fail(own, BI_PT_ERROR_DEVICE_CHECK, device_error_text(hc[kCntError]));
    }
    return BI_PT_OK;
}

bi_pt_status engine_take_errors(Owner* own, Engine& eng, hipStream_t s) {
    if (!eng.ready) return BI_PT_OK;
    uint32_t flags = 0;
    BI_HIP_TRY(own, hipMemcpyAsync(&flags, eng.counters.p + kCntError, sizeof flags, hipMemcpyDeviceToHost, s));
    BI_HIP_TRY(own, hipStreamSynchronize(s));
    if (flags == 0) return BI_PT_OK;
    launch_fill_u32(eng.counters.p + kCntError, 0u, 1u, s);
    BI_HIP_TRY(own, hipStreamSynchronize(s));
    return fail(own, BI_PT_ERROR_DEVICE_CHECK, device_error_text(flags));
}

bi_pt_status engine_inject_error(Owner* own, Engine& eng, uint32_t flags) {
    BI_TRY(engine_init(own, eng));
    BI_HIP_TRY(own, hipDeviceSynchronize());
    uint32_t w = 0;
    BI_HIP_TRY(own, hipMemcpy(&w, eng.counters.p + kCntError, sizeof w, hipMemcpyDeviceToHost));
    w |= flags;
    BI_HIP_TRY(own, hipMemcpy(eng.counters.p + kCntError, &w, sizeof w, hipMemcpyHostToDevice));
    return BI_PT_OK;
}

bi_pt_status validate_job(Owner* own, const bi_pt_camera* c, const bi_pt_settings* st) {
    if (!c || !st) return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "null camera or settings");
    if (c->width == 0 || c->height == 0 || st->spp == 0)
        return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "empty image or zero samples");
    if ((uint64_t)c->width * c->height > (1ull << 26))
        return fail(own, BI_PT_ERROR_UNSUPPORTED, "image larger than 2^26 pixels");
    if (!(st->ray_length > 0.0f) || !(st->ray_length <= kShadowTMax))
        return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "ray_length must be positive and at most 1.7e38");
    return BI_PT_OK;
}

uint64_t engine_bytes(uint32_t cap, uint32_t slots, uint32_t sh_mult) {
    const uint64_t c = cap, tiles = (cap + 255u) / 256u + 1u;
    uint64_t b = 0;
    b += 12 * c * 4 + 2 * c * 4;                 // queues A and B: 6 floats + path id each
    b += c * sizeof(HitRec);                     // hit records
    b += 4 * c * 4;                              // path state
    b += (uint64_t)slots * 3 * c * 4;            // radiance accumulators
    b += 2 * (uint64_t)sh_mult * c * (9 * 4 + 4);   // two shadow queue sets
    b += tiles * 8;                              // look-back words of the shade tiles
    return b;
}

namespace {

bi_pt_status ensure_wave_buffers(Owner* own, Engine& eng, uint32_t cap, uint32_t slots, uint32_t sh_mult) {
    if (cap <= eng.wave_cap && slots <= eng.wave_slots && sh_mult <= eng.wave_sh_mult) return BI_PT_OK;
    cap = std::max(cap, eng.wave_cap);
    slots = std::max(slots, eng.wave_slots);
    sh_mult = std::max(sh_mult, eng.wave_sh_mult);
    for (int k = 0; k < 6; ++k) {
        BI_TRY(dalloc(own, eng.qa[k], cap));
        BI_TRY(dalloc(own, eng.qb[k], cap));
    }
    BI_TRY(dalloc(own, eng.qa_path, cap));
    BI_TRY(dalloc(own, eng.qb_path, cap));
    BI_TRY(dalloc(own, eng.hit_rec, cap));
    for (int k = 0; k < 4; ++k) BI_TRY(dalloc(own, eng.ps_f[k], cap));
    BI_TRY(dalloc(own, eng.acc, (size_t)slots * 3u * cap));
    const size_t shcap = (size_t)cap * sh_mult;
    for (int t = 0; t < 2; ++t) {
        for (int k = 0; k < 9; ++k) BI_TRY(dalloc(own, eng.sq_f[t][k], shcap));
        BI_TRY(dalloc(own, eng.sq_pk[t], shcap));
    }
    BI_TRY(dalloc(own, eng.tile_state, (cap + 255u) / 256u + 1u));
    eng.wave_cap = cap;
    eng.wave_slots = slots;
    eng.wave_sh_mult = sh_mult;
    return BI_PT_OK;
}

RayQueue queue_of(DevBuf<float>* f, DevBuf<uint32_t>& path) {
    return RayQueue{f[0].p, f[1].p, f[2].p, f[3].p, f[4].p, f[5].p, path.p};
}

// Per-job facts of a call.
struct Prep {
    bi_pt_ctx* ctx;       // running image
    bi_pt_ctx* scene;     // scene, lights, sky, BVH
    CameraParams cam;
    uint32_t npix, M, D, nl, spp, sh, role;
    uint32_t nrect;       // rect lights (their accumulator follows the sky's)
    uint32_t rounds;      // treelet rounds if this job's call builds its scene
    bool sky_nee, reset, rebuild;
    float* sums;          // running sums (training paths: the cache's targets)
};

// One job's share of one wave.
struct Item {
    uint32_t job, sample_off, samples;
};

// Byte layout of the plan buffer: per wave, JobRec[k] then path_base[k+1] and
// pix_base[k+1]; then the output records JobRec[n] and pix_base[n+1].
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

bi_pt_status render_jobs(Owner* own, Engine& eng, const JobIn* jobs, uint32_t n, hipStream_t s,
                         const RenderOpts& opt, bi_pt_batch* graph_owner) {
    if (n == 0) return BI_PT_OK;
    if (!jobs) return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "null job array");
    BI_TRY(engine_init(own, eng));
    const bool timing = opt.timing && n == 1;
    const bool record = opt.record && n == 1;
    std::vector<Prep> prep(n);
    uint32_t max_npix = 0, slots = 2, sh_mult = 1;
    bi_pt_nrc* const nrc = opt.nrc;
    for (uint32_t j = 0; j < n; ++j) {
        const JobIn& J = jobs[j];
        bi_pt_ctx* ctx = J.ctx;
        const bool train = J.role == kRoleTrain;
        if (J.role != kRoleImage && !nrc)
            return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "a cached or training job needs a radiance cache");
        if (!ctx || !J.camera || !J.st || (!J.out_rgba && !train))
            return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "null context/camera/settings/output in job " + std::to_string(j));
        for (uint32_t k = 0; k < j && !train; ++k)
            if (jobs[k].ctx == ctx && jobs[k].role != kRoleTrain)
                return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "a context appears twice in one call");
        bi_pt_ctx* S = scene_of(ctx);
        if (!S->has_scene) return fail(own, BI_PT_ERROR_NO_SCENE, "no scene set (job " + std::to_string(j) + ")");
        if (S->tex_refs > S->num_tex)
            return fail(own, BI_PT_ERROR_INVALID_ARGUMENT,
                        "a material names texture index " + std::to_string(S->tex_refs - 1) + " but " +
                            std::to_string(S->num_tex) + " textures are set (job " + std::to_string(j) + ")");
        const bi_pt_camera* c = J.camera;
        const bi_pt_settings* st = J.st;
        if (bi_pt_status e = validate_job(own, c, st); e != BI_PT_OK)
            return fail(own, e, own->err + " (job " + std::to_string(j) + ")");
        const uint64_t npix64 = (uint64_t)c->width * c->height;
        if (!train && ctx->acc_count + st->spp > 0xffffffffull)
            return fail(own, BI_PT_ERROR_UNSUPPORTED, "more than 2^32 accumulated samples");
        Prep& p = prep[j];
        p.ctx = ctx;
        p.role = J.role;
        std::memcpy(p.cam.inv_view, c->matrix_inv_view, 64);
        std::memcpy(p.cam.inv_proj, c->matrix_inv_proj, 64);
        std::memcpy(p.cam.proj_view, c->matrix_proj_view, 64);
        p.cam.width = c->width;
        p.cam.height = c->height;
        p.npix = (uint32_t)npix64;
        p.M = std::min(std::max(st->max_bounces, 2u), 16u);
        p.D = p.M - 1;
        p.scene = S;
        p.nl = (uint32_t)S->lights.size();
        p.nrect = (uint32_t)S->rects_host.size();
        if (p.nrect && S->ltc_host.empty())
            return fail(own, BI_PT_ERROR_INVALID_ARGUMENT, "rect lights need the LTC tables (bi_pt_set_ltc_luts)");
        for (const RectRec& rr : S->rects_host)
            if (rr.texture_index >= (int32_t)S->num_rect_tex)
                return fail(own, BI_PT_ERROR_INVALID_ARGUMENT,
                            "rect light texture_index " + std::to_string(rr.texture_index) + " without that texture "
                            "(bi_pt_set_rect_light_textures)");
        p.spp = st->spp;
        p.sky_nee = !(st->flags & BI_PT_FLAG_NO_SKY_NEE) && S->sky.total > 0.0f;
        p.sh = 1;
        for (const LightRec& L : S->lights) p.sh += L.sm_index >= 0 ? 1u : 0u;
        // accumulation history: continue only with the same view, size and sampling settings
        const bool same = ctx->acc_count > 0 && st->accumulate && ctx->acc_w == p.cam.width &&
                          ctx->acc_h == p.cam.height && std::memcmp(&ctx->last_cam, &p.cam, sizeof p.cam) == 0 &&
                          ctx->last_seed == st->seed && ctx->last_flags == (st->flags & 3u) &&
                          ctx->last_bounces == p.M && ctx->last_ray_length == st->ray_length &&
                          ctx->last_version == S->scene_version && ctx->last_role == J.role;
        p.reset = !same || train;
        p.sums = nullptr;
        if (train) {
            p.npix = nrc->train_samples;
            p.spp = 1;
            p.sums = nrc->targets.p;
        }
        if (J.gbuf && !S->gbuf_tan) {   // the GBuffer reads leaf tangents: build them from now on
            S->gbuf_tan = true;
            S->bvh_built = false;
        }
        p.rebuild = !(opt.reuse_bvh && S->bvh_built);
        for (uint32_t k = 0; k < j; ++k)
            if (prep[k].scene == S) p.rebuild = false;   // one build per scene and call
        max_npix = std::max(max_npix, p.npix);
        slots = std::max(slots, p.nl + 2u + (p.nrect ? 1u : 0u));
        sh_mult = std::max(sh_mult, p.sh);
    }

    // wave plan: whole samples of whole images, jobs in call order, each job's samples in order.
    // Default capacity 2^25 paths: a 1920x1080 frame of 16 spp is one wave, so each kernel
    // launches once per segment and drains its tail once (measured: 23% more rays/s than
    // 2^22-path waves); buffers are sized to the largest wave actually planned (about
    // 270 B per path).  At most 2^28 paths, so packed shadow-ray ids fit in 32 bits.
    const Knobs& kn = opt.knobs;
    uint32_t cap = kn.wave_paths ? std::min(kn.wave_paths, kMaxWavePaths) : kDefaultWavePaths;
    cap = std::max(cap, max_npix);
    std::vector<std::vector<Item>> waves;
    {
        std::vector<uint32_t> left(n), done(n, 0);
        for (uint32_t j = 0; j < n; ++j) left[j] = prep[j].spp;
        for (;;) {
            std::vector<Item> w;
            uint64_t room = cap;
            for (uint32_t j = 0; j < n; ++j) {
                if (left[j] == 0) continue;
                const uint32_t ws = (uint32_t)std::min<uint64_t>(left[j], room / prep[j].npix);
                if (ws == 0) continue;
                w.push_back(Item{j, done[j], ws});
                room -= (uint64_t)ws * prep[j].npix;
                done[j] += ws;
                left[j] -= ws;
            }
            if (w.empty()) break;
            waves.push_back(std::move(w));
        }
    }
    uint32_t wave_max = 0;
    for (const auto& w : waves) {
        uint64_t paths = 0;
        for (const Item& it : w) paths += (uint64_t)it.samples * prep[it.job].npix;
        wave_max = std::max(wave_max, (uint32_t)paths);
    }
    // 12 LDS entries: the traversal kernels reach 6 (closest hit) and 7 (shadow) wavefronts per
    // SIMD, against 5 and 6 with 16 (C1: 71.6 -> 70.3 ms per frame)
    const uint32_t lds_stack = kn.stack_lds ? std::min(kn.stack_lds, 32u) : 12u;
    const uint32_t refill = kn.refill_lanes ? std::min(kn.refill_lanes, 64u) : 44u;
    const uint32_t parts = kn.work_parts ? std::min(kn.work_parts, kWorkPartsMax) : kWorkParts;
    const uint32_t block = kn.block_threads == 128 || kn.block_threads == 512 ? kn.block_threads : 256u;

    // non-opaque triangles anywhere in the call select the traversal kernels with any-hit tests
    bool any_hit = false, any_point = false;   // point lights: shadow rays of finite extent
    for (uint32_t j = 0; j < n; ++j) {
        any_hit = any_hit || prep[j].scene->non_opaque;
        for (const LightRec& L : prep[j].scene->lights) any_point = any_point || L.type == kLightPoint;
    }

    // device memory: per-job images, BVH builds, wave buffers, plan
    bool any_cached = false, any_train = false;
    for (uint32_t j = 0; j < n; ++j) {
        bi_pt_ctx* ctx = prep[j].ctx;
        bi_pt_status e = BI_PT_OK;
        any_cached |= prep[j].role == kRoleCached;
        any_train |= prep[j].role == kRoleTrain;
        if (prep[j].role != kRoleTrain) {
            if ((e = dalloc(ctx, ctx->sums, (size_t)3 * prep[j].npix)) != BI_PT_OK) return fail(own, e, ctx->err);
            prep[j].sums = ctx->sums.p;
        }
    }
    // treelet rounds of each scene's build: the configured number when the call traces at
    // least max(4 n, 2^20) paths of the scene's n triangles, else none (a round costs
    // ~0.1-0.5 ms and saves ~5% of the trace time, ~0.1 ns per path); a kept BVH with
    // fewer rounds than the call wants is rebuilt
    for (uint32_t j = 0; j < n; ++j) {
        Prep& p = prep[j];
        bool first = true;
        for (uint32_t k = 0; k < j; ++k) first = first && prep[k].scene != p.scene;
        p.rounds = 0;
        if (!first) continue;
        uint64_t paths = 0;
        for (uint32_t k = j; k < n; ++k)
            if (prep[k].scene == p.scene) paths += (uint64_t)prep[k].npix * prep[k].spp;
        const uint64_t need = std::max<uint64_t>(4ull * p.scene->num_tris, 1ull << 20);
        p.rounds = paths >= need ? p.scene->treelet_rounds : 0u;
        if (!p.rebuild && p.scene->built_rounds < p.rounds) p.rebuild = true;
        if (p.rebuild)
            if (bi_pt_status e = ensure_build_buffers(p.scene); e != BI_PT_OK) return fail(own, e, p.scene->err);
    }
    BI_TRY(ensure_wave_buffers(own, eng, wave_max, slots, sh_mult));
    if (any_cached) BI_TRY(nrc_ensure_query(nrc, eng.wave_cap));
    const size_t rec_align = alignof(JobRec);
    auto wave_bytes = [&](size_t k) {
        return align_up(k * sizeof(JobRec) + 2 * (k + 1) * sizeof(uint32_t), rec_align);
    };
    std::vector<size_t> wave_off(waves.size() + 1, 0);
    for (size_t w = 0; w < waves.size(); ++w) wave_off[w + 1] = wave_off[w] + wave_bytes(waves[w].size());
    const size_t plan_bytes = wave_off.back() + wave_bytes(n);
    BI_TRY(dalloc(own, eng.plan, plan_bytes));

    // host image of the plan, in pinned memory the previous call's copy has finished with
    if (eng.plan_pending) BI_HIP_TRY(own, hipEventSynchronize(eng.plan_ev));
    eng.plan_pending = false;
    if (eng.plan_host_n < plan_bytes) {
        if (eng.plan_host) BI_HIP_TRY(own, hipHostFree(eng.plan_host));
        eng.plan_host = nullptr;
        eng.plan_host_n = 0;
        BI_HIP_TRY(own, hipHostMalloc(reinterpret_cast<void**>(&eng.plan_host), plan_bytes, hipHostMallocDefault));
        eng.plan_host_n = plan_bytes;
    }
    unsigned char* const host = eng.plan_host;
    std::memset(host, 0, plan_bytes);
    auto fill_rec = [&](JobRec& r, uint32_t j, uint32_t sample_off, uint32_t samples, uint32_t pix_base) {
        const Prep& p = prep[j];
        bi_pt_ctx* ctx = p.ctx;
        bi_pt_ctx* S = p.scene;
        const bool empty = S->num_tris == 0;
        r.nodes = empty ? eng.empty_tree.p : S->nodes4.p;
        r.leaf_geo = empty ? nullptr : S->leaf_geo.p;
        r.leaf_shade = empty ? nullptr : S->leaf_shade.p;
        r.materials = S->mats.p;
        r.leaf_uv = empty ? nullptr : scene_leaf_uv(S);
        r.leaf_tan = empty ? nullptr : scene_leaf_tan(S);
        r.tex = scene_tex(S);
        r.num_tris = S->num_tris;
        r.num_lights = p.nl;
        r.sky_nee = p.sky_nee ? 1u : 0u;
        r.last_bounce = p.D - 1;
        for (uint32_t l = 0; l < p.nl; ++l) {
            r.lights[l] = S->lights[l];
            r.keep[l] = 1.0f - S->lights[l].shadow_strength;
        }
        r.sky = S->sky;
        r.rects = p.nrect ? S->rects.p : nullptr;
        r.num_rects = p.nrect;
        r.ltc = p.nrect ? S->ltc.p : nullptr;
        r.rect_tex = p.nrect && S->num_rect_tex ? S->rect_tex.p : nullptr;
        r.rect_tex_levels = SceneTex{S->rect_tex_levels.p, S->rect_tex_pool.p, S->srgb.p};
        r.cam = p.cam;
        r.npix = p.npix;
        r.sample0 = (uint32_t)(p.reset ? 0 : ctx->acc_count) + sample_off;
        r.samples = samples;
        r.seed = jobs[j].st->seed;
        r.flags = (jobs[j].st->flags & 3u) | (p.role == kRoleCached ? kJobCache : 0u) |
                  (p.role == kRoleTrain ? kJobTrain : 0u) |
                  ((jobs[j].st->flags & BI_PT_FLAG_OUTPUT_RGBA16F) ? kJobHalfOut : 0u);
        r.cache_bounce = p.role != kRoleImage ? nrc->cache_bounce : 0xffffffffu;
        r.cache_rec = p.role == kRoleTrain ? nrc->train_rec.p : nullptr;
        r.ray_length = jobs[j].st->ray_length;
        r.pix_base = pix_base;
        r.sums = p.sums;
        r.depth_out = jobs[j].out_depth;
        r.depth_sample = (uint32_t)(p.reset ? 0 : ctx->acc_count);
        r.out_rgba = jobs[j].out_rgba;
        r.out_den = (float)((p.reset ? 0 : ctx->acc_count) + p.spp);
        if (const bi_pt_gbuffer* g = jobs[j].gbuf) {
            r.gb_base = g->base_color;
            r.gb_nr = g->normal_roughness;
            r.gb_fres = g->fresnel;
            r.gb_m0 = g->material_0;
        }
    };
    std::vector<WaveJobs> wj(waves.size());
    for (size_t w = 0; w < waves.size(); ++w) {
        const size_t k = waves[w].size();
        unsigned char* base = host + wave_off[w];
        JobRec* recs = reinterpret_cast<JobRec*>(base);
        uint32_t* pb = reinterpret_cast<uint32_t*>(base + k * sizeof(JobRec));
        uint32_t* xb = pb + (k + 1);
        uint32_t paths = 0, pix = 0;
        for (size_t i = 0; i < k; ++i) {
            const Item& it = waves[w][i];
            new (&recs[i]) JobRec{};
            fill_rec(recs[i], it.job, it.sample_off, it.samples, pix);
            pb[i] = paths;
            xb[i] = pix;
            paths += it.samples * prep[it.job].npix;
            pix += prep[it.job].npix;
        }
        pb[k] = paths;
        xb[k] = pix;
        unsigned char* dbase = eng.plan.p + wave_off[w];
        wj[w] = WaveJobs{reinterpret_cast<const JobRec*>(dbase),
                         reinterpret_cast<const uint32_t*>(dbase + k * sizeof(JobRec)),
                         reinterpret_cast<const uint32_t*>(dbase + k * sizeof(JobRec)) + (k + 1), (uint32_t)k, paths};
    }
    WaveJobs out_jobs{};
    uint32_t out_pix = 0;
    {
        unsigned char* base = host + wave_off.back();
        JobRec* recs = reinterpret_cast<JobRec*>(base);
        uint32_t* pb = reinterpret_cast<uint32_t*>(base + n * sizeof(JobRec));
        uint32_t* xb = pb + (n + 1);
        for (uint32_t j = 0; j < n; ++j) {
            new (&recs[j]) JobRec{};
            fill_rec(recs[j], j, 0, 0, out_pix);
            pb[j] = 0;
            xb[j] = out_pix;
            out_pix += prep[j].npix;
        }
        pb[n] = 0;
        xb[n] = out_pix;
        unsigned char* dbase = eng.plan.p + wave_off.back();
        out_jobs = WaveJobs{reinterpret_cast<const JobRec*>(dbase),
                            reinterpret_cast<const uint32_t*>(dbase + n * sizeof(JobRec)),
                            reinterpret_cast<const uint32_t*>(dbase + n * sizeof(JobRec)) + (n + 1), n, 0};
    }
    BI_HIP_TRY(own, hipMemcpyAsync(eng.plan.p, host, plan_bytes, hipMemcpyHostToDevice, s));
    BI_HIP_TRY(own, hipEventRecord(eng.plan_ev, s));
    eng.plan_pending = true;

    // streams for concurrent LBVH builds: made once, before any capture, by the first call
    // that rebuilds two or more scenes
    if (!eng.build_fork) {
        uint32_t nbuild = 0;
        for (uint32_t j = 0; j < n; ++j) nbuild += prep[j].rebuild ? 1u : 0u;
        if (nbuild > 1) {
            BI_HIP_TRY(own, hipEventCreateWithFlags(&eng.build_fork, hipEventDisableTiming));
            for (int k = 0; k < Engine::kBuildStreams; ++k) {
                BI_HIP_TRY(own, hipStreamCreateWithFlags(&eng.build_s[k], hipStreamNonBlocking));
                BI_HIP_TRY(own, hipEventCreateWithFlags(&eng.build_join[k], hipEventDisableTiming));
            }
        }
    }

    // every call, outside any graph: the counters and the LBVH builds.  Queue lengths and
    // the persistent lanes' part counters start at zero (the ray statistics accumulate and
    // the error word is sticky until a check reads it).  The scenes' LBVHs are independent: build b goes to s (b % lanes == 0)
    // or to build stream b % lanes - 1, forked from s and joined back into s before
    // anything reads them.
    {
        uint32_t* cnt = eng.counters.p;
        launch_fill_u32(cnt, 0u, kCntError, s);
        launch_fill_u32(cnt + kCntError + 1, 0u, kCntRays - kCntError - 1, s);
        launch_fill_u32(cnt + kCntTraceWork, 0u, kCntSlots - kCntTraceWork, s);
        if (timing) BI_HIP_TRY(own, hipEventRecord(eng.ev[0], s));
        std::vector<uint32_t> builds;
        for (uint32_t j = 0; j < n; ++j)
            if (prep[j].rebuild) builds.push_back(j);
        const uint32_t lanes = std::min<uint32_t>((uint32_t)builds.size(), 1u + Engine::kBuildStreams);
        bi_pt_status berr = BI_PT_OK;
        std::string bmsg;
        if (lanes > 1) {
            berr = hipEventRecord(eng.build_fork, s) == hipSuccess ? BI_PT_OK : BI_PT_ERROR_HIP;
            for (uint32_t k = 1; k < lanes && berr == BI_PT_OK; ++k)
                if (hipStreamWaitEvent(eng.build_s[k - 1], eng.build_fork, 0) != hipSuccess) berr = BI_PT_ERROR_HIP;
            if (berr != BI_PT_OK) bmsg = "build stream fork failed";
        }
        for (size_t b = 0; b < builds.size() && berr == BI_PT_OK; ++b) {
            const uint32_t k = (uint32_t)(b % (lanes ? lanes : 1u));
            bi_pt_ctx* sc = prep[builds[b]].scene;
            berr = build_bvh(sc, k == 0 ? s : eng.build_s[k - 1], &cnt[kCntError], prep[builds[b]].rounds);
            if (berr != BI_PT_OK) bmsg = sc->err;
        }
        for (uint32_t k = 1; k < lanes; ++k) {   // join even after a failure
            if (hipEventRecord(eng.build_join[k - 1], eng.build_s[k - 1]) != hipSuccess ||
                hipStreamWaitEvent(s, eng.build_join[k - 1], 0) != hipSuccess) {
                if (berr == BI_PT_OK) { berr = BI_PT_ERROR_HIP; bmsg = "build stream join failed"; }
            }
        }
        if (berr != BI_PT_OK) return fail(own, berr, bmsg);
        if (timing) BI_HIP_TRY(own, hipEventRecord(eng.ev[1], s));
    }

    // graph replay: the same launch sequence with the same buffers (the plan's contents,
    // uploaded above, carry everything that changes between calls)
    const bool graph = graph_owner && graph_owner->use_graph && !timing && !record;
    // a capture that an error leaves open is ended and its graph dropped, so the caller's
    // stream is usable again
    struct CaptureGuard {
        hipStream_t s = nullptr;
        ~CaptureGuard() {
            if (!s) return;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
        }
    } capture;
    std::vector<uint64_t> key;
    if (graph) {
        key = {n, cap, lds_stack, refill, parts, block, (uint64_t)(uintptr_t)eng.plan.p, eng.wave_cap, eng.wave_slots,
               eng.wave_sh_mult, plan_bytes, (uint64_t)any_hit, (uint64_t)any_point};
        for (uint32_t j = 0; j < n; ++j) {
            const Prep& p = prep[j];
            key.insert(key.end(), {(uint64_t)(uintptr_t)p.ctx, (uint64_t)(uintptr_t)p.scene,
                                   (uint64_t)(uintptr_t)p.scene->nodes4.p, (uint64_t)(uintptr_t)p.ctx->sums.p,
                                   p.npix, p.spp, p.M, p.nl, p.sh, (uint64_t)p.reset, (uint64_t)p.rebuild,
                                   p.scene->num_tris, (uint64_t)(jobs[j].gbuf != nullptr)});
        }
        if (graph_owner->graph_exec && key == graph_owner->graph_key) {
            BI_HIP_TRY(own, hipGraphLaunch(graph_owner->graph_exec, s));
            goto finish;
        }
        if (graph_owner->graph_exec) { (void)hipGraphExecDestroy(graph_owner->graph_exec); graph_owner->graph_exec = nullptr; }
        if (graph_owner->graph) { (void)hipGraphDestroy(graph_owner->graph); graph_owner->graph = nullptr; }
        graph_owner->graph_key.clear();
        BI_HIP_TRY(own, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        capture.s = s;
    }
    {
        uint32_t* cnt = eng.counters.p;
        uint64_t* rays = reinterpret_cast<uint64_t*>(cnt + kCntRays);
        // the images of jobs that start over
        for (uint32_t j = 0; j < n; ++j) {
            if (prep[j].reset) launch_fill_u32(reinterpret_cast<uint32_t*>(prep[j].sums), 0u, 3 * prep[j].npix, s);
            if (prep[j].role == kRoleTrain)   // training samples start unrecorded
                launch_fill_u32(reinterpret_cast<uint32_t*>(nrc->train_rec.p), 0u, 16 * prep[j].npix, s);
        }
        if ((any_cached || any_train) && prep[0].scene->cb_out.p)   // the cache's position box
            BI_HIP_TRY(own, hipMemcpyAsync(nrc->bounds.p, prep[0].scene->cb_out.p, 6 * sizeof(float),
                                           hipMemcpyDeviceToDevice, s));

        PathState ps{eng.ps_f[0].p, eng.ps_f[1].p, eng.ps_f[2].p, eng.ps_f[3].p, eng.acc.p, eng.wave_cap};
        HitBuf hits{eng.hit_rec.p};
        RayQueue qA = queue_of(eng.qa, eng.qa_path), qB = queue_of(eng.qb, eng.qb_path);
        ShadowQueue sqs[2];
        for (int t = 0; t < 2; ++t)
            sqs[t] = ShadowQueue{eng.sq_f[t][0].p, eng.sq_f[t][1].p, eng.sq_f[t][2].p, eng.sq_f[t][3].p, eng.sq_f[t][4].p,
                                 eng.sq_f[t][5].p, eng.sq_f[t][6].p, eng.sq_f[t][7].p, eng.sq_f[t][8].p, eng.sq_pk[t].p};
        // shadow launches overlap the next segment's closest hit, except when timing kernel
        // classes, recording queues, or training the cache (whose reset of a training path's
        // accumulators at the cache segment must follow the earlier shadow additions)
        const bool overlap = !timing && !record && !any_train;

        if (record) { eng.rec_ext.clear(); eng.rec_sh.clear(); }

        // the launches of one kind are timed together: events around each class of kernel
        std::vector<hipEvent_t> tev;
        std::vector<int> tcls;
        auto mark = [&](int cls) -> bi_pt_status {
            if (!timing) return BI_PT_OK;
            hipEvent_t e;
            BI_HIP_TRY(own, hipEventCreate(&e));
            BI_HIP_TRY(own, hipEventRecord(e, s));
            tev.push_back(e);
            tcls.push_back(cls);
            return BI_PT_OK;
        };
        uint32_t launches_closest = 0, launches_shadow = 0;

        for (size_t w = 0; w < waves.size(); ++w) {
            const WaveJobs& J = wj[w];
            uint32_t D = 0, max_nl = 0;
            bool wave_cached = false, wave_rect = false;
            for (const Item& it : waves[w]) {
                D = std::max(D, prep[it.job].D);
                max_nl = std::max(max_nl, prep[it.job].nl);
                wave_rect |= prep[it.job].nrect > 0;
                wave_cached |= prep[it.job].role == kRoleCached;
            }
            if (record) {
                eng.rec_ext.emplace_back(D);
                eng.rec_sh.emplace_back(D);
                std::vector<uint32_t> all(J.num_paths);
                for (uint32_t k = 0; k < J.num_paths; ++k) all[k] = k;
                eng.rec_ext.back()[0] = all;
            }
            bool set_used[2] = {false, false};
            for (uint32_t b = 0; b < D; ++b) {
                const bool primary = b == 0;
                const uint32_t set = overlap ? (b & 1u) : 0u;   // shadow queue set of this segment
                const ShadowQueue& sq = sqs[set];
                uint32_t* cnt_sh = &cnt[kCntShadow + set];
                uint32_t* shadow_work = &cnt[set ? kCntShadowWork1 : kCntShadowWork];
                const RayQueue& qin = (b & 1u) ? qA : qB;      // segment b reads what segment b-1 wrote
                const RayQueue& qout = (b & 1u) ? qB : qA;
                uint32_t* cnt_in = (b & 1u) ? &cnt[kCntExtA] : &cnt[kCntExtB];
                uint32_t* cnt_out = (b & 1u) ? &cnt[kCntExtB] : &cnt[kCntExtA];

                TraceArgs ta{};
                ta.jobs = J;
                ta.q = qin;
                ta.count_in = primary ? nullptr : cnt_in;
                ta.hits = hits;
                ta.work = &cnt[kCntTraceWork];
                ta.parts = parts;
                ta.refill = refill;
                ta.any_hit = any_hit ? 1u : 0u;
                ta.error = &cnt[kCntError];
                BI_TRY(mark(0));
                launch_trace_closest(ta, J.num_paths, primary, lds_stack, block, s);
                ++launches_closest;
                if (primary) {   // GBuffers of the jobs whose first sample of the call is in this wave
                    const size_t k = waves[w].size();
                    const uint32_t* pbh = reinterpret_cast<const uint32_t*>(host + wave_off[w] + k * sizeof(JobRec));
                    for (size_t i = 0; i < k; ++i) {
                        const Item& it = waves[w][i];
                        if (jobs[it.job].gbuf && it.sample_off == 0)
                            launch_gbuffer(J.rec + i, hits.rec, pbh[i], prep[it.job].npix, s);
                    }
                }
                BI_TRY(mark(1));

                // the set's previous shadow launch (two segments back) has ended before its
                // queue and counters are written again
                if (overlap && set_used[set]) BI_HIP_TRY(own, hipStreamWaitEvent(s, eng.ev_shadow[set], 0));
                launch_shade_reset(eng.tile_state.p, primary ? nullptr : cnt_in, J.num_paths, cnt_out, cnt_sh,
                                   J.num_paths, s);
                ShadeArgs sa{};
                sa.jobs = J;
                sa.qin = qin;
                sa.count_in = primary ? nullptr : cnt_in;
                sa.hits = hits;
                sa.ps = ps;
                sa.qout = qout;
                sa.sq = sq;
                sa.count_ext = cnt_out;
                sa.count_shadow = cnt_sh;
                sa.ray_counters = rays;
                sa.tile_state = eng.tile_state.p;
                sa.error = &cnt[kCntError];
                sa.trace_work_reset = &cnt[kCntTraceWork];
                sa.shadow_work_reset = shadow_work;
                sa.bounce = b;
                const bool cache_now = wave_cached && b == nrc->cache_bounce;
                sa.nrc_query = cache_now ? nrc->query.p : nullptr;
                launch_shade(sa, J.num_paths, max_nl, wave_rect, s);
                if (cache_now) {
                    // paths that ended in the cache at this segment: their radiance estimate
                    NrcInferArgs na{};
                    na.rec = nrc->query.p;
                    na.count_in = primary ? nullptr : cnt_in;
                    na.num = J.num_paths;
                    na.fwd = nrc->fwd.p;
                    na.bounds = nrc->bounds.p;
                    na.ps = ps;
                    launch_nrc_infer(na, s);
                }

                BI_TRY(mark(2));

                ShadowArgs sha{};
                sha.jobs = J;
                sha.q = sq;
                sha.count_in = cnt_sh;
                sha.ps = ps;
                sha.ray_counters = rays;
                sha.work = shadow_work;
                sha.parts = parts;
                sha.refill = refill;
                sha.any_hit = (any_hit || any_point) ? 1u : 0u;
                sha.error = &cnt[kCntError];
                if (overlap) {   // fork: the shadow rays trace beside the next closest-hit launch
                    BI_HIP_TRY(own, hipEventRecord(eng.ev_shade, s));
                    BI_HIP_TRY(own, hipStreamWaitEvent(eng.aux_s, eng.ev_shade, 0));
                    launch_trace_shadow(sha, J.num_paths * sh_mult, lds_stack, block, eng.aux_s);
                    BI_HIP_TRY(own, hipEventRecord(eng.ev_shadow[set], eng.aux_s));
                    set_used[set] = true;
                } else {
                    launch_trace_shadow(sha, J.num_paths * sh_mult, lds_stack, block, s);
                }
                ++launches_shadow;
                BI_TRY(mark(3));
                BI_HIP_TRY(own, hipGetLastError());
                if (record) {
                    uint32_t hc[3];
                    BI_HIP_TRY(own, hipMemcpyAsync(hc, cnt, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                    BI_HIP_TRY(own, hipStreamSynchronize(s));
                    const uint32_t next = hc[(b & 1u) ? kCntExtB : kCntExtA];
                    auto& shq = eng.rec_sh.back()[b];
                    shq.resize(hc[kCntShadow]);
                    if (!shq.empty())
                        BI_HIP_TRY(own, hipMemcpy(shq.data(), eng.sq_pk[0].p, shq.size() * 4, hipMemcpyDeviceToHost));
                    if (b + 1 < D) {
                        auto& ex = eng.rec_ext.back()[b + 1];
                        ex.resize(next);
                        if (next) BI_HIP_TRY(own, hipMemcpy(ex.data(), qout.path, next * 4, hipMemcpyDeviceToHost));
                    }
                }
            }
            for (int t = 0; t < 2; ++t)   // join: the wave's shadow launches have ended
                if (set_used[t]) BI_HIP_TRY(own, hipStreamWaitEvent(s, eng.ev_shadow[t], 0));
            uint32_t wave_pix = 0;
            for (const Item& it : waves[w]) wave_pix += prep[it.job].npix;
            ResolveArgs ra{};
            ra.jobs = J;
            ra.ps = ps;
            ra.num_pix = wave_pix;
            launch_resolve(ra, s);
            BI_TRY(mark(4));
        }
        launch_output(out_jobs, out_pix, s);
        BI_HIP_TRY(own, hipGetLastError());
        if (any_train) {   // the cache learns from this frame's training paths
            if (bi_pt_status e = nrc_train_frame(nrc, s); e != BI_PT_OK) return fail(own, e, nrc->err);
        }
        eng.stats.launches_trace_closest = launches_closest;
        eng.stats.launches_trace_shadow = launches_shadow;
        if (timing) {
            BI_HIP_TRY(own, hipEventRecord(eng.ev[2], s));
            BI_HIP_TRY(own, hipEventSynchronize(eng.ev[2]));
            float ms_build = 0, total = 0;
            BI_HIP_TRY(own, hipEventElapsedTime(&ms_build, eng.ev[0], eng.ev[1]));
            BI_HIP_TRY(own, hipEventElapsedTime(&total, eng.ev[0], eng.ev[2]));
            float cls_ms[5] = {0, 0, 0, 0, 0};
            // class of the interval (tev[k-1], tev[k]) is tcls[k]: 1 trace, 2 shade, 3 shadow, 4 resolve
            for (size_t k = 1; k < tev.size(); ++k) {
                float ms = 0;
                BI_HIP_TRY(own, hipEventElapsedTime(&ms, tev[k - 1], tev[k]));
                int c = tcls[k];
                if (c >= 1 && c <= 4) cls_ms[c] += ms;
            }
            for (auto e : tev) (void)hipEventDestroy(e);
            eng.stats.ms_build = ms_build;
            eng.stats.ms_trace_closest = cls_ms[1];
            eng.stats.ms_shade = cls_ms[2];
            eng.stats.ms_trace_shadow = cls_ms[3];
            eng.stats.ms_resolve = cls_ms[4];
            eng.stats.ms_total = total;
        }
    }
    if (graph) {
        hipGraph_t g = nullptr;
        capture.s = nullptr;
        BI_HIP_TRY(own, hipStreamEndCapture(s, &g));
        graph_owner->graph = g;
        BI_HIP_TRY(own, hipGraphInstantiate(&graph_owner->graph_exec, g, nullptr, nullptr, 0));
        graph_owner->graph_key = key;
        BI_HIP_TRY(own, hipGraphLaunch(graph_owner->graph_exec, s));
    }
finish:
    // host bookkeeping of the running images
    uint32_t images = 0;
    bi_pt_ctx* image_ctx = nullptr;
    for (uint32_t j = 0; j < n; ++j) {
        const Prep& p = prep[j];
        bi_pt_ctx* ctx = p.ctx;
        const bi_pt_settings* st = jobs[j].st;
        eng.camera_rays += (uint64_t)p.npix * p.spp;
        if (p.rebuild) p.scene->bvh_built = true;
        if (p.role == kRoleTrain) continue;
        ++images;
        image_ctx = ctx;
        ctx->last_role = p.role;
        if (p.reset) ctx->acc_count = 0;
        ctx->acc_w = p.cam.width;
        ctx->acc_h = p.cam.height;
        ctx->last_cam = p.cam;
        ctx->last_seed = st->seed;
        ctx->last_flags = st->flags & 3u;
        ctx->last_bounces = p.M;
        ctx->last_ray_length = st->ray_length;
        ctx->acc_count += p.spp;
        ctx->last_version = p.scene->scene_version;
    }
    eng.frames++;
    eng.stats.frames = eng.frames;
    eng.stats.rays_camera = eng.camera_rays;
    eng.stats.samples_accumulated = images == 1 ? image_ctx->acc_count : 0;
    return BI_PT_OK;
}

}  // namespace bipt
