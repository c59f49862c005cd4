// bi_pt.hip -- host side of the C ABI (include/bi_pt.h): context, uploads and the
// per-frame orchestration of the wavefront pass on one HIP stream.
//
// Frame structure (PathTracingPass::render, src/renderer/pass/path_tracing.cpp:224-488,
// re-designed as compacted waves instead of full-screen passes):
//   LBVH build (every call: the reference rebuilds the TLAS every frame, accel.cpp:156-158)
//   for each wave of whole samples:
//     segment 0:   trace camera rays (generated in the kernel) -> shade -> shadow rays
//     segment b:   trace extension queue -> shade -> shadow rays          (b < max_bounces-1)
//     resolve:     per-pixel running sums += radiance of each sample, in sample order
//   output:        running mean -> RGBA32F (pt_accumulate.hlsl:9 as sum / count)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/bi_pt.h"
#include "pt_kernels.hpp"

using namespace bipt;

namespace {

struct DevArena;

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

}  // namespace

struct bi_pt_ctx {
    int device = 0;
    uint64_t budget = 0;
    uint64_t allocated = 0;
    std::string err;
    std::vector<void*> to_free;

    // scene
    bool has_scene = false;
    DevBuf<float> pos, nrm;
    DevBuf<uint32_t> idx;
    DevBuf<InstanceRec> inst;
    DevBuf<MaterialRec> mats;
    std::vector<InstanceRec> inst_host;
    uint32_t num_tris = 0, num_inst = 0, num_mats = 0;
    // lights
    std::vector<LightRec> lights;
    // sky
    DevBuf<float4> sky_tex;
    DevBuf<float> sky_w, sky_ccdf, sky_rows, sky_marg;
    SkyParams sky{};
    std::vector<float> sky_host;
    bool sky_set = false;
    // build buffers
    uint32_t build_cap = 0;
    DevBuf<float> tri_v, tri_n, cb_out;
    DevBuf<uint32_t> tri_mat, cb_ord, keys0, keys1, vals0, vals1, hist, refit_cnt;
    DevBuf<BvhNode> nodes;
    DevBuf<Bvh4Node> nodes4;
    DevBuf<int32_t> parent;
    DevBuf<LeafGeo> leaf_geo;
    DevBuf<LeafShade> leaf_shade;
    bool bvh_built = false;
    // wave buffers
    uint32_t wave_cap = 0, wave_slots = 0, wave_sh_mult = 0;
    DevBuf<float> qa[6], qb[6], ps_f[4], acc, sq_f[9];
    DevBuf<uint32_t> qa_path, qb_path, sq_pk;
    DevBuf<HitRec> hit_rec;
    DevBuf<float> st_f[9], st_c;      // shade staging: hit point, extension dir, sky dir; contributions
    DevBuf<uint32_t> st_path, st_mask;
    DevBuf<uint64_t> tile_sums;
    DevBuf<uint32_t> tile_sky;
    DevBuf<uint32_t> counters;   // kCntSlots uint32 (ray counters at kCntRays as 4 x u64)
    // accumulation
    DevBuf<float> sums;
    uint32_t acc_w = 0, acc_h = 0;
    uint64_t acc_count = 0;
    CameraParams last_cam{};
    uint32_t last_seed = 0, last_flags = 0, last_bounces = 0;
    float last_ray_length = 0.0f;
    // statistics
    bi_pt_stats stats{};
    uint64_t frames = 0;
    uint64_t camera_rays = 0;
    hipEvent_t ev[2 * 8] = {};
    bool events_ready = false;
    // recorded queues [wave][bounce]
    std::vector<std::vector<std::vector<uint32_t>>> rec_ext, rec_sh;
};

namespace {

bi_pt_status fail(bi_pt_ctx* c, bi_pt_status st, const std::string& msg) {
    if (c) c->err = msg;
    return st;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, BI_PT_ERROR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
bi_pt_status dalloc(bi_pt_ctx* ctx, DevBuf<T>& b, size_t n) {
    if (b.p && b.n >= n) return BI_PT_OK;
    if (b.p) {
        (void)hipFree(b.p);
        ctx->allocated -= b.n * sizeof(T);
        b.p = nullptr;
        b.n = 0;
    }
    if (n == 0) return BI_PT_OK;
    const size_t bytes = n * sizeof(T);
    if (ctx->budget && ctx->allocated + bytes > ctx->budget)
        return fail(ctx, BI_PT_ERROR_OUT_OF_MEMORY,
                    "HBM budget exceeded: need " + std::to_string(bytes) + " B more, budget " +
                        std::to_string(ctx->budget) + " B, in use " + std::to_string(ctx->allocated) + " B");
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(ctx, BI_PT_ERROR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    b.p = static_cast<T*>(p);
    b.n = n;
    ctx->allocated += bytes;
    return BI_PT_OK;
}

template <class T>
void dfree(bi_pt_ctx* ctx, DevBuf<T>& b) {
    if (b.p) {
        (void)hipFree(b.p);
        ctx->allocated -= b.n * sizeof(T);
    }
    b.p = nullptr;
    b.n = 0;
}

#define TRY(expr)                               \
    do {                                        \
        bi_pt_status s_ = (expr);               \
        if (s_ != BI_PT_OK) return s_;          \
    } while (0)

// Importance table of the sky (luminance * sin(theta) per texel), built in FP32
// in a fixed order on the host.
void build_sky_table(const float* rgba, uint32_t W, uint32_t H, const float color[3], std::vector<float>& w,
                     std::vector<float>& ccdf, std::vector<float>& rows, std::vector<float>& marg, float& total) {
    w.assign((size_t)W * H, 0.0f);
    ccdf.assign((size_t)(W + 1) * H, 0.0f);
    rows.assign(H, 0.0f);
    marg.assign(H + 1, 0.0f);
    const float pi_over_h = kPi / (float)H;
    for (uint32_t i = 0; i < H; ++i) {
        float st = pt_sin(((float)i + 0.5f) * pi_over_h);
        float run = 0.0f;
        for (uint32_t j = 0; j < W; ++j) {
            const float* p = &rgba[4 * ((size_t)i * W + j)];
            float r = p[0] * color[0], g = p[1] * color[1], b = p[2] * color[2];
            float lum = (0.2126f * r + 0.7152f * g) + 0.0722f * b;
            float x = lum * st;
            if (!(x > 0.0f)) x = 0.0f;
            w[(size_t)i * W + j] = x;
            ccdf[(size_t)i * (W + 1) + j] = run;
            run = run + x;
        }
        ccdf[(size_t)i * (W + 1) + W] = run;
        rows[i] = run;
        marg[i + 1] = marg[i] + run;
    }
    total = marg[H];
}

BvhView bvh_view(const bi_pt_ctx* ctx) {
    return BvhView{ctx->nodes4.p, ctx->leaf_geo.p, ctx->num_tris};
}

bi_pt_status ensure_build_buffers(bi_pt_ctx* ctx) {
    const uint32_t n = ctx->num_tris;
    if (n == 0) return BI_PT_OK;
    const size_t ni = n > 1 ? n - 1 : 1;
    TRY(dalloc(ctx, ctx->tri_v, (size_t)9 * n));
    TRY(dalloc(ctx, ctx->tri_n, (size_t)9 * n));
    TRY(dalloc(ctx, ctx->tri_mat, n));
    TRY(dalloc(ctx, ctx->cb_ord, 8));
    TRY(dalloc(ctx, ctx->cb_out, 8));
    TRY(dalloc(ctx, ctx->keys0, n));
    TRY(dalloc(ctx, ctx->keys1, n));
    TRY(dalloc(ctx, ctx->vals0, n));
    TRY(dalloc(ctx, ctx->vals1, n));
    TRY(dalloc(ctx, ctx->hist, sort_hist_entries(n)));
    TRY(dalloc(ctx, ctx->nodes, ni));
    TRY(dalloc(ctx, ctx->nodes4, ni));
    TRY(dalloc(ctx, ctx->parent, ni + n));
    TRY(dalloc(ctx, ctx->refit_cnt, ni));
    TRY(dalloc(ctx, ctx->leaf_geo, n));
    TRY(dalloc(ctx, ctx->leaf_shade, n));
    return BI_PT_OK;
}

bi_pt_status build_bvh(bi_pt_ctx* ctx, hipStream_t s) {
    TRY(ensure_build_buffers(ctx));
    if (ctx->num_tris == 0) { ctx->bvh_built = true; return BI_PT_OK; }
    BuildArgs a{};
    a.positions = ctx->pos.p;
    a.normals = ctx->nrm.p;
    a.indices = ctx->idx.p;
    a.instances = ctx->inst.p;
    a.num_instances = ctx->num_inst;
    a.num_tris = ctx->num_tris;
    a.tri_v = ctx->tri_v.p;
    a.tri_n = ctx->tri_n.p;
    a.tri_mat = ctx->tri_mat.p;
    a.cbounds_ord = ctx->cb_ord.p;
    a.cbounds_out = ctx->cb_out.p;
    a.keys[0] = ctx->keys0.p;
    a.keys[1] = ctx->keys1.p;
    a.vals[0] = ctx->vals0.p;
    a.vals[1] = ctx->vals1.p;
    a.hist = ctx->hist.p;
    a.nodes = ctx->nodes.p;
    a.nodes4 = ctx->nodes4.p;
    a.parent = ctx->parent.p;
    a.refit_cnt = ctx->refit_cnt.p;
    a.leaf_geo = ctx->leaf_geo.p;
    a.leaf_shade = ctx->leaf_shade.p;
    a.error = &ctx->counters.p[kCntError];
    launch_build_lbvh(a, s);
    HIP_TRY(hipGetLastError());
    ctx->bvh_built = true;
    return BI_PT_OK;
}

bi_pt_status ensure_wave_buffers(bi_pt_ctx* ctx, uint32_t cap, uint32_t slots, uint32_t sh_mult) {
    if (cap <= ctx->wave_cap && slots <= ctx->wave_slots && sh_mult <= ctx->wave_sh_mult) return BI_PT_OK;
    cap = std::max(cap, ctx->wave_cap);
    slots = std::max(slots, ctx->wave_slots);
    sh_mult = std::max(sh_mult, ctx->wave_sh_mult);
    for (int k = 0; k < 6; ++k) {
        TRY(dalloc(ctx, ctx->qa[k], cap));
        TRY(dalloc(ctx, ctx->qb[k], cap));
    }
    TRY(dalloc(ctx, ctx->qa_path, cap));
    TRY(dalloc(ctx, ctx->qb_path, cap));
    TRY(dalloc(ctx, ctx->hit_rec, cap));
    for (int k = 0; k < 4; ++k) TRY(dalloc(ctx, ctx->ps_f[k], cap));
    TRY(dalloc(ctx, ctx->acc, (size_t)slots * 3u * cap));
    const size_t shcap = (size_t)cap * sh_mult;
    for (int k = 0; k < 9; ++k) TRY(dalloc(ctx, ctx->sq_f[k], shcap));
    TRY(dalloc(ctx, ctx->sq_pk, shcap));
    for (int k = 0; k < 9; ++k) TRY(dalloc(ctx, ctx->st_f[k], cap));
    TRY(dalloc(ctx, ctx->st_c, (size_t)3 * (slots - 1u) * cap));
    TRY(dalloc(ctx, ctx->st_path, cap));
    TRY(dalloc(ctx, ctx->st_mask, cap));
    TRY(dalloc(ctx, ctx->tile_sums, (cap + 255u) / 256u + 1u));
    TRY(dalloc(ctx, ctx->tile_sky, (cap + 255u) / 256u + 1u));
    ctx->wave_cap = cap;
    ctx->wave_slots = slots;
    ctx->wave_sh_mult = sh_mult;
    return BI_PT_OK;
}

RayQueue queue_of(DevBuf<float>* f, DevBuf<uint32_t>& path) {
    return RayQueue{f[0].p, f[1].p, f[2].p, f[3].p, f[4].p, f[5].p, path.p};
}

}  // namespace

extern "C" {

bi_pt_status bi_pt_create(const bi_pt_device_desc* desc, bi_pt_ctx** out_ctx) {
    if (!out_ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    *out_ctx = nullptr;
    bi_pt_ctx* ctx = new (std::nothrow) bi_pt_ctx();
    if (!ctx) return BI_PT_ERROR_OUT_OF_MEMORY;
    ctx->device = desc ? desc->device : 0;
    ctx->budget = desc ? desc->hbm_budget_bytes : 0;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0 || ctx->device < 0 || ctx->device >= ndev) {
        delete ctx;
        return e != hipSuccess ? BI_PT_ERROR_HIP : BI_PT_ERROR_INVALID_ARGUMENT;
    }
    e = hipSetDevice(ctx->device);
    if (e != hipSuccess) { delete ctx; return BI_PT_ERROR_HIP; }
    if (dalloc(ctx, ctx->counters, kCntSlots) != BI_PT_OK) { delete ctx; return BI_PT_ERROR_OUT_OF_MEMORY; }
    if (hipMemset(ctx->counters.p, 0, kCntSlots * sizeof(uint32_t)) != hipSuccess) { delete ctx; return BI_PT_ERROR_HIP; }
    for (auto& ev : ctx->ev) {
        if (hipEventCreate(&ev) != hipSuccess) { delete ctx; return BI_PT_ERROR_HIP; }
    }
    ctx->events_ready = true;
    *out_ctx = ctx;
    return BI_PT_OK;
}

bi_pt_status bi_pt_destroy(bi_pt_ctx* ctx) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    dfree(ctx, ctx->pos); dfree(ctx, ctx->nrm); dfree(ctx, ctx->idx); dfree(ctx, ctx->inst); dfree(ctx, ctx->mats);
    dfree(ctx, ctx->sky_tex); dfree(ctx, ctx->sky_w); dfree(ctx, ctx->sky_ccdf); dfree(ctx, ctx->sky_rows);
    dfree(ctx, ctx->sky_marg);
    dfree(ctx, ctx->tri_v); dfree(ctx, ctx->tri_n); dfree(ctx, ctx->cb_out); dfree(ctx, ctx->tri_mat);
    dfree(ctx, ctx->cb_ord); dfree(ctx, ctx->keys0); dfree(ctx, ctx->keys1); dfree(ctx, ctx->vals0);
    dfree(ctx, ctx->vals1); dfree(ctx, ctx->hist); dfree(ctx, ctx->refit_cnt); dfree(ctx, ctx->nodes); dfree(ctx, ctx->nodes4);
    dfree(ctx, ctx->parent); dfree(ctx, ctx->leaf_geo); dfree(ctx, ctx->leaf_shade);
    for (int k = 0; k < 6; ++k) { dfree(ctx, ctx->qa[k]); dfree(ctx, ctx->qb[k]); }
    for (int k = 0; k < 4; ++k) dfree(ctx, ctx->ps_f[k]);
    for (int k = 0; k < 9; ++k) dfree(ctx, ctx->sq_f[k]);
    dfree(ctx, ctx->qa_path); dfree(ctx, ctx->qb_path); dfree(ctx, ctx->hit_rec); dfree(ctx, ctx->acc);
    dfree(ctx, ctx->sq_pk);
    for (auto& b : ctx->st_f) dfree(ctx, b);
    dfree(ctx, ctx->st_c); dfree(ctx, ctx->st_path); dfree(ctx, ctx->st_mask); dfree(ctx, ctx->tile_sums); dfree(ctx, ctx->tile_sky); dfree(ctx, ctx->counters); dfree(ctx, ctx->sums);
    if (ctx->events_ready)
        for (auto& ev : ctx->ev) (void)hipEventDestroy(ev);
    delete ctx;
    return BI_PT_OK;
}

const char* bi_pt_last_error(const bi_pt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

bi_pt_status bi_pt_set_scene(bi_pt_ctx* ctx, const bi_pt_mesh* meshes, uint32_t num_meshes,
                             const bi_pt_instance* instances, uint32_t num_instances,
                             const bi_pt_material* materials, uint32_t num_materials) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if ((num_meshes && !meshes) || (num_instances && !instances) || (num_materials && !materials))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null scene array");
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<uint32_t> vtx_off(num_meshes), idx_off(num_meshes);
    size_t nv = 0, ni = 0;
    for (uint32_t m = 0; m < num_meshes; ++m) {
        if (meshes[m].num_indices % 3 != 0)
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "mesh index count is not a multiple of 3");
        if ((meshes[m].num_vertices && (!meshes[m].positions || !meshes[m].normals)) ||
            (meshes[m].num_indices && !meshes[m].indices))
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null mesh buffer");
        for (uint32_t k = 0; k < meshes[m].num_indices; ++k)
            if (meshes[m].indices[k] >= meshes[m].num_vertices)
                return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "mesh index out of range");
        for (size_t k = 0; k < 3 * (size_t)meshes[m].num_vertices; ++k)
            if (!std::isfinite(meshes[m].positions[k]) || !std::isfinite(meshes[m].normals[k]))
                return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite vertex position or normal");
        vtx_off[m] = (uint32_t)nv;
        idx_off[m] = (uint32_t)ni;
        nv += meshes[m].num_vertices;
        ni += meshes[m].num_indices;
    }
    uint64_t tris = 0;
    std::vector<InstanceRec> ih(num_instances);
    for (uint32_t i = 0; i < num_instances; ++i) {
        const bi_pt_instance& I = instances[i];
        if (I.mesh >= num_meshes || I.material >= num_materials)
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "instance mesh/material out of range");
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(I.transform[k]))
                return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite instance transform");
        std::memcpy(ih[i].xform, I.transform, sizeof ih[i].xform);
        ih[i].first_tri = (uint32_t)tris;
        ih[i].idx_off = idx_off[I.mesh];
        ih[i].vtx_off = vtx_off[I.mesh];
        ih[i].material = I.material;
        tris += meshes[I.mesh].num_indices / 3;
    }
    if (tris >= (1ull << 29)) return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "more than 2^29 triangles");
    for (uint32_t m = 0; m < num_materials; ++m)
        if (!(materials[m].roughness >= 0.0f) || !(materials[m].anisotropy >= 0.0f && materials[m].anisotropy <= 1.0f))
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "material roughness/anisotropy out of range");
    std::vector<float> hp(3 * nv), hn(3 * nv);
    std::vector<uint32_t> hi(ni);
    for (uint32_t m = 0; m < num_meshes; ++m) {
        if (meshes[m].num_vertices) {
            std::memcpy(&hp[3 * (size_t)vtx_off[m]], meshes[m].positions, 12 * (size_t)meshes[m].num_vertices);
            std::memcpy(&hn[3 * (size_t)vtx_off[m]], meshes[m].normals, 12 * (size_t)meshes[m].num_vertices);
        }
        if (meshes[m].num_indices)
            std::memcpy(&hi[idx_off[m]], meshes[m].indices, 4 * (size_t)meshes[m].num_indices);
    }
    TRY(dalloc(ctx, ctx->pos, std::max<size_t>(hp.size(), 1)));
    TRY(dalloc(ctx, ctx->nrm, std::max<size_t>(hn.size(), 1)));
    TRY(dalloc(ctx, ctx->idx, std::max<size_t>(hi.size(), 1)));
    TRY(dalloc(ctx, ctx->inst, std::max<size_t>(ih.size(), 1)));
    TRY(dalloc(ctx, ctx->mats, std::max<size_t>(num_materials, 1)));
    if (!hp.empty()) HIP_TRY(hipMemcpy(ctx->pos.p, hp.data(), hp.size() * 4, hipMemcpyHostToDevice));
    if (!hn.empty()) HIP_TRY(hipMemcpy(ctx->nrm.p, hn.data(), hn.size() * 4, hipMemcpyHostToDevice));
    if (!hi.empty()) HIP_TRY(hipMemcpy(ctx->idx.p, hi.data(), hi.size() * 4, hipMemcpyHostToDevice));
    if (!ih.empty()) HIP_TRY(hipMemcpy(ctx->inst.p, ih.data(), ih.size() * sizeof(InstanceRec), hipMemcpyHostToDevice));
    if (num_materials)
        HIP_TRY(hipMemcpy(ctx->mats.p, materials, num_materials * sizeof(MaterialRec), hipMemcpyHostToDevice));
    ctx->inst_host = ih;
    ctx->num_tris = (uint32_t)tris;
    ctx->num_inst = num_instances;
    ctx->num_mats = num_materials;
    ctx->has_scene = true;
    ctx->bvh_built = false;
    ctx->acc_count = 0;
    return BI_PT_OK;
}

bi_pt_status bi_pt_update_transforms(bi_pt_ctx* ctx, const float* xf, uint32_t num_instances) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    if (num_instances != ctx->num_inst || (num_instances && !xf))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "instance count differs from the scene");
    for (size_t k = 0; k < 12 * (size_t)num_instances; ++k)
        if (!std::isfinite(xf[k])) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite instance transform");
    for (uint32_t i = 0; i < num_instances; ++i) std::memcpy(ctx->inst_host[i].xform, xf + 12 * (size_t)i, 48);
    HIP_TRY(hipSetDevice(ctx->device));
    if (num_instances)
        HIP_TRY(hipMemcpy(ctx->inst.p, ctx->inst_host.data(), num_instances * sizeof(InstanceRec),
                          hipMemcpyHostToDevice));
    ctx->bvh_built = false;
    ctx->acc_count = 0;
    return BI_PT_OK;
}

bi_pt_status bi_pt_set_lights(bi_pt_ctx* ctx, const bi_pt_dir_light* lights, uint32_t num_lights) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (num_lights > (uint32_t)kMaxDirLights)
        return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "at most 4 directional lights");
    if (num_lights && !lights) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null lights");
    std::vector<LightRec> next(num_lights);
    if (num_lights) std::memcpy(next.data(), lights, num_lights * sizeof(LightRec));
    const bool changed = next.size() != ctx->lights.size() ||
                         (num_lights && std::memcmp(next.data(), ctx->lights.data(), num_lights * sizeof(LightRec)) != 0);
    ctx->lights.swap(next);
    if (changed) ctx->acc_count = 0;   // unchanged lights (set every frame) keep the running image
    return BI_PT_OK;
}

bi_pt_status bi_pt_set_sky(bi_pt_ctx* ctx, const float* rgba, uint32_t w, uint32_t h, const float color[3],
                           const float rot[9]) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if ((w == 0) != (h == 0) || ((w || h) && !rgba) || !color || !rot)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "bad sky arguments");
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t ntex = (size_t)w * h * 4;
    if (ctx->sky_set && ctx->sky.width == w && ctx->sky.height == h &&
        std::memcmp(ctx->sky.color, color, sizeof ctx->sky.color) == 0 &&
        std::memcmp(ctx->sky.rot, rot, sizeof ctx->sky.rot) == 0 && ctx->sky_host.size() == ntex &&
        (ntex == 0 || std::memcmp(ctx->sky_host.data(), rgba, ntex * sizeof(float)) == 0))
        return BI_PT_OK;   // unchanged sky (set every frame): keep tables and the running image
    ctx->sky_host.assign(rgba, rgba + ntex);
    ctx->sky_set = true;
    SkyParams sp{};
    sp.width = w;
    sp.height = h;
    std::memcpy(sp.color, color, sizeof sp.color);
    std::memcpy(sp.rot, rot, sizeof sp.rot);
    if (w) {
        std::vector<float> wt, cc, rows, marg;
        float total = 0.0f;
        build_sky_table(rgba, w, h, color, wt, cc, rows, marg, total);
        TRY(dalloc(ctx, ctx->sky_tex, (size_t)w * h));
        TRY(dalloc(ctx, ctx->sky_w, wt.size()));
        TRY(dalloc(ctx, ctx->sky_ccdf, cc.size()));
        TRY(dalloc(ctx, ctx->sky_rows, rows.size()));
        TRY(dalloc(ctx, ctx->sky_marg, marg.size()));
        HIP_TRY(hipMemcpy(ctx->sky_tex.p, rgba, (size_t)w * h * 16, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->sky_w.p, wt.data(), wt.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->sky_ccdf.p, cc.data(), cc.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->sky_rows.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->sky_marg.p, marg.data(), marg.size() * 4, hipMemcpyHostToDevice));
        sp.texels = ctx->sky_tex.p;
        sp.weight = ctx->sky_w.p;
        sp.ccdf = ctx->sky_ccdf.p;
        sp.rows = ctx->sky_rows.p;
        sp.marg = ctx->sky_marg.p;
        sp.total = total;
        sp.wh_over_total = total > 0.0f ? (float)(w * h) / total : 0.0f;
    }
    ctx->sky = sp;
    ctx->acc_count = 0;
    return BI_PT_OK;
}

bi_pt_status bi_pt_reset_accumulation(bi_pt_ctx* ctx) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    ctx->acc_count = 0;
    return BI_PT_OK;
}

bi_pt_status bi_pt_render(bi_pt_ctx* ctx, const bi_pt_camera* camera, const bi_pt_settings* st, float* out_rgba,
                          float* out_depth, bi_pt_stream stream) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!camera || !st || !out_rgba) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null camera/settings/output");
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    if (camera->width == 0 || camera->height == 0 || st->spp == 0)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "empty image or zero samples");
    const uint64_t npix64 = (uint64_t)camera->width * camera->height;
    if (npix64 > (1ull << 26)) return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "image larger than 2^26 pixels");
    if (!(st->ray_length > 0.0f) || !(st->ray_length <= kShadowTMax))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "ray_length must be positive and at most 1.7e38");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint32_t npix = (uint32_t)npix64;
    const uint32_t M = std::min(std::max(st->max_bounces, 2u), 16u);
    const uint32_t D = M - 1;
    const uint32_t nl = (uint32_t)ctx->lights.size();
    const uint32_t nslots = nl + 2u;
    const bool sky_nee = !(st->flags & BI_PT_FLAG_NO_SKY_NEE) && ctx->sky.total > 0.0f;
    const bool timing = (st->flags & BI_PT_FLAG_TIMING) != 0;
    const bool record = (st->flags & BI_PT_FLAG_RECORD_QUEUES) != 0;

    CameraParams cam{};
    std::memcpy(cam.inv_view, camera->matrix_inv_view, 64);
    std::memcpy(cam.inv_proj, camera->matrix_inv_proj, 64);
    std::memcpy(cam.proj_view, camera->matrix_proj_view, 64);
    cam.width = camera->width;
    cam.height = camera->height;
    // accumulation history: continue only with the same view, size and sampling settings
    const bool same = ctx->acc_count > 0 && st->accumulate && ctx->acc_w == cam.width && ctx->acc_h == cam.height &&
                      std::memcmp(&ctx->last_cam, &cam, sizeof cam) == 0 && ctx->last_seed == st->seed &&
                      ctx->last_flags == (st->flags & 3u) && ctx->last_bounces == M &&
                      ctx->last_ray_length == st->ray_length;
    TRY(dalloc(ctx, ctx->sums, (size_t)3 * npix));
    if (!same) {
        HIP_TRY(hipMemsetAsync(ctx->sums.p, 0, (size_t)12 * npix, s));
        ctx->acc_count = 0;
    }
    ctx->acc_w = cam.width;
    ctx->acc_h = cam.height;
    ctx->last_cam = cam;
    ctx->last_seed = st->seed;
    ctx->last_flags = st->flags & 3u;
    ctx->last_bounces = M;
    ctx->last_ray_length = st->ray_length;
    if (ctx->acc_count + st->spp > 0xffffffffull)
        return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "more than 2^32 accumulated samples");

    // wave plan: whole samples of every pixel per wave
    uint32_t wave_paths = st->wave_paths ? st->wave_paths : (1u << 22);
    uint32_t wave_samples = std::max(1u, std::min(st->spp, wave_paths / npix));
    const uint32_t cap = wave_samples * npix;
    uint32_t sh_mult = 1;
    for (const LightRec& L : ctx->lights) sh_mult += L.sm_index >= 0 ? 1u : 0u;
    uint32_t lds_stack = st->stack_lds ? std::min(st->stack_lds, 32u) : 16u;
    const uint32_t refill = st->refill_lanes ? std::min(st->refill_lanes, 64u) : 32u;
    const uint32_t parts = st->work_parts ? std::min(st->work_parts, kWorkPartsMax) : kWorkParts;
    TRY(ensure_wave_buffers(ctx, cap, nslots, sh_mult));

    // device counters: queue counts, persistent-lane ray counters and error flag per call; the ray statistics
    // (slots kCntRays..) accumulate over the context's lifetime
    HIP_TRY(hipMemsetAsync(ctx->counters.p, 0, kCntRays * sizeof(uint32_t), s));
    HIP_TRY(hipMemsetAsync(ctx->counters.p + kCntTraceWork, 0, (kCntSlots - kCntTraceWork) * sizeof(uint32_t), s));
    if (timing) HIP_TRY(hipEventRecord(ctx->ev[0], s));
    TRY(build_bvh(ctx, s));
    if (timing) HIP_TRY(hipEventRecord(ctx->ev[1], s));
    uint32_t* cnt = ctx->counters.p;
    uint64_t* rays = reinterpret_cast<uint64_t*>(cnt + kCntRays);

    const BvhView bvh = bvh_view(ctx);
    PathState ps{ctx->ps_f[0].p, ctx->ps_f[1].p, ctx->ps_f[2].p, ctx->ps_f[3].p, ctx->acc.p, ctx->wave_cap};
    HitBuf hits{ctx->hit_rec.p};
    RayQueue qA = queue_of(ctx->qa, ctx->qa_path), qB = queue_of(ctx->qb, ctx->qb_path);
    ShadowQueue sq{ctx->sq_f[0].p, ctx->sq_f[1].p, ctx->sq_f[2].p, ctx->sq_f[3].p, ctx->sq_f[4].p,
                   ctx->sq_f[5].p, ctx->sq_f[6].p, ctx->sq_f[7].p, ctx->sq_f[8].p, ctx->sq_pk.p};
    const size_t wc = ctx->wave_cap;
    ShadeStage stage{ctx->st_f[0].p, ctx->st_f[1].p, ctx->st_f[2].p, ctx->st_path.p,
                     ctx->st_f[3].p, ctx->st_f[4].p, ctx->st_f[5].p,
                     ctx->st_f[6].p, ctx->st_f[7].p, ctx->st_f[8].p,
                     ctx->st_c.p, ctx->st_c.p + (size_t)(nl + 1u) * wc, ctx->st_c.p + (size_t)2 * (nl + 1u) * wc,
                     ctx->st_mask.p, ctx->tile_sums.p, ctx->tile_sky.p};
    if (record) { ctx->rec_ext.clear(); ctx->rec_sh.clear(); }

    // the launches of one kind are timed together: events around each class of kernel
    std::vector<hipEvent_t> tev;
    std::vector<int> tcls;
    std::vector<hipEvent_t> extra_events;
    auto mark = [&](int cls) -> bi_pt_status {
        if (!timing) return BI_PT_OK;
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventRecord(e, s));
        tev.push_back(e);
        tcls.push_back(cls);
        return BI_PT_OK;
    };
    uint32_t launches_closest = 0, launches_shadow = 0;

    uint32_t done = 0;
    while (done < st->spp) {
        const uint32_t ws = std::min(wave_samples, st->spp - done);
        WaveParams wp{};
        wp.cam = cam;
        wp.num_paths = ws * npix;
        wp.npix = npix;
        wp.sample0 = (uint32_t)ctx->acc_count + done;
        wp.seed = st->seed;
        wp.flags = st->flags & 3u;
        wp.ray_length = st->ray_length;
        if (record) {
            ctx->rec_ext.emplace_back(D);
            ctx->rec_sh.emplace_back(D);
            std::vector<uint32_t> all(wp.num_paths);
            for (uint32_t k = 0; k < wp.num_paths; ++k) all[k] = k;
            ctx->rec_ext.back()[0] = all;
        }
        for (uint32_t b = 0; b < D; ++b) {
            const bool primary = b == 0;
            const RayQueue& qin = (b & 1u) ? qA : qB;      // segment b reads what segment b-1 wrote
            const RayQueue& qout = (b & 1u) ? qB : qA;
            uint32_t* cnt_in = (b & 1u) ? &cnt[kCntExtA] : &cnt[kCntExtB];
            uint32_t* cnt_out = (b & 1u) ? &cnt[kCntExtB] : &cnt[kCntExtA];

            TraceArgs ta{};
            ta.bvh = bvh;
            ta.wave = wp;
            ta.q = qin;
            ta.count_in = primary ? nullptr : cnt_in;
            ta.hits = hits;
            ta.work = &cnt[kCntTraceWork];
            ta.parts = parts;
            ta.refill = refill;
            ta.error = &cnt[kCntError];
            TRY(mark(0));
            launch_trace_closest(ta, wp.num_paths, primary, lds_stack, s);
            ++launches_closest;
            TRY(mark(1));

            ShadeArgs sa{};
            sa.wave = wp;
            sa.leaf_shade = ctx->leaf_shade.p;
            sa.materials = ctx->mats.p;
            for (uint32_t l = 0; l < nl; ++l) sa.lights[l] = ctx->lights[l];
            sa.num_lights = nl;
            sa.sky = ctx->sky;
            sa.sky_nee = sky_nee ? 1u : 0u;
            sa.qin = qin;
            sa.count_in = primary ? nullptr : cnt_in;
            sa.hits = hits;
            sa.st = stage;
            sa.ps = ps;
            sa.trace_work_reset = &cnt[kCntTraceWork];
            sa.shadow_work_reset = &cnt[kCntShadowWork];
            sa.bounce = b;
            sa.last_bounce = D - 1;
            sa.depth_out = out_depth;
            sa.depth_sample = (uint32_t)ctx->acc_count;
            launch_shade(sa, wp.num_paths, s);

            ScanArgs sc{};
            sc.tile_sums = stage.tile_sums;
            sc.tile_sky = stage.tile_sky;
            sc.count_in = primary ? nullptr : cnt_in;
            sc.num_entries = wp.num_paths;
            sc.count_ext = cnt_out;
            sc.count_shadow = &cnt[kCntShadow];
            sc.ray_counters = rays;
            launch_scan_tiles(sc, s);

            CompactArgs ca{};
            ca.st = stage;
            ca.count_in = primary ? nullptr : cnt_in;
            ca.num_entries = wp.num_paths;
            ca.num_lights = nl;
            ca.cap = ctx->wave_cap;
            for (uint32_t l = 0; l < nl; ++l)
                for (int c = 0; c < 3; ++c) ca.ldir[l][c] = ctx->lights[l].direction[c];
            ca.qout = qout;
            ca.sq = sq;
            launch_compact(ca, wp.num_paths, s);
            TRY(mark(2));

            ShadowArgs sha{};
            sha.bvh = bvh;
            sha.q = sq;
            sha.count_in = &cnt[kCntShadow];
            sha.ps = ps;
            for (uint32_t l = 0; l < nl; ++l) sha.keep[l] = 1.0f - ctx->lights[l].shadow_strength;
            sha.num_lights = nl;
            sha.work = &cnt[kCntShadowWork];
            sha.parts = parts;
            sha.refill = refill;
            sha.error = &cnt[kCntError];
            launch_trace_shadow(sha, wp.num_paths * sh_mult, lds_stack, s);
            ++launches_shadow;
            TRY(mark(3));
            HIP_TRY(hipGetLastError());
            if (record) {
                uint32_t hc[3];
                HIP_TRY(hipMemcpyAsync(hc, cnt, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                HIP_TRY(hipStreamSynchronize(s));
                const uint32_t next = hc[(b & 1u) ? kCntExtB : kCntExtA];
                auto& sh = ctx->rec_sh.back()[b];
                sh.resize(hc[kCntShadow]);
                if (!sh.empty())
                    HIP_TRY(hipMemcpy(sh.data(), ctx->sq_pk.p, sh.size() * 4, hipMemcpyDeviceToHost));
                if (b + 1 < D) {
                    auto& ex = ctx->rec_ext.back()[b + 1];
                    ex.resize(next);
                    if (next) HIP_TRY(hipMemcpy(ex.data(), qout.path, next * 4, hipMemcpyDeviceToHost));
                }
            }
        }
        ResolveArgs ra{};
        ra.ps = ps;
        ra.num_slots = nslots;
        ra.npix = npix;
        ra.samples = ws;
        ra.width = cam.width;
        ra.height = cam.height;
        ra.sums = ctx->sums.p;
        launch_resolve(ra, s);
        TRY(mark(4));
        done += ws;
    }
    ctx->acc_count += st->spp;
    launch_output(ctx->sums.p, npix, (float)ctx->acc_count, 0, out_rgba, s);
    HIP_TRY(hipGetLastError());
    ctx->frames++;
    ctx->camera_rays += (uint64_t)npix * st->spp;
    ctx->stats.frames = ctx->frames;
    ctx->stats.rays_camera = ctx->camera_rays;
    ctx->stats.samples_accumulated = ctx->acc_count;
    ctx->stats.launches_trace_closest = launches_closest;
    ctx->stats.launches_trace_shadow = launches_shadow;
    if (timing) {
        HIP_TRY(hipEventRecord(ctx->ev[2], s));
        HIP_TRY(hipEventSynchronize(ctx->ev[2]));
        float ms_build = 0, total = 0;
        HIP_TRY(hipEventElapsedTime(&ms_build, ctx->ev[0], ctx->ev[1]));
        HIP_TRY(hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[2]));
        float cls_ms[5] = {0, 0, 0, 0, 0};
        // class of the interval (tev[k-1], tev[k]) is tcls[k]: 1 trace, 2 shade, 3 shadow, 4 resolve
        for (size_t k = 1; k < tev.size(); ++k) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, tev[k - 1], tev[k]));
            int c = tcls[k];
            if (c >= 1 && c <= 4) cls_ms[c] += ms;
        }
        for (auto e : tev) (void)hipEventDestroy(e);
        ctx->stats.ms_build = ms_build;
        ctx->stats.ms_trace_closest = cls_ms[1];
        ctx->stats.ms_shade = cls_ms[2];
        ctx->stats.ms_trace_shadow = cls_ms[3];
        ctx->stats.ms_resolve = cls_ms[4];
        ctx->stats.ms_total = total;
    }
    return BI_PT_OK;
}

bi_pt_status bi_pt_get_stats(const bi_pt_ctx* cctx, bi_pt_stats* out) {
    bi_pt_ctx* ctx = const_cast<bi_pt_ctx*>(cctx);
    if (!ctx || !out) return BI_PT_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    uint32_t hc[kCntSlots];
    HIP_TRY(hipMemcpy(hc, ctx->counters.p, sizeof hc, hipMemcpyDeviceToHost));
    uint64_t r[4];
    std::memcpy(r, &hc[kCntRays], sizeof r);
    ctx->stats.rays_extension = r[1];
    ctx->stats.rays_light_shadow = r[2];
    ctx->stats.rays_sky_shadow = r[3];
    *out = ctx->stats;
    if (hc[kCntError]) {
        *out = ctx->stats;
        return fail(ctx, BI_PT_ERROR_DEVICE_CHECK,
                    std::string("device check failed: ") + ((hc[kCntError] & 1u) ? "traversal stack overflow " : "") +
                        ((hc[kCntError] & 4u) ? "refit wait limit " : "") +
                        ((hc[kCntError] & 8u) ? "traversal step bound exceeded" : ""));
    }
    return BI_PT_OK;
}

bi_pt_status bi_pt_debug_get_bvh(bi_pt_ctx* ctx, uint32_t* morton, uint32_t* order, int32_t* children,
                                 float* child_bounds, int32_t* parents, float* centroid_bounds) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->bvh_built) {
        TRY(build_bvh(ctx, nullptr));
    }
    HIP_TRY(hipDeviceSynchronize());
    const uint32_t n = ctx->num_tris;
    if (n == 0) return BI_PT_OK;
    const size_t ni = n > 1 ? n - 1 : 1;
    // keys/values: the 4-pass sort leaves the sorted pair in buffer 0; the codes in
    // input order are recomputed by the Morton kernel into buffer... buffer 0 holds sorted
    // keys after the build, so input-order codes are gathered through the permutation.
    std::vector<uint32_t> sk(n), sv(n);
    HIP_TRY(hipMemcpy(sk.data(), ctx->keys0.p, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(sv.data(), ctx->vals0.p, n * 4, hipMemcpyDeviceToHost));
    if (order) std::memcpy(order, sv.data(), n * 4);
    if (morton)
        for (uint32_t k = 0; k < n; ++k) morton[sv[k]] = sk[k];
    std::vector<BvhNode> nodes(ni);
    HIP_TRY(hipMemcpy(nodes.data(), ctx->nodes.p, ni * sizeof(BvhNode), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ni; ++i) {
        if (children) { children[2 * i] = nodes[i].child[0]; children[2 * i + 1] = nodes[i].child[1]; }
        if (child_bounds) std::memcpy(&child_bounds[12 * i], nodes[i].box, 48);
    }
    if (parents) HIP_TRY(hipMemcpy(parents, ctx->parent.p, (ni + n) * 4, hipMemcpyDeviceToHost));
    if (centroid_bounds) HIP_TRY(hipMemcpy(centroid_bounds, ctx->cb_out.p, 24, hipMemcpyDeviceToHost));
    return BI_PT_OK;
}

bi_pt_status bi_pt_debug_get_bvh4(bi_pt_ctx* ctx, int32_t* child4, float* box4) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->bvh_built) TRY(build_bvh(ctx, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    const uint32_t n = ctx->num_tris;
    if (n == 0) return BI_PT_OK;
    const size_t ni = n > 1 ? n - 1 : 1;
    std::vector<Bvh4Node> nodes(ni);
    HIP_TRY(hipMemcpy(nodes.data(), ctx->nodes4.p, ni * sizeof(Bvh4Node), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ni; ++i) {
        if (child4) std::memcpy(&child4[4 * i], nodes[i].child, 16);
        if (box4) std::memcpy(&box4[24 * i], nodes[i].lox, 96);
    }
    return BI_PT_OK;
}

uint32_t bi_pt_debug_num_tris(const bi_pt_ctx* ctx) { return ctx ? ctx->num_tris : 0; }

bi_pt_status bi_pt_debug_queue(bi_pt_ctx* ctx, uint32_t wave, uint32_t bounce, uint32_t kind, uint32_t* out,
                               uint32_t capacity, uint32_t* count) {
    if (!ctx || !count) return BI_PT_ERROR_INVALID_ARGUMENT;
    auto& rec = kind == 0 ? ctx->rec_ext : ctx->rec_sh;
    if (wave >= rec.size() || bounce >= rec[wave].size())
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "no recorded queue for this wave/bounce");
    const auto& v = rec[wave][bounce];
    *count = (uint32_t)v.size();
    if (out) std::memcpy(out, v.data(), std::min<size_t>(capacity, v.size()) * 4);
    return BI_PT_OK;
}

bi_pt_status bi_pt_debug_get_sums(bi_pt_ctx* ctx, float* sums) {
    if (!ctx || !sums) return BI_PT_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    if (!ctx->sums.p) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "nothing rendered yet");
    HIP_TRY(hipMemcpy(sums, ctx->sums.p, (size_t)12 * ctx->acc_w * ctx->acc_h, hipMemcpyDeviceToHost));
    return BI_PT_OK;
}

bi_pt_status bi_pt_debug_trace(bi_pt_ctx* ctx, uint32_t n, const float* rays, int any_hit, float* hit_t,
                               uint32_t* hit_leaf, float* hit_u, float* hit_v) {
    if (!ctx || (n && (!rays || !hit_leaf))) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    if (n == 0) return BI_PT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->bvh_built) TRY(build_bvh(ctx, nullptr));
    // all rays share tmin = 0.001 and the first ray's tmax
    const float tmax = rays[7];
    if (!(tmax > 0.001f) || !(tmax <= kShadowTMax))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "debug trace needs 0.001 < tmax <= 1.7e38");
    for (uint32_t i = 0; i < n; ++i)
        if (rays[8 * i + 3] != 0.001f || rays[8 * i + 7] != tmax)
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "debug trace needs tmin = 0.001 and one tmax");
    DevBuf<float> f[6];
    DevBuf<uint32_t> path, cnt, out;
    DevBuf<HitRec> hrec;
    for (int k = 0; k < 6; ++k) TRY(dalloc(ctx, f[k], n));
    TRY(dalloc(ctx, hrec, n));
    TRY(dalloc(ctx, path, n));
    TRY(dalloc(ctx, cnt, 4 + kWorkParts * kWorkStride));
    TRY(dalloc(ctx, out, n));
    std::vector<float> col(n);
    const int src[6] = {0, 1, 2, 4, 5, 6};
    bi_pt_status status = BI_PT_OK;
    for (int k = 0; k < 6 && status == BI_PT_OK; ++k) {
        for (uint32_t i = 0; i < n; ++i) col[i] = rays[8 * i + src[k]];
        if (hipMemcpy(f[k].p, col.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess) status = BI_PT_ERROR_HIP;
    }
    std::vector<uint32_t> hn(4 + kWorkParts * kWorkStride, 0u);
    hn[0] = n;
    if (status == BI_PT_OK && hipMemcpy(cnt.p, hn.data(), hn.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        status = BI_PT_ERROR_HIP;
    RayQueue q{f[0].p, f[1].p, f[2].p, f[3].p, f[4].p, f[5].p, path.p};
    if (status == BI_PT_OK) {
        if (any_hit) {
            launch_debug_anyhit(bvh_view(ctx), q, n, tmax, out.p, &ctx->counters.p[kCntError], nullptr);
            if (hipMemcpy(hit_leaf, out.p, n * 4, hipMemcpyDeviceToHost) != hipSuccess) status = BI_PT_ERROR_HIP;
        } else {
            TraceArgs ta{};
            ta.bvh = bvh_view(ctx);
            ta.wave.ray_length = tmax;
            ta.q = q;
            ta.count_in = cnt.p;
            ta.work = cnt.p + 4;   // zeroed part counters
            ta.parts = kWorkParts;
            ta.refill = 32;
            ta.hits = HitBuf{hrec.p};
            ta.error = &ctx->counters.p[kCntError];
            launch_trace_closest(ta, n, false, 16, nullptr);
            std::vector<HitRec> h(n);
            if (hipMemcpy(h.data(), hrec.p, n * sizeof(HitRec), hipMemcpyDeviceToHost) != hipSuccess) {
                status = BI_PT_ERROR_HIP;
            } else {
                for (uint32_t i = 0; i < n; ++i) {
                    hit_leaf[i] = h[i].leaf;
                    if (hit_t) hit_t[i] = h[i].t;
                    if (hit_u) hit_u[i] = h[i].u;
                    if (hit_v) hit_v[i] = h[i].v;
                }
            }
        }
    }
    for (int k = 0; k < 6; ++k) dfree(ctx, f[k]);
    dfree(ctx, hrec);
    dfree(ctx, path);
    dfree(ctx, cnt);
    dfree(ctx, out);
    if (status != BI_PT_OK) return fail(ctx, status, "debug trace failed");
    return BI_PT_OK;
}

}  // extern "C"
