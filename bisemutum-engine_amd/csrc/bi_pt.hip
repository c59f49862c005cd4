// bi_pt.hip -- host side of the C ABI (include/bi_pt.h): contexts, scene/light/sky
// uploads, the per-frame LBVH build, single-job rendering (through the wave engine of
// engine.hip), statistics and the inspection hooks used by the parity tests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pt_host.hpp"

using namespace bipt;

namespace {

// Importance table of the sky, built in FP32 in a fixed order on the host: per texel
// luminance(texel * colour) times sin(theta) of its row (equirect) or 1/r^3 of its
// centre (cube face texel at face coordinates (a, b), r^2 = 1 + a^2 + b^2: the
// texel's solid angle up to a constant), then the sum of each row and their total.
void build_sky_table(const float* rgba, uint32_t W, uint32_t H, uint32_t layout, const float color[3],
                     std::vector<float>& w, std::vector<float>& rows, float& total) {
    w.assign((size_t)W * H, 0.0f);
    rows.assign(H, 0.0f);
    total = 0.0f;
    const float pi_over_h = kPi / (float)H;
    for (uint32_t i = 0; i < H; ++i) {
        const uint32_t r = layout == kSkyCube ? i % W : i;
        const float b = (float)(2 * r + 1) / (float)W - 1.0f;
        const float st = layout == kSkyCube ? 0.0f : pt_sin(((float)i + 0.5f) * pi_over_h);
        float run = 0.0f;
        for (uint32_t j = 0; j < W; ++j) {
            const float* p = &rgba[4 * ((size_t)i * W + j)];
            float rr = p[0] * color[0], g = p[1] * color[1], bb = p[2] * color[2];
            float lum = (0.2126f * rr + 0.7152f * g) + 0.0722f * bb;
            float x;
            if (layout == kSkyCube) {
                const float a = (float)(2 * j + 1) / (float)W - 1.0f;
                const float q = (1.0f + a * a) + b * b;
                x = lum * (1.0f / (q * std::sqrt(q)));
            } else {
                x = lum * st;
            }
            if (!(x > 0.0f)) x = 0.0f;
            w[(size_t)i * W + j] = x;
            run = run + x;
        }
        rows[i] = run;
        total = total + run;
    }
}

// Vose's alias tables (Vose 1991, "A linear algorithm for generating random numbers with
// a given distribution") over n weights of sum `sum`: entry k keeps itself with probability
// prob[k], else goes to alias[k].  The weights scaled by n / sum are split into entries
// below 1 and the rest, each a stack in index order; the top small entry takes its
// probability and the top large one as its alias, which keeps what is left over.  Entries
// left in either stack at the end keep themselves.  The CPU restatement takes the same steps.
void build_alias(const float* w, uint32_t n, float sum, float* prob, uint32_t* alias) {
    std::vector<float> q(n);
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    const float scale = sum > 0.0f ? (float)n / sum : 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
        q[k] = w[k] * scale;
        (q[k] < 1.0f ? small : large).push_back(k);
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t l = small.back();
        small.pop_back();
        const uint32_t g = large.back();
        large.pop_back();
        prob[l] = q[l];
        alias[l] = g;
        q[g] = (q[g] + q[l]) - 1.0f;
        (q[g] < 1.0f ? small : large).push_back(g);
    }
    for (uint32_t k : large) { prob[k] = 1.0f; alias[k] = k; }
    for (uint32_t k : small) { prob[k] = 1.0f; alias[k] = k; }
}

}  // namespace

namespace bipt {

BvhView bvh_view(const bi_pt_ctx* ctx) {
    return BvhView{ctx->nodes4.p, ctx->leaf_geo.p, ctx->num_tris, ctx->leaf_shade.p, ctx->mats.p,
                   scene_leaf_uv(ctx), scene_tex(ctx), 0u};
}

bi_pt_status ensure_build_buffers(bi_pt_ctx* ctx) {
    const uint32_t n = ctx->num_tris;
    if (n == 0) return BI_PT_OK;
    const size_t ni = n > 1 ? n - 1 : 1;
    BI_TRY(dalloc(ctx, ctx->tri_v, (size_t)9 * n));
    BI_TRY(dalloc(ctx, ctx->tri_n, (size_t)9 * n));
    BI_TRY(dalloc(ctx, ctx->tri_mat, n));
    BI_TRY(dalloc(ctx, ctx->cb_ord, 8));
    BI_TRY(dalloc(ctx, ctx->cb_out, 8));
    BI_TRY(dalloc(ctx, ctx->keys0, n));
    BI_TRY(dalloc(ctx, ctx->keys1, n));
    BI_TRY(dalloc(ctx, ctx->vals0, n));
    BI_TRY(dalloc(ctx, ctx->vals1, n));
    BI_TRY(dalloc(ctx, ctx->hist, sort_hist_entries(n)));
    BI_TRY(dalloc(ctx, ctx->nodes, ni));
    BI_TRY(dalloc(ctx, ctx->nodes4, ni));
    BI_TRY(dalloc(ctx, ctx->parent, ni + n));
    BI_TRY(dalloc(ctx, ctx->refit_cnt, ni));
    BI_TRY(dalloc(ctx, ctx->leaf_geo, n));
    BI_TRY(dalloc(ctx, ctx->leaf_shade, n));
    BI_TRY(dalloc(ctx, ctx->tl_cost, ni));
    BI_TRY(dalloc(ctx, ctx->tl_height, ni));
    BI_TRY(dalloc(ctx, ctx->tl_snap, ni));
    BI_TRY(dalloc(ctx, ctx->tl_starts, 1 + treelet_start_cap(n)));
    if (ctx->need_uv) BI_TRY(dalloc(ctx, ctx->leaf_uv, n));
    else dfree(ctx, ctx->leaf_uv);
    if (scene_wants_tan(ctx)) BI_TRY(dalloc(ctx, ctx->leaf_tan, n));
    else dfree(ctx, ctx->leaf_tan);
    return BI_PT_OK;
}

// Enqueues the LBVH build of the context's scene on s (buffers must exist: see
// ensure_build_buffers); error flags of the refit go to `error`.
bi_pt_status build_bvh(bi_pt_ctx* ctx, hipStream_t s, uint32_t* error, uint32_t rounds) {
    ctx->built_rounds = rounds;
    if (ctx->num_tris == 0) return BI_PT_OK;
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
    a.materials = ctx->mats.p;
    a.uvs = ctx->need_uv ? ctx->uvs.p : nullptr;
    a.tangents = scene_wants_tan(ctx) ? ctx->tans.p : nullptr;
    a.leaf_uv = ctx->need_uv ? ctx->leaf_uv.p : nullptr;
    a.leaf_tan = scene_wants_tan(ctx) ? ctx->leaf_tan.p : nullptr;
    a.error = error;
    a.treelet_rounds = rounds;
    a.tl_cost = ctx->tl_cost.p;
    a.tl_height = ctx->tl_height.p;
    a.tl_snap = ctx->tl_snap.p;
    a.tl_starts = ctx->tl_starts.p;
    launch_build_lbvh(a, s);
    BI_HIP_TRY(ctx, hipGetLastError());
    return BI_PT_OK;
}

}  // namespace bipt

namespace {

// Builds the context's BVH now (inspection hooks), on the default stream.
bi_pt_status build_bvh_now(bi_pt_ctx* ctx) {
    BI_TRY(engine_init(ctx, ctx->eng));
    BI_TRY(ensure_build_buffers(ctx));
    BI_TRY(build_bvh(ctx, nullptr, &ctx->eng.counters.p[kCntError], ctx->treelet_rounds));
    ctx->bvh_built = true;
    return BI_PT_OK;
}

}  // namespace

namespace {

// Ray queries: the job record of the context's scene and the path bases {0, n} (one job),
// written by a kernel so the record travels as a launch argument.
__global__ void k_put_query_job(bipt::JobRec rec, uint32_t n, unsigned char* dst) {
    if (threadIdx.x != 0) return;
    *reinterpret_cast<bipt::JobRec*>(dst) = rec;
    uint32_t* bases = reinterpret_cast<uint32_t*>(dst + sizeof(bipt::JobRec));
    bases[0] = 0u;
    bases[1] = n;
    bases[2] = 0u;
    bases[3] = 0u;
}

// Closest-hit records -> caller's arrays: t (+inf on a miss), the triangle's index in the
// scene's instance order (leaf -> sorted order), barycentrics.
__global__ __launch_bounds__(256) void k_query_hits(const bipt::HitRec* __restrict__ hits,
                                                    const uint32_t* __restrict__ order, uint32_t n, float* t,
                                                    uint32_t* tri, float* u, float* v) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const bipt::HitRec h = hits[i];
    const bool hit = h.leaf != bipt::kLeafMiss;
    if (t) t[i] = hit ? h.t : __builtin_inff();
    tri[i] = hit ? order[h.leaf] : bipt::kLeafMiss;
    if (u) u[i] = hit ? h.u : 0.0f;
    if (v) v[i] = hit ? h.v : 0.0f;
}

// Checks a query and makes sure the context's BVH is current (built on s).
bi_pt_status query_prepare(bi_pt_ctx*& ctx, uint32_t n, const bi_pt_rays* rays, float tmax, hipStream_t s) {
    if (ctx && ctx->scene_src) ctx = ctx->scene_src;   // a view traces its scene
    if (!ctx || (n && !rays)) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (n && (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null ray array");
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    if (!(tmax > kTMin) || !(tmax <= kShadowTMax))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "ray queries need 0.001 < tmax <= 1.7e38");
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    BI_TRY(engine_init(ctx, ctx->eng));
    if (!ctx->bvh_built) {
        BI_TRY(ensure_build_buffers(ctx));
        BI_TRY(build_bvh(ctx, s, &ctx->eng.counters.p[kCntError], ctx->treelet_rounds));
        ctx->bvh_built = true;
    }
    return BI_PT_OK;
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
    if (bi_pt_status st = engine_init(ctx, ctx->eng); st != BI_PT_OK) {
        engine_free(ctx, ctx->eng);
        delete ctx;
        return st;
    }
    *out_ctx = ctx;
    return BI_PT_OK;
}

bi_pt_status bi_pt_destroy(bi_pt_ctx* ctx) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    dfree(ctx, ctx->pos); dfree(ctx, ctx->nrm); dfree(ctx, ctx->idx); dfree(ctx, ctx->inst); dfree(ctx, ctx->mats);
    dfree(ctx, ctx->rects); dfree(ctx, ctx->ltc);
    dfree(ctx, ctx->rect_tex); dfree(ctx, ctx->rect_tex_levels); dfree(ctx, ctx->rect_tex_pool);
    dfree(ctx, ctx->sky_tex); dfree(ctx, ctx->sky_w);
    dfree(ctx, ctx->sky_row_alias); dfree(ctx, ctx->sky_tex_alias);
    dfree(ctx, ctx->tri_v); dfree(ctx, ctx->tri_n); dfree(ctx, ctx->cb_out); dfree(ctx, ctx->tri_mat);
    dfree(ctx, ctx->cb_ord); dfree(ctx, ctx->keys0); dfree(ctx, ctx->keys1); dfree(ctx, ctx->vals0);
    dfree(ctx, ctx->vals1); dfree(ctx, ctx->hist); dfree(ctx, ctx->refit_cnt); dfree(ctx, ctx->nodes);
    dfree(ctx, ctx->nodes4); dfree(ctx, ctx->parent); dfree(ctx, ctx->leaf_geo); dfree(ctx, ctx->leaf_shade);
    dfree(ctx, ctx->tl_cost); dfree(ctx, ctx->tl_height); dfree(ctx, ctx->tl_snap); dfree(ctx, ctx->tl_starts);
    dfree(ctx, ctx->rq_job); dfree(ctx, ctx->rq_cnt); dfree(ctx, ctx->rq_path); dfree(ctx, ctx->rq_hits);
    dfree(ctx, ctx->sums);
    engine_free(ctx, ctx->eng);
    delete ctx;
    return BI_PT_OK;
}

const char* bi_pt_last_error(const bi_pt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

bi_pt_status bi_pt_set_scene(bi_pt_ctx* ctx, const bi_pt_mesh* meshes, uint32_t num_meshes,
                             const bi_pt_instance* instances, uint32_t num_instances,
                             const bi_pt_material* materials, uint32_t num_materials) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if ((num_meshes && !meshes) || (num_instances && !instances) || (num_materials && !materials))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null scene array");
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
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
    bool need_uv = false, need_tan = false, non_opaque = false;
    uint32_t tex_refs = 0;
    for (uint32_t m = 0; m < num_materials; ++m) {
        const bi_pt_material& M = materials[m];
        if (M.pattern > BI_PT_PATTERN_CHECKER_XZ)
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "unknown material pattern");
        if (M.blend_mode > BI_PT_BLEND_TRANSLUCENT || M.opacity_channel > 3 || M.roughness_channel > 3 ||
            !std::isfinite(M.opacity) || !std::isfinite(M.step_at) || !std::isfinite(M.step_below) ||
            !std::isfinite(M.step_above))
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "material blend mode / opacity / channel out of range");
        for (uint32_t t : {M.base_color_tex, M.f0_tex, M.normal_tex, M.roughness_tex, M.opacity_tex}) {
            if (t == BI_PT_NO_TEXTURE) continue;
            need_uv = true;
            tex_refs = std::max(tex_refs, t);   // references are index + 1
        }
        need_tan = need_tan || M.normal_tex != BI_PT_NO_TEXTURE;
        non_opaque = non_opaque || M.blend_mode != BI_PT_BLEND_OPAQUE;
        if (M.pattern == BI_PT_PATTERN_CHECKER_XZ &&
            (M.pattern_material >= num_materials || materials[M.pattern_material].pattern != BI_PT_PATTERN_NONE))
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT,
                        "checkerboard pattern_material must name a material without a pattern");
    }
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
    // texture coordinates (default 0) of every vertex when a material samples a texture;
    // tangents (default (1, 0, 0, 1)) always: normal maps and the GBuffer read them
    std::vector<float> huv(need_uv ? 2 * nv : 0, 0.0f), htan;
    {
        htan.resize(4 * nv);
        for (size_t v = 0; v < nv; ++v) {
            htan[4 * v] = 1.0f; htan[4 * v + 1] = 0.0f; htan[4 * v + 2] = 0.0f; htan[4 * v + 3] = 1.0f;
        }
    }
    for (uint32_t m = 0; m < num_meshes; ++m) {
        const size_t nvm = meshes[m].num_vertices;
        if (need_uv && meshes[m].texcoords) {
            for (size_t k = 0; k < 2 * nvm; ++k)
                if (!std::isfinite(meshes[m].texcoords[k]))
                    return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite texture coordinate");
            std::memcpy(&huv[2 * (size_t)vtx_off[m]], meshes[m].texcoords, 8 * nvm);
        }
        if (meshes[m].tangents) {
            for (size_t k = 0; k < 4 * nvm; ++k)
                if (!std::isfinite(meshes[m].tangents[k]))
                    return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite tangent");
            std::memcpy(&htan[4 * (size_t)vtx_off[m]], meshes[m].tangents, 16 * nvm);
        }
    }
    if (need_uv) {
        BI_TRY(dalloc(ctx, ctx->uvs, std::max<size_t>(huv.size(), 1)));
        if (!huv.empty()) BI_HIP_TRY(ctx, hipMemcpy(ctx->uvs.p, huv.data(), huv.size() * 4, hipMemcpyHostToDevice));
    } else {
        dfree(ctx, ctx->uvs);
    }
    BI_TRY(dalloc(ctx, ctx->tans, std::max<size_t>(htan.size(), 1)));
    if (!htan.empty()) BI_HIP_TRY(ctx, hipMemcpy(ctx->tans.p, htan.data(), htan.size() * 4, hipMemcpyHostToDevice));
    ctx->need_uv = need_uv;
    ctx->need_tan = need_tan;
    ctx->non_opaque = non_opaque;
    ctx->tex_refs = tex_refs;
    BI_TRY(dalloc(ctx, ctx->pos, std::max<size_t>(hp.size(), 1)));
    BI_TRY(dalloc(ctx, ctx->nrm, std::max<size_t>(hn.size(), 1)));
    BI_TRY(dalloc(ctx, ctx->idx, std::max<size_t>(hi.size(), 1)));
    BI_TRY(dalloc(ctx, ctx->inst, std::max<size_t>(ih.size(), 1)));
    BI_TRY(dalloc(ctx, ctx->mats, std::max<size_t>(num_materials, 1)));
    if (!hp.empty()) BI_HIP_TRY(ctx, hipMemcpy(ctx->pos.p, hp.data(), hp.size() * 4, hipMemcpyHostToDevice));
    if (!hn.empty()) BI_HIP_TRY(ctx, hipMemcpy(ctx->nrm.p, hn.data(), hn.size() * 4, hipMemcpyHostToDevice));
    if (!hi.empty()) BI_HIP_TRY(ctx, hipMemcpy(ctx->idx.p, hi.data(), hi.size() * 4, hipMemcpyHostToDevice));
    if (!ih.empty()) BI_HIP_TRY(ctx, hipMemcpy(ctx->inst.p, ih.data(), ih.size() * sizeof(InstanceRec), hipMemcpyHostToDevice));
    if (num_materials)
        BI_HIP_TRY(ctx, hipMemcpy(ctx->mats.p, materials, num_materials * sizeof(MaterialRec), hipMemcpyHostToDevice));
    ctx->inst_host = ih;
    ctx->num_tris = (uint32_t)tris;
    ctx->num_inst = num_instances;
    ctx->num_mats = num_materials;
    ctx->has_scene = true;
    ctx->bvh_built = false;
    ctx->acc_count = 0;
    ctx->scene_version++;
    return BI_PT_OK;
}

bi_pt_status bi_pt_update_transforms(bi_pt_ctx* ctx, const float* xf, uint32_t num_instances) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    if (num_instances != ctx->num_inst || (num_instances && !xf))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "instance count differs from the scene");
    for (size_t k = 0; k < 12 * (size_t)num_instances; ++k)
        if (!std::isfinite(xf[k])) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite instance transform");
    for (uint32_t i = 0; i < num_instances; ++i) std::memcpy(ctx->inst_host[i].xform, xf + 12 * (size_t)i, 48);
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (num_instances)
        BI_HIP_TRY(ctx, hipMemcpy(ctx->inst.p, ctx->inst_host.data(), num_instances * sizeof(InstanceRec),
                          hipMemcpyHostToDevice));
    ctx->bvh_built = false;
    ctx->acc_count = 0;
    ctx->scene_version++;
    return BI_PT_OK;
}

namespace {
// sRGB byte -> linear (IEC 61966-2-1), evaluated in double and rounded once; shared by the
// material and rect-light textures.
double srgb_decode(int c) {
    const double x = c / 255.0;
    return x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4);
}
bi_pt_status upload_srgb_table(bi_pt_ctx* ctx) {
    if (ctx->srgb.p) return BI_PT_OK;
    float lut[256];
    for (int c = 0; c < 256; ++c) lut[c] = (float)srgb_decode(c);
    BI_TRY(dalloc(ctx, ctx->srgb, 256));
    BI_HIP_TRY(ctx, hipMemcpy(ctx->srgb.p, lut, sizeof lut, hipMemcpyHostToDevice));
    return BI_PT_OK;
}
uint32_t texel_words(uint32_t format) { return format == BI_PT_TEXEL_RGBA32F ? 4u : 1u; }
uint32_t max_mip_levels(uint32_t w, uint32_t h) {
    uint32_t n = 1, m = std::max(w, h);
    while (m > 1) { m >>= 1; ++n; }
    return n;
}
}  // namespace

bi_pt_status bi_pt_set_textures(bi_pt_ctx* ctx, const bi_pt_texture* textures, uint32_t num_textures) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (num_textures && !textures) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null textures");
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<TexRec> recs(num_textures);
    size_t words = 0;
    for (uint32_t k = 0; k < num_textures; ++k) {
        const bi_pt_texture& T = textures[k];
        if (!T.texels || T.width == 0 || T.height == 0 || T.width > 16384 || T.height > 16384 ||
            T.format > BI_PT_TEXEL_RGBA32F || T.filter > 1 || T.address_u > BI_PT_ADDRESS_CLAMP_TO_EDGE ||
            T.address_v > BI_PT_ADDRESS_CLAMP_TO_EDGE)
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "texture " + std::to_string(k) + ": bad description");
        words = (words + 3u) & ~(size_t)3u;   // RGBA32F texels are read as 16-B words
        recs[k] = TexRec{(uint32_t)words, T.width, T.height, T.format, T.filter, T.address_u, T.address_v, 0u};
        words += (size_t)T.width * T.height * (T.format == BI_PT_TEXEL_RGBA32F ? 4u : 1u);
        if (words >= (1ull << 32)) return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "textures larger than 16 GiB");
    }
    std::vector<uint32_t> pool(words, 0u);
    for (uint32_t k = 0; k < num_textures; ++k) {
        const bi_pt_texture& T = textures[k];
        const size_t n = (size_t)T.width * T.height * (T.format == BI_PT_TEXEL_RGBA32F ? 4u : 1u);
        if (T.format == BI_PT_TEXEL_RGBA32F) {
            const float* f = static_cast<const float*>(T.texels);
            for (size_t i = 0; i < n; ++i)
                if (!std::isfinite(f[i]))
                    return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "texture " + std::to_string(k) + ": non-finite texel");
        }
        std::memcpy(&pool[recs[k].offset], T.texels, 4 * n);
    }
    BI_TRY(dalloc(ctx, ctx->tex_recs, std::max<size_t>(recs.size(), 1)));
    BI_TRY(dalloc(ctx, ctx->tex_pool, std::max<size_t>(pool.size(), 4)));
    BI_TRY(upload_srgb_table(ctx));
    if (!recs.empty())
        BI_HIP_TRY(ctx, hipMemcpy(ctx->tex_recs.p, recs.data(), recs.size() * sizeof(TexRec), hipMemcpyHostToDevice));
    if (!pool.empty()) BI_HIP_TRY(ctx, hipMemcpy(ctx->tex_pool.p, pool.data(), pool.size() * 4, hipMemcpyHostToDevice));
    ctx->num_tex = num_textures;
    ctx->acc_count = 0;
    ctx->scene_version++;
    return BI_PT_OK;
}

namespace {

// The kernels' light records: the directional lights, then the point lights.  A change
// restarts the running image (unchanged lights, set every frame, keep it).
bi_pt_status commit_lights(bi_pt_ctx* ctx, std::vector<LightRec>& dir, std::vector<LightRec>& point) {
    if (dir.size() + point.size() > (size_t)kMaxDirLights)
        return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "at most 4 lights (directional and point together)");
    std::vector<LightRec> all(dir);
    all.insert(all.end(), point.begin(), point.end());
    const bool changed = all.size() != ctx->lights.size() ||
                         (!all.empty() && std::memcmp(all.data(), ctx->lights.data(), all.size() * sizeof(LightRec)) != 0);
    ctx->dir_lights.swap(dir);
    ctx->point_lights.swap(point);
    ctx->lights.swap(all);
    if (changed) {
        ctx->acc_count = 0;
        ctx->scene_version++;
    }
    return BI_PT_OK;
}

}  // namespace

bi_pt_status bi_pt_set_lights(bi_pt_ctx* ctx, const bi_pt_dir_light* lights, uint32_t num_lights) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (num_lights > (uint32_t)kMaxDirLights)
        return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "at most 4 directional lights");
    if (num_lights && !lights) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null lights");
    std::vector<LightRec> dir(num_lights);
    for (uint32_t l = 0; l < num_lights; ++l) {
        const bi_pt_dir_light& D = lights[l];
        LightRec& r = dir[l];
        r = LightRec{};
        std::memcpy(r.emission, D.emission, sizeof r.emission);
        r.sm_index = D.sm_index;
        std::memcpy(r.direction, D.direction, sizeof r.direction);
        r.shadow_strength = D.shadow_strength;
        r.type = kLightDirectional;
    }
    std::vector<LightRec> point(ctx->point_lights);
    return commit_lights(ctx, dir, point);
}

bi_pt_status bi_pt_set_point_lights(bi_pt_ctx* ctx, const bi_pt_point_light* lights, uint32_t num_lights) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (num_lights > (uint32_t)kMaxDirLights)
        return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "at most 4 lights (directional and point together)");
    if (num_lights && !lights) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null lights");
    std::vector<LightRec> point(num_lights);
    for (uint32_t l = 0; l < num_lights; ++l) {
        const bi_pt_point_light& P = lights[l];
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(P.position[k]) || !std::isfinite(P.emission[k]) || !std::isfinite(P.direction[k]))
                return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite point light");
        LightRec& r = point[l];
        r = LightRec{};
        std::memcpy(r.emission, P.emission, sizeof r.emission);
        r.sm_index = P.sm_index;
        std::memcpy(r.direction, P.direction, sizeof r.direction);
        r.shadow_strength = P.shadow_strength;
        std::memcpy(r.position, P.position, sizeof r.position);
        r.range_sqr_inv = P.range_sqr_inv;
        r.cos_inner = P.cos_inner;
        r.cos_outer = P.cos_outer;
        r.type = kLightPoint;
    }
    std::vector<LightRec> dir(ctx->dir_lights);
    return commit_lights(ctx, dir, point);
}

bi_pt_status bi_pt_set_rect_lights(bi_pt_ctx* ctx, const bi_pt_rect_light* lights, uint32_t num_lights) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (num_lights > kMaxRectLights) return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "at most 16 rect lights");
    if (num_lights && !lights) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null rect lights");
    static_assert(sizeof(bi_pt_rect_light) == sizeof(RectRec), "bi_pt_rect_light is the kernels' record");
    std::vector<RectRec> recs(num_lights);
    for (uint32_t l = 0; l < num_lights; ++l) {
        std::memcpy(&recs[l], &lights[l], sizeof(RectRec));
        const RectRec& r = recs[l];
        if (r.texture_index >= (int32_t)kMaxRectTextures)
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "rect light texture_index beyond 16 textures");
        if (r.texture_index >= 0 && !(std::isfinite(r.inv_texel_size) && r.inv_texel_size >= 0.0f &&
                                      std::isfinite(r.inv_width_sqr) && std::isfinite(r.inv_height_sqr)))
            return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "textured rect light with non-finite texel or size terms");
        const float* f[] = {r.emission, r.p0, r.p1, r.p2, r.p3};
        for (const float* v : f)
            for (int k = 0; k < 3; ++k)
                if (!std::isfinite(v[k])) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite rect light");
    }
    const bool changed = recs.size() != ctx->rects_host.size() ||
                         (!recs.empty() && std::memcmp(recs.data(), ctx->rects_host.data(), recs.size() * sizeof(RectRec)) != 0);
    if (!changed) return BI_PT_OK;
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (num_lights) {
        BI_TRY(dalloc(ctx, ctx->rects, num_lights));
        BI_HIP_TRY(ctx, hipMemcpy(ctx->rects.p, recs.data(), num_lights * sizeof(RectRec), hipMemcpyHostToDevice));
    }
    ctx->rects_host.swap(recs);
    ctx->acc_count = 0;
    ctx->scene_version++;
    return BI_PT_OK;
}

bi_pt_status bi_pt_set_rect_light_textures(bi_pt_ctx* ctx, const bi_pt_rect_texture* textures, uint32_t num_textures) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (num_textures > kMaxRectTextures) return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "at most 16 rect-light textures");
    if (num_textures && !textures) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null rect-light textures");
    std::vector<RectTexRec> recs(num_textures);
    std::vector<TexRec> levels;
    size_t words = 0;
    for (uint32_t k = 0; k < num_textures; ++k) {
        const bi_pt_rect_texture& T = textures[k];
        const bool ok = T.texels && T.width > 0 && T.height > 0 && T.width <= 16384 && T.height <= 16384 &&
                        T.num_levels >= 1 && T.num_levels <= max_mip_levels(T.width, T.height) &&
                        T.format <= BI_PT_TEXEL_RGBA32F && T.mag_filter <= 1 && T.min_filter <= 1 &&
                        T.mipmap_mode <= 1 && T.address_u <= BI_PT_ADDRESS_CLAMP_TO_EDGE &&
                        T.address_v <= BI_PT_ADDRESS_CLAMP_TO_EDGE && std::isfinite(T.lod_bias) &&
                        std::isfinite(T.lod_min) && std::isfinite(T.lod_max) && T.lod_min <= T.lod_max;
        if (!ok) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "rect-light texture " + std::to_string(k) + ": bad description");
        recs[k] = RectTexRec{(uint32_t)levels.size(), T.num_levels, T.mag_filter, T.min_filter, T.mipmap_mode,
                             T.lod_bias, T.lod_min, T.lod_max};
        uint32_t w = T.width, h = T.height;
        for (uint32_t l = 0; l < T.num_levels; ++l) {
            words = (words + 3u) & ~(size_t)3u;   // RGBA32F texels are read as 16-B words
            levels.push_back(TexRec{(uint32_t)words, w, h, T.format, 0u, T.address_u, T.address_v, 0u});
            words += (size_t)w * h * texel_words(T.format);
            if (words >= (1ull << 32)) return fail(ctx, BI_PT_ERROR_UNSUPPORTED, "rect-light textures larger than 16 GiB");
            w = std::max(w / 2u, 1u);
            h = std::max(h / 2u, 1u);
        }
    }
    std::vector<uint32_t> pool(words, 0u);
    for (uint32_t k = 0; k < num_textures; ++k) {
        const bi_pt_rect_texture& T = textures[k];
        const uint32_t* src = static_cast<const uint32_t*>(T.texels);
        for (uint32_t l = 0; l < T.num_levels; ++l) {
            const TexRec& L = levels[recs[k].first_level + l];
            const size_t n = (size_t)L.width * L.height * texel_words(T.format);
            if (T.format == BI_PT_TEXEL_RGBA32F)
                for (size_t i = 0; i < n; ++i) {
                    float f;
                    std::memcpy(&f, &src[i], 4);
                    if (!std::isfinite(f))
                        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT,
                                    "rect-light texture " + std::to_string(k) + ": non-finite texel");
                }
            std::memcpy(&pool[L.offset], src, 4 * n);
            src += n;
        }
    }
    auto same_bytes = [](const auto& a, const auto& b) {
        return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(a[0])) == 0);
    };
    if (same_bytes(recs, ctx->rect_tex_host) && same_bytes(levels, ctx->rect_tex_levels_host) &&
        same_bytes(pool, ctx->rect_tex_pool_host))
        return BI_PT_OK;   // unchanged (set every frame): the running image continues
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    BI_TRY(dalloc(ctx, ctx->rect_tex, std::max<size_t>(recs.size(), 1)));
    BI_TRY(dalloc(ctx, ctx->rect_tex_levels, std::max<size_t>(levels.size(), 1)));
    BI_TRY(dalloc(ctx, ctx->rect_tex_pool, std::max<size_t>(pool.size(), 4)));
    BI_TRY(upload_srgb_table(ctx));
    if (!recs.empty()) {
        BI_HIP_TRY(ctx, hipMemcpy(ctx->rect_tex.p, recs.data(), recs.size() * sizeof(RectTexRec), hipMemcpyHostToDevice));
        BI_HIP_TRY(ctx, hipMemcpy(ctx->rect_tex_levels.p, levels.data(), levels.size() * sizeof(TexRec),
                                  hipMemcpyHostToDevice));
        BI_HIP_TRY(ctx, hipMemcpy(ctx->rect_tex_pool.p, pool.data(), pool.size() * 4, hipMemcpyHostToDevice));
    }
    ctx->num_rect_tex = num_textures;
    ctx->rect_tex_host.swap(recs);
    ctx->rect_tex_levels_host.swap(levels);
    ctx->rect_tex_pool_host.swap(pool);
    ctx->acc_count = 0;
    ctx->scene_version++;
    return BI_PT_OK;
}

bi_pt_status bi_pt_rect_texture_mips(uint32_t format, uint32_t width, uint32_t height, uint32_t num_levels,
                                     const void* level0, void* chain) {
    if (!level0 || !chain || format > BI_PT_TEXEL_RGBA32F || width == 0 || height == 0 || width > 16384 ||
        height > 16384 || num_levels == 0 || num_levels > max_mip_levels(width, height))
        return BI_PT_ERROR_INVALID_ARGUMENT;
    // texels as floats (RGBA8 decoded: c / 255, sRGB colour through the curve), averaged in
    // FP32 in the shader's order, stored back (RGBA8: round to nearest; sRGB colour encoded)
    auto decode = [format](const uint32_t* p, size_t i, float v[4]) {
        if (format == BI_PT_TEXEL_RGBA32F) { std::memcpy(v, &p[4 * i], 16); return; }
        const uint32_t w = p[i];
        for (int c = 0; c < 4; ++c) {
            const int b = (int)((w >> (8 * c)) & 255u);
            v[c] = (format == BI_PT_TEXEL_RGBA8_SRGB && c < 3) ? (float)srgb_decode(b) : (float)b / 255.0f;
        }
    };
    auto encode = [format](float v[4], uint32_t* p, size_t i) {
        if (format == BI_PT_TEXEL_RGBA32F) { std::memcpy(&p[4 * i], v, 16); return; }
        uint32_t w = 0;
        for (int c = 0; c < 4; ++c) {
            double x = std::min(std::max((double)v[c], 0.0), 1.0);
            if (format == BI_PT_TEXEL_RGBA8_SRGB && c < 3)
                x = x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
            w |= (uint32_t)std::lrint(x * 255.0) << (8 * c);
        }
        p[i] = w;
    };
    const size_t tw = texel_words(format);
    const uint32_t* src = static_cast<const uint32_t*>(level0);
    uint32_t* dst = static_cast<uint32_t*>(chain);
    std::memmove(dst, src, (size_t)width * height * tw * 4);
    uint32_t w = width, h = height;
    for (uint32_t l = 1; l < num_levels; ++l) {
        const uint32_t* s = dst;
        uint32_t* d = dst + (size_t)w * h * tw;
        const uint32_t dw = std::max(w / 2u, 1u), dh = std::max(h / 2u, 1u);
        const bool odd_x = dw & 1u, odd_y = dh & 1u;   // of the destination, as mipmap.hlsl tests it
        for (uint32_t y = 0; y < dh; ++y)
            for (uint32_t x = 0; x < dw; ++x) {
                // reads beyond the source level give zero (robust image access)
                auto at = [&](uint32_t i, uint32_t j, float v[4]) {
                    if (i >= w || j >= h) { v[0] = v[1] = v[2] = v[3] = 0.0f; return; }
                    decode(s, (size_t)j * w + i, v);
                };
                float a[4], b[4], c[4], e[4], r[4];
                const uint32_t sx = 2u * x, sy = 2u * y;
                at(sx, sy, a); at(sx, sy + 1u, b); at(sx + 1u, sy, c); at(sx + 1u, sy + 1u, e);
                for (int k = 0; k < 4; ++k) r[k] = (a[k] + b[k]) + (c[k] + e[k]);
                float n = 4.0f;
                if (odd_x) {
                    at(sx + 2u, sy, a); at(sx + 2u, sy + 1u, b);
                    for (int k = 0; k < 4; ++k) r[k] = r[k] + (a[k] + b[k]);
                    n = n + 2.0f;
                }
                if (odd_y) {
                    at(sx, sy + 2u, a); at(sx + 1u, sy + 2u, b);
                    for (int k = 0; k < 4; ++k) r[k] = r[k] + (a[k] + b[k]);
                    n = n + 2.0f;
                }
                if (odd_x && odd_y) {
                    at(sx + 2u, sy + 2u, a);
                    for (int k = 0; k < 4; ++k) r[k] = r[k] + a[k];
                    n = n + 1.0f;
                }
                for (int k = 0; k < 4; ++k) r[k] = r[k] / n;
                encode(r, d, (size_t)y * dw + x);
            }
        dst = d;
        w = dw;
        h = dh;
    }
    return BI_PT_OK;
}

bi_pt_status bi_pt_set_ltc_luts(bi_pt_ctx* ctx, const float* matrix_lut0, const float* matrix_lut1,
                                const float* matrix_lut2, const float* norm_lut) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (!matrix_lut0 || !matrix_lut1 || !matrix_lut2 || !norm_lut)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null LTC table");
    // four float4 tables: the matrix rows as given (RGBA32F), the norm table's RG32F pairs padded
    std::vector<float> t((size_t)4 * kLtcTexels * 4, 0.0f);
    const float* m[3] = {matrix_lut0, matrix_lut1, matrix_lut2};
    for (int k = 0; k < 3; ++k) std::memcpy(&t[(size_t)k * kLtcTexels * 4], m[k], (size_t)kLtcTexels * 4 * sizeof(float));
    for (uint32_t i = 0; i < kLtcTexels; ++i) {
        t[((size_t)3 * kLtcTexels + i) * 4] = norm_lut[2 * i];
        t[((size_t)3 * kLtcTexels + i) * 4 + 1] = norm_lut[2 * i + 1];
    }
    for (float v : t)
        if (!std::isfinite(v)) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "non-finite LTC table entry");
    if (t == ctx->ltc_host) return BI_PT_OK;
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    BI_TRY(dalloc(ctx, ctx->ltc, (size_t)4 * kLtcTexels));
    BI_HIP_TRY(ctx, hipMemcpy(ctx->ltc.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->ltc_host.swap(t);
    if (!ctx->rects_host.empty()) {
        ctx->acc_count = 0;
        ctx->scene_version++;
    }
    return BI_PT_OK;
}

namespace {

// The sky of a context: texels (equirect w x h, or cube faces w x w with h = 6 w) and
// their importance table.
bi_pt_status set_sky(bi_pt_ctx* ctx, const float* rgba, uint32_t w, uint32_t h, uint32_t layout,
                     const float color[3], const float rot[9]) {
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t ntex = (size_t)w * h * 4;
    if (ctx->sky_set && ctx->sky.width == w && ctx->sky.height == h && ctx->sky.layout == layout &&
        std::memcmp(ctx->sky.color, color, sizeof ctx->sky.color) == 0 &&
        std::memcmp(ctx->sky.rot, rot, sizeof ctx->sky.rot) == 0 && ctx->sky_host.size() == ntex &&
        (ntex == 0 || std::memcmp(ctx->sky_host.data(), rgba, ntex * sizeof(float)) == 0))
        return BI_PT_OK;   // unchanged sky (set every frame): keep tables and the running image
    ctx->sky_host.assign(rgba, rgba + ntex);
    ctx->sky_set = true;
    SkyParams sp{};
    sp.width = w;
    sp.height = h;
    sp.layout = layout;
    std::memcpy(sp.color, color, sizeof sp.color);
    std::memcpy(sp.rot, rot, sizeof sp.rot);
    if (w) {
        std::vector<float> wt, rows;
        float total = 0.0f;
        build_sky_table(rgba, w, h, layout, color, wt, rows, total);
        // alias tables: rows by their sums, each row's texels by their weights
        std::vector<SkyRowAlias> ra(h);
        std::vector<SkyTexAlias> ta((size_t)w * h);
        {
            std::vector<float> prob(std::max(w, h));
            std::vector<uint32_t> al(std::max(w, h));
            build_alias(rows.data(), h, total, prob.data(), al.data());
            for (uint32_t i = 0; i < h; ++i) ra[i] = SkyRowAlias{prob[i], al[i]};
            for (uint32_t i = 0; i < h; ++i) {
                const float* wr = &wt[(size_t)i * w];
                build_alias(wr, w, rows[i], prob.data(), al.data());
                for (uint32_t j = 0; j < w; ++j) ta[(size_t)i * w + j] = SkyTexAlias{prob[j], al[j], wr[j], wr[al[j]]};
            }
        }
        BI_TRY(dalloc(ctx, ctx->sky_row_alias, ra.size()));
        BI_TRY(dalloc(ctx, ctx->sky_tex_alias, ta.size()));
        BI_HIP_TRY(ctx, hipMemcpy(ctx->sky_row_alias.p, ra.data(), ra.size() * sizeof(SkyRowAlias), hipMemcpyHostToDevice));
        BI_HIP_TRY(ctx, hipMemcpy(ctx->sky_tex_alias.p, ta.data(), ta.size() * sizeof(SkyTexAlias), hipMemcpyHostToDevice));
        BI_TRY(dalloc(ctx, ctx->sky_tex, (size_t)w * h));
        BI_TRY(dalloc(ctx, ctx->sky_w, wt.size()));
        BI_HIP_TRY(ctx, hipMemcpy(ctx->sky_tex.p, rgba, (size_t)w * h * 16, hipMemcpyHostToDevice));
        BI_HIP_TRY(ctx, hipMemcpy(ctx->sky_w.p, wt.data(), wt.size() * 4, hipMemcpyHostToDevice));
        sp.texels = ctx->sky_tex.p;
        sp.weight = ctx->sky_w.p;
        sp.row_alias = ctx->sky_row_alias.p;
        sp.tex_alias = ctx->sky_tex_alias.p;
        sp.total = total;
        if (total > 0.0f)
            sp.pdf_scale = layout == kSkyCube ? ((float)w * (float)w / total) * 0.25f : (float)(w * h) / total;
    }
    ctx->sky = sp;
    ctx->acc_count = 0;
    ctx->scene_version++;
    return BI_PT_OK;
}

}  // namespace

bi_pt_status bi_pt_set_sky(bi_pt_ctx* ctx, const float* rgba, uint32_t w, uint32_t h, const float color[3],
                           const float rot[9]) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if ((w == 0) != (h == 0) || ((w || h) && !rgba) || !color || !rot)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "bad sky arguments");
    return set_sky(ctx, rgba, w, h, kSkyEquirect, color, rot);
}

bi_pt_status bi_pt_set_sky_cube(bi_pt_ctx* ctx, const float* rgba_faces, uint32_t face_size, const float color[3],
                                const float rot[9]) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if ((face_size && !rgba_faces) || !color || !rot)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "bad sky arguments");
    if (face_size > (1u << 13))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "cube face larger than 8192 texels");
    return set_sky(ctx, rgba_faces, face_size, 6u * face_size, face_size ? kSkyCube : kSkyEquirect, color, rot);
}

bi_pt_status bi_pt_create_view(bi_pt_ctx* scene, bi_pt_ctx** out_view) {
    if (!scene || !out_view) return BI_PT_ERROR_INVALID_ARGUMENT;
    *out_view = nullptr;
    if (scene->scene_src) return fail(scene, BI_PT_ERROR_INVALID_ARGUMENT, "a view of a view");
    bi_pt_device_desc d{};
    d.device = scene->device;
    d.hbm_budget_bytes = scene->budget;
    bi_pt_ctx* v = nullptr;
    BI_TRY(bi_pt_create(&d, &v));
    v->scene_src = scene;
    v->has_scene = true;
    *out_view = v;
    return BI_PT_OK;
}

bi_pt_status bi_pt_set_build_options(bi_pt_ctx* ctx, const bi_pt_build_options* opt) {
    if (!ctx || !opt) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (ctx->scene_src) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "a view renders another context's scene");
    if (opt->treelet_rounds > 16) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "at most 16 treelet rounds");
    if (opt->treelet_rounds != ctx->treelet_rounds) {
        ctx->treelet_rounds = opt->treelet_rounds;
        ctx->bvh_built = false;   // the next render builds again (its images do not change)
    }
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
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    JobIn job{ctx, camera, st, out_rgba, out_depth};
    RenderOpts opt;
    opt.knobs = Knobs{st->wave_paths, st->stack_lds, st->refill_lanes, st->work_parts, st->block_threads};
    opt.timing = (st->flags & BI_PT_FLAG_TIMING) != 0;
    opt.record = (st->flags & BI_PT_FLAG_RECORD_QUEUES) != 0;
    opt.reuse_bvh = (st->flags & BI_PT_FLAG_REUSE_BVH) != 0;
    return render_jobs(ctx, ctx->eng, &job, 1, reinterpret_cast<hipStream_t>(stream), opt, nullptr);
}

bi_pt_status bi_pt_drawables_layout(const bi_pt_gpu_scene* gs, const bi_pt_drawable* d, uint32_t n,
                                    bi_pt_submesh_range* ranges, uint32_t* num_ranges, bi_pt_instance* inst) {
    if (!gs || (n && (!d || !ranges || !num_ranges || !inst)) || !gs->positions || !gs->indices)
        return BI_PT_ERROR_INVALID_ARGUMENT;
    constexpr uint64_t kAbsent = ~0ull;
    uint32_t nr = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const bi_pt_drawable& D = d[k];
        if (!(D.attributes & BI_PT_VA_POSITION) || D.index_count % 3 != 0) return BI_PT_ERROR_INVALID_ARGUMENT;
        if ((uint64_t)D.index_offset + D.index_count > gs->num_indices) return BI_PT_ERROR_INVALID_ARGUMENT;
        bi_pt_submesh_range r{};
        r.position_offset = D.position_offset;
        r.normal_offset = (D.attributes & BI_PT_VA_NORMAL) ? D.normal_offset : kAbsent;
        r.tangent_offset = (D.attributes & BI_PT_VA_TANGENT) ? D.tangent_offset : kAbsent;
        r.texcoord_offset = (D.attributes & BI_PT_VA_TEXCOORD) ? D.texcoord_offset : kAbsent;
        r.index_offset = D.index_offset;
        r.index_count = D.index_count;
        uint32_t nv = 0;
        for (uint32_t i = 0; i < D.index_count; ++i) nv = std::max(nv, gs->indices[D.index_offset + i] + 1u);
        if (D.index_count && nv == 0) return BI_PT_ERROR_INVALID_ARGUMENT;   // index 0xffffffff
        r.num_vertices = nv;
        auto fits = [&](uint64_t off, uint32_t width, const float* buf, uint64_t len) {
            return off == kAbsent || (buf && off + (uint64_t)width * nv <= len);
        };
        if (!fits(r.position_offset, 3, gs->positions, gs->num_positions) ||
            !fits(r.normal_offset, 3, gs->normals, gs->num_normals) ||
            !fits(r.tangent_offset, 4, gs->tangents, gs->num_tangents) ||
            !fits(r.texcoord_offset, 2, gs->texcoords, gs->num_texcoords))
            return BI_PT_ERROR_INVALID_ARGUMENT;
        uint32_t m = 0;
        while (m < nr && std::memcmp(&ranges[m], &r, sizeof r) != 0) ++m;   // drawables that share a submesh
        if (m == nr) ranges[nr++] = r;
        // the TLAS instance's 3x4 rows are the matrix's rows: element [c][r] of glm's
        // column-major storage goes to row r, column c (accel.cpp:119-131)
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col < 4; ++col) inst[k].transform[4 * row + col] = D.matrix[4 * col + row];
        inst[k].mesh = m;
        inst[k].material = D.material;
    }
    *num_ranges = nr;
    return BI_PT_OK;
}

bi_pt_status bi_pt_set_scene_drawables(bi_pt_ctx* ctx, const bi_pt_gpu_scene* gs, const bi_pt_drawable* d,
                                       uint32_t n, const bi_pt_material* materials, uint32_t num_materials) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    std::vector<bi_pt_submesh_range> ranges(n);
    std::vector<bi_pt_instance> inst(n);
    uint32_t nr = 0;
    if (bi_pt_drawables_layout(gs, d, n, ranges.data(), &nr, inst.data()) != BI_PT_OK)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT,
                    "drawable records do not fit the GPU-scene buffers (offsets, index counts or attributes)");
    // the meshes point into the flat buffers; missing normals read (0, 0, 1) (hit.hlsl:54)
    std::vector<bi_pt_mesh> meshes(nr);
    std::vector<std::vector<float>> flat_normals;
    for (uint32_t m = 0; m < nr; ++m) {
        const bi_pt_submesh_range& r = ranges[m];
        bi_pt_mesh& M = meshes[m];
        M.positions = gs->positions + r.position_offset;
        if (r.normal_offset != ~0ull) {
            M.normals = gs->normals + r.normal_offset;
        } else {
            flat_normals.emplace_back((size_t)3 * r.num_vertices, 0.0f);
            for (uint32_t v = 0; v < r.num_vertices; ++v) flat_normals.back()[3 * (size_t)v + 2] = 1.0f;
            M.normals = flat_normals.back().data();
        }
        M.indices = gs->indices + r.index_offset;
        M.num_vertices = r.num_vertices;
        M.num_indices = r.index_count;
        M.texcoords = r.texcoord_offset != ~0ull ? gs->texcoords + r.texcoord_offset : nullptr;
        M.tangents = r.tangent_offset != ~0ull ? gs->tangents + r.tangent_offset : nullptr;
    }
    return bi_pt_set_scene(ctx, meshes.data(), nr, inst.data(), n, materials, num_materials);
}

bi_pt_status bi_pt_render_gbuffer(bi_pt_ctx* ctx, const bi_pt_camera* camera, const bi_pt_settings* st,
                                  float* out_rgba, float* out_depth, const bi_pt_gbuffer* gbuffer, bi_pt_stream stream) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!camera || !st || !out_rgba || !gbuffer)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null camera/settings/output/gbuffer");
    if (st->flags & BI_PT_FLAG_RECORD_QUEUES)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "queue recording and the GBuffer are not combined");
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    JobIn job{ctx, camera, st, out_rgba, out_depth};
    const bool any = gbuffer->base_color || gbuffer->normal_roughness || gbuffer->fresnel || gbuffer->material_0;
    job.gbuf = any ? gbuffer : nullptr;
    RenderOpts opt;
    opt.knobs = Knobs{st->wave_paths, st->stack_lds, st->refill_lanes, st->work_parts, st->block_threads};
    opt.timing = (st->flags & BI_PT_FLAG_TIMING) != 0;
    opt.reuse_bvh = (st->flags & BI_PT_FLAG_REUSE_BVH) != 0;
    return render_jobs(ctx, ctx->eng, &job, 1, reinterpret_cast<hipStream_t>(stream), opt, nullptr);
}

bi_pt_status bi_pt_get_stats(const bi_pt_ctx* cctx, bi_pt_stats* out) {
    bi_pt_ctx* ctx = const_cast<bi_pt_ctx*>(cctx);
    if (!ctx || !out) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const bi_pt_status st = engine_check(ctx, ctx->eng, out);
    out->samples_accumulated = ctx->acc_count;
    return st;
}

bi_pt_status bi_pt_debug_inject_error(bi_pt_ctx* ctx, uint32_t flags) {
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return engine_inject_error(ctx, ctx->eng, flags);
}

bi_pt_status bi_pt_debug_get_bvh(bi_pt_ctx* ctx, uint32_t* morton, uint32_t* order, int32_t* children,
                                 float* child_bounds, int32_t* parents, float* centroid_bounds) {
    if (ctx && ctx->scene_src) ctx = ctx->scene_src;   // a view is inspected through its scene
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!ctx->bvh_built || ctx->built_rounds != ctx->treelet_rounds) {   // the tree as configured
        BI_TRY(build_bvh_now(ctx));
    }
    BI_HIP_TRY(ctx, hipDeviceSynchronize());
    const uint32_t n = ctx->num_tris;
    if (n == 0) return BI_PT_OK;
    const size_t ni = n > 1 ? n - 1 : 1;
    // keys/values: the 4-pass sort leaves the sorted pair in buffer 0; the codes in
    // input order are recomputed by the Morton kernel into buffer... buffer 0 holds sorted
    // keys after the build, so input-order codes are gathered through the permutation.
    std::vector<uint32_t> sk(n), sv(n);
    BI_HIP_TRY(ctx, hipMemcpy(sk.data(), ctx->keys0.p, n * 4, hipMemcpyDeviceToHost));
    BI_HIP_TRY(ctx, hipMemcpy(sv.data(), ctx->vals0.p, n * 4, hipMemcpyDeviceToHost));
    if (order) std::memcpy(order, sv.data(), n * 4);
    if (morton)
        for (uint32_t k = 0; k < n; ++k) morton[sv[k]] = sk[k];
    std::vector<BvhNode> nodes(ni);
    BI_HIP_TRY(ctx, hipMemcpy(nodes.data(), ctx->nodes.p, ni * sizeof(BvhNode), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ni; ++i) {
        if (children) { children[2 * i] = nodes[i].child[0]; children[2 * i + 1] = nodes[i].child[1]; }
        if (child_bounds) std::memcpy(&child_bounds[12 * i], nodes[i].box, 48);
    }
    if (parents) BI_HIP_TRY(ctx, hipMemcpy(parents, ctx->parent.p, (ni + n) * 4, hipMemcpyDeviceToHost));
    if (centroid_bounds) BI_HIP_TRY(ctx, hipMemcpy(centroid_bounds, ctx->cb_out.p, 24, hipMemcpyDeviceToHost));
    return BI_PT_OK;
}

bi_pt_status bi_pt_debug_get_bvh4(bi_pt_ctx* ctx, int32_t* child4, float* box4) {
    if (ctx && ctx->scene_src) ctx = ctx->scene_src;   // a view is inspected through its scene
    if (!ctx) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!ctx->bvh_built || ctx->built_rounds != ctx->treelet_rounds) BI_TRY(build_bvh_now(ctx));
    BI_HIP_TRY(ctx, hipDeviceSynchronize());
    const uint32_t n = ctx->num_tris;
    if (n == 0) return BI_PT_OK;
    const size_t ni = n > 1 ? n - 1 : 1;
    std::vector<Bvh4Node> nodes(ni);
    BI_HIP_TRY(ctx, hipMemcpy(nodes.data(), ctx->nodes4.p, ni * sizeof(Bvh4Node), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ni; ++i) {
        if (child4) std::memcpy(&child4[4 * i], nodes[i].child, 16);
        if (box4) std::memcpy(&box4[24 * i], nodes[i].lox, 96);
    }
    return BI_PT_OK;
}

uint32_t bi_pt_debug_num_tris(const bi_pt_ctx* ctx) {
    if (ctx && ctx->scene_src) ctx = ctx->scene_src;
    return ctx ? ctx->num_tris : 0;
}

bi_pt_status bi_pt_debug_queue(bi_pt_ctx* ctx, uint32_t wave, uint32_t bounce, uint32_t kind, uint32_t* out,
                               uint32_t capacity, uint32_t* count) {
    if (!ctx || !count) return BI_PT_ERROR_INVALID_ARGUMENT;
    auto& rec = kind == 0 ? ctx->eng.rec_ext : ctx->eng.rec_sh;
    if (wave >= rec.size() || bounce >= rec[wave].size())
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "no recorded queue for this wave/bounce");
    const auto& v = rec[wave][bounce];
    *count = (uint32_t)v.size();
    if (out) std::memcpy(out, v.data(), std::min<size_t>(capacity, v.size()) * 4);
    return BI_PT_OK;
}

bi_pt_status bi_pt_debug_get_sums(bi_pt_ctx* ctx, float* sums) {
    if (!ctx || !sums) return BI_PT_ERROR_INVALID_ARGUMENT;
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    BI_HIP_TRY(ctx, hipDeviceSynchronize());
    if (!ctx->sums.p) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "nothing rendered yet");
    BI_HIP_TRY(ctx, hipMemcpy(sums, ctx->sums.p, (size_t)12 * ctx->acc_w * ctx->acc_h, hipMemcpyDeviceToHost));
    return BI_PT_OK;
}

bi_pt_status bi_pt_debug_trace(bi_pt_ctx* ctx, uint32_t n, const float* rays, int any_hit, float* hit_t,
                               uint32_t* hit_leaf, float* hit_u, float* hit_v) {
    if (ctx && ctx->scene_src) ctx = ctx->scene_src;   // a view is inspected through its scene
    if (!ctx || (n && (!rays || !hit_leaf))) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    if (n == 0) return BI_PT_OK;
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!ctx->bvh_built) BI_TRY(build_bvh_now(ctx));
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
    for (int k = 0; k < 6; ++k) BI_TRY(dalloc(ctx, f[k], n));
    BI_TRY(dalloc(ctx, hrec, n));
    BI_TRY(dalloc(ctx, path, n));
    BI_TRY(dalloc(ctx, cnt, 4 + kWorkParts * kWorkStride));
    BI_TRY(dalloc(ctx, out, n));
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
            launch_debug_anyhit(bvh_view(ctx), q, n, tmax, out.p, &ctx->eng.counters.p[kCntError], nullptr);
            if (hipMemcpy(hit_leaf, out.p, n * 4, hipMemcpyDeviceToHost) != hipSuccess) status = BI_PT_ERROR_HIP;
        } else {
            // one job: this scene, extent tmax; path ids are only used to find the job
            JobRec rec{};
            rec.nodes = ctx->num_tris ? ctx->nodes4.p : ctx->eng.empty_tree.p;
            rec.leaf_geo = ctx->leaf_geo.p;
            rec.leaf_shade = ctx->leaf_shade.p;
            rec.materials = ctx->mats.p;
            rec.leaf_uv = scene_leaf_uv(ctx);
            rec.tex = scene_tex(ctx);
            rec.num_tris = ctx->num_tris;
            rec.ray_length = tmax;
            const uint32_t bases[4] = {0u, n, 0u, 0u};
            DevBuf<unsigned char> jb;
            BI_TRY(dalloc(ctx, jb, sizeof(JobRec) + sizeof bases));
            if (hipMemcpy(jb.p, &rec, sizeof rec, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(jb.p + sizeof(JobRec), bases, sizeof bases, hipMemcpyHostToDevice) != hipSuccess)
                status = BI_PT_ERROR_HIP;
            TraceArgs ta{};
            ta.jobs = WaveJobs{reinterpret_cast<const JobRec*>(jb.p),
                               reinterpret_cast<const uint32_t*>(jb.p + sizeof(JobRec)),
                               reinterpret_cast<const uint32_t*>(jb.p + sizeof(JobRec)) + 2, 1u, n};
            ta.q = q;
            ta.count_in = cnt.p;
            ta.work = cnt.p + 4;   // zeroed part counters
            ta.parts = kWorkParts;
            ta.refill = 32;
            ta.any_hit = ctx->non_opaque ? 1u : 0u;
            ta.hits = HitBuf{hrec.p};
            ta.error = &ctx->eng.counters.p[kCntError];
            if (status == BI_PT_OK) launch_trace_closest(ta, n, false, 16, 256, nullptr);
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
            dfree(ctx, jb);
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

bi_pt_status bi_pt_trace_closest(bi_pt_ctx* ctx, uint32_t n, const bi_pt_rays* rays, float tmax, float* hit_t,
                                 uint32_t* hit_triangle, float* hit_u, float* hit_v, bi_pt_stream stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    BI_TRY(query_prepare(ctx, n, rays, tmax, s));
    if (n == 0) return BI_PT_OK;
    if (!hit_triangle) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null hit_triangle");
    BI_TRY(dalloc(ctx, ctx->rq_job, sizeof(JobRec) + 4 * sizeof(uint32_t)));
    BI_TRY(dalloc(ctx, ctx->rq_cnt, 4 + kWorkParts * kWorkStride));
    BI_TRY(dalloc(ctx, ctx->rq_path, n));
    BI_TRY(dalloc(ctx, ctx->rq_hits, n));
    JobRec rec{};
    rec.nodes = ctx->num_tris ? ctx->nodes4.p : ctx->eng.empty_tree.p;
    rec.leaf_geo = ctx->leaf_geo.p;
    rec.leaf_shade = ctx->leaf_shade.p;
    rec.materials = ctx->mats.p;
    rec.leaf_uv = scene_leaf_uv(ctx);
    rec.tex = scene_tex(ctx);
    rec.num_tris = ctx->num_tris;
    rec.ray_length = tmax;
    hipLaunchKernelGGL(k_put_query_job, dim3(1), dim3(64), 0, s, rec, n, ctx->rq_job.p);
    launch_fill_u32(ctx->rq_cnt.p, 0u, 4 + kWorkParts * kWorkStride, s);   // count, then zeroed part counters
    launch_fill_u32(ctx->rq_cnt.p, n, 1, s);
    launch_fill_u32(ctx->rq_path.p, 0u, n, s);
    TraceArgs ta{};
    ta.jobs = WaveJobs{reinterpret_cast<const JobRec*>(ctx->rq_job.p),
                       reinterpret_cast<const uint32_t*>(ctx->rq_job.p + sizeof(JobRec)),
                       reinterpret_cast<const uint32_t*>(ctx->rq_job.p + sizeof(JobRec)) + 2, 1u, n};
    ta.q = RayQueue{const_cast<float*>(rays->ox), const_cast<float*>(rays->oy), const_cast<float*>(rays->oz),
                    const_cast<float*>(rays->dx), const_cast<float*>(rays->dy), const_cast<float*>(rays->dz),
                    ctx->rq_path.p};
    ta.count_in = ctx->rq_cnt.p;
    ta.work = ctx->rq_cnt.p + 4;
    ta.parts = kWorkParts;
    ta.refill = 32;
    ta.any_hit = ctx->non_opaque ? 1u : 0u;
    ta.hits = HitBuf{ctx->rq_hits.p};
    ta.error = &ctx->eng.counters.p[kCntError];
    launch_trace_closest(ta, n, false, 16, 256, s);
    hipLaunchKernelGGL(k_query_hits, dim3((n + 255u) / 256u), dim3(256), 0, s, ctx->rq_hits.p,
                       ctx->num_tris ? ctx->vals0.p : ctx->rq_path.p, n, hit_t, hit_triangle, hit_u, hit_v);
    BI_HIP_TRY(ctx, hipGetLastError());
    return BI_PT_OK;
}

}  // extern "C"

// ---- closest-hit queries without device allocation ------------------------------
// The scratch of bi_pt_trace_closest (the job record and its path bases, the ray count
// and the persistent lanes' part counters, one hit record per ray) lives in this static
// device array instead of context buffers.  Each GPU holds its own copy for the life of
// the process; every call on a GPU uses that GPU's copy, so calls that overlap on one GPU
// race (callers serialise them, e.g. on one stream).
namespace {

constexpr size_t kNoAllocJobBytes = 1024;     // JobRec + 4 path/pixel bases
constexpr size_t kNoAllocCntBytes = 2048;     // count + part counters
constexpr size_t kNoAllocBytes =
    kNoAllocJobBytes + kNoAllocCntBytes + (size_t)BI_PT_NOALLOC_MAX_RAYS * sizeof(HitRec);
static_assert(sizeof(JobRec) + 4 * sizeof(uint32_t) <= kNoAllocJobBytes, "job record fits its scratch slot");
static_assert((4 + kWorkParts * kWorkStride) * sizeof(uint32_t) <= kNoAllocCntBytes, "counters fit");
static_assert(kNoAllocJobBytes % 256 == 0 && kNoAllocCntBytes % 256 == 0, "slots keep 256-B alignment");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch (looked up once per device, with the device
// current; host code uses this address, never the array's name).
bi_pt_status noalloc_scratch(Owner* own, int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(own, BI_PT_ERROR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(own, BI_PT_ERROR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[device];
    return BI_PT_OK;
}

}  // namespace

extern "C" {

bi_pt_status bi_pt_trace_closest_noalloc(bi_pt_ctx* ctx, uint32_t n, const bi_pt_rays* rays, float tmax,
                                         float* hit_t, uint32_t* hit_triangle, float* hit_u, float* hit_v,
                                         bi_pt_stream stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (ctx && ctx->scene_src) ctx = ctx->scene_src;   // a view traces its scene
    if (!ctx || (n && !rays)) return BI_PT_ERROR_INVALID_ARGUMENT;
    if (n > BI_PT_NOALLOC_MAX_RAYS)
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT,
                    "bi_pt_trace_closest_noalloc takes at most " + std::to_string(BI_PT_NOALLOC_MAX_RAYS) +
                        " rays per call (got " + std::to_string(n) + ")");
    if (n && (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null ray array");
    if (!ctx->has_scene) return fail(ctx, BI_PT_ERROR_NO_SCENE, "no scene set");
    if (!(tmax > kTMin) || !(tmax <= kShadowTMax))
        return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "ray queries need 0.001 < tmax <= 1.7e38");
    // nothing is built or allocated here: the LBVH must be current
    if (!ctx->eng.ready || !ctx->bvh_built)
        return fail(ctx, BI_PT_ERROR_UNSUPPORTED,
                    "bi_pt_trace_closest_noalloc needs the scene's LBVH built (render or query the scene once "
                    "after its last change)");
    if (n == 0) return BI_PT_OK;
    if (!hit_triangle) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null hit_triangle");
    BI_HIP_TRY(ctx, hipSetDevice(ctx->device));
    unsigned char* base = nullptr;
    BI_TRY(noalloc_scratch(ctx, ctx->device, &base));
    unsigned char* job = base;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(base + kNoAllocJobBytes);
    HitRec* hits = reinterpret_cast<HitRec*>(base + kNoAllocJobBytes + kNoAllocCntBytes);
    JobRec rec{};
    rec.nodes = ctx->num_tris ? ctx->nodes4.p : ctx->eng.empty_tree.p;
    rec.leaf_geo = ctx->leaf_geo.p;
    rec.leaf_shade = ctx->leaf_shade.p;
    rec.materials = ctx->mats.p;
    rec.leaf_uv = scene_leaf_uv(ctx);
    rec.tex = scene_tex(ctx);
    rec.num_tris = ctx->num_tris;
    rec.ray_length = tmax;
    // the previous call's contents are stale: the record, the count and the part
    // counters are written afresh (fill kernels, no memset nodes under capture)
    hipLaunchKernelGGL(k_put_query_job, dim3(1), dim3(64), 0, s, rec, n, job);
    launch_fill_u32(cnt, 0u, 4 + kWorkParts * kWorkStride, s);
    launch_fill_u32(cnt, n, 1, s);
    TraceArgs ta{};
    ta.jobs = WaveJobs{reinterpret_cast<const JobRec*>(job), reinterpret_cast<const uint32_t*>(job + sizeof(JobRec)),
                       reinterpret_cast<const uint32_t*>(job + sizeof(JobRec)) + 2, 1u, n};
    // one job: the kernel never reads path ids
    ta.q = RayQueue{const_cast<float*>(rays->ox), const_cast<float*>(rays->oy), const_cast<float*>(rays->oz),
                    const_cast<float*>(rays->dx), const_cast<float*>(rays->dy), const_cast<float*>(rays->dz), nullptr};
    ta.count_in = cnt;
    ta.work = cnt + 4;
    ta.parts = kWorkParts;
    ta.refill = 32;
    ta.any_hit = ctx->non_opaque ? 1u : 0u;
    ta.hits = HitBuf{hits};
    ta.error = &ctx->eng.counters.p[kCntError];
    launch_trace_closest(ta, n, false, 16, 256, s);
    // a miss never reads the order; an empty scene has no order, any valid address serves
    hipLaunchKernelGGL(k_query_hits, dim3((n + 255u) / 256u), dim3(256), 0, s, hits,
                       ctx->num_tris ? ctx->vals0.p : cnt, n, hit_t, hit_triangle, hit_u, hit_v);
    BI_HIP_TRY(ctx, hipGetLastError());
    return BI_PT_OK;
}

bi_pt_status bi_pt_trace_occlusion(bi_pt_ctx* ctx, uint32_t n, const bi_pt_rays* rays, float tmax,
                                   uint32_t* occluded, bi_pt_stream stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    BI_TRY(query_prepare(ctx, n, rays, tmax, s));
    if (n == 0) return BI_PT_OK;
    if (!occluded) return fail(ctx, BI_PT_ERROR_INVALID_ARGUMENT, "null occluded");
    if (ctx->num_tris == 0) {
        launch_fill_u32(occluded, 0u, n, s);
        BI_HIP_TRY(ctx, hipGetLastError());
        return BI_PT_OK;
    }
    RayQueue q{const_cast<float*>(rays->ox), const_cast<float*>(rays->oy), const_cast<float*>(rays->oz),
               const_cast<float*>(rays->dx), const_cast<float*>(rays->dy), const_cast<float*>(rays->dz), nullptr};
    launch_debug_anyhit(bvh_view(ctx), q, n, tmax, occluded, &ctx->eng.counters.p[kCntError], s);
    BI_HIP_TRY(ctx, hipGetLastError());
    return BI_PT_OK;
}

}  // extern "C"
