// pt_host.hpp -- host-side state of the C ABI: contexts (one scene + its running
// image), the wave engine (queues and per-wave buffers, shared by every job a call
// renders) and the batch object (an engine that renders several contexts' jobs in one
// set of waves).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/bi_pt.h"
#include "pt_kernels.hpp"

namespace bipt {

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

// Device memory owner: HBM budget accounting and the last error message.
struct Owner {
    int device = 0;
    uint64_t budget = 0;      // 0 = unlimited
    uint64_t allocated = 0;
    std::string err;
};

inline bi_pt_status fail(Owner* o, bi_pt_status st, const std::string& msg) {
    if (o) o->err = msg;
    return st;
}

#define BI_HIP_TRY(own, expr)                                                                       \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return ::bipt::fail((own), BI_PT_ERROR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BI_TRY(expr)                            \
    do {                                        \
        bi_pt_status s_ = (expr);               \
        if (s_ != BI_PT_OK) return s_;          \
    } while (0)

template <class T>
bi_pt_status dalloc(Owner* o, DevBuf<T>& b, size_t n) {
    if (b.p && b.n >= n) return BI_PT_OK;
    if (b.p) {
        (void)hipFree(b.p);
        o->allocated -= b.n * sizeof(T);
        b.p = nullptr;
        b.n = 0;
    }
    if (n == 0) return BI_PT_OK;
    const size_t bytes = n * sizeof(T);
    if (o->budget && o->allocated + bytes > o->budget)
        return fail(o, BI_PT_ERROR_OUT_OF_MEMORY,
                    "HBM budget exceeded: need " + std::to_string(bytes) + " B more, budget " +
                        std::to_string(o->budget) + " B, in use " + std::to_string(o->allocated) + " B");
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(o, BI_PT_ERROR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    b.p = static_cast<T*>(p);
    b.n = n;
    o->allocated += bytes;
    return BI_PT_OK;
}

template <class T>
void dfree(Owner* o, DevBuf<T>& b) {
    if (b.p) {
        (void)hipFree(b.p);
        o->allocated -= b.n * sizeof(T);
    }
    b.p = nullptr;
    b.n = 0;
}

// Tuning knobs of the wave machinery (0 = automatic).
struct Knobs {
    uint32_t wave_paths = 0;
    uint32_t stack_lds = 0;
    uint32_t refill_lanes = 0;
    uint32_t work_parts = 0;
    uint32_t block_threads = 0;   // traversal workgroup size: 128, 256 or 512 (0 = 256)
};

// Queues and per-wave buffers, sized for the largest wave rendered so far.
struct Engine {
    uint32_t wave_cap = 0, wave_slots = 0, wave_sh_mult = 0;
    DevBuf<float> qa[6], qb[6], ps_f[4], acc, sq_f[2][9];
    DevBuf<uint32_t> qa_path, qb_path, sq_pk[2];
    DevBuf<HitRec> hit_rec;
    DevBuf<uint64_t> tile_state;           // look-back words of the shade launch's tiles
    DevBuf<uint32_t> counters;             // kCntSlots (ray counters at kCntRays as 4 x u64)
    DevBuf<unsigned char> plan;            // job records and bases of every wave of a call
    unsigned char* plan_host = nullptr;    // pinned staging of the plan (copied asynchronously)
    size_t plan_host_n = 0;
    bool plan_pending = false;             // plan_ev marks the end of the last plan copy
    DevBuf<Bvh4Node> empty_tree;           // one node of empty slots (jobs without triangles)
    hipEvent_t ev[3] = {};
    hipEvent_t plan_ev = nullptr;
    // A segment's shadow launch runs on aux_s, forked after its shading (ev_shade) and
    // overlapping the next segment's closest-hit launch; segments alternate between two
    // shadow queue sets, and a set is reused only after its last shadow launch has ended
    // (ev_shadow[set]); the wave joins aux_s before its resolve.
    hipStream_t aux_s = nullptr;
    hipEvent_t ev_shade = nullptr, ev_shadow[2] = {};
    // LBVH builds of independent scenes run concurrently: forked from the call's stream
    // onto these streams (default priority) and joined back by events
    static constexpr int kBuildStreams = 3;
    hipStream_t build_s[kBuildStreams] = {};
    hipEvent_t build_fork = nullptr, build_join[kBuildStreams] = {};
    bool ready = false;
    // statistics
    bi_pt_stats stats{};
    uint64_t frames = 0, camera_rays = 0;
    // queues recorded by a single-job call with BI_PT_FLAG_RECORD_QUEUES: [wave][segment]
    std::vector<std::vector<std::vector<uint32_t>>> rec_ext, rec_sh;
};

bi_pt_status engine_init(Owner* own, Engine& eng);
// Device bytes of the wave buffers for waves of `cap` paths, `slots` radiance kinds and
// `sh_mult` shadow rays per path (what ensure_wave_buffers allocates).
uint64_t engine_bytes(uint32_t cap, uint32_t slots, uint32_t sh_mult);
void engine_free(Owner* own, Engine& eng);
// Statistics and the error flags raised since the last check (synchronises the device;
// clears the flags it reports).
bi_pt_status engine_check(Owner* own, Engine& eng, bi_pt_stats* out);
// The error flags raised since the last check, after the work on stream s (synchronises
// s; clears the flags it reports).
bi_pt_status engine_take_errors(Owner* own, Engine& eng, hipStream_t s);
std::string device_error_text(uint32_t flags);
// ORs flags into the error word from the host (error-path tests).
bi_pt_status engine_inject_error(Owner* own, Engine& eng, uint32_t flags);
// Rejects a camera / settings pair that a render call would reject (image size, samples,
// ray length).
bi_pt_status validate_job(Owner* own, const bi_pt_camera* c, const bi_pt_settings* st);

}  // namespace bipt

// One scene, its per-frame BVH and its running image.  A view (bi_pt_create_view)
// renders the scene of another context (scene_src: scene, lights, sky and BVH are
// read from there) into a running image of its own.
struct bi_pt_ctx : bipt::Owner {
    bi_pt_ctx* scene_src = nullptr;
    uint64_t scene_version = 0;     // bumped by every change of scene, transforms, lights or sky
    uint64_t last_version = 0;      // scene version of the running image
    // scene
    bool has_scene = false;
    bipt::DevBuf<float> pos, nrm;
    bipt::DevBuf<uint32_t> idx;
    bipt::DevBuf<bipt::InstanceRec> inst;
    bipt::DevBuf<bipt::MaterialRec> mats;
    std::vector<bipt::InstanceRec> inst_host;
    uint32_t num_tris = 0, num_inst = 0, num_mats = 0;
    // texture coordinates and tangents (vertex pools, present when the materials sample
    // textures / normal maps) and the material textures
    bool need_uv = false, need_tan = false;
    bool non_opaque = false;        // some material is alpha-tested or translucent (any-hit tests)
    bool gbuf_tan = false;          // a GBuffer was asked for: leaf tangents are built from then on
    bipt::DevBuf<float> uvs, tans;
    bipt::DevBuf<bipt::TexRec> tex_recs;
    bipt::DevBuf<uint32_t> tex_pool;
    bipt::DevBuf<float> srgb;
    uint32_t num_tex = 0;
    uint32_t tex_refs = 0;          // 1 + the largest texture index a material names (0: none)
    // lights: the directional ones, then the point ones (the records the kernels read)
    std::vector<bipt::LightRec> lights, dir_lights, point_lights;
    // rect lights and the LTC tables they are integrated with (4 * kLtcTexels texels)
    std::vector<bipt::RectRec> rects_host;
    bipt::DevBuf<bipt::RectRec> rects;
    bipt::DevBuf<float4> ltc;
    std::vector<float> ltc_host;
    // the rect lights' textures: one record each, their mip levels' records and texels
    bipt::DevBuf<bipt::RectTexRec> rect_tex;
    bipt::DevBuf<bipt::TexRec> rect_tex_levels;
    bipt::DevBuf<uint32_t> rect_tex_pool;
    uint32_t num_rect_tex = 0;
    std::vector<bipt::RectTexRec> rect_tex_host;      // what the device holds, to skip unchanged sets
    std::vector<bipt::TexRec> rect_tex_levels_host;
    std::vector<uint32_t> rect_tex_pool_host;
    // sky
    bipt::DevBuf<float4> sky_tex;
    bipt::DevBuf<float> sky_w;
    bipt::DevBuf<bipt::SkyRowAlias> sky_row_alias;
    bipt::DevBuf<bipt::SkyTexAlias> sky_tex_alias;
    bipt::SkyParams sky{};
    std::vector<float> sky_host;
    bool sky_set = false;
    // build buffers
    bipt::DevBuf<float> tri_v, tri_n, cb_out;
    bipt::DevBuf<uint32_t> tri_mat, cb_ord, keys0, keys1, vals0, vals1, hist, refit_cnt;
    bipt::DevBuf<bipt::BvhNode> nodes;
    bipt::DevBuf<bipt::Bvh4Node> nodes4;
    bipt::DevBuf<int32_t> parent;
    bipt::DevBuf<bipt::LeafGeo> leaf_geo;
    bipt::DevBuf<bipt::LeafShade> leaf_shade;
    bipt::DevBuf<bipt::LeafUv> leaf_uv;
    bipt::DevBuf<bipt::LeafTan> leaf_tan;
    bipt::DevBuf<float> tl_cost;
    bipt::DevBuf<int32_t> tl_height, tl_snap;
    bipt::DevBuf<uint32_t> tl_starts;
    uint32_t treelet_rounds = 3;    // bi_pt_set_build_options
    bool bvh_built = false;
    uint32_t built_rounds = 0;      // treelet rounds of the BVH built last
    // ray queries (bi_pt_trace_closest / bi_pt_trace_occlusion): job record + bases,
    // queue-part counters, path ids, hit records
    bipt::DevBuf<unsigned char> rq_job;
    bipt::DevBuf<uint32_t> rq_cnt, rq_path;
    bipt::DevBuf<bipt::HitRec> rq_hits;
    // accumulation
    bipt::DevBuf<float> sums;
    uint32_t acc_w = 0, acc_h = 0;
    uint64_t acc_count = 0;
    bipt::CameraParams last_cam{};
    uint32_t last_seed = 0, last_flags = 0, last_bounces = 0, last_role = 0;
    float last_ray_length = 0.0f;
    // wave machinery of bi_pt_render
    bipt::Engine eng;
};

// Several contexts' jobs rendered in one set of waves.
struct bi_pt_batch : bipt::Owner {
    bipt::Engine eng;
    bipt::Knobs knobs;
    bool use_graph = false;
    // captured call (hipGraph) and the plan it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
};

namespace bipt {

// The context that holds the scene a context renders.
inline bi_pt_ctx* scene_of(bi_pt_ctx* c) { return c->scene_src ? c->scene_src : c; }

bi_pt_status ensure_build_buffers(bi_pt_ctx* ctx);
// rounds: treelet rounds of this build (the context's setting, or fewer for a call that
// traces too few paths to repay them)
bi_pt_status build_bvh(bi_pt_ctx* ctx, hipStream_t s, uint32_t* error, uint32_t rounds);
BvhView bvh_view(const bi_pt_ctx* ctx);
// Texture inputs of a scene context's kernels.
inline SceneTex scene_tex(const bi_pt_ctx* c) { return SceneTex{c->tex_recs.p, c->tex_pool.p, c->srgb.p}; }
inline const LeafUv* scene_leaf_uv(const bi_pt_ctx* c) { return c->need_uv ? c->leaf_uv.p : nullptr; }
// Leaf tangents are built for normal maps and, once one is asked for, for the GBuffer.
inline bool scene_wants_tan(const bi_pt_ctx* c) { return c->need_tan || c->gbuf_tan; }
inline const LeafTan* scene_leaf_tan(const bi_pt_ctx* c) { return scene_wants_tan(c) ? c->leaf_tan.p : nullptr; }

// One job of a call.  Roles: a plain image; an image whose paths end in the call's
// radiance cache; the cache's training paths (one sample per slot into the cache's
// targets, no image, no accumulation; its context only names the scene).
enum : uint32_t { kRoleImage = 0, kRoleCached = 1, kRoleTrain = 2 };
struct JobIn {
    bi_pt_ctx* ctx;
    const bi_pt_camera* camera;
    const bi_pt_settings* st;
    float* out_rgba;
    float* out_depth;
    uint32_t role = kRoleImage;
    const bi_pt_gbuffer* gbuf = nullptr;   // primary-hit GBuffer of the call's first sample
};

struct RenderOpts {
    Knobs knobs;
    bool timing = false;      // per-kernel-class events (single-job calls)
    bool record = false;      // record queue contents (single-job calls)
    bool reuse_bvh = false;   // keep a built BVH whose scene is unchanged
    bi_pt_nrc* nrc = nullptr; // radiance cache of kRoleCached / kRoleTrain jobs
};

// Renders every job (spp samples each, added to its context's running image) in shared
// waves on one stream.  Errors are reported on `own`.
bi_pt_status render_jobs(Owner* own, Engine& eng, const JobIn* jobs, uint32_t n, hipStream_t s,
                         const RenderOpts& opt, bi_pt_batch* graph_owner);

}  // namespace bipt
