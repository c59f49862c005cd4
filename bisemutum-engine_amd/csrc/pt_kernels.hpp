// pt_kernels.hpp -- kernel argument blocks and launchers of the path-tracing pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_device.hpp"
#include "pt_layout.hpp"

namespace bipt {

PT_HD uint32_t expand10(uint32_t v) {
    v &= 1023u;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
PT_HD uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
}

struct BuildArgs {
    const float* positions;
    const float* normals;
    const uint32_t* indices;
    const InstanceRec* instances;
    uint32_t num_instances;
    uint32_t num_tris;
    float* tri_v;
    float* tri_n;
    uint32_t* tri_mat;
    uint32_t* cbounds_ord;   // 6 order-preserving uint32 (min xyz, max xyz)
    float* cbounds_out;      // 6 floats (may be null)
    uint32_t* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;
    BvhNode* nodes;
    Bvh4Node* nodes4;
    int32_t* parent;
    uint32_t* refit_cnt;
    LeafGeo* leaf_geo;
    LeafShade* leaf_shade;
    const MaterialRec* materials;   // blend modes: non-opaque leaves are flagged
    // per-vertex texture coordinates (2 floats) and tangents (4), and the per-leaf copies
    // made from them (all null unless the scene's materials need them)
    const float* uvs;
    const float* tangents;
    LeafUv* leaf_uv;
    LeafTan* leaf_tan;
    uint32_t* error;
    // treelet restructuring after the refit (0 rounds: the Morton/Karras tree as built)
    uint32_t treelet_rounds;
    float* tl_cost;
    int32_t* tl_height;
    int32_t* tl_snap;    // per internal node: height at the start of a round
    uint32_t* tl_starts; // 1 + treelet_start_cap(num_tris) entries
};
void launch_build_lbvh(const BuildArgs& a, hipStream_t s);
uint32_t sort_hist_entries(uint32_t n);
uint32_t treelet_start_cap(uint32_t n);

// Jobs of a wave: records, path bases (num_jobs + 1 entries, the last = paths in the
// wave) and pixel bases (num_jobs + 1 entries) for the resolve pass.
struct WaveJobs {
    const JobRec* rec;
    const uint32_t* path_base;
    const uint32_t* pix_base;
    uint32_t num_jobs;
    uint32_t num_paths;
};

// Job of a path (or pixel slot): the last j with base[j] <= x.
PT_HD uint32_t job_of(const uint32_t* base, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;   // base[lo] <= x < base[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Slot, pixel and sample of a job's local path: images walk their pixels in 8x8 tile
// order, sample after sample; the cache's training paths take one pixel per slot drawn
// uniformly from the image (PCG32 keyed by the job seed and slot) and use the slot as
// their sample index, so no two training paths share a random stream.
constexpr uint32_t kTrainPixelBounce = 254u;
struct PathKey {
    uint32_t q, pixel, sample;
};
PT_HD PathKey path_key(const JobRec& J, uint32_t local) {
    const uint32_t q = local % J.npix, sl = local / J.npix;
    if (J.flags & kJobTrain) {
        Pcg r = pcg_keyed(J.seed, q, 0u, kTrainPixelBounce);
        const uint64_t npix = (uint64_t)J.cam.width * J.cam.height;
        return PathKey{q, (uint32_t)(((uint64_t)pcg_next(r) * npix) >> 32), q};
    }
    return PathKey{q, pixel_of(q, J.cam.width, J.cam.height), J.sample0 + sl};
}

struct TraceArgs {
    WaveJobs jobs;
    RayQueue q;                 // input queue (unused for primary rays)
    const uint32_t* count_in;   // device count (null: primary, count = jobs.num_paths)
    HitBuf hits;
    uint32_t* work;             // part counters of the persistent lanes (zero at launch)
    uint32_t parts;             // queue parts (1..kWorkPartsMax)
    uint32_t refill;            // idle lanes before a refill (1..64)
    uint32_t any_hit;           // nonzero: some scene of the call has non-opaque triangles
    uint32_t* error;
};
// block: threads per workgroup of the persistent grid (128, 256 or 512; else 256).
void launch_trace_closest(const TraceArgs& a, uint32_t capacity, bool primary, uint32_t lds_stack, uint32_t block,
                          hipStream_t s);

void launch_debug_anyhit(const BvhView& bvh, const RayQueue& q, uint32_t n, float tmax, uint32_t* out,
                         uint32_t* error, hipStream_t s);

struct ShadowArgs {
    WaveJobs jobs;
    ShadowQueue q;
    const uint32_t* count_in;
    PathState ps;
    uint64_t* ray_counters;     // sky shadow rays move from [2] (all shadow rays) to [3]
    uint32_t* work;             // part counters of the persistent lanes (zero at launch)
    uint32_t parts;             // queue parts (1..kWorkPartsMax)
    uint32_t refill;            // idle lanes before a refill (1..64)
    uint32_t any_hit;           // nonzero: some scene of the call has non-opaque triangles or point lights
    uint32_t* error;
};
void launch_trace_shadow(const ShadowArgs& a, uint32_t capacity, uint32_t lds_stack, uint32_t block, hipStream_t s);

// Shading of one segment writes its outputs straight into the next queues, in entry
// order: each 256-entry tile publishes its totals and finds its offset by decoupled
// look-back over the tiles before it.
struct ShadeArgs {
    WaveJobs jobs;
    RayQueue qin;
    const uint32_t* count_in;   // null at bounce 0
    HitBuf hits;
    PathState ps;
    RayQueue qout;              // extension rays of the next segment
    ShadowQueue sq;             // shadow rays of this segment (kinds: lights, then the sky)
    uint32_t* count_ext;        // written by the last tile (zeroed by launch_shade_reset)
    uint32_t* count_shadow;
    uint64_t* ray_counters;     // [camera, extension, light shadow, sky shadow]
    uint64_t* tile_state;       // per tile: look-back word (zero at launch)
    uint32_t* error;            // 32: a look-back wait ran out
    uint32_t* trace_work_reset;   // zeroed here for the next closest-hit launch
    uint32_t* shadow_work_reset;  // zeroed here for this segment's shadow launch
    uint32_t bounce;
    NrcRecord* nrc_query;         // non-null: the cache kernel runs after this segment's
                                  // shading; entry i's query (or kNoPath) goes to [i]
};
// Clears the output counts and the look-back words of the tiles the next shade launch
// covers (count_in entries, or num_entries if count_in is null).
void launch_shade_reset(uint64_t* tile_state, const uint32_t* count_in, uint32_t num_entries, uint32_t* count_ext,
                        uint32_t* count_shadow, uint32_t capacity, hipStream_t s);
// max_lights: the most lights of any job in the launch (at most kMaxDirLights); rect: some
// job of the launch has rect lights.
void launch_shade(const ShadeArgs& a, uint32_t capacity, uint32_t max_lights, bool rect, hipStream_t s);

struct ResolveArgs {
    WaveJobs jobs;
    PathState ps;
    uint32_t num_pix;     // resolve slots of the wave (sum of the jobs' pixels)
};
void launch_resolve(const ResolveArgs& a, hipStream_t s);
// Running mean of every job of the wave -> its RGBA32F output.
void launch_output(const WaveJobs& jobs, uint32_t num_pix, hipStream_t s);
// Primary-hit GBuffer of the job record `job` (device) from the segment-0 hit records of its
// paths [path0, path0 + npix): its first sample in the wave, which must be the call's first.
void launch_gbuffer(const JobRec* job, const HitRec* hits, uint32_t path0, uint32_t npix, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s);

}  // namespace bipt
