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
    uint32_t* error;
};
void launch_build_lbvh(const BuildArgs& a, hipStream_t s);
uint32_t sort_hist_entries(uint32_t n);

// Wave description shared by the trace and shade kernels.
struct WaveParams {
    CameraParams cam;
    uint32_t num_paths;   // paths in this wave: (samples in wave) * npix
    uint32_t npix;
    uint32_t sample0;     // global sample index of the wave's first sample
    uint32_t seed;
    uint32_t flags;
    float ray_length;
};

struct TraceArgs {
    BvhView bvh;
    WaveParams wave;
    RayQueue q;                 // input queue (unused for primary rays)
    const uint32_t* count_in;   // device count (null: primary, count = wave.num_paths)
    HitBuf hits;
    uint32_t* work;             // part counters of the persistent lanes (zero at launch)
    uint32_t parts;             // queue parts (1..kWorkPartsMax)
    uint32_t refill;            // idle lanes before a refill (1..64)
    uint32_t* error;
};
void launch_trace_closest(const TraceArgs& a, uint32_t capacity, bool primary, uint32_t lds_stack, hipStream_t s);

void launch_debug_anyhit(const BvhView& bvh, const RayQueue& q, uint32_t n, float tmax, uint32_t* out,
                         uint32_t* error, hipStream_t s);

struct ShadowArgs {
    BvhView bvh;
    ShadowQueue q;
    const uint32_t* count_in;
    PathState ps;
    float keep[kMaxDirLights];  // 1 - shadow_strength per light
    uint32_t num_lights;
    uint32_t* work;             // part counters of the persistent lanes (zero at launch)
    uint32_t parts;             // queue parts (1..kWorkPartsMax)
    uint32_t refill;            // idle lanes before a refill (1..64)
    uint32_t* error;
};
void launch_trace_shadow(const ShadowArgs& a, uint32_t capacity, uint32_t lds_stack, hipStream_t s);

// Per-entry staging written by the shade kernel and read by the compaction kernel.
struct ShadeStage {
    float *px, *py, *pz;    // hit point (entries with any output)
    uint32_t* path;
    float *ex, *ey, *ez;    // extension direction
    float *sx, *sy, *sz;    // sky shadow-ray direction
    float *cx, *cy, *cz;    // shadow contribution of kind k (lights 0..n-1, sky n) at k * cap + entry
    uint32_t* mask;         // bit 0 extension ray, bit 1 + l light l, bit 1 + kMaxDirLights sky
    uint64_t* tile_sums;    // per 256-entry tile: extension << 32 | shadow (exclusive prefix after the scan)
    uint32_t* tile_sky;     // per tile: sky shadow rays
};

struct ShadeArgs {
    WaveParams wave;
    const LeafShade* leaf_shade;
    const MaterialRec* materials;
    LightRec lights[kMaxDirLights];
    uint32_t num_lights;
    SkyParams sky;
    uint32_t sky_nee;
    RayQueue qin;
    const uint32_t* count_in;   // null at bounce 0
    HitBuf hits;
    ShadeStage st;
    PathState ps;
    uint32_t* trace_work_reset;   // zeroed here for the next closest-hit launch
    uint32_t* shadow_work_reset;  // zeroed here for this segment's shadow launch
    uint32_t bounce;
    uint32_t last_bounce;       // segments - 1
    float* depth_out;           // nullable: reverse-Z depth of the primary hit
    uint32_t depth_sample;
};
void launch_shade(const ShadeArgs& a, uint32_t capacity, hipStream_t s);

struct ScanArgs {
    uint64_t* tile_sums;
    const uint32_t* tile_sky;
    const uint32_t* count_in;   // null: num_entries
    uint32_t num_entries;
    uint32_t* count_ext;
    uint32_t* count_shadow;
    uint64_t* ray_counters;     // [camera, extension, light shadow, sky shadow]
};
void launch_scan_tiles(const ScanArgs& a, hipStream_t s);

struct CompactArgs {
    ShadeStage st;
    const uint32_t* count_in;   // null: num_entries
    uint32_t num_entries;
    uint32_t num_lights;
    size_t cap;
    float ldir[kMaxDirLights][3];
    RayQueue qout;
    ShadowQueue sq;
};
void launch_compact(const CompactArgs& a, uint32_t capacity, hipStream_t s);

struct ResolveArgs {
    PathState ps;
    uint32_t num_slots;   // accumulator kinds in use: miss, lights, sky
    uint32_t npix;
    uint32_t samples;     // samples in this wave
    uint32_t width, height;
    float* sums;          // running per-pixel sums (3 floats)
};
void launch_resolve(const ResolveArgs& a, hipStream_t s);
void launch_output(const float* sums, uint32_t npix, float inv_count_den, uint32_t count, float* out_rgba,
                   hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s);

}  // namespace bipt
