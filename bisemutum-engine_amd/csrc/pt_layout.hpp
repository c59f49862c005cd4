// pt_layout.hpp -- device data layout of the path-tracing pass (host and device).
//
// HBM layout (one context):
//   scene      mesh pool (float xyz positions/normals, uint32 indices; texture
//              coordinates and tangents where materials sample textures), instance
//              records, material table (128 B each), texture records + texel pool + sRGB
//              table, sky texels (equirect or six cube faces) + importance and guide tables
//   per frame  world triangles (36 B) + normals (36 B) in instance order, Morton keys,
//              sorted order, BVH2 nodes (64 B: both child boxes + child links; the build
//              structure), BVH4 nodes (128 B, one 128-B line: four child boxes as six
//              4-wide planes + four child links; the traversal structure, stored at the
//              index of the BVH2 node it collapses), leaf intersection records (48 B:
//              v0, e1, e2, a non-opaque mark), leaf shading records (48 B: n0, n1, n2,
//              material) and, where needed, leaf texture coordinates (32 B) and tangents
//              (48 B) in sorted leaf order
//   per wave   SoA ray queues (o.xyz, d.xyz, path id), 16-B hit records (t, leaf, u, v),
//              SoA path state (throughput, last BSDF pdf), per-kind radiance accumulators,
//              SoA shadow queue (o.xyz, d.xyz, contribution.xyz, path<<3|kind)
// Nodes are read whole by one lane (7 x 16-B loads); rays, hits and path state are
// read by consecutive lanes at consecutive addresses (coalesced).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bipt {

constexpr int kMaxDirLights = 4;
constexpr int kMaxKinds = kMaxDirLights + 3;   // miss, lights..., sky, rect lights
// Traversal stack bound: a Karras tree over 30-bit Morton codes with index tie-breaks
// is at most 63 levels deep, a BVH4 level spans at least one of them, and a step pushes
// at most 3 siblings.
constexpr int kStackMax = 192;
constexpr uint32_t kLeafMiss = 0xffffffffu;
constexpr uint32_t kKindSky = 7u;

struct alignas(16) BvhNode {
    float box[12];     // child0 lo.xyz hi.xyz, child1 lo.xyz hi.xyz
    int32_t child[2];  // internal k >= 0, leaf j as ~j
    int32_t pad[2];
};
static_assert(sizeof(BvhNode) == 64, "node record is 64 bytes");

// BVH4 node: child k's box is (lox[k], loy[k], loz[k]) - (hix[k], hiy[k], hiz[k]);
// child k is internal (BVH4 node index >= 0) or leaf j (stored as ~j).  An empty slot
// has the box (+inf,+inf,+inf)-(+inf,+inf,+inf), which every slab test rejects.
struct alignas(128) Bvh4Node {
    float lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "BVH4 node record is 128 bytes");

struct alignas(16) LeafGeo {
    float v[12];       // v0.xyz, e1.xyz, e2.xyz; v[9]: flag bits (kLeafNonOpaque), then unused
};
constexpr uint32_t kLeafNonOpaque = 1u;   // the triangle's material is alpha-tested or translucent
struct alignas(16) LeafShade {
    float n[9];        // n0, n1, n2 (cofactor-transformed, not normalised)
    uint32_t material;
    uint32_t pad[2];
};
static_assert(sizeof(LeafGeo) == 48 && sizeof(LeafShade) == 48, "leaf records are 48 bytes");

struct MaterialRec {   // = bi_pt_material
    float base_color[3];
    float roughness;
    float f0_color[3];
    float anisotropy;
    float f90_color[3];
    float ior;
    uint32_t two_sided;
    uint32_t pattern;           // kPatternCheckerXZ: odd cells this record, even cells pattern_material
    uint32_t pattern_material;
    uint32_t blend_mode;        // kBlendOpaque / kBlendAlphaTest / kBlendTranslucent
    float opacity;              // surface.opacity unless opacity_tex names a texture
    uint32_t opacity_tex;       // texture index + 1 whose channel gives the opacity (kNoTexture: none)
    uint32_t opacity_channel;   // 0..3
    uint32_t opacity_step;      // 1: opacity = channel < step_at ? step_below : step_above
    float step_at, step_below, step_above;
    uint32_t base_color_tex;    // texture index + 1: surface.base_color = texel.rgb
    uint32_t f0_tex;            // texture index + 1: surface.f0_color = texel.rgb
    uint32_t normal_tex;        // texture index + 1: surface.normal_map_value = texel.rgb
    uint32_t roughness_tex;     // texture index + 1: surface.roughness = texel channel roughness_channel
    uint32_t roughness_channel;
    uint32_t pad[4];
};
constexpr uint32_t kPatternCheckerXZ = 1u;
constexpr uint32_t kBlendOpaque = 0u, kBlendAlphaTest = 1u, kBlendTranslucent = 2u;
constexpr uint32_t kNoTexture = 0u;   // texture references are index + 1, so a zeroed record has none
constexpr uint32_t kSurfaceModelLit = 1u;   // SurfaceModel::lit (include/bisemutum/graphics/material.hpp:13-16)
constexpr float kAlphaTestThreshold = 0.01f;   // ALPHA_TEST_OPACITY_THRESHOLD, material/utils.hlsl:38
static_assert(sizeof(MaterialRec) == 128, "material record is 128 bytes");

// Texture of the material table: texels in the scene's texture pool from word `offset`
// (RGBA8: one word per texel, RGBA32F: four), level 0 only (the reference's ray-tracing
// materials sample level 0: material.cpp rewrites Sample into SampleLevel(..., 0)).
struct TexRec {
    uint32_t offset;
    uint32_t width, height;
    uint32_t format;            // kTexRgba8Unorm / kTexRgba8Srgb / kTexRgba32f
    uint32_t filter;            // 0 nearest, 1 linear
    uint32_t address_u, address_v;   // 0 repeat, 1 mirrored repeat, 2 clamp to edge
    uint32_t pad;
};
constexpr uint32_t kTexRgba8Unorm = 0u, kTexRgba8Srgb = 1u, kTexRgba32f = 2u;
static_assert(sizeof(TexRec) == 32, "texture record is 32 bytes");

// Textures of a scene as the shading and any-hit code read them.
struct SceneTex {
    const TexRec* recs;
    const uint32_t* pool;
    const float* srgb;          // 256 entries: sRGB-encoded byte -> linear
};

// Per-leaf texture coordinates (uv0, uv1, uv2, unused) and world tangents (xyz, w = the
// bitangent sign) of the three vertices; present when the scene's meshes carry them.
struct alignas(16) LeafUv {
    float uv[8];
};
struct alignas(16) LeafTan {
    float t[12];
};

// A light of the job: directional (bi_pt_dir_light = DirLightData) or point / spot
// (bi_pt_point_light = PointLightData), the fields the shading reads.  Directional
// lights come first, then point lights; at most kMaxDirLights in all.
constexpr uint32_t kLightDirectional = 0u, kLightPoint = 1u;
struct LightRec {
    float emission[3];
    int32_t sm_index;       // < 0: no shadow ray
    float direction[3];     // directional: towards the light; point: the spot axis, towards the light
    float shadow_strength;
    float position[3];      // point lights
    float range_sqr_inv;
    float cos_inner, cos_outer;
    uint32_t type;          // kLightDirectional / kLightPoint
    float pad;
};
static_assert(sizeof(LightRec) == 64, "light record is 64 bytes");

// A rect light (bi_pt_rect_light = RectLightData, lights_struct.hlsl:38-53): corners
// position0..3 (centre + (w, h), (-w, h), (-w, -h), (w, -h) in the light's frame),
// emission, one- or two-sided.  Rect lights cast no shadows; each shaded vertex adds
// their linearly-transformed-cosine integrals.
constexpr uint32_t kMaxRectLights = 16;
struct RectRec {
    float emission[3];
    int32_t texture_index;  // < 0: untextured; else the rect-light texture it samples
    float center[3];
    uint32_t two_sided;
    float p0[3];
    float inv_width_sqr;
    float p1[3];
    float inv_height_sqr;
    float p2[3];
    float inv_texel_size;
    float p3[3];
    float pad1;
    float normal[3];
    float pad2;
};
static_assert(sizeof(RectRec) == 112, "rect light record is 112 bytes");
// A rect light's texture (LightsContext::rect_light_textures, lights.cpp:229-238, with its
// sampler): num_levels mip levels, level l a TexRec at first_level + l of the rect-texture
// level table; filters 0 nearest / 1 linear; the sampler's LOD bias and clamp.
constexpr uint32_t kMaxRectTextures = 16;
struct RectTexRec {
    uint32_t first_level, num_levels;
    uint32_t mag_filter, min_filter, mipmap_mode;
    float lod_bias, lod_min, lod_max;
};
static_assert(sizeof(RectTexRec) == 32, "rect texture record is 32 bytes");
// The LTC tables (lights.cpp:26-37): 8 x 8 x 64 texels each, x fastest; three matrix rows
// (xyz of RGBA32F) and the norm / Fresnel pair (xy of RG32F), held as four float4 tables.
constexpr uint32_t kLtcSize = 8, kLtcDepth = 64, kLtcTexels = kLtcSize * kLtcSize * kLtcDepth;

struct InstanceRec {
    float xform[12];
    uint32_t first_tri;   // first world triangle of this instance
    uint32_t idx_off;     // first index of its mesh in the index pool
    uint32_t vtx_off;     // first vertex of its mesh in the vertex pool
    uint32_t material;
};
static_assert(sizeof(InstanceRec) == 64, "instance record is 64 bytes");

struct CameraParams {
    float inv_view[16];
    float inv_proj[16];
    float proj_view[16];
    uint32_t width, height;
};

// Sky texel layouts: an equirect (width x height, row 0 = +Y pole) or a cube map of six
// width x width faces stored one after the other (+X, -X, +Y, -Y, +Z, -Z; height =
// 6 * width), the layers of the engine's skybox texture.
constexpr uint32_t kSkyEquirect = 0u, kSkyCube = 1u;
// Alias tables of the sky's importance sampling (Vose's method): row r keeps itself with
// probability prob, else goes to row alias; a texel entry also carries its own weight and
// its alias's, so the sampled texel's weight comes with the entry.
struct SkyRowAlias {
    float prob;
    uint32_t alias;
};
struct SkyTexAlias {
    float prob;
    uint32_t alias;
    float w, w_alias;
};
struct SkyParams {
    const float4* texels;
    const float* weight;   // per texel: luminance * sin(theta) (equirect) or * 1/r^3 (cube)
    const SkyRowAlias* row_alias;   // per row (rows weighted by their sums)
    const SkyTexAlias* tex_alias;   // per texel (each row's texels by their weights)
    uint32_t width, height;
    uint32_t layout;
    float total;
    float pdf_scale;       // W*H/total (equirect), S*S/(4*total) (cube)
    float color[3];
    float rot[9];
};

struct BvhView {
    const Bvh4Node* nodes;
    const LeafGeo* leaf_geo;
    uint32_t num_tris;
    // any-hit inputs of non-opaque leaves
    const LeafShade* leaf_shade;
    const MaterialRec* materials;
    const LeafUv* leaf_uv;
    SceneTex tex;
    uint32_t seed;
};

// Job kinds beside plain images (JobRec::flags).
constexpr uint32_t kJobCache = 1u << 2;   // image paths end in the radiance cache at cache_bounce
constexpr uint32_t kJobHalfOut = 1u << 4;   // the colour output is RGBA16F
constexpr uint32_t kJobTrain = 1u << 3;   // the cache's training paths: one per slot, random pixels,
                                          // radiance after cache_bounce goes to the slot's target
constexpr uint32_t kNoPath = 0xffffffffu;

// Surface sample of the radiance cache (16 floats, 64 B): a query of an image path
// (tag = path, kNoPath when the entry has none) or a training sample (tag = 1 when the
// path reached its record segment, else 0).
struct alignas(16) NrcRecord {
    float p[3];       // position (world)
    float wo[3];      // direction towards the previous vertex
    float n[3];       // shading normal
    float roughness;
    float albedo[3];  // base colour
    uint32_t tag;
    float pad[2];
};
static_assert(sizeof(NrcRecord) == 64, "cache record is 64 bytes");

// One render job of a wave: a scene (BVH, shading records, materials, lights, sky), a
// camera and sampling parameters, and where its paths sit in the wave.  A wave may hold
// paths of several jobs (batched rendering); job j owns the paths [path_base[j],
// path_base[j+1]) and the pixel slots [pix_base, pix_base + npix) of the resolve pass.
// Kernels find a path's job by a binary search over the wave's path bases.
struct JobRec {
    const Bvh4Node* nodes;
    const LeafGeo* leaf_geo;
    const LeafShade* leaf_shade;
    const MaterialRec* materials;
    const LeafUv* leaf_uv;          // null: the scene has no texture coordinates
    const LeafTan* leaf_tan;        // null: no tangents (normal maps are not applied)
    SceneTex tex;
    uint32_t num_tris;
    uint32_t num_lights;
    uint32_t sky_nee;       // sky next-event estimation on
    uint32_t last_bounce;   // segments - 1
    float keep[kMaxDirLights];   // 1 - shadow_strength per light
    LightRec lights[kMaxDirLights];
    SkyParams sky;
    CameraParams cam;
    uint32_t npix;
    uint32_t sample0;       // sample index of the job's first sample in this wave
    uint32_t samples;       // samples of the job in this wave
    uint32_t seed;
    uint32_t flags;         // bit 0: pixel jitter; kJobCache, kJobTrain, kJobHalfOut
    uint32_t cache_bounce;  // segment at which the radiance cache takes over (kJobCache) or
                            // at which a training path records its sample (kJobTrain)
    NrcRecord* cache_rec;   // kJobTrain: one record per slot
    float ray_length;
    uint32_t pix_base;      // first resolve slot of the job in this wave
    float* sums;            // running per-pixel sums (3 floats per pixel)
    float* depth_out;       // nullable: reverse-Z depth of the primary hit
    uint32_t depth_sample;  // the sample whose primary hit gives the depth
    float* out_rgba;        // output mean (RGBA32F, or RGBA16F with kJobHalfOut)
    float out_den;          // samples accumulated after this call
    // primary-hit GBuffer of sample depth_sample (each nullable): rgba16f base colour,
    // rgba16f normal / tangent / roughness, rgba16 unorm fresnel, rgba8 unorm material_0
    uint16_t* gb_base;
    uint16_t* gb_nr;
    uint16_t* gb_fres;
    uint8_t* gb_m0;
    const RectRec* rects;   // num_rects rect lights (their slot follows the sky's)
    uint32_t num_rects;
    const float4* ltc;      // 4 * kLtcTexels: matrix rows 0, 1, 2, then the norm table
    const RectTexRec* rect_tex;   // the rect lights' textures (null: none)
    SceneTex rect_tex_levels;     // their mip levels: level records, texel pool, sRGB table
};

// Ray queue in SoA form.
struct RayQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    uint32_t* path;
};
// Hit record of one ray, one 16-B store: a lane finishes its ray at its own time, so
// the records of a wavefront land scattered; one record per ray keeps that to one
// partial cache line instead of four.
struct alignas(16) HitRec {
    float t;
    uint32_t leaf;
    float u, v;
};
struct HitBuf {
    HitRec* rec;
};
// A shadow ray's direction is stored for sky rays only; a light's ray takes it from the
// light (a directional light's direction, or from the origin towards a point light).
struct ShadowQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz, *cx, *cy, *cz;
    uint32_t* pk;   // path << 3 | kind
};
struct PathState {
    float *bx, *by, *bz, *pdf;
    float* acc;       // [kind][3][cap]
    uint32_t cap;
};

// Device-side counters of one wave (uint32 slots).
constexpr uint32_t kDefaultWavePaths = 1u << 25;   // paths per wave unless a job asks otherwise
constexpr uint32_t kMaxWavePaths = 1u << 28;       // path << 3 | kind stays within 32 bits
constexpr uint32_t kWorkParts = 8;     // default ray-queue parts of the persistent lanes (one per XCD)
constexpr uint32_t kWorkPartsMax = 32;
constexpr uint32_t kWorkStride = 32;   // uint32 between part counters: one 128-B line each
enum CounterSlot : uint32_t {
    kCntExtA = 0,       // entries in extension queue A
    kCntExtB = 1,       // entries in extension queue B
    kCntShadow = 2,     // entries in shadow queue set 0 (set 1: kCntShadow + 1)
    kCntError = 4,      // nonzero: a kernel hit an internal limit (1 stack, 4 refit wait, 8 step bound,
                        // 16 treelet start list, 32 shade look-back wait)
    kCntRays = 8,       // 4 x 64-bit ray counters (camera, extension, light shadow, sky shadow)
    kCntTraceWork = 32,                                       // part counters of the closest-hit lanes
    kCntShadowWork = kCntTraceWork + kWorkPartsMax * kWorkStride, // part counters of the shadow lanes (set 0)
    kCntShadowWork1 = kCntShadowWork + kWorkPartsMax * kWorkStride, // ... of shadow queue set 1
    kCntSlots = kCntShadowWork1 + kWorkPartsMax * kWorkStride
};

}  // namespace bipt
