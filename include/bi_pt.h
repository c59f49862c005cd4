/*
 * bi_pt.h -- C ABI of the MI355X-native wavefront path-tracing pass.
 *
 * Drop-in boundary for bisemutum's builtin-renderer path-tracing pass
 * (paths relative to the reference's bisemutum/ directory):
 *   PathTracingPass::update_params   src/renderer/pass/path_tracing.cpp:183-222
 *   PathTracingPass::render          src/renderer/pass/path_tracing.hpp:27-35, .cpp:224-488
 *   InputData / OutputData           src/renderer/pass/path_tracing.hpp:13-23
 *   PathTracingSettings              include/bisemutum/renderer/basic.hpp:76-81
 * The pass's render-graph inputs become explicit calls: scene geometry and
 * materials (bi_pt_set_scene, replacing the TLAS/BLAS inputs of
 * src/graphics/accel.cpp:11-159 and the scene buffers of
 * src/graphics/gpu_scene_data.hpp:10-24), directional lights (bi_pt_set_lights,
 * LightsContext, src/renderer/context/lights.cpp:48-95), the sky
 * (bi_pt_set_sky, SkyboxContext / SkyboxSystem, src/scene_basic/skybox.cpp:43-73)
 * and the camera (bi_pt_render, gfx::Camera matrices, src/graphics/camera.cpp:96-110).
 * The output colour texture (rgba16f in the reference) is an RGBA32F device
 * buffer; the optional depth output is reverse-Z device depth as pt_depth.hlsl.
 *
 * Conventions: matrices are row-major (m[row*4 + col]); device pointers are
 * HIP allocations on the context's device; calls on one context are not
 * thread-safe (externally serialised, as the reference records on one thread);
 * work is stream-ordered on the caller's hipStream_t.  Every call returns a
 * bi_pt_status; bi_pt_last_error(ctx) describes the last failure.
 */
#ifndef BI_PT_H
#define BI_PT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* bi_pt_stream;  /* hipStream_t; NULL = default stream */

typedef enum bi_pt_status {
    BI_PT_OK = 0,
    BI_PT_ERROR_INVALID_ARGUMENT = 1,
    BI_PT_ERROR_OUT_OF_MEMORY = 2,     /* device allocation or HBM budget exhausted */
    BI_PT_ERROR_HIP = 3,               /* a HIP runtime call failed */
    BI_PT_ERROR_NO_SCENE = 4,
    BI_PT_ERROR_DEVICE_CHECK = 5,      /* a kernel reported an internal-limit violation */
    BI_PT_ERROR_UNSUPPORTED = 6
} bi_pt_status;

typedef struct bi_pt_ctx bi_pt_ctx;

/* Device selection and memory policy. */
typedef struct bi_pt_device_desc {
    int32_t device;              /* HIP device ordinal */
    uint32_t _pad;
    uint64_t hbm_budget_bytes;   /* 0 = no limit; allocations beyond it fail with OUT_OF_MEMORY */
} bi_pt_device_desc;

/* One mesh (host memory, copied at bi_pt_set_scene): float xyz positions and
 * normals (flat float buffers of 3 per vertex, as the reference's
 * positions_buffer / normals_buffer), uint32 triangle-list indices, and optionally
 * texture coordinates (2 floats per vertex, texcoords_buffer) and tangents (4 floats per
 * vertex: xyz and the bitangent sign w, tangents_buffer; gpu_scene_data.hpp:10-24).
 * Missing texture coordinates read as 0 and missing tangents as (1, 0, 0, 1), the
 * defaults of fetch_vertex_attributes (hit.hlsl:74, 116). */
typedef struct bi_pt_mesh {
    const float* positions;
    const float* normals;
    const uint32_t* indices;
    uint32_t num_vertices;
    uint32_t num_indices;
    const float* texcoords;      /* nullable */
    const float* tangents;       /* nullable */
} bi_pt_mesh;

/* One drawable: 3x4 row-major object-to-world transform, laid out like
 * rhi::AccelerationStructureInstanceDesc::transform (include/bisemutum/rhi/accel.hpp:48-55). */
typedef struct bi_pt_instance {
    float transform[12];
    uint32_t mesh;
    uint32_t material;
} bi_pt_instance;

/* Lit material (SurfaceData of shaders/core/material/utils.hlsl:5-31, as the
 * material_function of a material asset fills it).  128 bytes.  anisotropy follows
 * get_anisotropic_roughness; ior is only used for cos < 0 (never for reflection).
 * pattern BI_PT_PATTERN_CHECKER_XZ: the world-space checkerboard of the example
 * project's material (examples/scene_basic/materials/checkerboard.toml): at a hit
 * point P, cells with (int(floor(P.x)) ^ int(floor(P.z))) odd use this record and
 * even cells use record pattern_material (which must have no pattern itself);
 * floor(P) is clamped to [-2^31, 2^31 - 128] before the conversion.
 * Texture-driven fields (material functions that sample textures at vertex.texcoord,
 * examples/scene_basic/materials/{textured,cage}.toml): base_color_tex / f0_tex set the
 * colour from the texel's rgb, roughness_tex the roughness from one channel,
 * normal_tex the tangent-space normal map value (rgb; N = normalize(n.x T + n.y B +
 * n.z N), n = 2 rgb - 1, hits/rt_gbuffer_hit.hlsl:8-11).  Textures are sampled at level
 * 0 with their own filter and addressing (bi_pt_set_textures).
 * blend_mode (accel.cpp:108-116, hits/rt_gbuffer_hit.hlsl:20-33): an alpha-tested
 * surface is absent where its opacity is below 0.01; a translucent one is absent for a
 * ray whose uniform number (one per ray) is below 1 - opacity.  The opacity is
 * `opacity`, or channel opacity_channel of opacity_tex, stepped to step_below /
 * step_above at step_at when opacity_step is set (cage.toml: value.w < 0.5 ? 0 : 1).
 * Blend mode and opacity come from the triangle's own record (a checkerboard's cells
 * share them).  Texture references hold the texture's index + 1 (BI_PT_NO_TEXTURE = 0:
 * none), so a zero-initialised record is an opaque parameter material. */
enum {
    BI_PT_PATTERN_NONE = 0,
    BI_PT_PATTERN_CHECKER_XZ = 1
};
enum {
    BI_PT_BLEND_OPAQUE = 0,
    BI_PT_BLEND_ALPHA_TEST = 1,
    BI_PT_BLEND_TRANSLUCENT = 2
};
#define BI_PT_NO_TEXTURE 0u
typedef struct bi_pt_material {
    float base_color[3];
    float roughness;
    float f0_color[3];
    float anisotropy;
    float f90_color[3];
    float ior;
    uint32_t two_sided;
    uint32_t pattern;            /* BI_PT_PATTERN_* (0 for a parameter material) */
    uint32_t pattern_material;   /* the record of even cells */
    uint32_t blend_mode;         /* BI_PT_BLEND_* */
    float opacity;
    uint32_t opacity_tex;
    uint32_t opacity_channel;    /* 0..3 = r, g, b, a */
    uint32_t opacity_step;
    float step_at, step_below, step_above;
    uint32_t base_color_tex;
    uint32_t f0_tex;
    uint32_t normal_tex;
    uint32_t roughness_tex;
    uint32_t roughness_channel;  /* 0..3 */
    uint32_t _pad[4];
} bi_pt_material;

/* A material texture (host memory, copied at bi_pt_set_textures): level 0 of the
 * engine's texture asset (TextureAsset, src/scene_basic/texture.cpp:92-136) with its
 * sampler's level-0 filter and addressing (rhi::SamplerDesc, include/bisemutum/rhi/sampler.hpp:37-53).
 * texels: width * height, row-major from v = 0; RGBA8 as 4 bytes per texel (UNORM,
 * or sRGB decoded to linear before filtering), RGBA32F as 4 floats. */
enum {
    BI_PT_TEXEL_RGBA8_UNORM = 0,
    BI_PT_TEXEL_RGBA8_SRGB = 1,
    BI_PT_TEXEL_RGBA32F = 2
};
enum {
    BI_PT_ADDRESS_REPEAT = 0,
    BI_PT_ADDRESS_MIRRORED_REPEAT = 1,
    BI_PT_ADDRESS_CLAMP_TO_EDGE = 2
};
typedef struct bi_pt_texture {
    const void* texels;
    uint32_t width, height;
    uint32_t format;             /* BI_PT_TEXEL_* */
    uint32_t filter;             /* 0 nearest, 1 linear */
    uint32_t address_u, address_v;   /* BI_PT_ADDRESS_* */
} bi_pt_texture;

/* Byte-identical to DirLightData (shaders/renderer/lights_struct.hlsl:3-14,
 * src/renderer/context/lights.hpp:12-22).  64 bytes.  emission = colour*strength,
 * direction points TO the light.  sm_index < 0: casts no shadow (no shadow ray);
 * otherwise an exact shadow ray replaces the shadow map and an occluded sample
 * keeps (1 - shadow_strength) of the light (lights.hlsl:62-63).  The cascade and
 * bias fields are accepted and ignored. */
typedef struct bi_pt_dir_light {
    float emission[3];
    int32_t sm_index;
    float direction[3];
    float shadow_strength;
    float cascade_shadow_radius_sqr[4];
    float shadow_depth_bias;
    float shadow_normal_bias;
    float _pad1[2];
} bi_pt_dir_light;

/* Camera: matrices as gfx::Camera computes them (camera.cpp:96-110). */
typedef struct bi_pt_camera {
    float matrix_inv_view[16];
    float matrix_inv_proj[16];
    float matrix_proj_view[16];
    uint32_t width;
    uint32_t height;
} bi_pt_camera;

enum {
    BI_PT_FLAG_JITTER = 1u << 0,         /* jitter primary rays inside the pixel (reference: pixel centre) */
    BI_PT_FLAG_NO_SKY_NEE = 1u << 1,     /* sky only on BSDF-sampled misses, as the reference */
    BI_PT_FLAG_OUTPUT_RGBA16F = 1u << 2, /* the colour output is RGBA16F (4 IEEE halves per pixel, rounded
                                          * to nearest even), the reference's OutputData.color format
                                          * (path_tracing.hpp:19-23); else RGBA32F */
    BI_PT_FLAG_RECORD_QUEUES = 1u << 8,  /* keep every bounce's queue contents for bi_pt_debug_queue */
    BI_PT_FLAG_TIMING = 1u << 9,         /* time each kernel class with HIP events (bi_pt_get_stats) */
    BI_PT_FLAG_REUSE_BVH = 1u << 10      /* keep the built LBVH while the scene and transforms are unchanged */
};

/* PathTracingSettings (basic.hpp:76-81) plus the sample schedule.
 * max_bounces keeps the reference's meaning: clamped to [2,16], the pass traces
 * max_bounces-1 path segments.  accumulate: keep summing into the running image
 * while the camera is unchanged (pt_accumulate.hlsl); otherwise restart. */
typedef struct bi_pt_settings {
    float ray_length;         /* TMax of camera/extension rays; farther geometry reads as sky */
    uint32_t max_bounces;
    uint32_t accumulate;
    uint32_t spp;             /* samples per pixel rendered by this call */
    uint32_t seed;            /* PCG32 streams keyed by (seed, pixel, sample, bounce) */
    uint32_t flags;
    uint32_t wave_paths;      /* paths in flight per wave; 0 = 2^25 (about 270 B of queues per path), at most 2^28 */
    uint32_t stack_lds;       /* traversal stack entries kept in LDS (0 = automatic, <= 32) */
    uint32_t refill_lanes;    /* idle lanes before a wavefront fetches new rays (0 = automatic, <= 64) */
    uint32_t work_parts;      /* ray-queue parts with their own fetch counter (0 = automatic, <= 32) */
    uint32_t block_threads;   /* traversal workgroup size: 128, 256 or 512 (0 = 256) */
} bi_pt_settings;

typedef struct bi_pt_stats {
    /* traced rays, cumulative over the context's lifetime */
    uint64_t rays_camera;
    uint64_t rays_extension;
    uint64_t rays_light_shadow;
    uint64_t rays_sky_shadow;
    uint64_t frames;
    uint64_t samples_accumulated;  /* per pixel, in the running image */
    /* milliseconds of the last render call (BI_PT_FLAG_TIMING), per kernel class */
    float ms_build;
    float ms_trace_closest;
    float ms_shade;
    float ms_trace_shadow;
    float ms_resolve;
    float ms_total;
    uint32_t launches_trace_closest;
    uint32_t launches_trace_shadow;
} bi_pt_stats;

/* ---- lifecycle (PathTracingPass ctor, path_tracing.cpp:128-181) ---- */
bi_pt_status bi_pt_create(const bi_pt_device_desc* desc, bi_pt_ctx** out_ctx);
bi_pt_status bi_pt_destroy(bi_pt_ctx* ctx);
const char* bi_pt_last_error(const bi_pt_ctx* ctx);
/* A view renders the scene (meshes, instances, materials, lights, sky and LBVH) of
 * `scene` into a running image of its own: several cameras or render jobs of one
 * resident scene.  The view's scene inputs cannot be set; changing the scene context's
 * inputs restarts its views' running images.  The scene context must outlive its views.
 * (No reference counterpart: the reference keeps one history per camera,
 * camera.cpp:176-204.) */
bi_pt_status bi_pt_create_view(bi_pt_ctx* scene, bi_pt_ctx** out_view);

/* ---- inputs (update_params, path_tracing.cpp:183-222, and the TLAS/scene inputs) ---- */
bi_pt_status bi_pt_set_scene(bi_pt_ctx* ctx, const bi_pt_mesh* meshes, uint32_t num_meshes,
                             const bi_pt_instance* instances, uint32_t num_instances,
                             const bi_pt_material* materials, uint32_t num_materials);
/* The engine's GPU scene as its ray-tracing shaders read it (GpuSceneData,
 * src/graphics/gpu_scene_data.hpp:10-24): vertex attributes of every mesh in shared flat
 * buffers (positions and normals 3 floats per vertex, tangents 4, texture coordinates 2;
 * uint32 indices), here host copies, with their lengths in elements.  Nullable but
 * positions and indices. */
typedef struct bi_pt_gpu_scene {
    const float* positions;
    uint64_t num_positions;
    const float* normals;
    uint64_t num_normals;
    const float* tangents;
    uint64_t num_tangents;
    const float* texcoords;
    uint64_t num_texcoords;
    const uint32_t* indices;
    uint64_t num_indices;
} bi_pt_gpu_scene;
/* Vertex attributes a mesh has (VertexAttributesType, include/bisemutum/graphics/
 * vertex_attributes_type.hpp:7-19); the hit shader reads only those (hit.hlsl:55-136). */
enum {
    BI_PT_VA_POSITION = 0x1,
    BI_PT_VA_NORMAL = 0x4,
    BI_PT_VA_TANGENT = 0x8,
    BI_PT_VA_TEXCOORD = 0x20
};
/* One drawable: its hit-group shader record (DrawableSbtData, src/graphics/
 * drawable_stb_data.hpp:7-17, as graphics_manager.cpp:1314-1330 fills it: element offsets
 * into the flat buffers with the submesh's base vertex included, the submesh's first
 * index), the submesh's index count (its BLAS build input, graphics_manager.cpp:620-640),
 * its mesh's vertex attributes, its material's record in the table passed alongside
 * (material_offset locates the engine's parameter struct, which this pass reads as a
 * bi_pt_material), and its object-to-world matrix as glm holds it (column-major,
 * drawable->transform.matrix()), transposed into the 3x4 instance transform as the TLAS
 * build does (accel.cpp:119-131).  A mesh without normals reads (0, 0, 1), without
 * tangents (1, 0, 0, 1), without texture coordinates (0, 0) (hit.hlsl:54, 74, 117). */
typedef struct bi_pt_drawable {
    uint32_t drawable_index;
    uint32_t position_offset;
    uint32_t normal_offset;
    uint32_t tangent_offset;
    uint32_t color_offset;
    uint32_t texcoord_offset;
    uint32_t texcoord2_offset;
    uint32_t index_offset;
    uint32_t material_offset;
    uint32_t index_count;
    uint32_t attributes;         /* BI_PT_VA_* */
    uint32_t material;           /* index into the material table */
    float matrix[16];            /* glm::mat4, column-major */
} bi_pt_drawable;
/* Where one mesh of the converted scene sits in the flat buffers (element offsets;
 * UINT64_MAX: attribute absent) and its index and vertex counts (vertices: one past the
 * largest index). */
typedef struct bi_pt_submesh_range {
    uint64_t position_offset, normal_offset, tangent_offset, texcoord_offset, index_offset;
    uint32_t index_count, num_vertices;
} bi_pt_submesh_range;
/* The scene layout of drawables (host only, no device work): one range per distinct
 * submesh record (drawables that share one share it) and one instance per drawable
 * (transposed transform, range index, material).  ranges and instances hold n entries;
 * *num_ranges receives the count used.  INVALID_ARGUMENT when a record reaches past its
 * buffer, an index count is not a multiple of 3, or an attribute the drawable names has no
 * buffer. */
bi_pt_status bi_pt_drawables_layout(const bi_pt_gpu_scene* scene, const bi_pt_drawable* drawables, uint32_t n,
                                    bi_pt_submesh_range* ranges, uint32_t* num_ranges, bi_pt_instance* instances);
/* bi_pt_set_scene from the engine's GPU-scene buffers and drawables (the inputs the
 * reference's TLAS and hit shaders read, gpu_scene_data.hpp:10-24, accel.cpp:103-132). */
bi_pt_status bi_pt_set_scene_drawables(bi_pt_ctx* ctx, const bi_pt_gpu_scene* scene, const bi_pt_drawable* drawables,
                                       uint32_t num_drawables, const bi_pt_material* materials,
                                       uint32_t num_materials);
/* Replace instance transforms only (same count/meshes/materials); the LBVH is rebuilt by the next render. */
bi_pt_status bi_pt_update_transforms(bi_pt_ctx* ctx, const float* transforms_3x4, uint32_t num_instances);
/* The material textures of the scene (the material table names them by index); replaces
 * the previous set.  A render fails (INVALID_ARGUMENT) while a material names a texture
 * that is not set. */
bi_pt_status bi_pt_set_textures(bi_pt_ctx* ctx, const bi_pt_texture* textures, uint32_t num_textures);
bi_pt_status bi_pt_set_lights(bi_pt_ctx* ctx, const bi_pt_dir_light* lights, uint32_t num_lights);
/* Byte-identical to PointLightData (shaders/renderer/lights_struct.hlsl:16-28, filled by
 * LightsContext::collect_all_lights, src/renderer/context/lights.cpp:120-150).  64 bytes.
 * Radiance at a point at distance d: emission * saturate(1 - (d^2 range_sqr_inv)^2) /
 * max(d^2, 0.001), times the spot factor (clamp(dot(l, direction), cos_outer, cos_inner) -
 * cos_outer) / max(cos_inner - cos_outer, 0.001) when cos_inner > cos_outer (l: unit
 * vector to the light; direction: the spot axis pointing to the light) (point_light_eval,
 * lights.hlsl:15-25).  sm_index < 0: no shadow; otherwise an exact shadow ray to the
 * light's position replaces the cube shadow map, and an occluded sample keeps
 * (1 - shadow_strength). */
typedef struct bi_pt_point_light {
    float emission[3];
    float range_sqr_inv;
    float position[3];
    float cos_inner;
    float direction[3];
    float cos_outer;
    int32_t sm_index;
    float shadow_strength;
    float shadow_depth_bias;
    float shadow_normal_bias;
} bi_pt_point_light;
/* Point and spot lights; directional and point lights together are at most 4. */
bi_pt_status bi_pt_set_point_lights(bi_pt_ctx* ctx, const bi_pt_point_light* lights, uint32_t num_lights);

/* RectLightData (shaders/renderer/lights_struct.hlsl:38-53), 112 bytes, as
 * LightsContext::collect_all_lights fills it (src/renderer/context/lights.cpp:204-242):
 * corners position0..3 = centre + (w, h), (-w, h), (-w, -h), (w, -h) along the light's x and
 * y axes (w, h the half width and height), normal its z axis, emission colour * strength.
 * Rect lights cast no shadows (deferred_lighting_secondary.hlsl:73-96 has no shadow term);
 * each shaded vertex adds their linearly-transformed-cosine lighting. */
typedef struct bi_pt_rect_light {
    float emission[3];
    int32_t texture_index;    /* < 0: untextured; else a texture of bi_pt_set_rect_light_textures */
    float center_position[3];
    uint32_t two_sided;
    float position0[3];
    float inv_width_sqr;
    float position1[3];
    float inv_height_sqr;
    float position2[3];
    float inv_texel_size;
    float position3[3];
    float pad1;
    float normal[3];
    float pad2;
} bi_pt_rect_light;

/* Replaces LightsContext's rect_lights buffer (lights.cpp:204-243): at most 16; rendering
 * with rect lights needs the LTC tables (bi_pt_set_ltc_luts), and a textured light
 * (texture_index >= 0, inv_texel_size = max(texture width / light width, texture height /
 * light height) as lights.cpp:233-236 sets it) needs that texture at render time.
 * BI_PT_ERROR_UNSUPPORTED for more than 16 lights. */
bi_pt_status bi_pt_set_rect_lights(bi_pt_ctx* ctx, const bi_pt_rect_light* lights, uint32_t num_lights);

/* A rect light's texture with its sampler (LightsContext's rect_light_textures and
 * rect_light_samplers, lights.cpp:229-238,261-268; rhi::SamplerDesc,
 * include/bisemutum/rhi/sampler.hpp:37-51): num_levels mip levels one after the other in
 * host memory, level 0 first, level l max(1, width >> l) x max(1, height >> l) texels
 * (num_levels <= floor(log2(max(width, height))) + 1).  Filters 0 nearest / 1 linear. */
typedef struct bi_pt_rect_texture {
    const void* texels;
    uint32_t width, height, num_levels;
    uint32_t format;                 /* BI_PT_TEXEL_* */
    uint32_t mag_filter, min_filter, mipmap_mode;
    uint32_t address_u, address_v;   /* BI_PT_ADDRESS_* */
    float lod_bias, lod_min, lod_max;
} bi_pt_rect_texture;
/* The textures rect lights sample along their integrals' directions (rect_light_eval_ltc,
 * shaders/renderer/lights.hlsl:495-511: the diffuse term with roughness 1, the specular one
 * with sqrt(rx ry), at the level log2(max(1, distance * roughness * inv_texel_size))
 * (rect_light_sample_texture, lights.hlsl:429-445)); at most 16 (max_num_rect_light_textures,
 * src/renderer/context/lights.hpp:65). */
bi_pt_status bi_pt_set_rect_light_textures(bi_pt_ctx* ctx, const bi_pt_rect_texture* textures, uint32_t num_textures);
/* Host only: the mip chain the engine generates for a texture asset with levels
 * (TextureAsset::update_gpu_data, src/scene_basic/texture.cpp:176-178 ->
 * CommandHelpers::generate_mipmaps_2d, src/graphics/command_helpers.cpp:66-199 running
 * shaders/core/mipmap.hlsl): each level max(1, w / 2) x max(1, h / 2) texels, the average of
 * the 2 x 2 source texels, widened to 3 where the destination's width or height is odd,
 * source texels beyond the level read as zero (robust image access); averaged in FP32,
 * RGBA8 stored rounded to nearest, sRGB colour decoded before and encoded after.  chain
 * receives all num_levels levels (level 0 copied), laid out as bi_pt_rect_texture reads. */
bi_pt_status bi_pt_rect_texture_mips(uint32_t format, uint32_t width, uint32_t height, uint32_t num_levels,
                                     const void* level0, void* chain);

/* The LTC look-up tables LightsContext loads (lights.cpp:26-37, the engine's
 * ltc_matrix_lut0..2 and ltc_norm_lut texture assets): 8 x 8 x 64 texels each, x fastest,
 * then y, then z; the three matrix tables RGBA32F (4 floats a texel, rows of the transform
 * in xyz), the norm table RG32F (2 floats a texel).  Sampled with linear filtering and
 * clamp to edge, as the engine's sampler. */
bi_pt_status bi_pt_set_ltc_luts(bi_pt_ctx* ctx, const float* matrix_lut0, const float* matrix_lut1,
                                const float* matrix_lut2, const float* norm_lut);
/* Equirect RGBA32F sky (host memory, row 0 = +Y pole), colour multiplier
 * (SkyboxSystem color*strength) and 3x3 row-major rotation applied to directions
 * before lookup (deferred_lighting_secondary.hlsl:25). */
bi_pt_status bi_pt_set_sky(bi_pt_ctx* ctx, const float* rgba32f, uint32_t width, uint32_t height,
                           const float color[3], const float rotation_3x3[9]);
/* Cube-map sky: the engine's skybox texture (SkyboxSystem::current_skybox,
 * src/scene_basic/skybox.cpp:56-67: a 2D texture of width == height with six RGBA32F
 * layers, made by shaders/core/equitangular_to_cubemap.hlsl:20-34), host memory, six
 * face_size x face_size faces one after the other in the layer order +X, -X, +Y, -Y,
 * +Z, -Z, rows from v = 0, orientation of shaders/core/utils/cubemap.hlsl:3-21.
 * Looked up as deferred_lighting_secondary.hlsl:25-26 samples its TextureCube with the
 * linear sampler of src/renderer/context/skybox.cpp:31-37: level 0, bilinear, across
 * face edges seamlessly as Vulkan filters cube maps (an out-of-face tap reads the
 * adjacent face's texel; beyond a corner the mean of the three corner texels).  Importance sampling
 * uses a per-texel table (luminance times solid angle).  face_size 0 clears the sky;
 * at most 8192.  Replaces any equirect sky. */
bi_pt_status bi_pt_set_sky_cube(bi_pt_ctx* ctx, const float* rgba32f_faces, uint32_t face_size,
                                const float color[3], const float rotation_3x3[9]);

/* ---- the pass (PathTracingPass::render, path_tracing.cpp:224-488) ----
 * Rebuilds the LBVH, traces settings->spp samples per pixel and writes the
 * running mean to out_rgba32f (device, width*height*4 floats; with
 * BI_PT_FLAG_OUTPUT_RGBA16F width*height*4 halves at the same pointer).  out_depth
 * (device, width*height floats, nullable) receives the primary hit's reverse-Z
 * depth (0 on miss).  Stream-ordered; returns once the work is enqueued
 * (unless statistics or queue recording require a synchronisation). */
bi_pt_status bi_pt_render(bi_pt_ctx* ctx, const bi_pt_camera* camera, const bi_pt_settings* settings,
                          float* out_rgba32f, float* out_depth, bi_pt_stream stream);

/* Primary-hit GBuffer of a render (OutputData.gbuffer, path_tracing.cpp:264-265,353-358,
 * 482-487): for the first sample the call traces per pixel (the sample whose depth
 * out_depth holds), the surface its camera ray hits, packed as pack_surface_to_gbuffer
 * (shaders/renderer/gbuffer.hlsl:18-34) packs it after the closest-hit shader's normal
 * map and two-sided flip (hits/rt_gbuffer_hit.hlsl:7-17), in the formats of
 * GBufferTextures (src/renderer/context/gbuffer.hpp:14-17).  Device images of
 * width*height texels, row-major; each pointer nullable.  A pixel whose ray misses holds
 * zeros.
 *   base_color        rgba16f (4 half words): base colour, 1
 *   normal_roughness  rgba16f: octahedral normal (x, y), the tangent's direction in the
 *                     normal's frame (frame.hlsl:9-20) as pack_normal_and_tangent's third
 *                     component (pack.hlsl:102-110), roughness
 *   fresnel           rgba16 unorm (4 uint16): f0 then f90, each an rg11b10 word split into
 *                     two unorm16 halves (pack.hlsl:6-16; the word's halves over 65536,
 *                     stored as the exact unorm16 value)
 *   material_0        rgba8 unorm (4 bytes): anisotropy, 1/ior, 0, surface model (lit) / 256
 * Tangents are the meshes' (default (1, 0, 0, 1) in object space), so the first call that
 * asks for a GBuffer builds the scene's per-leaf tangents and keeps building them. */
typedef struct bi_pt_gbuffer {
    uint16_t* base_color;
    uint16_t* normal_roughness;
    uint16_t* fresnel;
    uint8_t* material_0;
} bi_pt_gbuffer;
bi_pt_status bi_pt_render_gbuffer(bi_pt_ctx* ctx, const bi_pt_camera* camera, const bi_pt_settings* settings,
                                  float* out_rgba32f, float* out_depth, const bi_pt_gbuffer* gbuffer,
                                  bi_pt_stream stream);
bi_pt_status bi_pt_reset_accumulation(bi_pt_ctx* ctx);
/* How the per-frame BVH is built (no reference counterpart: the reference's BVH is the
 * driver's).  After the LBVH (Morton codes, radix sort, Karras hierarchy, refit), rounds
 * of treelet restructuring lower the tree's surface-area cost (7-leaf treelets, optimal
 * topology by dynamic programming; Karras & Aila 2013).  Closest hits do not depend on
 * the tree, so images are the same for any setting; only the speed changes. */
typedef struct bi_pt_build_options {
    uint32_t treelet_rounds;   /* 0 = the Morton/Karras tree as built; default 3; at most 16 */
    uint32_t _reserved[3];
} bi_pt_build_options;
bi_pt_status bi_pt_set_build_options(bi_pt_ctx* ctx, const bi_pt_build_options* options);
/* Statistics; returns BI_PT_ERROR_DEVICE_CHECK when a kernel of any call on this context
 * since the last bi_pt_get_stats raised an internal-limit flag (the flags are kept until
 * this call reports them, then cleared). */
bi_pt_status bi_pt_get_stats(const bi_pt_ctx* ctx, bi_pt_stats* out);

/* ---- batched rendering (no reference counterpart: the reference renders one view per
 * pass invocation; this serves several render jobs at once) ----
 * A batch renders the jobs of several contexts in one set of waves: their ray queues
 * are shared, every kernel launch carries rays of all jobs.  Each job is exactly what
 * bi_pt_render(job.ctx, job.camera, job.settings, ...) would produce, bit for bit
 * (samples are added to the context's running image in the same order).  A context
 * may appear once per call.  The queue and traversal knobs of the jobs' settings are
 * ignored: the batch's tuning applies. */
typedef struct bi_pt_batch bi_pt_batch;
typedef struct bi_pt_job {
    bi_pt_ctx* ctx;                    /* scene, lights, sky and running image of the job */
    const bi_pt_camera* camera;
    const bi_pt_settings* settings;
    float* out_rgba32f;                /* device, width*height*4 */
    float* out_depth;                  /* device, width*height, nullable */
} bi_pt_job;
typedef struct bi_pt_tuning {
    uint32_t wave_paths;    /* paths in flight per wave (0 = automatic) */
    uint32_t stack_lds;     /* traversal stack entries in LDS (0 = automatic, <= 32) */
    uint32_t refill_lanes;  /* idle lanes before a wavefront fetches rays (0 = automatic, <= 64) */
    uint32_t work_parts;    /* ray-queue parts with their own fetch counter (0 = automatic, <= 32) */
    uint32_t use_graph;     /* capture a call's launches in a hipGraph and replay it while the
                               jobs' shapes (sizes, spp, bounces, lights, scenes) repeat */
    uint32_t block_threads; /* traversal workgroup size: 128, 256 or 512 (0 = 256) */
} bi_pt_tuning;
bi_pt_status bi_pt_batch_create(const bi_pt_device_desc* desc, bi_pt_batch** out_batch);
bi_pt_status bi_pt_batch_destroy(bi_pt_batch* batch);
const char* bi_pt_batch_last_error(const bi_pt_batch* batch);
bi_pt_status bi_pt_batch_set_tuning(bi_pt_batch* batch, const bi_pt_tuning* tuning);
bi_pt_status bi_pt_batch_get_tuning(const bi_pt_batch* batch, bi_pt_tuning* tuning);
/* Renders every job; stream-ordered.  BI_PT_FLAG_REUSE_BVH in a job's flags keeps that
 * context's built LBVH; BI_PT_FLAG_TIMING / RECORD_QUEUES are ignored here. */
bi_pt_status bi_pt_render_batch(bi_pt_batch* batch, const bi_pt_job* jobs, uint32_t num_jobs, bi_pt_stream stream);
/* Ray statistics of the batch (all its calls) and the device error flags raised by any of
 * its calls since the last check (which clears them). */
bi_pt_status bi_pt_batch_get_stats(bi_pt_batch* batch, bi_pt_stats* out);
/* Picks the batch's tuning for these jobs by timing candidate knob settings (wave
 * capacity, traversal block shape, LDS stack depth, refill threshold) on them, `rounds` calls each; writes the
 * best to *best (nullable) and installs it.  The jobs' running images are restarted. */
bi_pt_status bi_pt_batch_autotune(bi_pt_batch* batch, const bi_pt_job* jobs, uint32_t num_jobs, uint32_t rounds,
                                  bi_pt_stream stream, bi_pt_tuning* best);

/* ---- serving loop (no reference counterpart: the reference renders one camera per frame,
 * basic.cpp:157-161) ----
 * Render jobs of uneven scene size, resolution and spp are submitted at any time.  A
 * job is admitted once the HBM budget has room for its running image and output (the
 * resident scenes and the wave buffers are counted too); each bi_pt_server_step renders
 * the next samples of every admitted job in one batched call (shared ray queues), so
 * jobs join and leave at round boundaries.  A finished job's image equals, bit for bit,
 * a bi_pt_render of the same camera and settings on a context holding its scene. */
typedef struct bi_pt_server bi_pt_server;
typedef struct bi_pt_server_desc {
    int32_t device;
    uint32_t max_jobs_per_round;   /* 0 = 64 */
    uint64_t hbm_budget_bytes;     /* 0 = unlimited */
    uint32_t spp_per_round;        /* samples of each job per round; 0 = 16 */
    uint32_t wave_paths;           /* paths in flight per wave; 0 = 2^25 (the budget reserves its queues) */
} bi_pt_server_desc;
typedef struct bi_pt_request {
    uint32_t scene_id;             /* from bi_pt_server_add_scene */
    uint32_t _pad;
    double arrival_s;              /* arrival time on the server clock; < 0 = now */
    bi_pt_camera camera;
    bi_pt_settings settings;       /* spp = the job's total samples; accumulate and queue knobs ignored;
                                      checked at submission as bi_pt_render checks them, and
                                      BI_PT_FLAG_OUTPUT_RGBA16F is rejected (results are RGBA32F) */
} bi_pt_request;
/* A job whose round failed (a HIP error or a kernel's internal-limit flag) leaves with
 * BI_PT_JOB_FAILED; bi_pt_server_step returns that round's error, the other jobs go on. */
enum { BI_PT_JOB_WAITING = 0, BI_PT_JOB_RUNNING = 1, BI_PT_JOB_DONE = 2, BI_PT_JOB_FAILED = 3 };
typedef struct bi_pt_job_status {
    uint32_t state;
    uint32_t spp_done;
    double submit_s, admit_s, done_s;   /* server clock, seconds */
} bi_pt_job_status;
typedef struct bi_pt_server_stats {
    uint64_t rays;                 /* all traced rays */
    uint64_t rounds;
    uint32_t waiting, running;
    uint64_t hbm_used;
} bi_pt_server_stats;
bi_pt_status bi_pt_server_create(const bi_pt_server_desc* desc, bi_pt_server** out_server);
bi_pt_status bi_pt_server_destroy(bi_pt_server* server);
const char* bi_pt_server_last_error(const bi_pt_server* server);
double bi_pt_server_clock(const bi_pt_server* server);
/* Makes a context's scene resident (its LBVH is built now and kept); the context must
 * outlive the server and keep its scene unchanged. */
bi_pt_status bi_pt_server_add_scene(bi_pt_server* server, bi_pt_ctx* scene, uint32_t* scene_id);
bi_pt_status bi_pt_server_submit(bi_pt_server* server, const bi_pt_request* request, uint64_t* job_id);
/* One scheduling round (admission, one batched call, retirement); waits for the round.
 * *active (nullable) = jobs still waiting or running. */
bi_pt_status bi_pt_server_step(bi_pt_server* server, bi_pt_stream stream, uint32_t* active);
bi_pt_status bi_pt_server_poll(bi_pt_server* server, uint64_t job_id, bi_pt_job_status* status);
/* Copies a finished job's RGBA32F image to host memory (nullable: discard) and frees the job.
 * For a failed job: returns the status of the round that failed it and frees the job. */
bi_pt_status bi_pt_server_take_result(bi_pt_server* server, uint64_t job_id, float* host_rgba,
                                      uint64_t capacity_floats);
bi_pt_status bi_pt_server_get_stats(bi_pt_server* server, bi_pt_server_stats* out);

/* ---- neural radiance cache (no reference counterpart: the reference traces every
 * path to its last segment, path_tracing.cpp:290-343) ----
 * A small MLP (64 inputs, 4 hidden layers of 64 with ReLU, RGB out; BF16 on the matrix
 * cores with FP32 accumulation) predicts the radiance leaving a surface point.  In
 * bi_pt_render_nrc the image's paths end in the cache at segment `cache_bounce`; the
 * same waves trace `train_samples` extra paths from random pixels to the full depth,
 * whose radiance gathered after that segment (per unit throughput) is the target of
 * their sample there; after the frame the cache takes `train_steps` Adam steps on those
 * samples (relative L2 loss), so it trains online on the tracer's own path suffixes.
 * Until `use_after_steps` steps have been taken the image is rendered plainly. */
typedef struct bi_pt_nrc bi_pt_nrc;
typedef struct bi_pt_nrc_desc {
    uint32_t train_samples;       /* training paths per frame; 0 = 65,536 (a multiple of 64 * train_steps) */
    uint32_t train_steps;         /* Adam steps per frame; 0 = 4 */
    uint32_t cache_bounce;        /* segment at which paths end in the cache; 0 = 2 */
    float learning_rate;          /* 0 = 3e-3 (beta1 0.9, beta2 0.99, eps 1e-15) */
    uint32_t use_after_steps;     /* plain rendering before this many training steps */
    uint32_t seed;                /* initial weights */
    uint32_t flags;               /* BI_PT_NRC_FREEZE: render with the cache, no training */
    float target_clamp;           /* a frame's training targets are clamped, per channel, to this
                                     multiple of their mean luminance, so that a single firefly
                                     does not dominate the loss; 0 = 16; < 0: no clamp */
} bi_pt_nrc_desc;
enum { BI_PT_NRC_FREEZE = 1u << 0 };
typedef struct bi_pt_nrc_stats {
    uint64_t frames, steps;
    float last_loss;              /* mean relative L2 of the last frame's (or step's) valid samples */
    uint32_t _pad;
    uint64_t last_valid_samples;  /* training paths that reached their sample */
} bi_pt_nrc_stats;
/* The cache renders the scene of `ctx` into ctx's running image. */
bi_pt_status bi_pt_nrc_create(bi_pt_ctx* ctx, const bi_pt_nrc_desc* desc, bi_pt_nrc** out_nrc);
bi_pt_status bi_pt_nrc_destroy(bi_pt_nrc* nrc);
const char* bi_pt_nrc_last_error(const bi_pt_nrc* nrc);
/* bi_pt_render with the cache (settings and outputs as there), then one frame of training. */
bi_pt_status bi_pt_render_nrc(bi_pt_nrc* nrc, const bi_pt_camera* camera, const bi_pt_settings* settings,
                              float* out_rgba32f, float* out_depth, bi_pt_stream stream);
bi_pt_status bi_pt_nrc_get_stats(bi_pt_nrc* nrc, bi_pt_nrc_stats* out);
/* FP32 master weights, 17,408 floats: W_l[k][i] row-major for l = 0..3 (64 x 64), then
 * W_4 (64 x 16, outputs 3..15 zero).  Setting them restarts Adam. */
bi_pt_status bi_pt_nrc_get_params(bi_pt_nrc* nrc, float* params, uint32_t count);
bi_pt_status bi_pt_nrc_set_params(bi_pt_nrc* nrc, const float* params, uint32_t count);
/* Box that normalises positions (lo.xyz, hi.xyz); each render sets it to the scene's
 * triangle-centroid bounds. */
bi_pt_status bi_pt_nrc_set_bounds(bi_pt_nrc* nrc, const float lo_hi[6]);
/* Direct use of the network on device records (64 B each: position, direction towards
 * the viewer, normal, roughness, base colour, tag; a training sample counts when tag
 * == 1): raw outputs (3 floats per record, device), or one Adam step on count samples
 * with device targets (3 floats each). */
bi_pt_status bi_pt_nrc_infer(bi_pt_nrc* nrc, const void* records, uint32_t count, float* out_rgb,
                             bi_pt_stream stream);
bi_pt_status bi_pt_nrc_train(bi_pt_nrc* nrc, const void* records, const float* targets, uint32_t count,
                             bi_pt_stream stream);
/* Gradient of the last step's loss (the sum of its chunk partials), 17,408 floats. */
bi_pt_status bi_pt_nrc_debug_gradient(bi_pt_nrc* nrc, float* grad, uint32_t count);
/* The last frame's first `count` training samples: records (16 floats each) and
 * targets (3 floats each), to host memory. */
bi_pt_status bi_pt_nrc_debug_samples(bi_pt_nrc* nrc, float* records, float* targets, uint32_t count);

/* ---- ray queries for the engine's other ray-tracing consumers ----
 * The pass's traversal kernels with rays made by the caller: RT reflections
 * (src/renderer/pass/reflection.cpp:161-170,384-404), DDGI probe tracing
 * (src/renderer/pass/ddgi_update.cpp:229-256, shaders/renderer/ddgi/trace_gbuffer.hlsl:10-51)
 * and RTAO (shaders/renderer/ambient_occlusion/ambient_occlusion_rt.hlsl:43-63) each build
 * rays and call TraceRay against the scene's TLAS; here they pass device arrays instead.
 * Rays: n origins and directions as six device arrays (directions need not be unit
 * length; t is in units of the direction), TMin = 0.001 as in the reference's passes, one
 * TMax for all rays (0.001 < tmax <= 1.7e38).  The context's LBVH is built on `stream` if
 * the scene changed since the last build (order queries after other streams' renders of
 * the same context).  Closest hits are those of the path tracer (independent of the tree
 * and of traversal order).  Errors of the traversal (stack bound) are reported by
 * bi_pt_get_stats like a render's.  Stream-ordered, no host synchronisation. */
typedef struct bi_pt_rays {
    const float *ox, *oy, *oz;   /* device, n each */
    const float *dx, *dy, *dz;   /* device, n each */
} bi_pt_rays;
/* Closest hit: hit_t (+inf on a miss, nullable), hit_triangle (the triangle's index in the
 * scene's instance order, as bi_pt_set_scene lays them out; 0xffffffff on a miss),
 * barycentrics hit_u, hit_v of the second and third vertex (nullable).  All device, n each. */
bi_pt_status bi_pt_trace_closest(bi_pt_ctx* ctx, uint32_t n, const bi_pt_rays* rays, float tmax, float* hit_t,
                                 uint32_t* hit_triangle, float* hit_u, float* hit_v, bi_pt_stream stream);
/* bi_pt_trace_closest for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  Same arguments and results; its scratch (job record, part counters,
 * one 16-B hit record per ray) is a static device array of the library (4 MiB + 3 KiB,
 * held by each GPU the library runs on until the process ends), so at most
 * BI_PT_NOALLOC_MAX_RAYS rays per call (more: INVALID_ARGUMENT), and it builds nothing:
 * the scene's LBVH must be current (rendered or queried since the scene last changed;
 * else UNSUPPORTED).  Calls on one GPU share the scratch: calls that overlap on one GPU
 * race, so serialise them (for example on one stream). */
#define BI_PT_NOALLOC_MAX_RAYS (1u << 18)
bi_pt_status bi_pt_trace_closest_noalloc(bi_pt_ctx* ctx, uint32_t n, const bi_pt_rays* rays, float tmax,
                                         float* hit_t, uint32_t* hit_triangle, float* hit_u, float* hit_v,
                                         bi_pt_stream stream);
/* Occlusion (any hit in [0.001, tmax]): occluded[i] = 1 or 0 (device, n). */
bi_pt_status bi_pt_trace_occlusion(bi_pt_ctx* ctx, uint32_t n, const bi_pt_rays* rays, float tmax,
                                   uint32_t* occluded, bi_pt_stream stream);

/* ---- inspection hooks for parity tests (no reference counterpart) ----
 * Copy the last-built LBVH to host memory: morton[n] (input triangle order),
 * order[n] (sorted permutation), children[2*max(n-1,1)] (internal k >= 0, leaf j as ~j),
 * child_bounds[12*max(n-1,1)] (per node: child0 lo.xyz hi.xyz, child1 lo.xyz hi.xyz),
 * parents[max(n-1,1)+n] (internal nodes, then leaves; root -1), centroid_bounds[6]. */
bi_pt_status bi_pt_debug_get_bvh(bi_pt_ctx* ctx, uint32_t* morton, uint32_t* order, int32_t* children,
                                 float* child_bounds, int32_t* parents, float* centroid_bounds);
/* Copy the BVH4 traversal nodes built from the last LBVH: for each BVH2 internal node
 * k (max(n-1,1) of them), child4[4k..4k+3] (internal: node index, leaf j as ~j, empty
 * slot 0) and box4[24k..24k+23] (lo.x[4], lo.y[4], lo.z[4], hi.x[4], hi.y[4], hi.z[4];
 * an empty slot holds +inf). */
bi_pt_status bi_pt_debug_get_bvh4(bi_pt_ctx* ctx, int32_t* child4, float* box4);
uint32_t bi_pt_debug_num_tris(const bi_pt_ctx* ctx);
/* Queue contents recorded by the last render with BI_PT_FLAG_RECORD_QUEUES:
 * kind 0 = extension queue entering segment `bounce` (path ids), kind 1 = shadow
 * queue of segment `bounce` (path_id << 3 | ray kind; kinds 0-3 lights, 7 sky).
 * wave selects the wave of the call.  *count receives the length; at most capacity entries copied. */
bi_pt_status bi_pt_debug_queue(bi_pt_ctx* ctx, uint32_t wave, uint32_t bounce, uint32_t kind,
                               uint32_t* out, uint32_t capacity, uint32_t* count);
/* Running per-pixel radiance sums (device -> host, 3 floats per pixel). */
bi_pt_status bi_pt_debug_get_sums(bi_pt_ctx* ctx, float* sums_rgb);
/* Batched traversal of the current LBVH: rays are host memory, 8 floats each
 * (o.xyz, tmin, d.xyz, tmax); closest -> (t, leaf, u, v), any-hit -> leaf = occluded. */
bi_pt_status bi_pt_debug_trace(bi_pt_ctx* ctx, uint32_t n, const float* rays, int any_hit, float* hit_t,
                               uint32_t* hit_leaf, float* hit_u, float* hit_v);
/* Raises device error flags from the host (tests of the error paths: the flags a kernel
 * would raise on an internal-limit violation, without making any kernel violate one).
 * The context's next check (bi_pt_get_stats) or the server's next round reports them. */
bi_pt_status bi_pt_debug_inject_error(bi_pt_ctx* ctx, uint32_t flags);
bi_pt_status bi_pt_server_debug_inject_error(bi_pt_server* server, uint32_t flags);

#ifdef __cplusplus
}
#endif
#endif
