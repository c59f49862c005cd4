/*
 * oracle.h -- CPU restatement of the wavefront path-tracing pass.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and never as the thing measured or shipped.
 *
 * Two restatements live behind this header:
 *   ref32.cpp  scalar FP32 transcription.  Every floating-point operation is
 *              written in the same order as the HIP kernels and built with
 *              -ffp-contract=off, so integer stages (Morton codes, sorted
 *              order, BVH topology, queue contents) and radiance are expected
 *              to match the GPU bit for bit.
 *   ref64.cpp  FP64 oracle of the same integrand (libm transcendentals, its
 *              own median-split BVH, double-precision intersection), used for
 *              convergence checks at high sample counts.
 *
 * Reference files restated (paths relative to bisemutum/):
 *   shaders/renderer/raytracing/direction_sample/generate_camera_ray.hlsl:4-16
 *   shaders/renderer/raytracing/direction_sample/sample_secondary_ray.hlsl:11-69
 *   shaders/renderer/raytracing/rt_gbuffer.hlsl:7-36, hits/rt_gbuffer_hit.hlsl:6-24
 *   shaders/renderer/raytracing/deferred_lighting_secondary.hlsl:11-60
 *   shaders/core/material/lit.hlsl:5-35, material/utils.hlsl:40-121
 *   shaders/core/utils/frame.hlsl:9-20, shaders/core/equitangular_to_cubemap.hlsl:28-33
 *   shaders/core/raytracing/hit.hlsl:27-164 (attribute interpolation)
 *   shaders/renderer/lights.hlsl:10-13,69-101 (directional light, shadow strength)
 *   src/renderer/pass/path_tracing.cpp:224-488 (bounce loop), src/math/math.cpp:5-11
 */
#ifndef BI_PT_ORACLE_H
#define BI_PT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_ctx ora_ctx;

/* Material texture (host memory): level 0, width * height texels from v = 0, format
 * 0 RGBA8 UNORM, 1 RGBA8 sRGB, 2 RGBA32F; filter 0 nearest / 1 linear; addressing 0
 * repeat, 1 mirrored repeat, 2 clamp to edge (as bi_pt_texture). */
typedef struct ora_texture {
    const void* texels;
    uint32_t width, height, format, filter, address_u, address_v;
} ora_texture;

/* Rect-light texture with its mip levels and sampler (as bi_pt_rect_texture): num_levels
 * levels one after the other, level l max(1, width >> l) x max(1, height >> l) texels. */
typedef struct ora_rect_texture {
    const void* texels;
    uint32_t width, height, num_levels, format, mag_filter, min_filter, mipmap_mode, address_u, address_v;
    float lod_bias, lod_min, lod_max;
} ora_rect_texture;

/* Scene description, all host memory.
 * meshes are concatenated: mesh m owns vertices [vtx_off[m], vtx_off[m]+vtx_cnt[m])
 * of positions/normals (3 floats each) and indices [idx_off[m], idx_off[m]+idx_cnt[m])
 * (mesh-local vertex numbers).  Instance i places mesh inst_mesh[i] with the
 * 3x4 row-major transform inst_xform[12*i..] and material inst_mat[i].
 * materials: 32 floats each, layout of bi_pt_material (include/bi_pt.h).
 * dir_lights: 16 floats each, layout of bi_pt_dir_light (= DirLightData).
 * sky: sky_w*sky_h RGBA32F texels (row 0 = theta 0, the +Y pole), or with sky_layout 1
 * six sky_w x sky_w cube faces (+X, -X, +Y, -Y, +Z, -Z, rows from v = 0).
 * texcoords (2 floats per vertex) and tangents (4 per vertex) run parallel to the
 * positions (nullable: 0 and (1, 0, 0, 1)). */
typedef struct ora_scene_desc {
    uint32_t num_meshes;
    const float* positions;
    const float* normals;
    const uint32_t* indices;
    const uint32_t* vtx_off;
    const uint32_t* idx_off;
    const uint32_t* idx_cnt;
    uint32_t num_instances;
    const float* inst_xform;
    const uint32_t* inst_mesh;
    const uint32_t* inst_mat;
    uint32_t num_materials;
    const float* materials;
    uint32_t num_dir_lights;
    const float* dir_lights;
    uint32_t sky_w, sky_h;
    const float* sky_rgba;
    float sky_color[3];
    float sky_rot[9];
    const float* texcoords;
    const float* tangents;
    uint32_t num_textures;
    const ora_texture* textures;
    uint32_t sky_layout;    /* 0 equirect; 1 cube map: sky_w x sky_w faces, sky_h = 6 sky_w */
    uint32_t num_point_lights;
    const float* point_lights;   /* 16 floats each, layout of bi_pt_point_light; lights are the
                                    directional ones then these, at most 4 in all */
    uint32_t num_rect_lights;
    const float* rect_lights;    /* 28 floats each, layout of bi_pt_rect_light */
    const float* ltc_matrix[3];  /* LTC tables as bi_pt_set_ltc_luts: 8 x 8 x 64 texels, 4 floats each */
    const float* ltc_norm;       /* 2 floats a texel */
    uint32_t num_rect_textures;
    const ora_rect_texture* rect_textures;   /* the textures texture_index >= 0 names */
} ora_scene_desc;

/* Camera: row-major 4x4 matrices (m[row*4+col]), as bi_pt_camera. */
typedef struct ora_camera {
    float inv_view[16];
    float inv_proj[16];
    float proj_view[16];
    uint32_t width, height;
} ora_camera;

typedef struct ora_settings {
    float ray_length;
    uint32_t max_bounces;   /* reference semantics: clamped to [2,16]; segments = max_bounces-1 */
    uint32_t seed;
    uint32_t flags;         /* bit0: jitter primary rays, bit1: disable sky NEE,
                               ORA_FLAG_GBUFFER_QUANTIZE: drift mode (below) */
} ora_settings;
/* Drift mode of the FP32 render: every shaded surface goes through the reference's GBuffer
 * packing and unpacking first (half base colour and roughness, rg11b10 f0/f90 in unorm16
 * words, oct-encoded half normal, unorm8 anisotropy), measuring how far the reference's
 * quantisation moves the image from the FP32 path the kernels follow. */
#define ORA_FLAG_GBUFFER_QUANTIZE (1u << 30)

ora_ctx* ora_create(const ora_scene_desc* desc);
/* Rebuilds the BVH with `rounds` rounds of treelet restructuring after the LBVH (the
 * default is 3, as bi_pt_set_build_options; 0 keeps the Morton/Karras tree). */
void ora_set_treelet_rounds(ora_ctx* ctx, uint32_t rounds);
void ora_destroy(ora_ctx* ctx);
uint32_t ora_num_tris(const ora_ctx* ctx);

/* Flattened world-space triangles in instance order: 9 floats (v0,v1,v2),
 * 9 floats (n0,n1,n2, un-normalised cofactor-transformed), material id. */
void ora_get_world(const ora_ctx* ctx, float* tri_v, float* tri_n, uint32_t* tri_mat);

/* LBVH of the FP32 transcription.  morton[n] per triangle (input order),
 * order[n] sorted permutation, children[2*(n-1)] (internal k >= 0, leaf j as ~j),
 * child_bounds[12*(n-1)]: per node, child0 lo.xyz hi.xyz then child1 lo.xyz hi.xyz,
 * parents[2n-1]: parent of internal nodes [0,n-1) then of leaves [n-1, 2n-1). */
void ora_get_bvh(const ora_ctx* ctx, uint32_t* morton, uint32_t* order,
                 int32_t* children, float* child_bounds, int32_t* parents,
                 float* scene_centroid_bounds /* 6 */);

/* BVH4 traversal nodes collapsed from the LBVH, one per internal BVH2 node k
 * (max(n-1,1) of them): child4[4k+j] (internal: node index, leaf j as ~j, empty slot 0),
 * box4[24k..]: lo.x[4], lo.y[4], lo.z[4], hi.x[4], hi.y[4], hi.z[4] (empty slot +inf). */
void ora_get_bvh4(const ora_ctx* ctx, int32_t* child4, float* box4);

/* Batched traversal of the FP32 BVH.  rays: 8 floats each (o.xyz, tmin, d.xyz, tmax).
 * closest: hit_t, hit_leaf (sorted position, 0xffffffff = miss), hit_u, hit_v.
 * anyhit (any_hit != 0): hit_leaf[i] = 1 if occluded else 0. */
void ora_trace(const ora_ctx* ctx, uint32_t n, const float* rays, int any_hit,
               float* hit_t, uint32_t* hit_leaf, float* hit_u, float* hit_v);
/* Brute force FP64 closest hit (all triangles, Moller-Trumbore in double). */
void ora_trace_brute64(const ora_ctx* ctx, uint32_t n, const float* rays,
                       double* hit_t, uint32_t* hit_tri /* input order */);

/* FP32 render of samples [sample_base, sample_base+spp) for every pixel of the
 * camera.  sum_rgb (3*W*H floats) is accumulated in sample order: the caller
 * passes the running sums (zeros for a fresh image).
 * path_segments (optional, spp*W*H bytes, index s_local*npix + q where q is the
 * pixel's position in path order): number of traced extension segments.
 * shadow_bits (optional, spp*W*H*16 bytes): per path and segment, bit l set when
 * a shadow ray toward directional light l was queued, bit 7 for the sky ray.
 * ray_counts (4): camera, extension, light shadow, sky shadow rays traced. */
int ora_render32(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st,
                 uint32_t sample_base, uint32_t spp, float* sum_rgb,
                 uint8_t* path_segments, uint8_t* shadow_bits,
                 uint64_t* ray_counts, int num_threads);

/* ora_render32 restricted to a list of pixels (image indices, row-major): sum_rgb[3*i..]
 * is pixel pixels[i]'s running sum, accumulated over samples [sample_base,
 * sample_base+spp) in sample order, exactly as ora_render32 sums that pixel. */
/* Primary-hit GBuffer of sample `sample` of the listed pixels (bi_pt_render_gbuffer), by
 * list position: base_color, normal_roughness 4 half words each, fresnel 4 uint16 (unorm),
 * material_0 4 bytes (unorm); each nullable; zeros where the camera ray misses. */
int ora_gbuffer32(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t sample, uint32_t n,
                  const uint32_t* pixels, uint16_t* base, uint16_t* nr, uint16_t* fres, uint8_t* m0);
/* float -> IEEE half bits, round to nearest even. */
uint16_t ora_half32(float f);
/* One rect light's lighting at a lit surface (the kernels' rect_light_lit): light record
   rect (28 floats), LTC tables ltc (4 x 4096 texels x 4 floats: matrix rows 0-2, norm),
   point P, shading normal N (frame from N), direction V towards the viewer, roughness
   rx, ry, base colour, f0, f90.  Writes rgb and terms = (diffuse integral, specular
   integral, norm, Fresnel term). */
void ora_rect_light32(const float* rect, const float* ltc, const float* P, const float* N, const float* V, float rx,
                      float ry, const float* base, const float* f0, const float* f90, float* rgb, float* terms);
/* Rect light `light` of the scene at a lit surface, with the scene's LTC tables and
   rect-light textures (the kernels' rect_light_lit): as ora_rect_light32, plus dirs (6
   floats, nullable): the world directions the diffuse and specular integrals look up the
   light's texture along (zero where the integral is zero or the light is untextured). */
void ora_rect_light_scene32(const ora_ctx* ctx, uint32_t light, const float* P, const float* N, const float* V,
                            float rx, float ry, const float* base, const float* f0, const float* f90, float* rgb,
                            float* terms, float* dirs);
/* rect_light_sample_texture of light `light` (its texture): the texture seen from P along
   dir at lobe roughness r; rgb, plus uvl = (u, v, level) (nullable). */
void ora_rect_tex_lookup32(const ora_ctx* ctx, uint32_t light, const float* dir, const float* P, float r, float* rgb,
                           float* uvl);
/* SampleLevel of rect-light texture `tex` at (u, v) and level (4 floats out). */
void ora_rect_tex_sample32(const ora_ctx* ctx, uint32_t tex, float u, float v, float level, float* out4);
/* The kernels' log2 (inputs >= 1). */
void ora_log2_32(uint32_t n, const float* x, float* out);
/* The LTC tables' transform (9 floats, rows) and pair (2 floats) at (u3, u2, u1, u0). */
void ora_ltc_fetch32(const float* ltc, float u3, float u2, float u1, float u0, float* m, float* b);
int ora_render32_pixels(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t sample_base,
                        uint32_t spp, uint32_t n, const uint32_t* pixels, float* sum_rgb, int num_threads);

/* Radiance-cache modes of bi_pt_render_nrc (src: the build's cache; no reference
 * counterpart), for a cache whose network output is the constant cache_rgb:
 * ora_render32_cached renders like ora_render32 but every path that hits at segment
 * cache_bounce ends there with throughput x max(cache_rgb, 0) as its radiance of that
 * vertex and beyond.  ora_train_samples traces training slot q = 0..n-1 (pixel drawn by
 * PCG32 keyed (seed, q, 0, 254) uniformly over the image, sample q, jittered): rec[16q..]
 * receives the record of the hit at segment cache_bounce (all zero when the path does
 * not reach it) and target[3q..] the radiance gathered after it with throughput 1 --
 * with suffixes > 1, the mean over that many continuations from the recorded vertex
 * (continuation k draws with the seed xor k * 0x9E3779B9; k = 0 is the path's own). */
int ora_render32_cached(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t cache_bounce,
                        const float* cache_rgb, uint32_t sample_base, uint32_t spp, float* sum_rgb,
                        int num_threads);
int ora_train_samples(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t cache_bounce,
                      uint32_t n, uint32_t suffixes, float* rec, float* target);

/* Path-order pixel index of path-order position q (tile order when W and H are
 * multiples of 8, else row-major). */
uint32_t ora_pixel_of(uint32_t q, uint32_t width, uint32_t height);

/* FP64 oracle: for each listed pixel, mean and variance of the per-sample
 * radiance over samples [0, spp) (RNG seed seed64).  mean_rgb/var_rgb: 3 per pixel. */
int ora_render64(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st,
                 uint32_t num_pixels, const uint32_t* pixels, uint32_t spp, uint32_t seed64,
                 double* mean_rgb, double* var_rgb, int num_threads);

/* ---- unit-level entry points (for known-answer and invariant tests) ---- */
/* Level-0 samples of scene texture `tex` at n (u, v) pairs: 4 floats each. */
void ora_tex_sample32(const ora_ctx* ctx, uint32_t tex, uint32_t n, const float* uv, float* out4);
void ora_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out);
void ora_pcg32_keyed(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce,
                     uint32_t n, float* out);
uint32_t ora_morton3(uint32_t x, uint32_t y, uint32_t z);
void ora_math32(int fn, uint32_t n, const float* x, const float* y, float* out); /* 0 sin 1 cos 2 acos 3 atan2(x,y) */
/* Lit BSDF (lit.hlsl) value f(V,L)*cos for a material (16 floats), N,V,L world. */
void ora_bsdf32(const float* material, const float* n, const float* v, const float* l, float* out3);
void ora_bsdf64(const float* material, const double* n, const double* v, const double* l, double* out3);
/* VNDF sample in the local frame: given local v, alpha, (u0,u1) -> local h, pdf of reflected L. */
void ora_vndf32(const float* v_local, float alpha, float u0, float u1, float* h_out, float* pdf_out);
double ora_vndf_pdf64(const double* v_local, const double* l_local, double alpha);
/* Sky: radiance lookup, sample and pdf (FP32 transcription). */
void ora_sky_eval32(const ora_ctx* ctx, const float* dir, float* rgb);
void ora_sky_sample32(const ora_ctx* ctx, float u0, float u1, float* dir, float* pdf);
/* As ora_sky_sample32, plus the sky's radiance at the sample (looked up at the sampled
   texture coordinates, as the next-event estimation uses it). */
void ora_sky_sample_le32(const ora_ctx* ctx, float u0, float u1, float* dir, float* pdf, float* le);
float ora_sky_pdf32(const ora_ctx* ctx, const float* dir);
/* Batched forms for quadrature tests. */
void ora_bsdf64_batch(const float* material, const double* n, const double* v, uint32_t count, const double* l,
                      double* out3);
void ora_vndf_pdf64_batch(const double* v_local, uint32_t count, const double* l_local, double alpha, double* out);
void ora_sky_batch32(const ora_ctx* ctx, uint32_t count, const float* dirs, float* rgb, float* pdf);
/* Camera ray for pixel (px,py) and optional jitter (jx,jy in [0,1), 0.5 = centre). */
void ora_camera_ray32(const ora_camera* cam, uint32_t px, uint32_t py, float jx, float jy,
                      float* o, float* d);

#ifdef __cplusplus
}
#endif
#endif
