// pt_device.hpp -- device-side arithmetic of the path-tracing pass (gfx950).
//
// Every function rounds each float operation once, in the order written: the
// library is compiled with -ffp-contract=off, correctly rounded division and
// square root (HIP's default), and no fast-math, so the results are a fixed
// function of the inputs.  Where the pass uses min/max on values that may be
// NaN, comparisons are written out (a < b ? a : b) so NaN propagation is fixed too.
//
// Reference formulas (paths relative to bisemutum/):
//   lit BSDF                shaders/core/material/lit.hlsl:5-35
//   Fresnel / GGX / VNDF    shaders/core/material/utils.hlsl:40-121
//   shading frame           shaders/core/utils/frame.hlsl:9-20 (Duff et al. 2017)
//   camera ray              shaders/renderer/raytracing/direction_sample/generate_camera_ray.hlsl:9-15
//   equirect mapping        shaders/core/equitangular_to_cubemap.hlsl:28-33
//   cube map faces          shaders/core/utils/cubemap.hlsl:3-21 (face order and orientation)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_layout.hpp"

namespace bipt {

#define PT_HD __host__ __device__ __forceinline__

struct V3 { float x, y, z; };
PT_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
PT_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_HD V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
PT_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
PT_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
PT_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_HD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
PT_HD float sel_max(float a, float b) { return a > b ? a : b; }
PT_HD float sel_min(float a, float b) { return a < b ? a : b; }
PT_HD V3 normalize(V3 v) { float inv = 1.0f / __builtin_sqrtf(dot(v, v)); return v * inv; }
PT_HD float max3(V3 v) { return sel_max(sel_max(v.x, v.y), v.z); }
PT_HD bool any_pos(V3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }
PT_HD bool finite1(float x) { return __builtin_isfinite(x); }
PT_HD bool finite3(V3 v) { return finite1(v.x) && finite1(v.y) && finite1(v.z); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPiO2 = 1.57079632679489661923f;
constexpr float kPiO4 = 0.785398163397448309616f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvTwoPi = 0.159154943091895335769f;
constexpr float kTwoPiSq = 19.7392088021787172376f;
constexpr float kTMin = 0.001f;            // rt_gbuffer.hlsl:18 (TMin)
constexpr float kRobust = 1.00000036f;     // 1 + 2*gamma(3): conservative slab test (Ize 2013)
constexpr float kThroughputMin = 0.001f;   // sample_secondary_ray.hlsl:24
// Largest ray extent: half of FLT_MAX, so tmax * kRobust stays finite and the empty
// BVH4 slot (+inf box) is rejected by every slab test.
constexpr float kShadowTMax = 1.70141173e38f;

// ------------------------------------------------------------ transcendentals
// sin/cos/asin/acos/atan in single precision after the Cephes library (S. L.
// Moshier): Cody-Waite reduction by pi/4 in three parts, then the published
// minimax polynomials, evaluated by Horner without fused multiply-add.
PT_HD void pt_sincos(float x, float* s_out, float* c_out) {
    float ax = x < 0.0f ? -x : x;
    int q = (int)(ax * 1.27323954473516268615f);
    float y = (float)q;
    if (q & 1) { q += 1; y += 1.0f; }
    int j = q & 7;
    float r = ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    float z = r * r;
    float sk = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
    float ck = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
        - 0.5f * z + 1.0f;
    int sign_s = x < 0.0f ? -1 : 1;
    int sign_c = 1;
    if (j > 3) { j -= 4; sign_s = -sign_s; sign_c = -sign_c; }
    if (j > 1) sign_c = -sign_c;
    float s, c;
    if (j == 1 || j == 2) { s = ck; c = sk; } else { s = sk; c = ck; }
    *s_out = sign_s < 0 ? -s : s;
    *c_out = sign_c < 0 ? -c : c;
}
PT_HD float pt_sin(float x) { float s, c; pt_sincos(x, &s, &c); return s; }
PT_HD float pt_asin(float x) {
    float a = x < 0.0f ? -x : x;
    float z, s;
    bool big = a > 0.5f;
    if (big) { z = 0.5f * (1.0f - a); s = __builtin_sqrtf(z); }
    else { s = a; z = a * a; }
    float p = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z
                + 7.4953002686e-2f) * z + 1.6666752422e-1f) * z * s + s;
    if (big) { p = p + p; p = kPiO2 - p; }
    return x < 0.0f ? -p : p;
}
PT_HD float pt_acos(float x) {
    if (x < -0.5f) return kPi - 2.0f * pt_asin(__builtin_sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * pt_asin(__builtin_sqrtf(0.5f * (1.0f - x)));
    return kPiO2 - pt_asin(x);
}
PT_HD float pt_atan(float x) {
    float a = x < 0.0f ? -x : x;
    float y0;
    if (a > 2.414213562373095f) { y0 = kPiO2; a = -(1.0f / a); }
    else if (a > 0.4142135623730950f) { y0 = kPiO4; a = (a - 1.0f) / (a + 1.0f); }
    else { y0 = 0.0f; }
    float z = a * a;
    float r = y0 + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z
                     - 3.33329491539e-1f) * z * a + a);
    return x < 0.0f ? -r : r;
}
PT_HD float pt_atan2(float y, float x) {
    if (x == 0.0f) {
        if (y > 0.0f) return kPiO2;
        if (y < 0.0f) return -kPiO2;
        return 0.0f;
    }
    float r = pt_atan(y / x);
    if (x < 0.0f) r = (y < 0.0f) ? r - kPi : r + kPi;
    return r;
}

// -------------------------------------------------------------------- PCG32
// O'Neill 2014 (pcg32_random_r / pcg32_srandom_r); one stream per
// (seed, pixel, sample, bounce): initstate = pixel << 32 | sample,
// initseq = seed << 8 | bounce.  Floats are the top 24 bits times 2^-24.
struct Pcg { uint64_t state, inc; };
PT_HD uint32_t pcg_next(Pcg& r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
PT_HD Pcg pcg_keyed(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    Pcg r{0u, ((((uint64_t)seed << 8) | (bounce & 0xffu)) << 1u) | 1u};
    pcg_next(r);
    r.state += ((uint64_t)pixel << 32) | sample;
    pcg_next(r);
    return r;
}
PT_HD float pcg_float(Pcg& r) { return (float)(pcg_next(r) >> 8) * 0x1p-24f; }
constexpr uint32_t kJitterBounce = 255u;

// ------------------------------------------------------------------ camera
PT_HD void camera_ray(const CameraParams& c, uint32_t px, uint32_t py, float jx, float jy, V3* o, V3* d) {
    float u = ((float)px + jx) / (float)c.width;
    float v = ((float)py + jy) / (float)c.height;
    float nx = u * 2.0f - 1.0f, ny = 1.0f - v * 2.0f;
    const float* ip = c.inv_proj;
    V3 q = v3(((ip[0] * nx + ip[1] * ny) + ip[2]) + ip[3],
              ((ip[4] * nx + ip[5] * ny) + ip[6]) + ip[7],
              ((ip[8] * nx + ip[9] * ny) + ip[10]) + ip[11]);
    V3 dl = normalize(q);
    const float* iv = c.inv_view;
    V3 dw = v3(iv[0] * dl.x + iv[1] * dl.y + iv[2] * dl.z,
               iv[4] * dl.x + iv[5] * dl.y + iv[6] * dl.z,
               iv[8] * dl.x + iv[9] * dl.y + iv[10] * dl.z);
    *d = normalize(dw);
    *o = v3(iv[3], iv[7], iv[11]);
}

// Position in path order -> pixel index (8x8 tiles when both sides are multiples of 8).
PT_HD uint32_t pixel_of(uint32_t q, uint32_t W, uint32_t H) {
    if ((W & 7u) == 0 && (H & 7u) == 0) {
        uint32_t tile = q >> 6, within = q & 63u;
        uint32_t tiles_x = W >> 3;
        uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
        return ((ty << 3) + (within >> 3)) * W + (tx << 3) + (within & 7u);
    }
    return q;
}

// ---------------------------------------------------------------- material
struct Mat {
    V3 base, f0, f90;
    float rx, ry;
    bool two_sided;
};
// get_anisotropic_roughness (material/utils.hlsl:54-59) with the 0.001 floor.
PT_HD void mat_alpha(Mat& r, float roughness, float anisotropy) {
    float aniso = __builtin_sqrtf(1.0f - anisotropy * 0.9f);
    float r2 = roughness * roughness;
    r.rx = sel_max(r2 / aniso, 0.001f);
    r.ry = sel_max(r2 * aniso, 0.001f);
}
PT_HD Mat load_mat(const MaterialRec& m) {
    Mat r;
    r.base = v3(m.base_color[0], m.base_color[1], m.base_color[2]);
    r.f0 = v3(m.f0_color[0], m.f0_color[1], m.f0_color[2]);
    r.f90 = v3(m.f90_color[0], m.f90_color[1], m.f90_color[2]);
    r.two_sided = m.two_sided != 0u;
    mat_alpha(r, m.roughness, m.anisotropy);
    return r;
}
PT_HD bool mat_textured(const MaterialRec& m) {
    return (m.base_color_tex | m.f0_tex | m.normal_tex | m.roughness_tex) != kNoTexture;
}

// ---------------------------------------------------------------- textures
// Level-0 sampling of the material textures (SampleLevel(..., 0) with the texture's
// sampler): texel centres at (i + 0.5) / W, bilinear weights in full FP32 (hardware
// quantises them to 8 bits), repeat / mirrored-repeat / clamp-to-edge addressing
// (Vulkan's rules), RGBA8 decoded as c / 255 (UNORM) or through the sRGB table.
struct V4 { float x, y, z, w; };
PT_HD float comp4(const V4& v, uint32_t c) { return c == 0u ? v.x : c == 1u ? v.y : c == 2u ? v.z : v.w; }
PT_HD int tex_wrap(int i, int n, uint32_t mode) {
    if (mode == 2u) return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    if (mode == 1u) {
        int m = i % (2 * n);
        if (m < 0) m += 2 * n;
        return m >= n ? 2 * n - 1 - m : m;
    }
    int m = i % n;
    return m < 0 ? m + n : m;
}
PT_HD V4 tex_texel(const SceneTex& T, const TexRec& r, int i, int j) {
    const size_t k = (size_t)j * r.width + (size_t)i;
    if (r.format == kTexRgba32f) {
        const float* f = reinterpret_cast<const float*>(T.pool + r.offset) + 4 * k;
        return V4{f[0], f[1], f[2], f[3]};
    }
    const uint32_t w = T.pool[r.offset + k];
    const uint32_t c0 = w & 255u, c1 = (w >> 8) & 255u, c2 = (w >> 16) & 255u, c3 = w >> 24;
    if (r.format == kTexRgba8Srgb) return V4{T.srgb[c0], T.srgb[c1], T.srgb[c2], (float)c3 / 255.0f};
    return V4{(float)c0 / 255.0f, (float)c1 / 255.0f, (float)c2 / 255.0f, (float)c3 / 255.0f};
}
// x in texel units; coordinates that are not finite or beyond 2^24 texels read texel 0
PT_HD float tex_coord(float x) { return (x > -16777216.0f && x < 16777216.0f) ? x : 0.0f; }
PT_HD V4 tex_sample_rec(const SceneTex& T, const TexRec& r, uint32_t filter, float u, float v) {
    const int W = (int)r.width, H = (int)r.height;
    if (filter == 0u) {
        const int i = tex_wrap((int)__builtin_floorf(tex_coord(u * (float)W)), W, r.address_u);
        const int j = tex_wrap((int)__builtin_floorf(tex_coord(v * (float)H)), H, r.address_v);
        return tex_texel(T, r, i, j);
    }
    const float x = tex_coord(u * (float)W - 0.5f), y = tex_coord(v * (float)H - 0.5f);
    const float x0 = __builtin_floorf(x), y0 = __builtin_floorf(y);
    const float ax = x - x0, ay = y - y0;
    const int i0 = tex_wrap((int)x0, W, r.address_u), i1 = tex_wrap((int)x0 + 1, W, r.address_u);
    const int j0 = tex_wrap((int)y0, H, r.address_v), j1 = tex_wrap((int)y0 + 1, H, r.address_v);
    const V4 t00 = tex_texel(T, r, i0, j0), t10 = tex_texel(T, r, i1, j0);
    const V4 t01 = tex_texel(T, r, i0, j1), t11 = tex_texel(T, r, i1, j1);
    auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
    const V4 top{lerp(t00.x, t10.x, ax), lerp(t00.y, t10.y, ax), lerp(t00.z, t10.z, ax), lerp(t00.w, t10.w, ax)};
    const V4 bot{lerp(t01.x, t11.x, ax), lerp(t01.y, t11.y, ax), lerp(t01.z, t11.z, ax), lerp(t01.w, t11.w, ax)};
    return V4{lerp(top.x, bot.x, ay), lerp(top.y, bot.y, ay), lerp(top.z, bot.z, ay), lerp(top.w, bot.w, ay)};
}
PT_HD V4 tex_sample(const SceneTex& T, uint32_t id, float u, float v) {
    const TexRec r = T.recs[id];
    return tex_sample_rec(T, r, r.filter, u, v);
}
// Texture coordinates of a leaf at barycentrics (u, v): uv0 + (uv1 - uv0) u + (uv2 - uv0) v.
PT_HD void leaf_texcoord(const LeafUv& L, float bu, float bv, float* s, float* t) {
    *s = (L.uv[0] + (L.uv[2] - L.uv[0]) * bu) + (L.uv[4] - L.uv[0]) * bv;
    *t = (L.uv[1] + (L.uv[3] - L.uv[1]) * bu) + (L.uv[5] - L.uv[1]) * bv;
}

// surface.opacity of a material at texture coordinates (s, t).
PT_HD float surface_opacity(const MaterialRec& M, const SceneTex& T, float s, float t) {
    if (M.opacity_tex == kNoTexture) return M.opacity;
    const float c = comp4(tex_sample(T, M.opacity_tex - 1u, s, t), M.opacity_channel);
    return M.opacity_step ? (c < M.step_at ? M.step_below : M.step_above) : c;
}
// The per-ray uniform number of stochastic translucency: one draw per ray, keyed by the
// bits of its origin and direction (as gen_opacity_random_for, rt_gbuffer.hlsl:14-19,
// keys it by the ray; the stream is PCG32 keyed by the job seed).
constexpr uint32_t kOpacityBounce = 253u;
PT_HD float opacity_random(uint32_t seed, V3 o, V3 d) {
    const uint32_t h = __builtin_bit_cast(uint32_t, o.x) ^ __builtin_bit_cast(uint32_t, o.y) ^
                       __builtin_bit_cast(uint32_t, o.z) ^ __builtin_bit_cast(uint32_t, d.x) ^
                       __builtin_bit_cast(uint32_t, d.y) ^ __builtin_bit_cast(uint32_t, d.z);
    Pcg r = pcg_keyed(seed, h, 0u, kOpacityBounce);
    return pcg_float(r);
}
// Whether a non-opaque triangle counts as a hit of a ray (rt_gbuffer_rahit,
// hits/rt_gbuffer_hit.hlsl:20-33): alpha test drops it below the 0.01 threshold,
// translucency when the ray's number is below 1 - opacity.
PT_HD bool surface_present(const MaterialRec& M, const SceneTex& T, const LeafUv* uv, uint32_t leaf, float bu,
                           float bv, float rnd) {
    float s = 0.0f, t = 0.0f;
    if (uv) leaf_texcoord(uv[leaf], bu, bv, &s, &t);
    const float op = surface_opacity(M, T, s, t);
    if (M.blend_mode == kBlendAlphaTest) return !(op < kAlphaTestThreshold);
    return !(rnd < 1.0f - op);
}

// Checkerboard cell of a world coordinate: int(floor(x)), floor clamped to
// [-2^31, 2^31 - 128] (NaN to the low end) so the conversion is always defined.
PT_HD int checker_cell(float x) {
    float f = __builtin_floorf(x);
    if (!(f >= -2147483648.0f)) f = -2147483648.0f;
    if (f > 2147483520.0f) f = 2147483520.0f;
    return (int)f;
}
// The record that shades hit point P: a checkerboard material hands even cells of
// (int(floor(P.x)) ^ int(floor(P.z))) to its pattern_material.
PT_HD uint32_t resolve_material(const MaterialRec* mats, uint32_t id, V3 P) {
    if (mats[id].pattern == kPatternCheckerXZ && ((checker_cell(P.x) ^ checker_cell(P.z)) & 1) == 0)
        return mats[id].pattern_material;
    return id;
}

struct Frame { V3 t, b, n; };
PT_HD Frame make_frame(V3 n) {
    float sign = n.z > 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    Frame f;
    f.t = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f.b = v3(b, sign + n.y * n.y * a, -n.y);
    f.n = n;
    return f;
}
PT_HD V3 to_local(const Frame& f, V3 v) { return v3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
PT_HD V3 to_world(const Frame& f, V3 v) { return (f.t * v.x + f.b * v.y) + f.n * v.z; }

PT_HD float sq(float x) { return x * x; }
PT_HD float ggx_ndf(V3 h, float rx, float ry) {
    float a = (sq(h.x / rx) + sq(h.y / ry)) + sq(h.z);
    return kInvPi / (((rx * ry) * a) * a);
}
PT_HD float ggx_g1(V3 v, float rx, float ry) {
    float a = (sq(rx * v.x) + sq(ry * v.y)) / sel_max(v.z * v.z, 0.0001f);
    return 2.0f / (1.0f + __builtin_sqrtf(1.0f + a));
}
PT_HD float ggx_vis_hc(V3 v, V3 l, float rx, float ry) {
    float a = l.z * __builtin_sqrtf((sq(rx * v.x) + sq(ry * v.y)) + sq(v.z));
    float b = v.z * __builtin_sqrtf((sq(rx * l.x) + sq(ry * l.y)) + sq(l.z));
    return 0.5f / sel_max(a + b, 0.0001f);
}
// surface_eval_lit: diffuse + specular, both including cos(theta_l).
PT_HD V3 bsdf_eval(const Frame& fr, const Mat& m, V3 V, V3 L) {
    V3 H = normalize(V + L);
    V3 lh = to_local(fr, H), lv = to_local(fr, V), ll = to_local(fr, L);
    if (lv.z <= 0.0f || ll.z <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    float vh = dot(V, H);
    float c = sel_max(vh, 0.0f);
    float x = 1.0f - c;
    float p5 = ((x * x) * (x * x)) * x;
    V3 fres = m.f0 + (m.f90 - m.f0) * p5;
    float lz = sel_max(ll.z, 0.0f);
    V3 one = v3(1.0f, 1.0f, 1.0f);
    V3 diffuse = (((one - fres) * m.base) * kInvPi) * lz;
    float ndf = ggx_ndf(lh, m.rx, m.ry);
    float vis = ggx_vis_hc(lv, ll, m.rx, m.ry);
    V3 spec = ((fres * ndf) * vis) * lz;
    return diffuse + spec;
}
// ggx_vndf_sample (Heitz 2018)
PT_HD V3 vndf_sample(V3 v, float rx, float ry, float u0, float u1) {
    if (v.z < 0.0f) v = -v;
    V3 vh = normalize(v3(rx * v.x, ry * v.y, v.z));
    float len_sqr = vh.x * vh.x + vh.y * vh.y;
    V3 t1v;
    if (len_sqr > 0.0f) {
        float s = __builtin_sqrtf(len_sqr);
        t1v = v3(-vh.y / s, vh.x / s, 0.0f / s);
    } else {
        t1v = v3(1.0f, 0.0f, 0.0f);
    }
    V3 t2v = cross(vh, t1v);
    float r = __builtin_sqrtf(u0);
    float phi = kTwoPi * u1;
    float sp, cp;
    pt_sincos(phi, &sp, &cp);
    float t1 = r * cp;
    float t2 = r * sp;
    float s = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s) * __builtin_sqrtf(1.0f - t1 * t1) + s * t2;
    V3 nh = (t1v * t1 + t2v * t2) + vh * __builtin_sqrtf(sel_max((1.0f - t1 * t1) - t2 * t2, 0.0f));
    return normalize(v3(rx * nh.x, ry * nh.y, sel_max(nh.z, 0.0f)));
}
// ggx_vndf_sample_pdf / (4 |h.v|): density of the reflected direction.
PT_HD float vndf_pdf_l(V3 h, V3 v, float rx, float ry) {
    float hv = dot(h, v);
    float pdf_h = ((ggx_g1(v, rx, ry) * ggx_ndf(h, rx, ry)) * sel_max(hv, 0.0f)) / sel_max(v.z, 0.0001f);
    float ahv = hv < 0.0f ? -hv : hv;
    return pdf_h / (4.0f * ahv);
}
PT_HD float power_heuristic(float a, float b) {
    float a2 = a * a;
    return a2 / (a2 + b * b);
}

// -------------------------------------------------------------------- lights
// Radiance arriving at P from light L and the unit direction towards it: a directional
// light's emission along its direction; a point light's emission attenuated by distance
// and spot cone (point_light_eval, lights.hlsl:15-25).
PT_HD V3 light_eval(const LightRec& L, V3 P, V3* ldir) {
    const V3 e = v3(L.emission[0], L.emission[1], L.emission[2]);
    if (L.type != kLightPoint) {
        *ldir = v3(L.direction[0], L.direction[1], L.direction[2]);
        return e;
    }
    const V3 lv = v3(L.position[0], L.position[1], L.position[2]) - P;
    const float d2 = dot(lv, lv);
    const float dist = __builtin_sqrtf(d2);
    *ldir = v3(lv.x / dist, lv.y / dist, lv.z / dist);
    const float q = d2 * L.range_sqr_inv;
    float fall = 1.0f - q * q;
    fall = fall < 0.0f ? 0.0f : (fall > 1.0f ? 1.0f : fall);
    float att = fall / sel_max(d2, 0.001f);
    if (L.cos_inner > L.cos_outer) {
        float ct = dot(*ldir, v3(L.direction[0], L.direction[1], L.direction[2]));
        ct = ct < L.cos_outer ? L.cos_outer : (ct > L.cos_inner ? L.cos_inner : ct);
        att = att * ((ct - L.cos_outer) / sel_max(L.cos_inner - L.cos_outer, 0.001f));
    }
    return e * att;
}
// Shadow ray from P towards light L: the direction light_eval gives, and the extent:
// unbounded for a directional light, the distance to a point light.
PT_HD void light_shadow_ray(const LightRec& L, V3 P, V3* dir, float* tmax) {
    if (L.type != kLightPoint) {
        *dir = v3(L.direction[0], L.direction[1], L.direction[2]);
        *tmax = kShadowTMax;
        return;
    }
    const V3 lv = v3(L.position[0], L.position[1], L.position[2]) - P;
    const float dist = __builtin_sqrtf(dot(lv, lv));
    *dir = v3(lv.x / dist, lv.y / dist, lv.z / dist);
    *tmax = dist;
}

// --------------------------------------------------------------------- rect lights
// Linearly transformed cosines over a quad (lights.hlsl:161-470, after Heitz et al. 2016,
// "Real-Time Polygonal-Light Shading with Linearly Transformed Cosines", with anisotropic
// 4D tables): the diffuse integral of the clipped quad under the plain cosine, the
// specular one under the table's transform, both in the shading frame (t, b, n).
struct M3 { V3 r0, r1, r2; };   // rows
PT_HD V3 m3_mul(const M3& m, V3 v) { return v3(dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)); }
// Inverse by the adjugate: inv[i][j] = cofactor[j][i] / det, det expanded along row 0.
PT_HD M3 m3_inverse(const M3& m) {
    const float a = m.r0.x, b = m.r0.y, c = m.r0.z, d = m.r1.x, e = m.r1.y, f = m.r1.z, g = m.r2.x, h = m.r2.y,
                k = m.r2.z;
    const float A = e * k - f * h, B = f * g - d * k, Cc = d * h - e * g;
    const float det = (a * A + b * B) + c * Cc;
    const float s = 1.0f / det;
    M3 r;
    r.r0 = v3(A * s, (c * h - b * k) * s, (b * f - c * e) * s);
    r.r1 = v3(B * s, (a * k - c * g) * s, (c * d - a * f) * s);
    r.r2 = v3(Cc * s, (b * g - a * h) * s, (a * e - b * d) * s);
    return r;
}

// Texel coordinate of a normalised coordinate on an axis of n texels (linear filter,
// clamp to edge): the two texels and the weight of the second.
PT_HD void ltc_axis(float u, int n, int* i0, int* i1, float* f) {
    const float t = u * (float)n - 0.5f;
    const float fl = __builtin_floorf(t);
    *f = t - fl;
    const int a = (int)fl, b = a + 1;
    *i0 = a < 0 ? 0 : (a > n - 1 ? n - 1 : a);
    *i1 = b < 0 ? 0 : (b > n - 1 ? n - 1 : b);
}
PT_HD float lerpf(float a, float b, float t) { return a + (b - a) * t; }
PT_HD V4 lerp4(V4 a, V4 b, float t) {
    return V4{lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t), lerpf(a.w, b.w, t)};
}
PT_HD V4 ld4(const float4* t, int i) { const float4 v = t[i]; return V4{v.x, v.y, v.z, v.w}; }
// Trilinear sample of an 8 x 8 x 64 table (x fastest) at normalised (x, y, z).
PT_HD V4 ltc_sample(const float4* t, float x, float y, float z) {
    int x0, x1, y0, y1, z0, z1;
    float fx, fy, fz;
    ltc_axis(x, (int)kLtcSize, &x0, &x1, &fx);
    ltc_axis(y, (int)kLtcSize, &y0, &y1, &fy);
    ltc_axis(z, (int)kLtcDepth, &z0, &z1, &fz);
    const int s = (int)kLtcSize, s2 = s * s;
    V4 e[2];
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {   // the two depth slices, each bilinear in x and y
        const int zo = s2 * (k ? z1 : z0);
        const V4 c0 = lerp4(ld4(t, x0 + s * y0 + zo), ld4(t, x1 + s * y0 + zo), fx);
        const V4 c1 = lerp4(ld4(t, x0 + s * y1 + zo), ld4(t, x1 + s * y1 + zo), fx);
        e[k] = lerp4(c0, c1, fy);
    }
    return lerp4(e[0], e[1], fz);
}
// The tables' transform and (norm, Fresnel) pair at table coordinate (u3, u2, u1, u0):
// u0 (the major roughness) selects two groups of 8 depth slices, blended by its
// fraction (get_ltc_tex3d_coord, fetch_ltc_matrix, fetch_ltc_brdf: lights.hlsl:164-201).
PT_HD void ltc_fetch(const float4* ltc, float u3, float u2, float u1, float u0, M3* m, float* bx, float* by) {
    const float ws = u0 * 7.0f;
    const float ws_f = __builtin_floorf(ws);
    const float wc = __builtin_floorf(ws + 1.0f);
    const float ws_c = wc < 7.0f ? wc : 7.0f;
    const float w = ws - ws_f;
    const float x = (u3 * 7.0f + 0.5f) / 8.0f, y = (u2 * 7.0f + 0.5f) / 8.0f;
    const float z1 = ((u1 * 7.0f + 8.0f * ws_f) + 0.5f) / 64.0f, z2 = ((u1 * 7.0f + 8.0f * ws_c) + 0.5f) / 64.0f;
    V3 r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const V4 a = ltc_sample(ltc + k * (int)kLtcTexels, x, y, z1), b = ltc_sample(ltc + k * (int)kLtcTexels, x, y, z2);
        r[k] = v3(lerpf(a.x, b.x, w), lerpf(a.y, b.y, w), lerpf(a.z, b.z, w));
    }
    m->r0 = r[0]; m->r1 = r[1]; m->r2 = r[2];
    const V4 n1 = ltc_sample(ltc + 3 * (int)kLtcTexels, x, y, z1), n2 = ltc_sample(ltc + 3 * (int)kLtcTexels, x, y, z2);
    *bx = lerpf(n1.x, n2.x, w);
    *by = lerpf(n1.y, n2.y, w);
}
PT_HD void quad_reverse(V3 q[4]) {
    const V3 a = q[0], b = q[1];
    q[0] = q[3]; q[1] = q[2]; q[2] = b; q[3] = a;
}
// get_ltc_matrix_and_brdf (lights.hlsl:203-255): the table coordinate of the view
// direction and roughness, folded into the first quadrant of azimuth and the larger
// roughness first, the transform unfolded by row flips; reflections reverse the quad.
PT_HD void ltc_matrix_brdf(const float4* ltc, V3 lv, float rx, float ry, V3 q[4], M3* m, float* bx, float* by) {
    const float theta = pt_acos(lv.z);
    const bool flip_r = ry > rx;
    float phi = pt_atan2(lv.y, lv.x);
    phi = flip_r ? kPi * 0.5f - phi : phi;
    phi = phi >= 0.0f ? phi : phi + 2.0f * kPi;
    const float u0 = sel_max((flip_r ? ry : rx) - 0.001f, 0.0f) / (1.0f - 0.001f);
    const float u1 = flip_r ? rx / ry : ry / rx;
    const float u2 = theta / (kPi * 0.5f);
    if (phi < kPi * 0.5f) {
        ltc_fetch(ltc, phi / (kPi * 0.5f), u2, u1, u0, m, bx, by);
    } else if (phi < kPi) {
        ltc_fetch(ltc, (kPi - phi) / (kPi * 0.5f), u2, u1, u0, m, bx, by);
        quad_reverse(q);
        m->r0 = -m->r0;
    } else if (phi < 1.5f * kPi) {
        ltc_fetch(ltc, (phi - kPi) / (kPi * 0.5f), u2, u1, u0, m, bx, by);
        m->r0 = -m->r0;
        m->r1 = -m->r1;
    } else {
        ltc_fetch(ltc, (2.0f * kPi - phi) / (kPi * 0.5f), u2, u1, u0, m, bx, by);
        quad_reverse(q);
        m->r1 = -m->r1;
    }
    if (flip_r) {
        quad_reverse(q);
        const V3 t = m->r0;
        m->r0 = m->r1;
        m->r1 = t;
    }
}
// theta / sin(theta) of the edge between unit vectors v1 and v2 by the rational fit of
// ltc_integrate_edge (lights.hlsl:365-381); the edge's term is cross(v1, v2) times it.
PT_HD float ltc_edge_scale(V3 v1, V3 v2) {
    const float x = dot(v1, v2);
    const float y = x < 0.0f ? -x : x;
    const float a = 5.42031f + (3.12829f + 0.0902326f * y) * y;
    const float b = 3.45068f + (4.18814f + y) * y;
    float t = a / b;
    if (x < 0.0f) t = kPi / __builtin_sqrtf(1.0f - x * x) - t;
    return t;
}
// ltc_integrate (lights.hlsl:382-426): the quad q relative to P in the frame (t, b, n),
// transformed by minv (none when null), clipped to z > 0 (a vertex pair across the plane
// contributes the point in.z * out - out.z * in), projected onto the sphere and integrated
// edge by edge; one-sided lights count only positive integrals.  dir (nullable) receives
// the sum of the edges' cross(v1, v2) theta / sin(theta) in the transformed space (the
// vector whose image is the light texture's direction); it is left alone when the quad
// lies below the plane.
PT_HD float ltc_integrate(V3 P, const Frame& fr, const M3* minv, const V3 q[4], bool two_sided, V3* dir = nullptr) {
    V3 l[4];
    uint32_t inside = 0;
    for (int k = 0; k < 4; ++k) {
        const V3 r = q[k] - P;
        V3 v = v3(dot(fr.t, r), dot(fr.b, r), dot(fr.n, r));
        if (minv) v = m3_mul(*minv, v);
        l[k] = v;
        inside |= (v.z > 0.0f ? 1u : 0u) << k;
    }
    if (inside == 0u || inside == 5u || inside == 10u) return 0.0f;   // below, or a diagonal pair
    V3 c[5];
    int n = 0;
    for (int k = 0; k < 4; ++k) {
        const V3 a = l[k], b = l[(k + 1) & 3];
        const bool ai = (inside >> k) & 1u, bi = (inside >> ((k + 1) & 3)) & 1u;
        if (ai) c[n++] = a;
        if (ai != bi) {
            const V3 in = ai ? a : b, out = ai ? b : a;
            c[n++] = v3(in.z * out.x - out.z * in.x, in.z * out.y - out.z * in.y, in.z * out.z - out.z * in.z);
        }
    }
    for (int k = 0; k < n; ++k) c[k] = normalize(c[k]);
    float sum = 0.0f, sx = 0.0f, sy = 0.0f;
    for (int k = 0; k < n; ++k) {
        const V3 a = c[k], b = c[k + 1 < n ? k + 1 : 0];
        const float t = ltc_edge_scale(a, b);
        sum = sum + (a.x * b.y - a.y * b.x) * t;
        if (dir) {
            sx = sx + (a.y * b.z - a.z * b.y) * t;
            sy = sy + (a.z * b.x - a.x * b.z) * t;
        }
    }
    if (dir) *dir = v3(sx, sy, sum);
    float r = two_sided ? (sum < 0.0f ? -sum : sum) : sel_max(sum, 0.0f);
    if (!(r - r == 0.0f)) r = 0.0f;   // NaN or infinite
    return r;
}
// log2 of x >= 1 (Moshier's Cephes log2f: x = m 2^e with m in [sqrt(1/2), sqrt(2)), a
// degree-8 minimax polynomial for log(1 + f), scaled by log2(e) in two parts); exact at
// powers of two; +inf gives 128.
PT_HD float pt_log2(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    float e = (float)((int)((b >> 23) & 255u) - 126);
    float m = __builtin_bit_cast(float, (b & 0x007fffffu) | 0x3f000000u);
    if (m < 0.70710678118654752440f) {
        e = e - 1.0f;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    const float z = m * m;
    float y = 7.0376836292e-2f;
    y = y * m - 1.1514610310e-1f;
    y = y * m + 1.1676998740e-1f;
    y = y * m - 1.2420140846e-1f;
    y = y * m + 1.4249322787e-1f;
    y = y * m - 1.6668057665e-1f;
    y = y * m + 2.0000714765e-1f;
    y = y * m - 2.4999993993e-1f;
    y = y * m + 3.3333331174e-1f;
    y = (y * m) * z;
    y = y - 0.5f * z;
    const float log2ea = 0.44269504088896340736f;
    float r = y * log2ea;
    r = r + m * log2ea;
    r = r + y;
    r = r + m;
    return r + e;
}
PT_HD float saturate_nan0(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }
// SampleLevel of a rect-light texture (Vulkan's explicit-LOD rules): lambda = level + the
// sampler's bias, clamped to [lod_min, lod_max]; the minification filter for lambda > 0,
// else the magnification one; the level lambda clamped to [0, levels - 1], nearest (halves
// to the lower level) or the two levels around it blended by its fraction.
PT_HD V4 rect_tex_sample(const RectTexRec& r, const SceneTex& T, float u, float v, float level) {
    float lam = level + r.lod_bias;
    lam = lam > r.lod_max ? r.lod_max : lam;
    lam = lam < r.lod_min ? r.lod_min : lam;
    const uint32_t filter = lam > 0.0f ? r.min_filter : r.mag_filter;
    const float q = (float)(r.num_levels - 1u);
    float d = lam > 0.0f ? lam : 0.0f;
    d = d < q ? d : q;
    if (r.mipmap_mode == 0u) {
        const uint32_t l = (uint32_t)(__builtin_ceilf(d + 0.5f) - 1.0f);
        return tex_sample_rec(T, T.recs[r.first_level + l], filter, u, v);
    }
    const float fl = __builtin_floorf(d);
    const uint32_t l = (uint32_t)fl;
    const float f = d - fl;
    const V4 a = tex_sample_rec(T, T.recs[r.first_level + l], filter, u, v);
    if (f == 0.0f) return a;
    const V4 b = tex_sample_rec(T, T.recs[r.first_level + l + 1u], filter, u, v);
    return V4{lerpf(a.x, b.x, f), lerpf(a.y, b.y, f), lerpf(a.z, b.z, f), lerpf(a.w, b.w, f)};
}
// rect_light_sample_texture (lights.hlsl:429-445): where the ray from P along direction
// dir meets the light's plane, as (u, v) across its width (position2 -> position3) and
// height (position2 -> position1), clamped to the light; the level log2 of the texels the
// lobe spans there (distance * roughness * texels per unit).  Directions within 1e-4 of
// the plane see no texture (0).
PT_HD V3 rect_tex_lookup(const RectRec& R, const RectTexRec& rt, const SceneTex& T, V3 dir, V3 P, float roughness) {
    const V3 n = v3(R.normal[0], R.normal[1], R.normal[2]);
    const V3 p2 = v3(R.p2[0], R.p2[1], R.p2[2]);
    const float st = dot(dir, n), step = st < 0.0f ? -st : st;
    if (step < 0.0001f) return v3(0.0f, 0.0f, 0.0f);
    const float dd = dot(P - p2, n), dist = dd < 0.0f ? -dd : dd;
    const float t = dist / step;
    const V3 rp = (P + dir * t) - p2;
    const float u = saturate_nan0(dot(rp, v3(R.p3[0], R.p3[1], R.p3[2]) - p2) * R.inv_width_sqr);
    const float v = saturate_nan0(dot(rp, v3(R.p1[0], R.p1[1], R.p1[2]) - p2) * R.inv_height_sqr);
    const float texels = (t * roughness) * R.inv_texel_size;
    const float level = pt_log2(texels > 1.0f ? texels : 1.0f);
    const V4 c = rect_tex_sample(rt, T, u, v, level);
    return v3(c.x, c.y, c.z);
}
// The world direction of a polygon integral's vector s (transformed space, m its forward
// transform, none when null): normalize((m s) in the frame (t, b, n)).
PT_HD V3 ltc_direction(const Frame& fr, const M3* m, V3 s) {
    const V3 w = m ? m3_mul(*m, s) : s;
    return normalize((fr.t * w.x + fr.b * w.y) + fr.n * w.z);
}
// A rect light's contribution at a lit surface (rect_light_eval_ltc, lights.hlsl:
// 471-513, then surface_eval_lut_lit, material/lit.hlsl:37-58): the diffuse integral
// times (1 - F) base / pi plus the specular one times f0 * norm + f90 * Fresnel term.
// (Called, not inlined: its table reads and two polygon integrals would otherwise share
// the shading kernel's registers.)
inline __device__ __noinline__ V3 rect_light_lit(const RectRec& R, const float4* ltc, const RectTexRec* rtex,
                                                 const SceneTex& rtl, V3 P, const Frame& fr, V3 V, const Mat& m) {
    const V3 lv = v3(dot(V, fr.t), dot(V, fr.b), dot(V, fr.n));
    if (lv.z <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    V3 q[4] = {v3(R.p3[0], R.p3[1], R.p3[2]), v3(R.p2[0], R.p2[1], R.p2[2]), v3(R.p1[0], R.p1[1], R.p1[2]),
               v3(R.p0[0], R.p0[1], R.p0[2])};
    const bool two = R.two_sided != 0u;
    const V3 e = v3(R.emission[0], R.emission[1], R.emission[2]);
    const bool textured = R.texture_index >= 0;
    V3 sd, ss;
    const float i_d = ltc_integrate(P, fr, nullptr, q, two, textured ? &sd : nullptr);
    M3 mt;
    float bx, by;
    ltc_matrix_brdf(ltc, lv, m.rx, m.ry, q, &mt, &bx, &by);
    const M3 mi = m3_inverse(mt);
    const float i_s = ltc_integrate(P, fr, &mi, q, two, textured ? &ss : nullptr);
    V3 ld = e * i_d, ls = e * i_s;
    if (textured) {
        // the light's texture seen along each integral's direction (lights.hlsl:495-511)
        const RectTexRec& rt = rtex[R.texture_index];
        if (i_d != 0.0f) ld = ld * rect_tex_lookup(R, rt, rtl, ltc_direction(fr, nullptr, sd), P, 1.0f);
        if (i_s != 0.0f) ls = ls * rect_tex_lookup(R, rt, rtl, ltc_direction(fr, &mt, ss), P, __builtin_sqrtf(m.rx * m.ry));
    }
    const float ndotv = dot(fr.n, V);
    if (ndotv <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    const float x = 1.0f - ndotv;
    const float p5 = ((x * x) * (x * x)) * x;
    const V3 fres = m.f0 + (m.f90 - m.f0) * p5;
    const V3 one = v3(1.0f, 1.0f, 1.0f);
    const V3 diffuse = ((one - fres) * m.base) * kInvPi;
    const V3 specular = m.f0 * bx + m.f90 * by;
    return diffuse * ld + specular * ls;
}

// --------------------------------------------------------------------- sky
PT_HD V3 sky_to_local(const SkyParams& s, V3 d) {
    const float* R = s.rot;
    return v3((R[0] * d.x + R[1] * d.y) + R[2] * d.z,
              (R[3] * d.x + R[4] * d.y) + R[5] * d.z,
              (R[6] * d.x + R[7] * d.y) + R[8] * d.z);
}
PT_HD void sky_uv(V3 ds, float* u, float* v, float* theta) {
    float y = sel_min(sel_max(ds.y, -1.0f), 1.0f);
    float th = pt_acos(y);
    float ph = ds.z == 0.0f ? 0.0f : pt_atan2(ds.x, ds.z);   // equitangular_to_cubemap.hlsl:29
    if (ph < 0.0f) ph += kTwoPi;
    *u = ph * kInvTwoPi;
    *v = th * kInvPi;
    *theta = th;
}
__device__ __forceinline__ V3 sky_texel(const SkyParams& s, int i, int j) {
    float4 t = s.texels[(size_t)i * s.width + j];
    return v3(t.x, t.y, t.z);
}
// ---- cube map (kSkyCube)
// Face of a direction and its in-face coordinates sc, tc over |ma| (Vulkan's cube face
// selection; ties go to X, then Y).  The face layout of cubemap.hlsl:3-21: face f,
// texel column from a = 2u - 1, row from b = 2v - 1, direction cube_point(f, a, b).
PT_HD uint32_t cube_face(V3 d, float* sc, float* tc, float* ma) {
    const float ax = d.x < 0.0f ? -d.x : d.x, ay = d.y < 0.0f ? -d.y : d.y, az = d.z < 0.0f ? -d.z : d.z;
    if (ax >= ay && ax >= az) {
        *ma = ax;
        *tc = -d.y;
        if (d.x >= 0.0f) { *sc = -d.z; return 0u; }
        *sc = d.z;
        return 1u;
    }
    if (ay >= az) {
        *ma = ay;
        *sc = d.x;
        if (d.y >= 0.0f) { *tc = d.z; return 2u; }
        *tc = -d.z;
        return 3u;
    }
    *ma = az;
    *tc = -d.y;
    if (d.z >= 0.0f) { *sc = d.x; return 4u; }
    *sc = -d.x;
    return 5u;
}
PT_HD V3 cube_point(uint32_t f, float a, float b) {
    switch (f) {
        case 0u: return v3(1.0f, -b, -a);
        case 1u: return v3(-1.0f, -b, a);
        case 2u: return v3(a, 1.0f, b);
        case 3u: return v3(a, -1.0f, -b);
        case 4u: return v3(a, -b, 1.0f);
        default: return v3(-a, -b, -1.0f);
    }
}
// Texel index of in-face coordinate c (= sc / ma) on a face of S texels.
PT_HD int cube_index(float c, int S) {
    const int k = (int)__builtin_floorf(((c + 1.0f) * 0.5f) * (float)S);
    return k < 0 ? 0 : (k > S - 1 ? S - 1 : k);
}
// Texel (row i, column j) of face f whose centre lies beyond one edge of the face (one
// of i, j is -1 or S): the texel of the adjacent face across that edge (Vulkan's
// seamless cube filtering).  The centre, on the face's extended plane, is one half
// texel outside; projected through the cube it lands in the adjacent face's edge row or
// column, at the same position along the edge.
PT_HD void cube_across(int S, uint32_t f, int i, int j, uint32_t* g, int* gi, int* gj) {
    const float a = (float)(2 * j + 1) / (float)S - 1.0f, b = (float)(2 * i + 1) / (float)S - 1.0f;
    float sc, tc, ma;
    *g = cube_face(cube_point(f, a, b), &sc, &tc, &ma);
    *gj = cube_index(sc / ma, S);
    *gi = cube_index(tc / ma, S);
}
__device__ __forceinline__ V3 cube_texel(const SkyParams& s, uint32_t f, int i, int j) {
    const float4 t = s.texels[((size_t)f * s.width + (size_t)i) * s.width + (size_t)j];
    return v3(t.x, t.y, t.z);
}
// A bilinear tap of face f: inside the face its texel; across one edge the adjacent
// face's texel; beyond a corner the mean of the three texels that meet there.
__device__ __forceinline__ V3 cube_tap(const SkyParams& s, uint32_t f, int i, int j) {
    const int S = (int)s.width;
    const bool oi = i < 0 || i >= S, oj = j < 0 || j >= S;
    if (!oi && !oj) return cube_texel(s, f, i, j);
    uint32_t g;
    int gi, gj;
    if (oi != oj) {
        cube_across(S, f, i, j, &g, &gi, &gj);
        return cube_texel(s, g, gi, gj);
    }
    const int ci = i < 0 ? 0 : S - 1, cj = j < 0 ? 0 : S - 1;
    const V3 c0 = cube_texel(s, f, ci, cj);
    cube_across(S, f, ci, j, &g, &gi, &gj);
    const V3 c1 = cube_texel(s, g, gi, gj);
    cube_across(S, f, i, cj, &g, &gi, &gj);
    const V3 c2 = cube_texel(s, g, gi, gj);
    const V3 m = (c0 + c1) + c2;
    return v3(m.x / 3.0f, m.y / 3.0f, m.z / 3.0f);
}
// Bilinear lookup of face f at face coordinates (a, b) in [-1, 1] (texel centres at
// (k+0.5)/S).
__device__ __forceinline__ V3 cube_eval_face(const SkyParams& s, uint32_t f, float a, float b) {
    const float S = (float)s.width;
    const float x = ((a + 1.0f) * 0.5f) * S - 0.5f, y = ((b + 1.0f) * 0.5f) * S - 0.5f;
    const float fx0 = __builtin_floorf(x), fy0 = __builtin_floorf(y);
    const float fx = x - fx0, fy = y - fy0;
    const int j0 = (int)fx0, i0 = (int)fy0;
    const V3 c00 = cube_tap(s, f, i0, j0), c10 = cube_tap(s, f, i0, j0 + 1);
    const V3 c01 = cube_tap(s, f, i0 + 1, j0), c11 = cube_tap(s, f, i0 + 1, j0 + 1);
    const V3 top = c00 + (c10 - c00) * fx;
    const V3 bot = c01 + (c11 - c01) * fx;
    return top + (bot - top) * fy;
}
// Bilinear lookup in the face the direction selects.
__device__ __forceinline__ V3 cube_eval(const SkyParams& s, V3 ds) {
    float sc, tc, ma;
    const uint32_t f = cube_face(ds, &sc, &tc, &ma);
    if (!(ma > 0.0f)) return v3(0.0f, 0.0f, 0.0f);
    return cube_eval_face(s, f, sc / ma, tc / ma);
}
// Solid-angle density of the cube's piecewise-constant texel sampling at face
// coordinates (a, b): the texel's share w / total, uniform over its S^-2 of the face's
// uv area, and dA_uv / d(omega) = r^3 / 4 with r^2 = 1 + a^2 + b^2.
PT_HD float cube_density(float w, float scale, float a, float b) {
    const float q = (1.0f + a * a) + b * b;
    return (w * scale) * (q * __builtin_sqrtf(q));
}

// Bilinear lookup times the sky colour: equirect (texel centres at (i+0.5)/W, wrap in
// u, clamp in v) or cube map.
// Bilinear lookup of the equirect at (u, v), times the sky colour.
__device__ __forceinline__ V3 equirect_eval_uv(const SkyParams& s, float u, float v) {
    const int W = (int)s.width, H = (int)s.height;
    float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    int ix0 = (int)fx0, iy0 = (int)fy0;
    int ix1 = ix0 + 1, iy1 = iy0 + 1;
    ix0 = ix0 < 0 ? ix0 + W : (ix0 >= W ? ix0 - W : ix0);
    ix1 = ix1 < 0 ? ix1 + W : (ix1 >= W ? ix1 - W : ix1);
    iy0 = iy0 < 0 ? 0 : (iy0 > H - 1 ? H - 1 : iy0);
    iy1 = iy1 < 0 ? 0 : (iy1 > H - 1 ? H - 1 : iy1);
    V3 c00 = sky_texel(s, iy0, ix0), c10 = sky_texel(s, iy0, ix1);
    V3 c01 = sky_texel(s, iy1, ix0), c11 = sky_texel(s, iy1, ix1);
    V3 top = c00 + (c10 - c00) * fx;
    V3 bot = c01 + (c11 - c01) * fx;
    V3 c = top + (bot - top) * fy;
    return c * v3(s.color[0], s.color[1], s.color[2]);
}
__device__ __forceinline__ V3 sky_eval(const SkyParams& s, V3 d) {
    if (s.width == 0) return v3(0.0f, 0.0f, 0.0f);
    if (s.layout == kSkyCube) return cube_eval(s, sky_to_local(s, d)) * v3(s.color[0], s.color[1], s.color[2]);
    float u, v, th;
    sky_uv(sky_to_local(s, d), &u, &v, &th);
    return equirect_eval_uv(s, u, v);
}
__device__ __forceinline__ float sky_pdf(const SkyParams& s, V3 d) {
    if (!(s.total > 0.0f)) return 0.0f;
    if (s.layout == kSkyCube) {
        float sc, tc, ma;
        const uint32_t f = cube_face(sky_to_local(s, d), &sc, &tc, &ma);
        if (!(ma > 0.0f)) return 0.0f;
        const int S = (int)s.width;
        const float a = sc / ma, b = tc / ma;
        const float w = s.weight[((size_t)f * S + cube_index(b, S)) * S + cube_index(a, S)];
        if (!(w > 0.0f)) return 0.0f;
        return cube_density(w, s.pdf_scale, a, b);
    }
    float u, v, th;
    sky_uv(sky_to_local(s, d), &u, &v, &th);
    const int W = (int)s.width, H = (int)s.height;
    int j = (int)(u * (float)W), i = (int)(v * (float)H);
    j = j < 0 ? 0 : (j > W - 1 ? W - 1 : j);
    i = i < 0 ? 0 : (i > H - 1 ? H - 1 : i);
    float w = s.weight[(size_t)i * W + j];
    float st = pt_sin(th);
    if (!(st > 0.0f) || !(w > 0.0f)) return 0.0f;
    return (w * s.pdf_scale) / (kTwoPiSq * st);
}
// Piecewise-constant importance sampling of the texel weights by alias tables: u0 picks a
// row entry and its fraction decides between the entry and its alias, the rescaled
// fraction placing the point inside the row; u1 the same within the row's texels.  The
// sky's radiance there (le) is looked up at the sampled texture coordinates.
__device__ __forceinline__ bool sky_sample(const SkyParams& s, float u0, float u1, V3* dir, float* pdf, V3* le) {
    *pdf = 0.0f;
    if (!(s.total > 0.0f)) return false;
    const int W = (int)s.width, H = (int)s.height;
    const float fr = u0 * (float)H;
    int rc = (int)fr;
    rc = rc > H - 1 ? H - 1 : rc;
    const float x = fr - (float)rc;
    const SkyRowAlias ra = s.row_alias[rc];
    int i;
    float dv;
    if (x < ra.prob) {
        i = rc;
        dv = x / ra.prob;
    } else {
        i = (int)ra.alias;
        dv = (x - ra.prob) / (1.0f - ra.prob);
    }
    dv = dv < 0.0f ? 0.0f : (dv >= 1.0f ? 0x1.fffffep-1f : dv);
    const float fc = u1 * (float)W;
    int c = (int)fc;
    c = c > W - 1 ? W - 1 : c;
    const float y = fc - (float)c;
    const SkyTexAlias ta = s.tex_alias[(size_t)i * W + c];
    int j;
    float du, w;
    if (y < ta.prob) {
        j = c;
        du = y / ta.prob;
        w = ta.w;
    } else {
        j = (int)ta.alias;
        du = (y - ta.prob) / (1.0f - ta.prob);
        w = ta.w_alias;
    }
    if (!(w > 0.0f)) return false;
    du = du < 0.0f ? 0.0f : (du >= 1.0f ? 0x1.fffffep-1f : du);
    V3 ds;
    if (s.layout == kSkyCube) {
        const uint32_t f = (uint32_t)i / (uint32_t)W;
        const int r = i - (int)f * W;
        const float a = (((float)j + du) / (float)W) * 2.0f - 1.0f;
        const float b = (((float)r + dv) / (float)W) * 2.0f - 1.0f;
        ds = normalize(cube_point(f, a, b));
        *pdf = cube_density(w, s.pdf_scale, a, b);
        *le = cube_eval_face(s, f, a, b) * v3(s.color[0], s.color[1], s.color[2]);
    } else {
        float u = ((float)j + du) / (float)W;
        float v = ((float)i + dv) / (float)H;
        float st, ct, sp, cp;
        pt_sincos(v * kPi, &st, &ct);
        pt_sincos(u * kTwoPi, &sp, &cp);
        if (!(st > 0.0f)) return false;
        ds = v3(st * sp, ct, st * cp);
        *pdf = (w * s.pdf_scale) / (kTwoPiSq * st);
        *le = equirect_eval_uv(s, u, v);
    }
    const float* R = s.rot;
    *dir = v3((R[0] * ds.x + R[3] * ds.y) + R[6] * ds.z,
              (R[1] * ds.x + R[4] * ds.y) + R[7] * ds.z,
              (R[2] * ds.x + R[5] * ds.y) + R[8] * ds.z);
    return true;
}

// --------------------------------------------------------------------- GBuffer
// The shaders' float -> uint conversion (saturating, NaN -> 0).
PT_HD uint32_t cvt_u32_sat(float x) {
    return x > 0.0f ? (x >= 4294967296.0f ? 0xffffffffu : (uint32_t)x) : 0u;
}
// pack_color_rg11b10 (pack.hlsl:14-16)
PT_HD uint32_t pack_rg11b10(V3 c) {
    return cvt_u32_sat(c.x * 2047.0f) | (cvt_u32_sat(c.y * 2047.0f) << 11) | (cvt_u32_sat(c.z * 1023.0f) << 22);
}
// A 16-bit half of pack_u32_to_unorm16x2 (pack.hlsl:6-8), h / 65536, as an rgba16 unorm
// store holds it: the exact round(h * 65535 / 65536).
PT_HD uint32_t unorm16_of_half_word(uint32_t h) { return (h * 65535u + 32768u) >> 16; }
// An rgba8 unorm store of v: round(clamp(v, 0, 1) * 255) in FP32, ties to even.
PT_HD uint32_t unorm8_of(float v) {
    const float c = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;
    return (uint32_t)__builtin_rintf(c * 255.0f);
}
// pack_normal_and_tangent (pack.hlsl:89-110): octahedral N (oct_encode) and the tangent's
// direction in N's frame (frame.hlsl:9-20), L1-normalised, its x mapped to [0, 1] and
// signed by its y.
PT_HD void pack_normal_tangent(V3 N, V3 T, float out[3]) {
    const float s = (__builtin_fabsf(N.x) + __builtin_fabsf(N.y)) + __builtin_fabsf(N.z);
    float x = N.x / s, y = N.y / s;
    const float z = N.z / s;
    if (!(z >= 0.0f)) {
        const float wx = (1.0f - __builtin_fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        const float wy = (1.0f - __builtin_fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = wx;
        y = wy;
    }
    const Frame f = make_frame(N);
    float px = dot(T, f.t), py = dot(T, f.b);
    const float l = __builtin_fabsf(px) + __builtin_fabsf(py);
    px = px / l;
    py = py / l;
    const float pxx = px * 0.5f + 0.5f;
    out[0] = x;
    out[1] = y;
    out[2] = py < 0.0f ? -pxx : pxx;
}

}  // namespace bipt
