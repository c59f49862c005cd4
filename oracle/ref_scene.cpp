// ref_scene.cpp -- scene set-up for the CPU restatement: instance flattening,
// sky importance table, LBVH build entry.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Flattening restates the two-level scene of src/graphics/accel.cpp:104-132 (per-instance
// 3x4 row-major transforms, include/bisemutum/rhi/accel.hpp:48-55) as one world-space
// triangle list; normals use the cofactor of the 3x3 part, which is proportional to the
// inverse transpose that hit.hlsl:155 applies (WorldToObject4x3 * n).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "oracle.h"
#include "ref_scene.hpp"

namespace ora {
namespace {

constexpr float kPiF = 3.14159265358979323846f;

void flatten(const ora_scene_desc& d, Scene& s) {
    uint32_t total = 0;
    for (uint32_t i = 0; i < d.num_instances; ++i) total += d.idx_cnt[d.inst_mesh[i]] / 3;
    s.num_tris = total;
    s.tri_v.resize((size_t)9 * total);
    s.tri_n.resize((size_t)9 * total);
    s.tri_mat.resize(total);
    if (s.need_uv) s.tri_uv.assign((size_t)6 * total, 0.0f);
    s.tri_tan.assign((size_t)12 * total, 0.0f);   // normal maps and the GBuffer read them
    uint32_t t = 0;
    for (uint32_t i = 0; i < d.num_instances; ++i) {
        const float* m = &d.inst_xform[12 * i];
        const uint32_t mesh = d.inst_mesh[i];
        // cofactor matrix of the linear part
        float c00 = m[5] * m[10] - m[6] * m[9], c01 = m[6] * m[8] - m[4] * m[10], c02 = m[4] * m[9] - m[5] * m[8];
        float c10 = m[2] * m[9] - m[1] * m[10], c11 = m[0] * m[10] - m[2] * m[8], c12 = m[1] * m[8] - m[0] * m[9];
        float c20 = m[1] * m[6] - m[2] * m[5], c21 = m[2] * m[4] - m[0] * m[6], c22 = m[0] * m[5] - m[1] * m[4];
        float det = (m[0] * c00 + m[1] * c01) + m[2] * c02;
        float sgn = det < 0.0f ? -1.0f : 1.0f;
        const uint32_t ntri = d.idx_cnt[mesh] / 3;
        const uint32_t* idx = &d.indices[d.idx_off[mesh]];
        const float* P = &d.positions[3 * (size_t)d.vtx_off[mesh]];
        const float* Nn = &d.normals[3 * (size_t)d.vtx_off[mesh]];
        for (uint32_t k = 0; k < ntri; ++k, ++t) {
            for (int c = 0; c < 3; ++c) {
                uint32_t vi = idx[3 * k + c];
                const size_t gv = (size_t)d.vtx_off[mesh] + vi;   // vertex in the concatenated pool
                if (s.need_uv && d.texcoords) {
                    s.tri_uv[6 * (size_t)t + 2 * c] = d.texcoords[2 * gv];
                    s.tri_uv[6 * (size_t)t + 2 * c + 1] = d.texcoords[2 * gv + 1];
                }
                {
                    // ObjectToWorld3x4 * (tangent.xyz, 0) (hit.hlsl:156); w = the bitangent sign
                    const float tx = d.tangents ? d.tangents[4 * gv] : 1.0f;
                    const float ty = d.tangents ? d.tangents[4 * gv + 1] : 0.0f;
                    const float tz = d.tangents ? d.tangents[4 * gv + 2] : 0.0f;
                    const float tw = d.tangents ? d.tangents[4 * gv + 3] : 1.0f;
                    float* g = &s.tri_tan[12 * (size_t)t + 4 * c];
                    g[0] = (m[0] * tx + m[1] * ty) + m[2] * tz;
                    g[1] = (m[4] * tx + m[5] * ty) + m[6] * tz;
                    g[2] = (m[8] * tx + m[9] * ty) + m[10] * tz;
                    g[3] = tw;
                }
                float x = P[3 * vi], y = P[3 * vi + 1], z = P[3 * vi + 2];
                s.tri_v[9 * t + 3 * c + 0] = ((m[0] * x + m[1] * y) + m[2] * z) + m[3];
                s.tri_v[9 * t + 3 * c + 1] = ((m[4] * x + m[5] * y) + m[6] * z) + m[7];
                s.tri_v[9 * t + 3 * c + 2] = ((m[8] * x + m[9] * y) + m[10] * z) + m[11];
                float nx = Nn[3 * vi], ny = Nn[3 * vi + 1], nz = Nn[3 * vi + 2];
                s.tri_n[9 * t + 3 * c + 0] = ((c00 * nx + c01 * ny) + c02 * nz) * sgn;
                s.tri_n[9 * t + 3 * c + 1] = ((c10 * nx + c11 * ny) + c12 * nz) * sgn;
                s.tri_n[9 * t + 3 * c + 2] = ((c20 * nx + c21 * ny) + c22 * nz) * sgn;
            }
            s.tri_mat[t] = d.inst_mat[i];
        }
    }
}

// Alias tables (Vose 1991): the weights scaled by n / sum split into entries below 1 and
// the rest, each a stack in index order; the top small entry keeps its scaled weight as its
// probability and takes the top large one as alias, which keeps what is left over (its
// scaled weight + the small one's - 1) and goes back to the stack it now belongs in.
// Entries left over keep themselves (probability 1).
void build_alias(const float* w, uint32_t n, float sum, float* prob, uint32_t* alias) {
    std::vector<float> q(n);
    std::vector<uint32_t> small, large;
    const float scale = sum > 0.0f ? (float)n / sum : 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
        q[k] = w[k] * scale;
        if (q[k] < 1.0f) small.push_back(k); else large.push_back(k);
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t l = small.back();
        small.pop_back();
        const uint32_t g = large.back();
        large.pop_back();
        prob[l] = q[l];
        alias[l] = g;
        q[g] = (q[g] + q[l]) - 1.0f;
        if (q[g] < 1.0f) small.push_back(g); else large.push_back(g);
    }
    for (uint32_t k : large) { prob[k] = 1.0f; alias[k] = k; }
    for (uint32_t k : small) { prob[k] = 1.0f; alias[k] = k; }
}

// Importance table over the sky texels: w = luminance(texel * sky_color) times
// sin(theta_row) (equirect) or 1/r^3 at the texel centre (cube face texel at face
// coordinates a = (2j+1)/S - 1, b = (2i+1)/S - 1, r^2 = 1 + a^2 + b^2, proportional to
// the texel's solid angle); running sums per row and over rows (cube rows: face-major).
void build_sky_table(Scene& s) {
    const uint32_t W = s.sky_w, H = s.sky_h;
    const bool cube = s.sky_layout == 1u;
    s.sky_w_tex.assign((size_t)W * H, 0.0f);
    s.sky_ccdf.assign((size_t)(W + 1) * H, 0.0f);
    s.sky_rows.assign(H, 0.0f);
    s.sky_marg.assign(H + 1, 0.0f);
    s.sky_pdf_scale = 0.0f;
    if (W == 0 || H == 0) { s.sky_total = 0.0f; return; }
    const float pi_over_h = kPiF / (float)H;
    for (uint32_t i = 0; i < H; ++i) {
        float st = 0.0f, b = 0.0f;
        if (cube) b = (float)(2 * (i % W) + 1) / (float)W - 1.0f;
        else st = sin32(((float)i + 0.5f) * pi_over_h);
        float run = 0.0f;
        for (uint32_t j = 0; j < W; ++j) {
            const float* p = &s.sky_rgba[4 * ((size_t)i * W + j)];
            float r = p[0] * s.sky_color[0], g = p[1] * s.sky_color[1], bl = p[2] * s.sky_color[2];
            float lum = (0.2126f * r + 0.7152f * g) + 0.0722f * bl;
            float w;
            if (cube) {
                float a = (float)(2 * j + 1) / (float)W - 1.0f;
                float q = (1.0f + a * a) + b * b;
                w = lum * (1.0f / (q * std::sqrt(q)));
            } else {
                w = lum * st;
            }
            if (!(w > 0.0f)) w = 0.0f;
            s.sky_w_tex[(size_t)i * W + j] = w;
            s.sky_ccdf[(size_t)i * (W + 1) + j] = run;
            run = run + w;
        }
        s.sky_ccdf[(size_t)i * (W + 1) + W] = run;
        s.sky_rows[i] = run;
        s.sky_marg[i + 1] = s.sky_marg[i] + run;
    }
    s.sky_total = s.sky_marg[H];
    s.sky_row_prob.assign(H, 1.0f);
    s.sky_row_alias.assign(H, 0u);
    s.sky_tex_prob.assign((size_t)W * H, 1.0f);
    s.sky_tex_alias.assign((size_t)W * H, 0u);
    build_alias(s.sky_rows.data(), H, s.sky_total, s.sky_row_prob.data(), s.sky_row_alias.data());
    for (uint32_t i = 0; i < H; ++i)
        build_alias(&s.sky_w_tex[(size_t)i * W], W, s.sky_rows[i], &s.sky_tex_prob[(size_t)i * W],
                    &s.sky_tex_alias[(size_t)i * W]);
    if (s.sky_total > 0.0f)
        s.sky_pdf_scale = cube ? ((float)W * (float)W / s.sky_total) * 0.25f : (float)(W * H) / s.sky_total;
}

}  // namespace
}  // namespace ora

using namespace ora;

extern "C" {

ora_ctx* ora_create(const ora_scene_desc* d) {
    Scene* s = new Scene();
    s->materials.assign(d->materials, d->materials + kMatFloats * (size_t)d->num_materials);
    // texture references (index + 1) in words 17 (opacity) and 23-26 (colour, f0, normal, roughness)
    for (uint32_t m = 0; m < d->num_materials; ++m) {
        uint32_t w[32];
        std::memcpy(w, &d->materials[kMatFloats * m], sizeof w);
        s->need_uv = s->need_uv || w[17] || w[23] || w[24] || w[25] || w[26];
        s->need_tan = s->need_tan || w[25];
    }
    flatten(*d, *s);
    for (uint32_t k = 0; k < d->num_textures; ++k) {
        const ora_texture& T = d->textures[k];
        Tex t;
        t.width = T.width; t.height = T.height; t.format = T.format; t.filter = T.filter;
        t.address_u = T.address_u; t.address_v = T.address_v;
        const size_t words = (size_t)T.width * T.height * (T.format == 2u ? 4u : 1u);
        t.words.resize(words);
        std::memcpy(t.words.data(), T.texels, 4 * words);
        s->textures.push_back(std::move(t));
    }
    // sRGB byte -> linear (IEC 61966-2-1), evaluated in double and rounded once
    for (int c = 0; c < 256; ++c) {
        const double x = c / 255.0;
        s->srgb[c] = (float)(x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4));
    }
    s->num_dir_lights = d->num_dir_lights;
    s->dir_lights.assign(d->dir_lights, d->dir_lights + 16 * (size_t)d->num_dir_lights);
    s->num_rect_lights = d->rect_lights ? d->num_rect_lights : 0;
    if (s->num_rect_lights) {
        s->rect_lights.assign(d->rect_lights, d->rect_lights + 28 * (size_t)d->num_rect_lights);
        s->ltc.assign((size_t)4 * 4096 * 4, 0.0f);
        for (int k = 0; k < 3; ++k)
            if (d->ltc_matrix[k]) std::copy(d->ltc_matrix[k], d->ltc_matrix[k] + 4096 * 4, s->ltc.begin() + (size_t)k * 4096 * 4);
        if (d->ltc_norm)
            for (size_t i = 0; i < 4096; ++i) {
                s->ltc[((size_t)3 * 4096 + i) * 4] = d->ltc_norm[2 * i];
                s->ltc[((size_t)3 * 4096 + i) * 4 + 1] = d->ltc_norm[2 * i + 1];
            }
    }
    for (uint32_t k = 0; d->rect_textures && k < d->num_rect_textures; ++k) {
        const ora_rect_texture& T = d->rect_textures[k];
        RectTex r;
        r.mag_filter = T.mag_filter; r.min_filter = T.min_filter; r.mipmap_mode = T.mipmap_mode;
        r.lod_bias = T.lod_bias; r.lod_min = T.lod_min; r.lod_max = T.lod_max;
        const uint32_t* src = static_cast<const uint32_t*>(T.texels);
        uint32_t w = T.width, h = T.height;
        for (uint32_t l = 0; l < T.num_levels; ++l) {
            Tex t;
            t.width = w; t.height = h; t.format = T.format; t.address_u = T.address_u; t.address_v = T.address_v;
            const size_t words = (size_t)w * h * (T.format == 2u ? 4u : 1u);
            t.words.assign(src, src + words);
            src += words;
            r.levels.push_back(std::move(t));
            w = w / 2 > 1 ? w / 2 : 1;
            h = h / 2 > 1 ? h / 2 : 1;
        }
        s->rect_textures.push_back(std::move(r));
    }
    s->num_point_lights = d->point_lights ? d->num_point_lights : 0;
    if (s->num_point_lights)
        s->point_lights.assign(d->point_lights, d->point_lights + 16 * (size_t)d->num_point_lights);
    s->sky_w = d->sky_rgba ? d->sky_w : 0;
    s->sky_h = d->sky_rgba ? d->sky_h : 0;
    s->sky_layout = d->sky_rgba ? d->sky_layout : 0;
    if (d->sky_rgba) s->sky_rgba.assign(d->sky_rgba, d->sky_rgba + 4 * (size_t)d->sky_w * d->sky_h);
    std::memcpy(s->sky_color, d->sky_color, sizeof s->sky_color);
    std::memcpy(s->sky_rot, d->sky_rot, sizeof s->sky_rot);
    build_sky_table(*s);
    build_bvh(*s);
    return reinterpret_cast<ora_ctx*>(s);
}

void ora_set_treelet_rounds(ora_ctx* ctx, uint32_t rounds) {
    Scene* s = reinterpret_cast<Scene*>(ctx);
    s->treelet_rounds = rounds;
    build_bvh(*s);
}

void ora_destroy(ora_ctx* ctx) {
    Scene* s = reinterpret_cast<Scene*>(ctx);
    free_bvh64(*s);
    delete s;
}

uint32_t ora_num_tris(const ora_ctx* ctx) { return reinterpret_cast<const Scene*>(ctx)->num_tris; }

void ora_get_world(const ora_ctx* ctx, float* tri_v, float* tri_n, uint32_t* tri_mat) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    std::memcpy(tri_v, s.tri_v.data(), s.tri_v.size() * sizeof(float));
    std::memcpy(tri_n, s.tri_n.data(), s.tri_n.size() * sizeof(float));
    std::memcpy(tri_mat, s.tri_mat.data(), s.tri_mat.size() * sizeof(uint32_t));
}

void ora_get_bvh(const ora_ctx* ctx, uint32_t* morton, uint32_t* order, int32_t* children,
                 float* child_bounds, int32_t* parents, float* scene_centroid_bounds) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    const uint32_t n = s.num_tris;
    if (n == 0) return;
    std::memcpy(morton, s.morton.data(), n * sizeof(uint32_t));
    std::memcpy(order, s.order.data(), n * sizeof(uint32_t));
    const uint32_t ni = n > 1 ? n - 1 : 1;
    for (uint32_t i = 0; i < ni; ++i) {
        children[2 * i] = s.nodes[i].child[0];
        children[2 * i + 1] = s.nodes[i].child[1];
        std::memcpy(&child_bounds[12 * i], s.nodes[i].box[0], 6 * sizeof(float));
        std::memcpy(&child_bounds[12 * i + 6], s.nodes[i].box[1], 6 * sizeof(float));
    }
    if (parents) std::memcpy(parents, s.parent.data(), s.parent.size() * sizeof(int32_t));
    if (scene_centroid_bounds) std::memcpy(scene_centroid_bounds, s.cbounds, 6 * sizeof(float));
}

void ora_get_bvh4(const ora_ctx* ctx, int32_t* child4, float* box4) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    for (size_t i = 0; i < s.nodes4.size(); ++i) {
        const Node4& q = s.nodes4[i];
        std::memcpy(&child4[4 * i], q.child, 4 * sizeof(int32_t));
        for (int a = 0; a < 3; ++a) {
            std::memcpy(&box4[24 * i + 4 * a], q.lo[a], 4 * sizeof(float));
            std::memcpy(&box4[24 * i + 12 + 4 * a], q.hi[a], 4 * sizeof(float));
        }
    }
}

}  // extern "C"
