// ref_scene.hpp -- scene state shared by the FP32 and FP64 restatements.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#pragma once

#include <cstdint>
#include <vector>

namespace ora {

struct Node {
    int32_t child[2] = {0, 0};   // internal k >= 0, leaf j encoded as ~j
    float box[2][6] = {};        // per child: lo.xyz, hi.xyz
};

// BVH4 traversal node collapsed from BVH2 node k (stored at index k): child j's box is
// (lo[0][j], lo[1][j], lo[2][j]) - (hi[0][j], hi[1][j], hi[2][j]); an empty slot has the
// box (+inf)^3 - (+inf)^3 and child 0.
struct Node4 {
    float lo[3][4];
    float hi[3][4];
    int32_t child[4];
};

constexpr size_t kMatFloats = 32;   // floats per material record (bi_pt_material)

// Level-0 material texture: one word per RGBA8 texel, four per RGBA32F texel.
struct Tex {
    uint32_t width = 0, height = 0, format = 0, filter = 0, address_u = 0, address_v = 0;
    std::vector<uint32_t> words;
};

// Rect-light texture: its mip levels (Tex records, filter unused) and sampler.
struct RectTex {
    std::vector<Tex> levels;
    uint32_t mag_filter = 0, min_filter = 0, mipmap_mode = 0;
    float lod_bias = 0.0f, lod_min = 0.0f, lod_max = 0.0f;
};

struct Scene {
    uint32_t num_tris = 0;
    std::vector<float> tri_v;      // 9 per triangle, world space, instance order
    std::vector<float> tri_n;      // 9 per triangle, cofactor-transformed normals
    std::vector<uint32_t> tri_mat;
    std::vector<float> materials;  // kMatFloats per material
    // texture coordinates (6 per triangle) and world tangents (12 per triangle: xyz +
    // bitangent sign per vertex), kept when the materials sample textures / normal maps
    bool need_uv = false, need_tan = false;
    std::vector<float> tri_uv, tri_tan;
    std::vector<Tex> textures;
    float srgb[256] = {};
    uint32_t num_dir_lights = 0;
    std::vector<float> dir_lights; // 16 per light
    uint32_t num_point_lights = 0;
    std::vector<float> point_lights;   // 16 per light (layout of bi_pt_point_light); after the directional ones
    uint32_t num_rect_lights = 0;
    std::vector<float> rect_lights;    // 28 per light (layout of bi_pt_rect_light)
    std::vector<float> ltc;            // 4 tables x 4096 texels x 4 floats: matrix rows 0-2 (xyz), norm (xy)
    std::vector<RectTex> rect_textures;
    // sky
    uint32_t sky_w = 0, sky_h = 0;
    uint32_t sky_layout = 0;       // 0 equirect (W x H), 1 cube map (six W x W faces, H = 6 W)
    std::vector<float> sky_rgba;
    float sky_color[3] = {0, 0, 0};
    float sky_rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::vector<float> sky_w_tex;  // per texel luminance * sin(theta_row) (cube: * 1/r^3)
    // alias tables (Vose): rows by their sums, each row's texels by their weights
    std::vector<float> sky_row_prob, sky_tex_prob;
    std::vector<uint32_t> sky_row_alias, sky_tex_alias;
    std::vector<float> sky_ccdf;   // per row W+1 running sums
    std::vector<float> sky_rows;   // per row sum
    std::vector<float> sky_marg;   // H+1 running sums of rows
    float sky_total = 0.0f;
    float sky_pdf_scale = 0.0f;    // W*H/total (equirect), W*W/(4*total) (cube)
    // LBVH (FP32 transcription)
    float cbounds[6] = {};
    std::vector<uint32_t> morton, order, leaf_tri;
    std::vector<Node> nodes;
    std::vector<Node4> nodes4;
    std::vector<int32_t> parent;
    uint32_t treelet_rounds = 3;   // restructuring rounds after the refit (as the GPU's default)
    std::vector<float> leaf_v0, leaf_e1, leaf_e2;
    // FP64 oracle BVH (built lazily by ref64.cpp)
    mutable void* bvh64 = nullptr;
};

uint32_t expand10(uint32_t v);
uint32_t morton3(uint32_t x, uint32_t y, uint32_t z);
void build_bvh(Scene& s);
uint32_t pixel_of(uint32_t q, uint32_t W, uint32_t H);
void sincos32(float x, float* s, float* c);
float sin32(float x);
float cos32(float x);
float acos32(float x);
float atan2_32(float y, float x);
void free_bvh64(const Scene& s);

}  // namespace ora
