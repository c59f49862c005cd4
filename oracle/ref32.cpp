// ref32.cpp -- scalar FP32 transcription of the wavefront path-tracing pass.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -O3 -ffp-contract=off and
// no fast-math: every float operation below rounds once, in the order written,
// exactly as the HIP kernels do, so the GPU is compared with this file bit for bit.
//
// What each part restates (reference paths relative to bisemutum/):
//   camera_ray        shaders/renderer/raytracing/direction_sample/generate_camera_ray.hlsl:9-15,
//                     camera_position_world shaders/core/shader_params/camera.hlsl:7-9
//   trace_closest     rt_gbuffer.hlsl:17-35 (TMin 0.001, TMax ray_length, closest hit).  The
//                     reference's BVH lives in the GPU driver; the LBVH here is the build's own
//                     (Karras 2012) and has no reference counterpart.
//   shade_surface     hits/rt_gbuffer_hit.hlsl:6-17 + core/raytracing/hit.hlsl:75-81,155-156
//                     (interpolated normal, two-sided flip), lit.hlsl:5-35 (BSDF),
//                     material/utils.hlsl:40-121 (Fresnel, GGX, VNDF), frame.hlsl:9-20
//   light loop        deferred_lighting_secondary.hlsl:45-60, lights.hlsl:10-13,69-101
//   sky               deferred_lighting_secondary.hlsl:24-29, equitangular_to_cubemap.hlsl:28-33,
//                     cube map faces utils/cubemap.hlsl:3-21
//   bounce loop       path_tracing.cpp:289-460, sample_secondary_ray.hlsl:17-68
//   accumulation      pt_accumulate.hlsl:9 as a running FP32 sum over samples
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "ref_scene.hpp"

namespace ora {

// ---------------------------------------------------------------- float math
float sin32(float x);
float cos32(float x);

namespace {

constexpr float kFourOverPi = 1.27323954473516268615f;
// Cody-Waite split of pi/4 (Cephes sinf.c, S. L. Moshier)
constexpr float kDP1 = 0.78515625f;
constexpr float kDP2 = 2.4187564849853515625e-4f;
constexpr float kDP3 = 3.77489497744594108e-8f;

// Cephes single-precision kernels on |x| <= pi/4.
inline float sin_kernel(float x, float z) {
    return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
inline float cos_kernel(float z) {
    return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
        - 0.5f * z + 1.0f;
}

// Reduce |x| to octant j in {0..3} (after folding) and remainder r.
inline float reduce_octant(float ax, int& j) {
    int q = (int)(ax * kFourOverPi);
    float y = (float)q;
    if (q & 1) { q += 1; y += 1.0f; }
    j = q & 7;
    return ((ax - y * kDP1) - y * kDP2) - y * kDP3;
}

}  // namespace

void sincos32(float x, float* s_out, float* c_out) {
    float ax = x < 0.0f ? -x : x;
    int j;
    float r = reduce_octant(ax, j);
    float z = r * r;
    float sk = sin_kernel(r, z);
    float ck = cos_kernel(z);
    float s, c;
    int sign_s = x < 0.0f ? -1 : 1;
    int sign_c = 1;
    if (j > 3) { j -= 4; sign_s = -sign_s; sign_c = -sign_c; }
    if (j > 1) sign_c = -sign_c;
    if (j == 1 || j == 2) { s = ck; c = sk; } else { s = sk; c = ck; }
    *s_out = sign_s < 0 ? -s : s;
    *c_out = sign_c < 0 ? -c : c;
}
float sin32(float x) { float s, c; sincos32(x, &s, &c); return s; }
float cos32(float x) { float s, c; sincos32(x, &s, &c); return c; }

namespace {
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kPiO2F = 1.57079632679489661923f;
constexpr float kPiO4F = 0.785398163397448309616f;
}

// Cephes asinf.c / acosf.c
float asin32(float x) {
    float a = x < 0.0f ? -x : x;
    float z, s;
    bool big = a > 0.5f;
    if (big) { z = 0.5f * (1.0f - a); s = std::sqrt(z); }
    else { s = a; z = a * a; }
    float p = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z
                + 7.4953002686e-2f) * z + 1.6666752422e-1f) * z * s + s;
    if (big) { p = p + p; p = kPiO2F - p; }
    return x < 0.0f ? -p : p;
}
float acos32(float x) {
    if (x < -0.5f) return kPiF - 2.0f * asin32(std::sqrt(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * asin32(std::sqrt(0.5f * (1.0f - x)));
    return kPiO2F - asin32(x);
}
// Cephes atanf.c
float atan32(float x) {
    float a = x < 0.0f ? -x : x;
    float y0;
    if (a > 2.414213562373095f) { y0 = kPiO2F; a = -(1.0f / a); }
    else if (a > 0.4142135623730950f) { y0 = kPiO4F; a = (a - 1.0f) / (a + 1.0f); }
    else { y0 = 0.0f; }
    float z = a * a;
    float r = y0 + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z
                     - 3.33329491539e-1f) * z * a + a);
    return x < 0.0f ? -r : r;
}
float atan2_32(float y, float x) {
    if (x == 0.0f) {
        if (y > 0.0f) return kPiO2F;
        if (y < 0.0f) return -kPiO2F;
        return 0.0f;
    }
    float r = atan32(y / x);
    if (x < 0.0f) r = (y < 0.0f) ? r - kPiF : r + kPiF;
    return r;
}

namespace {

constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvTwoPi = 0.159154943091895335769f;
constexpr float kTwoPiSq = 19.7392088021787172376f;
constexpr float kTMin = 0.001f;
constexpr float kShadowTMax = 1.70141173e38f;   // half of FLT_MAX: tmax * kRobust stays finite
constexpr float kRobust = 1.00000036f;   // 1 + 2*gamma(3), Ize 2013 robust slab test
constexpr float kThroughputMin = 0.001f; // sample_secondary_ray.hlsl:24

struct F3 { float x, y, z; };
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 operator-(F3 a) { return f3(-a.x, -a.y, -a.z); }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline F3 normalize(F3 v) { float inv = 1.0f / std::sqrt(dot(v, v)); return v * inv; }
inline float fmax_(float a, float b) { return a > b ? a : b; }
inline float fmin_(float a, float b) { return a < b ? a : b; }
inline float max3(F3 v) { return fmax_(fmax_(v.x, v.y), v.z); }
inline bool any_pos(F3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }
inline bool finite3(F3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// ---------------------------------------------------------------- PCG32
// O'Neill 2014, pcg32_random_r / pcg32_srandom_r.
struct Pcg { uint64_t state, inc; };
inline uint32_t pcg_next(Pcg& r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
inline Pcg pcg_seed(uint64_t initstate, uint64_t initseq) {
    Pcg r{0u, (initseq << 1u) | 1u};
    pcg_next(r);
    r.state += initstate;
    pcg_next(r);
    return r;
}
// Stream keyed by (seed, pixel, sample, bounce).
inline Pcg pcg_keyed(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    return pcg_seed(((uint64_t)pixel << 32) | sample, ((uint64_t)seed << 8) | (bounce & 0xffu));
}
inline float pcg_float(Pcg& r) { return (float)(pcg_next(r) >> 8) * 0x1p-24f; }
constexpr uint32_t kJitterBounce = 255u;
constexpr uint32_t kOraGbufferQuantize = ORA_FLAG_GBUFFER_QUANTIZE;   // drift mode (oracle.h)

}  // namespace
}  // namespace ora

// ====================================================================== scene
namespace ora {

uint32_t expand10(uint32_t v) {
    v &= 1023u;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
}

namespace {

inline int clz32(uint32_t v) { return v == 0 ? 32 : __builtin_clz(v); }

// Karras 2012 delta: common prefix of (code, index) keys; -1 outside [0, n).
inline int delta(const uint32_t* codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

}  // namespace

// Treelet restructuring (Karras & Aila 2013) of the refit tree, children before
// parents: node n's treelet starts as its two children; the treelet leaf that is an
// internal node with the largest box surface (ties: the lower position) is replaced in
// place by its first child and its second child appended, until 7 leaves or only
// triangles remain.  cost(s) = area(s) + the cheapest split of s (parts holding the
// lowest leaf first, p = (p - 1) & s, first minimum kept); a single leaf costs its
// subtree's cost (a triangle: its box's surface).  A cheaper topology that does not
// raise the subtree's height, or keeps it <= 48, replaces the treelet, its internal
// nodes assigned in pre-order (left part first).  Only nodes whose subtree is at least
// 5 levels tall at the start of the round are restructured.  The GPU processes them
// by that height, one launch per height, which gives the same result.
namespace {
constexpr int kTreeletLeaves = 7;
constexpr int kTreeletMaxHeight = 48;
constexpr int kTreeletMinHeight = 5;   // lower subtrees are left as built
inline float half_area6(const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return (dx * dy + dy * dz) + dz * dx;
}
void treelet_union(const float (*bx)[6], int k, int s, float* u) {
    bool first = true;
    for (int i = 0; i < k; ++i) {
        if (!((s >> i) & 1)) continue;
        for (int e = 0; e < 3; ++e) {
            u[e] = first ? bx[i][e] : fmin_(u[e], bx[i][e]);
            u[3 + e] = first ? bx[i][3 + e] : fmax_(u[3 + e], bx[i][3 + e]);
        }
        first = false;
    }
}
void treelet_process(Scene& s, std::vector<float>& tcost, std::vector<int>& theight, int n, bool optimize) {
    const int N = (int)s.num_tris;
    std::vector<Node>& nodes = s.nodes;
    int ref[kTreeletLeaves];
    float bx[kTreeletLeaves][6];
    int inner[kTreeletLeaves - 1];
    int k = 2, ninner = 1;
    inner[0] = n;
    for (int c = 0; c < 2; ++c) {
        ref[c] = nodes[n].child[c];
        std::memcpy(bx[c], nodes[n].box[c], sizeof bx[c]);
    }
    while (k < kTreeletLeaves) {
        int best = -1;
        float ba = -1.0f;
        for (int i = 0; i < k; ++i) {
            if (ref[i] < 0) continue;
            const float ar = half_area6(bx[i]);
            if (ar > ba) { ba = ar; best = i; }
        }
        if (best < 0) break;
        const int m = ref[best];
        inner[ninner++] = m;
        ref[best] = nodes[m].child[0];
        ref[k] = nodes[m].child[1];
        std::memcpy(bx[best], nodes[m].box[0], sizeof bx[best]);
        std::memcpy(bx[k], nodes[m].box[1], sizeof bx[k]);
        ++k;
    }
    auto leaf_cost = [&](int i) { return ref[i] >= 0 ? tcost[ref[i]] : half_area6(bx[i]); };
    auto leaf_height = [&](int i) { return ref[i] >= 0 ? theight[ref[i]] : 0; };
    float all[6];
    for (int e = 0; e < 3; ++e) {
        all[e] = fmin_(nodes[n].box[0][e], nodes[n].box[1][e]);
        all[3 + e] = fmax_(nodes[n].box[0][3 + e], nodes[n].box[1][3 + e]);
    }
    auto child_cost = [&](int c) {
        const int r = nodes[n].child[c];
        return r >= 0 ? tcost[r] : half_area6(nodes[n].box[c]);
    };
    auto child_height = [&](int c) {
        const int r = nodes[n].child[c];
        return r >= 0 ? theight[r] : 0;
    };
    const float cur = (half_area6(all) + child_cost(0)) + child_cost(1);
    const int curh = 1 + std::max(child_height(0), child_height(1));
    if (optimize && k >= 3) {
        const int full = (1 << k) - 1;
        float cost[1 << kTreeletLeaves];
        int split[1 << kTreeletLeaves], hh[1 << kTreeletLeaves];
        for (int sub = 1; sub <= full; ++sub) {
            if ((sub & (sub - 1)) == 0) {
                const int i = __builtin_ctz(sub);
                cost[sub] = leaf_cost(i);
                hh[sub] = leaf_height(i);
                split[sub] = 0;
                continue;
            }
            float u[6];
            treelet_union(bx, k, sub, u);
            const int low = sub & -sub;
            float best = 3.0e38f;
            int bp = 0;
            for (int p = (sub - 1) & sub; p > 0; p = (p - 1) & sub) {
                if (!(p & low)) continue;
                const float c = cost[p] + cost[sub ^ p];
                if (c < best) { best = c; bp = p; }
            }
            cost[sub] = half_area6(u) + best;
            split[sub] = bp;
            hh[sub] = 1 + std::max(hh[bp], hh[sub ^ bp]);
        }
        if (cost[full] < cur && (hh[full] <= curh || hh[full] <= kTreeletMaxHeight)) {
            int next = 0;
            // pre-order: the subset's node, then its left part, then its right part
            struct Rebuild {
                Scene& s; std::vector<float>& tc; std::vector<int>& th; const int* ref; const float (*bx)[6];
                int k, N; const int* inner; int& next; const float* cost; const int* split; const int* hh;
                int operator()(int sub) {
                    if ((sub & (sub - 1)) == 0) return ref[__builtin_ctz(sub)];
                    const int me = inner[next++];
                    tc[me] = cost[sub];
                    th[me] = hh[sub];
                    const int parts[2] = {split[sub], sub ^ split[sub]};
                    for (int side = 0; side < 2; ++side) {
                        const int c = (*this)(parts[side]);
                        s.nodes[me].child[side] = c;
                        treelet_union(bx, k, parts[side], s.nodes[me].box[side]);
                        s.parent[c >= 0 ? c : (N - 1) + ~c] = me;
                    }
                    return me;
                }
            } rebuild{s, tcost, theight, ref, bx, k, N, inner, next, cost, split, hh};
            rebuild(full);
            return;
        }
    }
    tcost[n] = cur;
    theight[n] = curh;
}
// Cost and height of every subtree as the refit publishes them (children first).
void subtree_costs(const Scene& s, std::vector<float>& tcost, std::vector<int>& theight, int n) {
    float c[2];
    int h[2];
    for (int k = 0; k < 2; ++k) {
        const int r = s.nodes[n].child[k];
        if (r >= 0) {
            subtree_costs(s, tcost, theight, r);
            c[k] = tcost[r];
            h[k] = theight[r];
        } else {
            c[k] = half_area6(s.nodes[n].box[k]);
            h[k] = 0;
        }
    }
    float all[6];
    for (int e = 0; e < 3; ++e) {
        all[e] = fmin_(s.nodes[n].box[0][e], s.nodes[n].box[1][e]);
        all[3 + e] = fmax_(s.nodes[n].box[0][3 + e], s.nodes[n].box[1][3 + e]);
    }
    tcost[n] = (half_area6(all) + c[0]) + c[1];
    theight[n] = 1 + std::max(h[0], h[1]);
}
// Children first; nodes whose height at the start of the round (h0) is below the
// threshold are left alone, their stored cost and height included.
void treelet_pass(Scene& s, std::vector<float>& tcost, std::vector<int>& theight, const std::vector<int>& h0,
                  int n) {
    for (int c = 0; c < 2; ++c)
        if (s.nodes[n].child[c] >= 0) treelet_pass(s, tcost, theight, h0, s.nodes[n].child[c]);
    if (h0[n] >= kTreeletMinHeight) treelet_process(s, tcost, theight, n, true);
}
}  // namespace

void build_bvh(Scene& s) {
    const uint32_t n = s.num_tris;
    s.morton.assign(n, 0);
    s.order.assign(n, 0);
    s.leaf_tri.assign(n, 0);
    if (n == 0) { s.nodes.clear(); return; }
    // triangle bounds and centroid bounds (min/max are exact: order free)
    std::vector<float> lo(3 * n), hi(3 * n);
    float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t t = 0; t < n; ++t) {
        const float* v = &s.tri_v[9 * t];
        for (int k = 0; k < 3; ++k) {
            float a = v[k], b = v[3 + k], c = v[6 + k];
            float l = fmin_(fmin_(a, b), c), h = fmax_(fmax_(a, b), c);
            lo[3 * t + k] = l + 0.0f; hi[3 * t + k] = h + 0.0f;  // -0 -> +0: box bits independent of order
            float ce = (l + h) * 0.5f;
            cmin[k] = fmin_(cmin[k], ce); cmax[k] = fmax_(cmax[k], ce);
        }
    }
    for (int k = 0; k < 3; ++k) { s.cbounds[k] = cmin[k]; s.cbounds[3 + k] = cmax[k]; }
    float inv_ext[3];
    for (int k = 0; k < 3; ++k) {
        float e = cmax[k] - cmin[k];
        inv_ext[k] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    for (uint32_t t = 0; t < n; ++t) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float ce = (lo[3 * t + k] + hi[3 * t + k]) * 0.5f;
            float f = (ce - cmin[k]) * inv_ext[k] * 1024.0f;
            int qi = (int)f;
            q[k] = (uint32_t)(qi < 0 ? 0 : (qi > 1023 ? 1023 : qi));
        }
        s.morton[t] = morton3(q[0], q[1], q[2]);
    }
    for (uint32_t t = 0; t < n; ++t) s.order[t] = t;
    std::stable_sort(s.order.begin(), s.order.end(),
                     [&](uint32_t a, uint32_t b) { return s.morton[a] < s.morton[b]; });
    std::vector<uint32_t> codes(n);
    for (uint32_t k = 0; k < n; ++k) { codes[k] = s.morton[s.order[k]]; s.leaf_tri[k] = s.order[k]; }

    const uint32_t ni = n > 1 ? n - 1 : 1;
    s.nodes.assign(ni, Node{});
    s.parent.assign(ni + n, -1);
    if (n == 1) {
        Node& nd = s.nodes[0];
        nd.child[0] = ~0; nd.child[1] = ~0;
        s.parent[ni + 0] = 0;
    } else {
        const int N = (int)n;
        for (int i = 0; i < N - 1; ++i) {
            int d = (delta(codes.data(), N, i, i + 1) - delta(codes.data(), N, i, i - 1)) >= 0 ? 1 : -1;
            int dmin = delta(codes.data(), N, i, i - d);
            int lmax = 2;
            while (delta(codes.data(), N, i, i + lmax * d) > dmin) lmax *= 2;
            int l = 0;
            for (int t = lmax / 2; t >= 1; t /= 2)
                if (delta(codes.data(), N, i, i + (l + t) * d) > dmin) l += t;
            int j = i + l * d;
            int dnode = delta(codes.data(), N, i, j);
            int sp = 0;
            int t = l;
            do {
                t = (t + 1) / 2;
                if (delta(codes.data(), N, i, i + (sp + t) * d) > dnode) sp += t;
            } while (t > 1);
            int gamma = i + sp * d + (d < 0 ? -1 : 0);
            int left = (std::min(i, j) == gamma) ? ~gamma : gamma;
            int right = (std::max(i, j) == gamma + 1) ? ~(gamma + 1) : gamma + 1;
            s.nodes[i].child[0] = left;
            s.nodes[i].child[1] = right;
            s.parent[left < 0 ? (N - 1) + ~left : left] = i;
            s.parent[right < 0 ? (N - 1) + ~right : right] = i;
        }
    }
    // refit (post-order; min/max exact)
    std::vector<float> nb(6 * ni);
    auto leaf_box = [&](uint32_t leaf, float* b) {
        uint32_t t = s.leaf_tri[leaf];
        for (int k = 0; k < 3; ++k) { b[k] = lo[3 * t + k]; b[3 + k] = hi[3 * t + k]; }
    };
    // iterative post-order over internal nodes
    std::vector<int> stack;
    std::vector<uint8_t> visited(ni, 0);
    stack.push_back(0);
    while (!stack.empty()) {
        int nd = stack.back();
        bool ready = true;
        for (int c = 0; c < 2; ++c) {
            int ch = s.nodes[nd].child[c];
            if (ch >= 0 && !visited[ch]) { stack.push_back(ch); ready = false; }
        }
        if (!ready) continue;
        stack.pop_back();
        if (visited[nd]) continue;
        visited[nd] = 1;
        float box[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int c = 0; c < 2; ++c) {
            int ch = s.nodes[nd].child[c];
            float cb[6];
            if (ch < 0) leaf_box((uint32_t)~ch, cb);
            else std::memcpy(cb, &nb[6 * ch], sizeof cb);
            std::memcpy(s.nodes[nd].box[c], cb, sizeof cb);
            for (int k = 0; k < 3; ++k) {
                box[k] = fmin_(box[k], cb[k]);
                box[3 + k] = fmax_(box[3 + k], cb[3 + k]);
            }
        }
        std::memcpy(&nb[6 * nd], box, sizeof box);
    }
    if (n >= 3) {
        std::vector<float> cost(ni);
        std::vector<int> height(ni);
        subtree_costs(s, cost, height, 0);
        for (uint32_t r = 0; r < s.treelet_rounds; ++r) {
            const std::vector<int> h0 = height;   // heights at the start of the round select the nodes
            treelet_pass(s, cost, height, h0, 0);
        }
    }
    // BVH4 collapse: every internal node k independently -- from its two children, the
    // internal child with the largest box surface (ties: the lower slot) is replaced in
    // place by its first child and its second child appended, until four slots are
    // filled or only leaves remain.  Internal links keep BVH2 indices, so they name the
    // BVH4 node collapsed from that node.
    s.nodes4.assign(ni, Node4{});
    for (uint32_t k = 0; k < ni; ++k) {
        int id[4];
        float bx[4][6];
        int cnt = 2;
        for (int c = 0; c < 2; ++c) {
            id[c] = s.nodes[k].child[c];
            std::memcpy(bx[c], s.nodes[k].box[c], sizeof bx[c]);
        }
        while (cnt < 4) {
            int pick = -1;
            float best = -1.0f;
            for (int j = 0; j < cnt; ++j) {
                if (id[j] < 0) continue;
                const float dx = bx[j][3] - bx[j][0], dy = bx[j][4] - bx[j][1], dz = bx[j][5] - bx[j][2];
                const float ar = (dx * dy + dy * dz) + dz * dx;
                if (ar > best) { best = ar; pick = j; }
            }
            if (pick < 0) break;
            const Node& m = s.nodes[id[pick]];
            id[cnt] = m.child[1];
            std::memcpy(bx[cnt], m.box[1], sizeof bx[cnt]);
            id[pick] = m.child[0];
            std::memcpy(bx[pick], m.box[0], sizeof bx[pick]);
            ++cnt;
        }
        Node4& q = s.nodes4[k];
        for (int j = 0; j < 4; ++j) {
            const bool used = j < cnt;
            for (int a = 0; a < 3; ++a) {
                q.lo[a][j] = used ? bx[j][a] : INFINITY;
                q.hi[a][j] = used ? bx[j][3 + a] : INFINITY;
            }
            q.child[j] = used ? id[j] : 0;
        }
    }
    // intersection records in leaf order: v0, e1 = v1 - v0, e2 = v2 - v0
    s.leaf_v0.resize(3 * n); s.leaf_e1.resize(3 * n); s.leaf_e2.resize(3 * n);
    for (uint32_t k = 0; k < n; ++k) {
        const float* v = &s.tri_v[9 * s.leaf_tri[k]];
        for (int a = 0; a < 3; ++a) {
            s.leaf_v0[3 * k + a] = v[a];
            s.leaf_e1[3 * k + a] = v[3 + a] - v[a];
            s.leaf_e2[3 * k + a] = v[6 + a] - v[a];
        }
    }
}

}  // namespace ora

// ================================================================ traversal
namespace ora {
// GBuffer packing helpers (defined with the GBuffer export below), used by the drift mode.
uint16_t half_rne(float f);
uint32_t pack_rg11b10(F3 c);
uint16_t unorm16_half_word(uint32_t h);
uint8_t unorm8(float v);
namespace {

// ---------------------------------------------------------------- textures
// Level-0 sampling of the material textures (SampleLevel(..., 0), material.cpp:325-349):
// texel centres at (i + 0.5) / W, bilinear weights in full FP32, Vulkan's repeat /
// mirrored-repeat / clamp-to-edge addressing, RGBA8 as c / 255 (UNORM) or through the
// sRGB table.  The same operations in the same order as the HIP kernels.
struct F4 { float x, y, z, w; };
inline float comp4(const F4& v, uint32_t c) { return c == 0u ? v.x : c == 1u ? v.y : c == 2u ? v.z : v.w; }
inline int tex_wrap(int i, int n, uint32_t mode) {
    if (mode == 2u) return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    if (mode == 1u) {
        int m = i % (2 * n);
        if (m < 0) m += 2 * n;
        return m >= n ? 2 * n - 1 - m : m;
    }
    int m = i % n;
    return m < 0 ? m + n : m;
}
inline F4 tex_texel(const Scene& s, const Tex& t, int i, int j) {
    const size_t k = (size_t)j * t.width + (size_t)i;
    if (t.format == 2u) {
        float f[4];
        std::memcpy(f, &t.words[4 * k], sizeof f);
        return F4{f[0], f[1], f[2], f[3]};
    }
    const uint32_t w = t.words[k];
    const uint32_t c0 = w & 255u, c1 = (w >> 8) & 255u, c2 = (w >> 16) & 255u, c3 = w >> 24;
    if (t.format == 1u) return F4{s.srgb[c0], s.srgb[c1], s.srgb[c2], (float)c3 / 255.0f};
    return F4{(float)c0 / 255.0f, (float)c1 / 255.0f, (float)c2 / 255.0f, (float)c3 / 255.0f};
}
inline float tex_coord(float x) { return (x > -16777216.0f && x < 16777216.0f) ? x : 0.0f; }
F4 tex_sample_t(const Scene& s, const Tex& t, uint32_t filter, float u, float v) {
    const int W = (int)t.width, H = (int)t.height;
    if (filter == 0u) {
        const int i = tex_wrap((int)std::floor(tex_coord(u * (float)W)), W, t.address_u);
        const int j = tex_wrap((int)std::floor(tex_coord(v * (float)H)), H, t.address_v);
        return tex_texel(s, t, i, j);
    }
    const float x = tex_coord(u * (float)W - 0.5f), y = tex_coord(v * (float)H - 0.5f);
    const float x0 = std::floor(x), y0 = std::floor(y);
    const float ax = x - x0, ay = y - y0;
    const int i0 = tex_wrap((int)x0, W, t.address_u), i1 = tex_wrap((int)x0 + 1, W, t.address_u);
    const int j0 = tex_wrap((int)y0, H, t.address_v), j1 = tex_wrap((int)y0 + 1, H, t.address_v);
    const F4 t00 = tex_texel(s, t, i0, j0), t10 = tex_texel(s, t, i1, j0);
    const F4 t01 = tex_texel(s, t, i0, j1), t11 = tex_texel(s, t, i1, j1);
    auto lerp = [](float a, float b, float w) { return a + (b - a) * w; };
    const F4 top{lerp(t00.x, t10.x, ax), lerp(t00.y, t10.y, ax), lerp(t00.z, t10.z, ax), lerp(t00.w, t10.w, ax)};
    const F4 bot{lerp(t01.x, t11.x, ax), lerp(t01.y, t11.y, ax), lerp(t01.z, t11.z, ax), lerp(t01.w, t11.w, ax)};
    return F4{lerp(top.x, bot.x, ay), lerp(top.y, bot.y, ay), lerp(top.z, bot.z, ay), lerp(top.w, bot.w, ay)};
}
F4 tex_sample(const Scene& s, uint32_t id, float u, float v) {
    const Tex& t = s.textures[id];
    return tex_sample_t(s, t, t.filter, u, v);
}
inline uint32_t mat_word(const Scene& s, uint32_t mat, int k) {
    uint32_t w;
    std::memcpy(&w, &s.materials[kMatFloats * mat + k], 4);
    return w;
}
// Texture coordinates of triangle `tri` at barycentrics (u, v).
inline void tri_texcoord(const Scene& s, uint32_t tri, float bu, float bv, float* su, float* sv) {
    if (!s.need_uv) { *su = 0.0f; *sv = 0.0f; return; }
    const float* q = &s.tri_uv[6 * (size_t)tri];
    *su = (q[0] + (q[2] - q[0]) * bu) + (q[4] - q[0]) * bv;
    *sv = (q[1] + (q[3] - q[1]) * bu) + (q[5] - q[1]) * bv;
}
// Whether a non-opaque triangle counts as a hit of a ray (rt_gbuffer_rahit,
// hits/rt_gbuffer_hit.hlsl:20-33): the alpha test drops it below opacity 0.01,
// translucency when the ray's number (PCG32 keyed by the job seed and the xor of the
// ray's origin and direction bits, stream 253) is below 1 - opacity.
constexpr uint32_t kOpacityBounce = 253u;
inline uint32_t fbits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
bool surface_present(const Scene& s, uint32_t tri, float bu, float bv, uint32_t seed, F3 o, F3 d) {
    const uint32_t mat = s.tri_mat[tri];
    const float* M = &s.materials[kMatFloats * mat];
    float op = M[16];
    const uint32_t otex = mat_word(s, mat, 17);
    if (otex) {
        float su, sv;
        tri_texcoord(s, tri, bu, bv, &su, &sv);
        const float c = comp4(tex_sample(s, otex - 1u, su, sv), mat_word(s, mat, 18));
        op = mat_word(s, mat, 19) ? (c < M[20] ? M[21] : M[22]) : c;
    }
    if (mat_word(s, mat, 15) == 1u) return !(op < 0.01f);
    Pcg r = pcg_keyed(seed, fbits(o.x) ^ fbits(o.y) ^ fbits(o.z) ^ fbits(d.x) ^ fbits(d.y) ^ fbits(d.z), 0u,
                      kOpacityBounce);
    return !(pcg_float(r) < 1.0f - op);
}
inline bool tri_opaque(const Scene& s, uint32_t tri) { return mat_word(s, s.tri_mat[tri], 15) == 0u; }

struct Ray { F3 o, d; float tmin, tmax; uint32_t seed = 0; };
struct Hit { float t; uint32_t leaf; float u, v; };
constexpr uint32_t kMiss = 0xffffffffu;
constexpr int kStackMax = 192;   // 3 pushes per BVH4 level over at most 63 levels

inline F3 safe_inv(F3 d) {
    auto one = [](float x) {
        float ax = x < 0.0f ? -x : x;
        if (ax < 1e-20f) x = x < 0.0f ? -1e-20f : 1e-20f;
        return 1.0f / x;
    };
    return f3(one(d.x), one(d.y), one(d.z));
}

// Slab test of one child box against [tmin, tmax]; returns entry distance.
inline bool box_hit(const float* b, F3 o, F3 inv, float tmin, float tmax, float* tnear) {
    float tx0 = (b[0] - o.x) * inv.x, tx1 = (b[3] - o.x) * inv.x;
    float ty0 = (b[1] - o.y) * inv.y, ty1 = (b[4] - o.y) * inv.y;
    float tz0 = (b[2] - o.z) * inv.z, tz1 = (b[5] - o.z) * inv.z;
    float tn = fmax_(fmax_(fmin_(tx0, tx1), fmin_(ty0, ty1)), fmax_(fmin_(tz0, tz1), tmin));
    float tf = fmin_(fmin_(fmax_(tx0, tx1), fmax_(ty0, ty1)), fmin_(fmax_(tz0, tz1), tmax));
    *tnear = tn;
    return tn <= tf * kRobust;
}

// Moller-Trumbore with division-free barycentric acceptance; t = tn / det.
inline bool tri_test(const Scene& s, uint32_t leaf, F3 o, F3 d, float* t_out, float* un_out,
                     float* vn_out, float* det_out) {
    F3 v0 = f3(s.leaf_v0[3 * leaf], s.leaf_v0[3 * leaf + 1], s.leaf_v0[3 * leaf + 2]);
    F3 e1 = f3(s.leaf_e1[3 * leaf], s.leaf_e1[3 * leaf + 1], s.leaf_e1[3 * leaf + 2]);
    F3 e2 = f3(s.leaf_e2[3 * leaf], s.leaf_e2[3 * leaf + 1], s.leaf_e2[3 * leaf + 2]);
    F3 pv = cross(d, e2);
    float det = dot(e1, pv);
    if (det == 0.0f) return false;
    F3 tv = o - v0;
    float un = dot(tv, pv);
    F3 qv = cross(tv, e1);
    float vn = dot(d, qv);
    if (det > 0.0f) {
        if (un < 0.0f || vn < 0.0f || un + vn > det) return false;
    } else {
        if (un > 0.0f || vn > 0.0f || un + vn < det) return false;
    }
    float tn = dot(e2, qv);
    *t_out = tn / det;
    *un_out = un; *vn_out = vn; *det_out = det;
    return true;
}

// Slab test of BVH4 child j against [tmin, tmax]; returns the entry distance.
inline bool box4_hit(const Node4& nd, int j, F3 o, F3 inv, float tmin, float tmax, float* tnear) {
    const float b[6] = {nd.lo[0][j], nd.lo[1][j], nd.lo[2][j], nd.hi[0][j], nd.hi[1][j], nd.hi[2][j]};
    return box_hit(b, o, inv, tmin, tmax, tnear);
}

// Node-centric BVH4 traversal: every step tests the four child boxes of one node
// against [tmin, tbest], pushes the hit internal children other than the nearest
// (ascending slot order; the nearest is the lowest slot among equal entry distances),
// intersects the hit leaves and enters the nearest hit internal child, or pops.
// A leaf's hit at distance t is a candidate only if the leaf's own box passes the slab
// test up to t (entry distance <= t * kRobust); a candidate replaces the best if nearer
// or, at equal distance, of lower leaf index.  Boxes nest and the slab arithmetic
// rounds monotonically, so every candidate's ancestors pass for any running distance
// >= t and every candidate is reached in any visiting order: the result is the minimum
// of (t, leaf) over all candidates, whatever the order (the HIP kernel defers and
// batches its leaf tests).  A ray enters each node at most once, so a step count above
// the triangle count ends the ray (it cannot happen on a valid tree).
Hit trace_closest(const Scene& s, const Ray& r) {
    Hit h{r.tmax, kMiss, 0.0f, 0.0f};
    if (s.num_tris == 0) return h;
    F3 inv = safe_inv(r.d);
    float best_un = 0.0f, best_vn = 0.0f, best_det = 1.0f;
    int stack[kStackMax];
    int sp = 0;
    int node = 0;
    for (uint32_t steps = 1;; ++steps) {
        const Node4& nd = s.nodes4[node];
        float tn[4];
        bool hit[4];
        for (int j = 0; j < 4; ++j) hit[j] = box4_hit(nd, j, r.o, inv, r.tmin, h.t, &tn[j]);
        int nearest = -1;
        for (int j = 0; j < 4; ++j)
            if (hit[j] && nd.child[j] >= 0 && (nearest < 0 || tn[j] < tn[nearest])) nearest = j;
        bool ok = true;
        for (int j = 0; j < 4; ++j) {
            if (!ok || j == nearest || !hit[j] || nd.child[j] < 0) continue;
            if (sp >= kStackMax) ok = false;
            else stack[sp++] = nd.child[j];
        }
        for (int j = 0; j < 4; ++j) {
            if (!hit[j] || nd.child[j] >= 0) continue;
            const uint32_t leaf = (uint32_t)~nd.child[j];
            float t, un, vn, det;
            if (tri_test(s, leaf, r.o, r.d, &t, &un, &vn, &det) && t >= r.tmin && tn[j] <= t * kRobust &&
                (t < h.t || (t == h.t && leaf < h.leaf)) &&
                (tri_opaque(s, s.leaf_tri[leaf]) ||
                 surface_present(s, s.leaf_tri[leaf], un / det, vn / det, r.seed, r.o, r.d))) {
                h.t = t; h.leaf = leaf; best_un = un; best_vn = vn; best_det = det;
            }
        }
        if (!ok || steps > s.num_tris) break;
        if (nearest >= 0) node = nd.child[nearest];
        else if (sp > 0) node = stack[--sp];
        else break;
    }
    if (h.leaf != kMiss) { h.u = best_un / best_det; h.v = best_vn / best_det; }
    return h;
}

// Any-hit BVH4 traversal in the shadow kernel's order: the hit internal child whose box
// the ray enters farthest away (the lowest slot among equals) is entered and the other hit
// internal children pushed in slot order, the hit leaves intersected in slot order.  The
// answer does not depend on the order.
bool trace_any(const Scene& s, const Ray& r) {
    if (s.num_tris == 0) return false;
    F3 inv = safe_inv(r.d);
    int stack[kStackMax];
    int sp = 0;
    int node = 0;
    for (uint32_t steps = 1;; ++steps) {
        const Node4& nd = s.nodes4[node];
        int next = -1;
        bool ok = true;
        bool hit[4];
        float tn[4];
        int far = -1;
        for (int j = 0; j < 4; ++j) {
            hit[j] = box4_hit(nd, j, r.o, inv, r.tmin, r.tmax, &tn[j]);
            if (hit[j] && nd.child[j] >= 0 && (far < 0 || tn[j] > tn[far])) far = j;
        }
        for (int j = 0; j < 4; ++j) {
            if (!hit[j] || nd.child[j] < 0) continue;
            if (j == far) next = nd.child[j];
            else if (!ok) continue;
            else if (sp >= kStackMax) ok = false;
            else stack[sp++] = nd.child[j];
        }
        for (int j = 0; j < 4; ++j) {
            if (!hit[j] || nd.child[j] >= 0) continue;
            float t, un, vn, det;
            const uint32_t leaf = (uint32_t)~nd.child[j];
            if (tri_test(s, leaf, r.o, r.d, &t, &un, &vn, &det) && t >= r.tmin && t <= r.tmax &&
                (tri_opaque(s, s.leaf_tri[leaf]) ||
                 surface_present(s, s.leaf_tri[leaf], un / det, vn / det, r.seed, r.o, r.d)))
                return true;
        }
        if (!ok || steps > s.num_tris) return false;
        if (next >= 0) node = next;
        else if (sp > 0) node = stack[--sp];
        else return false;
    }
}

// ------------------------------------------------------------------- camera
void camera_ray(const ora_camera& c, uint32_t px, uint32_t py, float jx, float jy, F3* o, F3* d) {
    float u = ((float)px + jx) / (float)c.width;
    float v = ((float)py + jy) / (float)c.height;
    float nx = u * 2.0f - 1.0f, ny = 1.0f - v * 2.0f;
    const float* ip = c.inv_proj;
    F3 q = f3(((ip[0] * nx + ip[1] * ny) + ip[2]) + ip[3],
              ((ip[4] * nx + ip[5] * ny) + ip[6]) + ip[7],
              ((ip[8] * nx + ip[9] * ny) + ip[10]) + ip[11]);
    F3 dl = normalize(q);
    const float* iv = c.inv_view;
    F3 dw = f3(iv[0] * dl.x + iv[1] * dl.y + iv[2] * dl.z,
               iv[4] * dl.x + iv[5] * dl.y + iv[6] * dl.z,
               iv[8] * dl.x + iv[9] * dl.y + iv[10] * dl.z);
    *d = normalize(dw);
    *o = f3(iv[3], iv[7], iv[11]);
}

// ------------------------------------------------------------------ material
// Checkerboard materials (word 13 = 1, the example project's checkerboard.toml
// select): cells of (int(floor(P.x)) ^ int(floor(P.z))) that are even shade with the
// record in word 14, odd cells with the material itself.  floor is clamped to
// [-2^31, 2^31 - 128] (NaN to the low end) before the conversion.
inline int checker_cell(float x) {
    float f = std::floor(x);
    if (!(f >= -2147483648.0f)) f = -2147483648.0f;
    if (f > 2147483520.0f) f = 2147483520.0f;
    return (int)f;
}
inline uint32_t resolve_mat(const std::vector<float>& mats, uint32_t id, F3 P) {
    uint32_t pattern, alt;
    std::memcpy(&pattern, &mats[kMatFloats * (size_t)id + 13], 4);
    std::memcpy(&alt, &mats[kMatFloats * (size_t)id + 14], 4);
    if (pattern == 1u && ((checker_cell(P.x) ^ checker_cell(P.z)) & 1) == 0) return alt;
    return id;
}

struct Mat {
    F3 base, f0, f90;
    float rx, ry;
    bool two_sided;
};
// get_anisotropic_roughness (utils.hlsl:54-59)
inline void mat_alpha(Mat& r, float rough, float aniso_p) {
    float aniso = std::sqrt(1.0f - aniso_p * 0.9f);
    float r2 = rough * rough;
    r.rx = fmax_(r2 / aniso, 0.001f);
    r.ry = fmax_(r2 * aniso, 0.001f);
}
inline Mat load_mat(const float* m) {
    Mat r;
    r.base = f3(m[0], m[1], m[2]);
    r.f0 = f3(m[4], m[5], m[6]);
    r.f90 = f3(m[8], m[9], m[10]);
    uint32_t ts;
    std::memcpy(&ts, &m[12], sizeof ts);
    r.two_sided = ts != 0u;
    mat_alpha(r, m[3], m[7]);
    return r;
}

struct Frame { F3 t, b, n; };
// create_frame(normal), frame.hlsl:9-20 (Duff et al. 2017)
inline Frame make_frame(F3 n) {
    float sign = n.z > 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    Frame f;
    f.t = f3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f.b = f3(b, sign + n.y * n.y * a, -n.y);
    f.n = n;
    return f;
}
inline F3 to_local(const Frame& f, F3 v) { return f3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
inline F3 to_world(const Frame& f, F3 v) { return (f.t * v.x + f.b * v.y) + f.n * v.z; }

inline float sq(float x) { return x * x; }
inline float ggx_ndf(F3 h, float rx, float ry) {
    float a = (sq(h.x / rx) + sq(h.y / ry)) + sq(h.z);
    return kInvPi / (((rx * ry) * a) * a);
}
inline float ggx_g1(F3 v, float rx, float ry) {
    float a = (sq(rx * v.x) + sq(ry * v.y)) / fmax_(v.z * v.z, 0.0001f);
    return 2.0f / (1.0f + std::sqrt(1.0f + a));
}
inline float ggx_vis_hc(F3 v, F3 l, float rx, float ry) {
    float a = l.z * std::sqrt((sq(rx * v.x) + sq(ry * v.y)) + sq(v.z));
    float b = v.z * std::sqrt((sq(rx * l.x) + sq(ry * l.y)) + sq(l.z));
    return 0.5f / fmax_(a + b, 0.0001f);
}

// surface_eval_lit (lit.hlsl:5-35): diffuse + specular, both including cos(theta_l).
F3 bsdf_eval(const Frame& fr, const Mat& m, F3 V, F3 L) {
    F3 H = normalize(V + L);
    F3 lh = to_local(fr, H), lv = to_local(fr, V), ll = to_local(fr, L);
    if (lv.z <= 0.0f || ll.z <= 0.0f) return f3(0, 0, 0);
    float vh = dot(V, H);
    float c = fmax_(vh, 0.0f);
    float x = 1.0f - c;
    float p5 = ((x * x) * (x * x)) * x;
    F3 fres = m.f0 + (m.f90 - m.f0) * p5;
    float lz = fmax_(ll.z, 0.0f);
    F3 one = f3(1.0f, 1.0f, 1.0f);
    F3 diffuse = (((one - fres) * m.base) * kInvPi) * lz;
    float ndf = ggx_ndf(lh, m.rx, m.ry);
    float vis = ggx_vis_hc(lv, ll, m.rx, m.ry);
    F3 spec = ((fres * ndf) * vis) * lz;
    return diffuse + spec;
}

// ggx_vndf_sample (utils.hlsl:98-116)
F3 vndf_sample(F3 v, float rx, float ry, float u0, float u1) {
    if (v.z < 0.0f) v = -v;
    F3 vh = normalize(f3(rx * v.x, ry * v.y, v.z));
    float len_sqr = vh.x * vh.x + vh.y * vh.y;
    F3 t1v;
    if (len_sqr > 0.0f) {
        float s = std::sqrt(len_sqr);
        t1v = f3(-vh.y / s, vh.x / s, 0.0f / s);
    } else {
        t1v = f3(1.0f, 0.0f, 0.0f);
    }
    F3 t2v = cross(vh, t1v);
    float r = std::sqrt(u0);
    float phi = kTwoPi * u1;
    float sp, cp;
    sincos32(phi, &sp, &cp);
    float t1 = r * cp;
    float t2 = r * sp;
    float s = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s) * std::sqrt(1.0f - t1 * t1) + s * t2;
    F3 nh = (t1v * t1 + t2v * t2) + vh * std::sqrt(fmax_((1.0f - t1 * t1) - t2 * t2, 0.0f));
    return normalize(f3(rx * nh.x, ry * nh.y, fmax_(nh.z, 0.0f)));
}
// ggx_vndf_sample_pdf (utils.hlsl:118-121) divided by 4|h.v| (sample_secondary_ray.hlsl:57)
inline float vndf_pdf_l(F3 h, F3 v, float rx, float ry) {
    float hv = dot(h, v);
    float pdf_h = ((ggx_g1(v, rx, ry) * ggx_ndf(h, rx, ry)) * fmax_(hv, 0.0f)) / fmax_(v.z, 0.0001f);
    float ahv = hv < 0.0f ? -hv : hv;
    return pdf_h / (4.0f * ahv);
}

// ---------------------------------------------------------------------- sky
inline F3 sky_to_local(const Scene& s, F3 d) {
    const float* R = s.sky_rot;
    return f3((R[0] * d.x + R[1] * d.y) + R[2] * d.z,
              (R[3] * d.x + R[4] * d.y) + R[5] * d.z,
              (R[6] * d.x + R[7] * d.y) + R[8] * d.z);
}
inline void sky_uv(F3 ds, float* u, float* v, float* theta) {
    float y = fmin_(fmax_(ds.y, -1.0f), 1.0f);
    float th = acos32(y);
    float ph = ds.z == 0.0f ? 0.0f : atan2_32(ds.x, ds.z);
    if (ph < 0.0f) ph += kTwoPi;
    *u = ph * kInvTwoPi;
    *v = th * kInvPi;
    *theta = th;
}
// ---- cube map sky (Scene::sky_layout 1): six W x W faces, face-major rows.
// Face selection of Vulkan's cube sampling (largest |component|, ties to X then Y) with
// the face orientation of cubemap.hlsl:3-21: face f holds the direction
// cube_dir(f, a, b) at column coordinate a = 2u - 1 and row coordinate b = 2v - 1.
int cube_select(F3 d, float* sc, float* tc, float* ma) {
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    if (ax >= ay && ax >= az) {
        *ma = ax; *tc = -d.y;
        if (d.x >= 0.0f) { *sc = -d.z; return 0; }
        *sc = d.z; return 1;
    }
    if (ay >= az) {
        *ma = ay; *sc = d.x;
        if (d.y >= 0.0f) { *tc = d.z; return 2; }
        *tc = -d.z; return 3;
    }
    *ma = az; *tc = -d.y;
    if (d.z >= 0.0f) { *sc = d.x; return 4; }
    *sc = -d.x; return 5;
}
F3 cube_dir(int f, float a, float b) {
    static const float sign_tab[6][3] = {{1, -1, -1}, {-1, -1, 1}, {1, 1, 1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, -1}};
    const float* g = sign_tab[f];
    if (f < 2) return f3(g[0], g[1] * b, g[2] * a);
    if (f < 4) return f3(g[0] * a, g[1], g[2] * b);
    return f3(g[0] * a, g[1] * b, g[2]);
}
// texel number of a face coordinate c in [-1, 1] on S texels
int cube_texel_index(float c, int S) {
    int k = (int)std::floor(((c + 1.0f) * 0.5f) * (float)S);
    return k < 0 ? 0 : (k > S - 1 ? S - 1 : k);
}
F3 cube_fetch(const Scene& s, int f, int i, int j) {
    const int S = (int)s.sky_w;
    const float* p = &s.sky_rgba[4 * (((size_t)f * S + i) * S + j)];
    return f3(p[0], p[1], p[2]);
}
// A texel one step outside face f across an edge (row i or column j is -1 or S) is the
// adjacent face's texel that contains the projection of its centre through the cube.
F3 cube_fetch_outside(const Scene& s, int f, int i, int j) {
    const int S = (int)s.sky_w;
    float a = (float)(2 * j + 1) / (float)S - 1.0f, b = (float)(2 * i + 1) / (float)S - 1.0f;
    float sc, tc, ma;
    int g = cube_select(cube_dir(f, a, b), &sc, &tc, &ma);
    return cube_fetch(s, g, cube_texel_index(tc / ma, S), cube_texel_index(sc / ma, S));
}
// Bilinear tap with seamless edges: beyond a corner, the mean of the face's corner
// texel and the two adjacent faces' texels at that corner.
F3 cube_tap_value(const Scene& s, int f, int i, int j) {
    const int S = (int)s.sky_w;
    bool out_i = i < 0 || i >= S, out_j = j < 0 || j >= S;
    if (!out_i && !out_j) return cube_fetch(s, f, i, j);
    if (out_i != out_j) return cube_fetch_outside(s, f, i, j);
    int ci = i < 0 ? 0 : S - 1, cj = j < 0 ? 0 : S - 1;
    F3 sum = (cube_fetch(s, f, ci, cj) + cube_fetch_outside(s, f, ci, j)) + cube_fetch_outside(s, f, i, cj);
    return f3(sum.x / 3.0f, sum.y / 3.0f, sum.z / 3.0f);
}
// Bilinear lookup of face f at face coordinates (a, b).
F3 cube_lookup_face(const Scene& s, int f, float a, float b) {
    const float S = (float)s.sky_w;
    float x = ((a + 1.0f) * 0.5f) * S - 0.5f;
    float y = ((b + 1.0f) * 0.5f) * S - 0.5f;
    float x0 = std::floor(x), y0 = std::floor(y);
    float fx = x - x0, fy = y - y0;
    int j = (int)x0, i = (int)y0;
    F3 c00 = cube_tap_value(s, f, i, j), c10 = cube_tap_value(s, f, i, j + 1);
    F3 c01 = cube_tap_value(s, f, i + 1, j), c11 = cube_tap_value(s, f, i + 1, j + 1);
    F3 top = c00 + (c10 - c00) * fx;
    F3 bot = c01 + (c11 - c01) * fx;
    return top + (bot - top) * fy;
}
F3 cube_lookup(const Scene& s, F3 dl) {
    float sc, tc, ma;
    int f = cube_select(dl, &sc, &tc, &ma);
    if (!(ma > 0.0f)) return f3(0, 0, 0);
    return cube_lookup_face(s, f, sc / ma, tc / ma);
}
// Solid-angle density of a point (a, b) of the texel with weight w: (w / total) * S^2
// per unit face uv area, times dA_uv / d(omega) = (1 + a^2 + b^2)^(3/2) / 4.
float cube_pdf_at(const Scene& s, float w, float a, float b) {
    float q = (1.0f + a * a) + b * b;
    return (w * s.sky_pdf_scale) * (q * std::sqrt(q));
}

// Bilinear equirect lookup at (u, v), times the sky colour.
F3 equirect_eval_uv(const Scene& s, float u, float v) {
    const int W = (int)s.sky_w, H = (int)s.sky_h;
    float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    float fx0 = std::floor(x), fy0 = std::floor(y);
    float fx = x - fx0, fy = y - fy0;
    int ix0 = (int)fx0, iy0 = (int)fy0;
    int ix1 = ix0 + 1, iy1 = iy0 + 1;
    ix0 = ix0 < 0 ? ix0 + W : (ix0 >= W ? ix0 - W : ix0);
    ix1 = ix1 < 0 ? ix1 + W : (ix1 >= W ? ix1 - W : ix1);
    iy0 = iy0 < 0 ? 0 : (iy0 > H - 1 ? H - 1 : iy0);
    iy1 = iy1 < 0 ? 0 : (iy1 > H - 1 ? H - 1 : iy1);
    auto tex = [&](int i, int j) {
        const float* p = &s.sky_rgba[4 * ((size_t)i * W + j)];
        return f3(p[0], p[1], p[2]);
    };
    F3 c00 = tex(iy0, ix0), c10 = tex(iy0, ix1), c01 = tex(iy1, ix0), c11 = tex(iy1, ix1);
    F3 top = c00 + (c10 - c00) * fx;
    F3 bot = c01 + (c11 - c01) * fx;
    F3 c = top + (bot - top) * fy;
    return c * f3(s.sky_color[0], s.sky_color[1], s.sky_color[2]);
}
F3 sky_eval(const Scene& s, F3 d) {
    if (s.sky_w == 0) return f3(0, 0, 0);
    if (s.sky_layout == 1u)
        return cube_lookup(s, sky_to_local(s, d)) * f3(s.sky_color[0], s.sky_color[1], s.sky_color[2]);
    float u, v, th;
    sky_uv(sky_to_local(s, d), &u, &v, &th);
    return equirect_eval_uv(s, u, v);
}
float sky_pdf(const Scene& s, F3 d) {
    if (!(s.sky_total > 0.0f)) return 0.0f;
    if (s.sky_layout == 1u) {
        float sc, tc, ma;
        int f = cube_select(sky_to_local(s, d), &sc, &tc, &ma);
        if (!(ma > 0.0f)) return 0.0f;
        const int S = (int)s.sky_w;
        float a = sc / ma, b = tc / ma;
        float w = s.sky_w_tex[((size_t)f * S + cube_texel_index(b, S)) * S + cube_texel_index(a, S)];
        if (!(w > 0.0f)) return 0.0f;
        return cube_pdf_at(s, w, a, b);
    }
    float u, v, th;
    sky_uv(sky_to_local(s, d), &u, &v, &th);
    const int W = (int)s.sky_w, H = (int)s.sky_h;
    int j = (int)(u * (float)W), i = (int)(v * (float)H);
    j = j < 0 ? 0 : (j > W - 1 ? W - 1 : j);
    i = i < 0 ? 0 : (i > H - 1 ? H - 1 : i);
    float w = s.sky_w_tex[(size_t)i * W + j];
    float st = sin32(th);
    if (!(st > 0.0f) || !(w > 0.0f)) return 0.0f;
    return (w * s.sky_pdf_scale) / (kTwoPiSq * st);
}
// Piecewise-constant importance sampling of the texel weights by alias tables: u0 scaled
// by the row count picks a row entry, whose fraction chooses the entry or its alias and,
// rescaled, places the point in the row; u1 the same over the row's texels.  le: the sky's
// radiance at the sampled texture coordinates.
bool sky_sample(const Scene& s, float u0, float u1, F3* dir, float* pdf, F3* le) {
    *pdf = 0.0f;
    if (!(s.sky_total > 0.0f)) return false;
    const int W = (int)s.sky_w, H = (int)s.sky_h;
    const float fr = u0 * (float)H;
    int rc = (int)fr;
    rc = rc > H - 1 ? H - 1 : rc;
    const float x = fr - (float)rc;
    const float rp = s.sky_row_prob[rc];
    int i;
    float dv;
    if (x < rp) {
        i = rc;
        dv = x / rp;
    } else {
        i = (int)s.sky_row_alias[rc];
        dv = (x - rp) / (1.0f - rp);
    }
    dv = dv < 0.0f ? 0.0f : (dv >= 1.0f ? 0x1.fffffep-1f : dv);
    const float fc = u1 * (float)W;
    int c = (int)fc;
    c = c > W - 1 ? W - 1 : c;
    const float y = fc - (float)c;
    const float tp = s.sky_tex_prob[(size_t)i * W + c];
    int j;
    float du;
    if (y < tp) {
        j = c;
        du = y / tp;
    } else {
        j = (int)s.sky_tex_alias[(size_t)i * W + c];
        du = (y - tp) / (1.0f - tp);
    }
    const float w = s.sky_w_tex[(size_t)i * W + j];
    if (!(w > 0.0f)) return false;
    du = du < 0.0f ? 0.0f : (du >= 1.0f ? 0x1.fffffep-1f : du);
    F3 ds;
    if (s.sky_layout == 1u) {
        int f = i / W, row = i - f * W;
        float a = (((float)j + du) / (float)W) * 2.0f - 1.0f;
        float b = (((float)row + dv) / (float)W) * 2.0f - 1.0f;
        ds = normalize(cube_dir(f, a, b));
        *pdf = cube_pdf_at(s, w, a, b);
        *le = cube_lookup_face(s, f, a, b) * f3(s.sky_color[0], s.sky_color[1], s.sky_color[2]);
    } else {
        float u = ((float)j + du) / (float)W;
        float v = ((float)i + dv) / (float)H;
        float st, ct, sp, cp;
        sincos32(v * kPiF, &st, &ct);
        sincos32(u * kTwoPi, &sp, &cp);
        if (!(st > 0.0f)) return false;
        ds = f3(st * sp, ct, st * cp);
        *pdf = (w * s.sky_pdf_scale) / (kTwoPiSq * st);
        *le = equirect_eval_uv(s, u, v);
    }
    const float* R = s.sky_rot;
    *dir = f3((R[0] * ds.x + R[3] * ds.y) + R[6] * ds.z,
              (R[1] * ds.x + R[4] * ds.y) + R[7] * ds.z,
              (R[2] * ds.x + R[5] * ds.y) + R[8] * ds.z);
    return true;
}

inline float power_heuristic(float a, float b) {
    float a2 = a * a;
    return a2 / (a2 + b * b);
}

// ------------------------------------------------------------- path tracing
struct PathOut {
    F3 radiance;
    uint32_t segments;
    uint8_t shadow_bits[16];
    uint32_t n_light_shadow, n_sky_shadow;
};

constexpr int kMaxDirLights = 4;

// Radiance-cache modes of a path (bi_pt_render_nrc): at the hit of segment `bounce` the
// path either ends in a cache whose output is the constant `rgb` (its radiance of that
// vertex and beyond = throughput x rgb, in the miss slot), or -- a training path --
// records the vertex (16 floats: position, direction to the previous vertex, normal,
// roughness, base colour, tag 1) and restarts with throughput 1 and no radiance, so
// what it gathers afterwards is the sample's target.
// A training path's segments from `bounce` on draw their random numbers with the seed
// xor suffix_key (0: the path's own stream, as the GPU traces it), so several runs share
// the recorded vertex and differ only after it.
struct CacheMode {
    int bounce = -1;
    bool train = false;
    F3 rgb = {0.0f, 0.0f, 0.0f};
    float* rec = nullptr;
    uint32_t suffix_key = 0;
};

// point_light_eval (lights.hlsl:15-25) of a bi_pt_point_light record L at P: the radiance
// arriving, the unit direction to the light and the distance to it (the shadow ray's extent).
F3 point_light_eval(const float* L, F3 P, F3* ldir, float* dist_out) {
    const F3 lv = f3(L[4] - P.x, L[5] - P.y, L[6] - P.z);
    const float d2 = dot(lv, lv);
    const float dist = std::sqrt(d2);
    *ldir = f3(lv.x / dist, lv.y / dist, lv.z / dist);
    *dist_out = dist;
    const float q = d2 * L[3];
    float fall = 1.0f - q * q;
    fall = fall < 0.0f ? 0.0f : (fall > 1.0f ? 1.0f : fall);
    float att = fall / fmax_(d2, 0.001f);
    const float cos_inner = L[7], cos_outer = L[11];
    if (cos_inner > cos_outer) {
        float ct = dot(*ldir, f3(L[8], L[9], L[10]));
        ct = ct < cos_outer ? cos_outer : (ct > cos_inner ? cos_inner : ct);
        att = att * ((ct - cos_outer) / fmax_(cos_inner - cos_outer, 0.001f));
    }
    return f3(L[0] * att, L[1] * att, L[2] * att);
}

// ---------------------------------------------------------------- rect lights (LTC)
// Linearly transformed cosines over a quad light (lights.hlsl:161-513; Heitz et al. 2016
// with anisotropic 4D tables): rect_light_eval_ltc and surface_eval_lut_lit
// (material/lit.hlsl:37-58), in the shading frame of make_frame(N).
struct M3 { F3 r0, r1, r2; };
inline F3 m3_mul(const M3& m, F3 v) { return f3(dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)); }
// adjugate / determinant (determinant along row 0)
inline M3 m3_inverse(const M3& m) {
    const float a = m.r0.x, b = m.r0.y, c = m.r0.z, d = m.r1.x, e = m.r1.y, f = m.r1.z, g = m.r2.x, h = m.r2.y,
                k = m.r2.z;
    const float A = e * k - f * h, B = f * g - d * k, Cc = d * h - e * g;
    const float det = (a * A + b * B) + c * Cc;
    const float s = 1.0f / det;
    M3 r;
    r.r0 = f3(A * s, (c * h - b * k) * s, (b * f - c * e) * s);
    r.r1 = f3(B * s, (a * k - c * g) * s, (c * d - a * f) * s);
    r.r2 = f3(Cc * s, (b * g - a * h) * s, (a * e - b * d) * s);
    return r;
}
// linear filter, clamp to edge, on an axis of n texels
inline void ltc_axis(float u, int n, int* i0, int* i1, float* f) {
    const float t = u * (float)n - 0.5f;
    const float fl = std::floor(t);
    *f = t - fl;
    const int a = (int)fl, b = a + 1;
    *i0 = a < 0 ? 0 : (a > n - 1 ? n - 1 : a);
    *i1 = b < 0 ? 0 : (b > n - 1 ? n - 1 : b);
}
inline float lerpf(float a, float b, float t) { return a + (b - a) * t; }
inline F4 lerp4(F4 a, F4 b, float t) {
    return F4{lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t), lerpf(a.w, b.w, t)};
}
inline F4 ltc_texel(const float* t, int i) { return F4{t[4 * i], t[4 * i + 1], t[4 * i + 2], t[4 * i + 3]}; }
// trilinear sample of one 8 x 8 x 64 table (x fastest) at normalised (x, y, z)
F4 ltc_sample(const float* t, float x, float y, float z) {
    int x0, x1, y0, y1, z0, z1;
    float fx, fy, fz;
    ltc_axis(x, 8, &x0, &x1, &fx);
    ltc_axis(y, 8, &y0, &y1, &fy);
    ltc_axis(z, 64, &z0, &z1, &fz);
    F4 e[2];
    for (int k = 0; k < 2; ++k) {
        const int zo = 64 * (k ? z1 : z0);
        const F4 c0 = lerp4(ltc_texel(t, x0 + 8 * y0 + zo), ltc_texel(t, x1 + 8 * y0 + zo), fx);
        const F4 c1 = lerp4(ltc_texel(t, x0 + 8 * y1 + zo), ltc_texel(t, x1 + 8 * y1 + zo), fx);
        e[k] = lerp4(c0, c1, fy);
    }
    return lerp4(e[0], e[1], fz);
}
// get_ltc_tex3d_coord + fetch_ltc_matrix + fetch_ltc_brdf (lights.hlsl:164-201)
void ltc_fetch(const float* ltc, float u3, float u2, float u1, float u0, M3* m, float* bx, float* by) {
    const float ws = u0 * 7.0f;
    const float ws_f = std::floor(ws);
    const float wc = std::floor(ws + 1.0f);
    const float ws_c = wc < 7.0f ? wc : 7.0f;
    const float w = ws - ws_f;
    const float x = (u3 * 7.0f + 0.5f) / 8.0f, y = (u2 * 7.0f + 0.5f) / 8.0f;
    const float z1 = ((u1 * 7.0f + 8.0f * ws_f) + 0.5f) / 64.0f, z2 = ((u1 * 7.0f + 8.0f * ws_c) + 0.5f) / 64.0f;
    F3 r[3];
    for (int k = 0; k < 3; ++k) {
        const F4 a = ltc_sample(ltc + (size_t)k * 4096 * 4, x, y, z1), b = ltc_sample(ltc + (size_t)k * 4096 * 4, x, y, z2);
        r[k] = f3(lerpf(a.x, b.x, w), lerpf(a.y, b.y, w), lerpf(a.z, b.z, w));
    }
    m->r0 = r[0]; m->r1 = r[1]; m->r2 = r[2];
    const F4 n1 = ltc_sample(ltc + (size_t)3 * 4096 * 4, x, y, z1), n2 = ltc_sample(ltc + (size_t)3 * 4096 * 4, x, y, z2);
    *bx = lerpf(n1.x, n2.x, w);
    *by = lerpf(n1.y, n2.y, w);
}
inline void quad_reverse(F3 q[4]) {
    const F3 a = q[0], b = q[1];
    q[0] = q[3]; q[1] = q[2]; q[2] = b; q[3] = a;
}
// get_ltc_matrix_and_brdf (lights.hlsl:203-255)
void ltc_matrix_brdf(const float* ltc, F3 lv, float rx, float ry, F3 q[4], M3* m, float* bx, float* by) {
    const float theta = acos32(lv.z);
    const bool flip_r = ry > rx;
    float phi = atan2_32(lv.y, lv.x);
    phi = flip_r ? kPiF * 0.5f - phi : phi;
    phi = phi >= 0.0f ? phi : phi + 2.0f * kPiF;
    const float u0 = fmax_((flip_r ? ry : rx) - 0.001f, 0.0f) / (1.0f - 0.001f);
    const float u1 = flip_r ? rx / ry : ry / rx;
    const float u2 = theta / (kPiF * 0.5f);
    if (phi < kPiF * 0.5f) {
        ltc_fetch(ltc, phi / (kPiF * 0.5f), u2, u1, u0, m, bx, by);
    } else if (phi < kPiF) {
        ltc_fetch(ltc, (kPiF - phi) / (kPiF * 0.5f), u2, u1, u0, m, bx, by);
        quad_reverse(q);
        m->r0 = -m->r0;
    } else if (phi < 1.5f * kPiF) {
        ltc_fetch(ltc, (phi - kPiF) / (kPiF * 0.5f), u2, u1, u0, m, bx, by);
        m->r0 = -m->r0;
        m->r1 = -m->r1;
    } else {
        ltc_fetch(ltc, (2.0f * kPiF - phi) / (kPiF * 0.5f), u2, u1, u0, m, bx, by);
        quad_reverse(q);
        m->r1 = -m->r1;
    }
    if (flip_r) {
        quad_reverse(q);
        const F3 t = m->r0;
        m->r0 = m->r1;
        m->r1 = t;
    }
}
// ltc_integrate_edge (lights.hlsl:365-381): theta / sin(theta) by its rational fit; the
// edge's term is cross(v1, v2) times it
inline float ltc_edge_scale(F3 v1, F3 v2) {
    const float x = dot(v1, v2);
    const float y = x < 0.0f ? -x : x;
    const float a = 5.42031f + (3.12829f + 0.0902326f * y) * y;
    const float b = 3.45068f + (4.18814f + y) * y;
    float t = a / b;
    if (x < 0.0f) t = kPiF / std::sqrt(1.0f - x * x) - t;
    return t;
}
// ltc_integrate (lights.hlsl:382-426): the quad relative to P in the shading frame,
// transformed by minv (none: the cosine itself), clipped to z > 0 (in.z * out - out.z *
// in for a vertex pair across the plane; a diagonal pair is the reference's empty case),
// projected on the sphere, edges summed in polygon order.  dir (nullable): the summed
// cross(v1, v2) theta / sin(theta) (sum.xyz of the reference), untouched when empty.
float ltc_integrate(F3 P, const Frame& fr, const M3* minv, const F3 q[4], bool two_sided, F3* dir = nullptr) {
    F3 l[4];
    uint32_t inside = 0;
    for (int k = 0; k < 4; ++k) {
        const F3 r = q[k] - P;
        F3 v = f3(dot(fr.t, r), dot(fr.b, r), dot(fr.n, r));
        if (minv) v = m3_mul(*minv, v);
        l[k] = v;
        inside |= (v.z > 0.0f ? 1u : 0u) << k;
    }
    if (inside == 0u || inside == 5u || inside == 10u) return 0.0f;
    F3 c[5];
    int n = 0;
    for (int k = 0; k < 4; ++k) {
        const F3 a = l[k], b = l[(k + 1) & 3];
        const bool ai = (inside >> k) & 1u, bi = (inside >> ((k + 1) & 3)) & 1u;
        if (ai) c[n++] = a;
        if (ai != bi) {
            const F3 in = ai ? a : b, out = ai ? b : a;
            c[n++] = f3(in.z * out.x - out.z * in.x, in.z * out.y - out.z * in.y, in.z * out.z - out.z * in.z);
        }
    }
    for (int k = 0; k < n; ++k) c[k] = normalize(c[k]);
    float sum = 0.0f, sx = 0.0f, sy = 0.0f;
    for (int k = 0; k < n; ++k) {
        const F3 a = c[k], b = c[k + 1 < n ? k + 1 : 0];
        const float t = ltc_edge_scale(a, b);
        sum = sum + (a.x * b.y - a.y * b.x) * t;
        sx = sx + (a.y * b.z - a.z * b.y) * t;
        sy = sy + (a.z * b.x - a.x * b.z) * t;
    }
    if (dir) *dir = f3(sx, sy, sum);
    float r = two_sided ? (sum < 0.0f ? -sum : sum) : fmax_(sum, 0.0f);
    if (!std::isfinite(r)) r = 0.0f;
    return r;
}
// log2 of x >= 1 by Moshier's Cephes log2f: x = m 2^e, m folded into [sqrt(1/2), sqrt(2)),
// log(1 + f) by its degree-8 minimax polynomial, times log2(e) split in two.
float log2_32(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    float e = (float)((int)((b >> 23) & 255u) - 126);
    const uint32_t mb = (b & 0x007fffffu) | 0x3f000000u;
    float m;
    std::memcpy(&m, &mb, 4);
    if (m < 0.70710678118654752440f) {
        e = e - 1.0f;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    const float z = m * m;
    static const float P[9] = {7.0376836292e-2f, -1.1514610310e-1f, 1.1676998740e-1f, -1.2420140846e-1f,
                               1.4249322787e-1f, -1.6668057665e-1f, 2.0000714765e-1f, -2.4999993993e-1f,
                               3.3333331174e-1f};
    float y = P[0];
    for (int k = 1; k < 9; ++k) y = y * m + P[k];
    y = (y * m) * z;
    y = y - 0.5f * z;
    const float log2ea = 0.44269504088896340736f;
    float r = y * log2ea;
    r = r + m * log2ea;
    r = r + y;
    r = r + m;
    return r + e;
}
// saturate with HLSL's NaN -> 0
inline float saturate0(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }
// SampleLevel by Vulkan's explicit-LOD rules: lambda = level + bias clamped to [lod_min,
// lod_max]; minification filter for lambda > 0; level clamped to [0, levels - 1]; nearest
// mip mode ceil(d + 1/2) - 1, linear the two levels around d blended by its fraction.
F4 rect_tex_sample(const Scene& s, const RectTex& r, float u, float v, float level) {
    float lam = level + r.lod_bias;
    lam = lam > r.lod_max ? r.lod_max : lam;
    lam = lam < r.lod_min ? r.lod_min : lam;
    const uint32_t filter = lam > 0.0f ? r.min_filter : r.mag_filter;
    const float q = (float)(r.levels.size() - 1);
    float d = lam > 0.0f ? lam : 0.0f;
    d = d < q ? d : q;
    if (r.mipmap_mode == 0u) return tex_sample_t(s, r.levels[(size_t)(std::ceil(d + 0.5f) - 1.0f)], filter, u, v);
    const float fl = std::floor(d);
    const size_t l = (size_t)fl;
    const float f = d - fl;
    const F4 a = tex_sample_t(s, r.levels[l], filter, u, v);
    if (f == 0.0f) return a;
    const F4 b = tex_sample_t(s, r.levels[l + 1], filter, u, v);
    auto lerp = [](float x, float y, float w) { return x + (y - x) * w; };
    return F4{lerp(a.x, b.x, f), lerp(a.y, b.y, f), lerp(a.z, b.z, f), lerp(a.w, b.w, f)};
}
// rect_light_sample_texture (lights.hlsl:429-445): the light plane's point along dir from
// P as (u, v) over position2 -> position3 and position2 -> position1, clamped; level
// log2(max(1, t roughness inv_texel_size)).  R: 28 floats (bi_pt_rect_light).
F3 rect_tex_lookup(const Scene& s, const float* R, F3 dir, F3 P, float roughness, float* uvl = nullptr) {
    int32_t ti;
    std::memcpy(&ti, &R[3], 4);
    const F3 n = f3(R[24], R[25], R[26]);
    const F3 p2 = f3(R[16], R[17], R[18]);
    const float st = dot(dir, n), step = st < 0.0f ? -st : st;
    if (step < 0.0001f) {
        if (uvl) uvl[0] = uvl[1] = uvl[2] = 0.0f;
        return f3(0.0f, 0.0f, 0.0f);
    }
    const float dd = dot(P - p2, n), dist = dd < 0.0f ? -dd : dd;
    const float t = dist / step;
    const F3 rp = (P + dir * t) - p2;
    const float u = saturate0(dot(rp, f3(R[20], R[21], R[22]) - p2) * R[11]);   // inv_width_sqr
    const float v = saturate0(dot(rp, f3(R[12], R[13], R[14]) - p2) * R[15]);   // inv_height_sqr
    const float texels = (t * roughness) * R[19];                               // inv_texel_size
    const float level = log2_32(texels > 1.0f ? texels : 1.0f);
    if (uvl) { uvl[0] = u; uvl[1] = v; uvl[2] = level; }
    const F4 c = rect_tex_sample(s, s.rect_textures[(size_t)ti], u, v, level);
    return f3(c.x, c.y, c.z);
}
// world direction of an integral's vector sv: normalize((m sv) in the frame (t, b, n))
F3 ltc_direction(const Frame& fr, const M3* m, F3 sv) {
    const F3 w = m ? m3_mul(*m, sv) : sv;
    return normalize((fr.t * w.x + fr.b * w.y) + fr.n * w.z);
}
// The lighting of rect light R (28 floats, bi_pt_rect_light) at P; terms (nullable):
// diffuse integral, specular integral, norm, Fresnel term.
F3 rect_light_lit(const float* R, const float* ltc, F3 P, const Frame& fr, F3 V, const Mat& m, float* terms,
                  const Scene* s = nullptr, float* dirs = nullptr) {
    const F3 lv = f3(dot(V, fr.t), dot(V, fr.b), dot(V, fr.n));
    if (terms) terms[0] = terms[1] = terms[2] = terms[3] = 0.0f;
    if (lv.z <= 0.0f) return f3(0, 0, 0);
    F3 q[4] = {f3(R[20], R[21], R[22]), f3(R[16], R[17], R[18]), f3(R[12], R[13], R[14]), f3(R[8], R[9], R[10])};
    uint32_t two;
    std::memcpy(&two, &R[7], 4);
    const F3 e = f3(R[0], R[1], R[2]);
    int32_t ti;
    std::memcpy(&ti, &R[3], 4);
    const bool textured = ti >= 0 && s;   // the scene's textures (untextured without a scene)
    F3 sd = f3(0, 0, 0), ss = f3(0, 0, 0);
    const float i_d = ltc_integrate(P, fr, nullptr, q, two != 0u, &sd);
    M3 mt;
    float bx, by;
    ltc_matrix_brdf(ltc, lv, m.rx, m.ry, q, &mt, &bx, &by);
    const M3 mi = m3_inverse(mt);
    const float i_s = ltc_integrate(P, fr, &mi, q, two != 0u, &ss);
    if (terms) { terms[0] = i_d; terms[1] = i_s; terms[2] = bx; terms[3] = by; }
    F3 ld = e * i_d, ls = e * i_s;
    if (dirs) for (int k = 0; k < 6; ++k) dirs[k] = 0.0f;
    if (textured) {
        // the light's texture along each integral's direction (lights.hlsl:495-511)
        if (i_d != 0.0f) {
            const F3 w = ltc_direction(fr, nullptr, sd);
            if (dirs) { dirs[0] = w.x; dirs[1] = w.y; dirs[2] = w.z; }
            ld = ld * rect_tex_lookup(*s, R, w, P, 1.0f);
        }
        if (i_s != 0.0f) {
            const F3 w = ltc_direction(fr, &mt, ss);
            if (dirs) { dirs[3] = w.x; dirs[4] = w.y; dirs[5] = w.z; }
            ls = ls * rect_tex_lookup(*s, R, w, P, std::sqrt(m.rx * m.ry));
        }
    }
    const float ndotv = dot(fr.n, V);
    if (ndotv <= 0.0f) return f3(0, 0, 0);
    const float x = 1.0f - ndotv;
    const float p5 = ((x * x) * (x * x)) * x;
    const F3 fres = m.f0 + (m.f90 - m.f0) * p5;
    const F3 one = f3(1.0f, 1.0f, 1.0f);
    const F3 diffuse = ((one - fres) * m.base) * kInvPi;
    const F3 specular = m.f0 * bx + m.f90 * by;
    return diffuse * ld + specular * ls;
}

// The surface a ray of direction d sees at hit h of triangle tri, point P
// (hits/rt_gbuffer_hit.hlsl:6-17): interpolated normal, the material (a checkerboard's
// cell by P), texture-driven fields, the normal map, and the two-sided flip.
void hit_surface(const Scene& s, uint32_t tri, const Hit& h, F3 P, F3 d, F3* N_out, Mat* m_out, float* rough_out,
                 uint32_t* mid_out) {
    const float* tn = &s.tri_n[9 * tri];
    F3 n0 = f3(tn[0], tn[1], tn[2]), n1 = f3(tn[3], tn[4], tn[5]), n2 = f3(tn[6], tn[7], tn[8]);
    F3 N = normalize((n0 + (n1 - n0) * h.u) + (n2 - n0) * h.v);
    const uint32_t mid = resolve_mat(s.materials, s.tri_mat[tri], P);
    Mat m = load_mat(&s.materials[kMatFloats * mid]);
    float rough = s.materials[kMatFloats * mid + 3];
    {
        // texture-driven fields (material_function samples at vertex.texcoord) and the
        // normal map in the interpolated tangent frame (hits/rt_gbuffer_hit.hlsl:8-11,
        // hit.hlsl:155-157)
        const uint32_t tb = mat_word(s, mid, 23), tf = mat_word(s, mid, 24), tn = mat_word(s, mid, 25),
                       tr = mat_word(s, mid, 26);
        if (tb | tf | tn | tr) {
            float su, sv;
            tri_texcoord(s, tri, h.u, h.v, &su, &sv);
            if (tb) { F4 c = tex_sample(s, tb - 1u, su, sv); m.base = f3(c.x, c.y, c.z); }
            if (tf) { F4 c = tex_sample(s, tf - 1u, su, sv); m.f0 = f3(c.x, c.y, c.z); }
            if (tr) {
                rough = comp4(tex_sample(s, tr - 1u, su, sv), mat_word(s, mid, 27));
                mat_alpha(m, rough, s.materials[kMatFloats * mid + 7]);
            }
            if (tn && s.need_tan) {
                F4 c = tex_sample(s, tn - 1u, su, sv);
                F3 nt = f3(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f);
                const float* g = &s.tri_tan[12 * (size_t)tri];
                F3 T = normalize(f3((g[0] + (g[4] - g[0]) * h.u) + (g[8] - g[0]) * h.v,
                                    (g[1] + (g[5] - g[1]) * h.u) + (g[9] - g[1]) * h.v,
                                    (g[2] + (g[6] - g[2]) * h.u) + (g[10] - g[2]) * h.v));
                float tw = (g[3] + (g[7] - g[3]) * h.u) + (g[11] - g[3]) * h.v;
                F3 B = normalize(cross(N, T)) * tw;
                N = normalize((T * nt.x + B * nt.y) + N * nt.z);
            }
        }
    }
    if (m.two_sided && dot(d, N) > 0.0f) N = -N;
    *N_out = N;
    *m_out = m;
    *rough_out = rough;
    *mid_out = mid;
}

// IEEE half bits -> float (exact).
float half_to_float(uint16_t h) {
    const uint32_t e = (h >> 10) & 31u, m = h & 1023u;
    float f;
    if (e == 0u) {
        f = (float)m * 0x1p-24f;
    } else {
        const uint32_t x = (e == 31u ? 0x7f800000u : (e + 112u) << 23) | (m << 13);
        std::memcpy(&f, &x, 4);
    }
    return (h & 0x8000u) ? -f : f;
}
// Drift mode (settings flag kOraGbufferQuantize): the surface as the reference's lighting
// reads it back from its GBuffer (pack_surface_to_gbuffer, then unpack_gbuffer_to_surface,
// shaders/renderer/gbuffer.hlsl:18-45; pack.hlsl:6-23,86-129): base colour and roughness
// through half floats, f0 and f90 through rg11b10 fixed point split into two unorm16
// words (read back as uint(w / 65535 * 65535.5)), the normal oct-encoded in halves and
// decoded, anisotropy through unorm8.  The kernels keep FP32 throughout.
void gbuffer_round_trip(F3* N, Mat* m, float* rough, float aniso) {
    auto qh = [](float x) { return half_to_float(half_rne(x)); };
    m->base = f3(qh(m->base.x), qh(m->base.y), qh(m->base.z));
    *rough = qh(*rough);
    auto rt = [](F3 c) {
        const uint32_t p = pack_rg11b10(c);
        const uint32_t w0 = unorm16_half_word(p & 0xffffu), w1 = unorm16_half_word(p >> 16);
        const uint32_t u0 = (uint32_t)(((float)w0 / 65535.0f) * 65535.5f);
        const uint32_t u1 = (uint32_t)(((float)w1 / 65535.0f) * 65535.5f);
        const uint32_t u = u0 | (u1 << 16);
        return f3((float)(u & 0x7ffu) / 2047.0f, (float)((u >> 11) & 0x7ffu) / 2047.0f,
                  (float)((u >> 22) & 0x3ffu) / 1023.0f);
    };
    m->f0 = rt(m->f0);
    m->f90 = rt(m->f90);
    // oct_encode (pack.hlsl:86-90) in halves, oct_decode (pack.hlsl:94-99)
    const float sa = (std::fabs(N->x) + std::fabs(N->y)) + std::fabs(N->z);
    float x = N->x / sa, y = N->y / sa;
    if (!(N->z / sa >= 0.0f)) {
        const float wx = (1.0f - std::fabs(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        const float wy = (1.0f - std::fabs(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = wx;
        y = wy;
    }
    x = qh(x);
    y = qh(y);
    F3 n = f3(x, y, (1.0f - std::fabs(x)) - std::fabs(y));
    const float t = std::min(std::max(-n.z, 0.0f), 1.0f);
    n.x = n.x + (n.x >= 0.0f ? -t : t);
    n.y = n.y + (n.y >= 0.0f ? -t : t);
    *N = normalize(n);
    mat_alpha(*m, *rough, (float)unorm8(aniso) / 255.0f);
}

PathOut trace_path(const Scene& s, const ora_camera& cam, const ora_settings& st,
                   uint32_t pixel, uint32_t sample, const CacheMode* cm = nullptr) {
    PathOut out{};
    const uint32_t M = st.max_bounces < 2 ? 2 : (st.max_bounces > 16 ? 16 : st.max_bounces);
    const uint32_t D = M - 1;
    const bool sky_nee = (st.flags & 2u) == 0 && s.sky_total > 0.0f;
    const uint32_t ndir = s.num_dir_lights < (uint32_t)kMaxDirLights ? s.num_dir_lights : kMaxDirLights;
    const uint32_t nl = ndir + s.num_point_lights < (uint32_t)kMaxDirLights ? ndir + s.num_point_lights : kMaxDirLights;
    uint32_t px = pixel % cam.width, py = pixel / cam.width;
    float jx = 0.5f, jy = 0.5f;
    if (st.flags & 1u) {
        Pcg jr = pcg_keyed(st.seed, pixel, sample, kJitterBounce);
        jx = pcg_float(jr);
        jy = pcg_float(jr);
    }
    Ray ray;
    camera_ray(cam, px, py, jx, jy, &ray.o, &ray.d);
    ray.tmin = kTMin;
    ray.tmax = st.ray_length;
    ray.seed = st.seed;
    F3 beta = f3(1.0f, 1.0f, 1.0f);
    float pdf_prev = 0.0f;
    F3 acc_miss = f3(0, 0, 0), acc_sky = f3(0, 0, 0), acc_rect = f3(0, 0, 0);
    F3 acc_dir[kMaxDirLights];
    for (int l = 0; l < kMaxDirLights; ++l) acc_dir[l] = f3(0, 0, 0);
    for (uint32_t b = 0; b < D; ++b) {
        out.segments = b + 1;
        Hit h = trace_closest(s, ray);
        if (h.leaf == kMiss) {
            F3 ls = sky_eval(s, ray.d);
            float w = 1.0f;
            if (b > 0 && sky_nee) w = power_heuristic(pdf_prev, sky_pdf(s, ray.d));
            acc_miss = (beta * ls) * w;
            break;
        }
        uint32_t tri = s.leaf_tri[h.leaf];
        F3 P = ray.o + ray.d * h.t;
        F3 N;
        Mat m;
        float rough;
        uint32_t mid;
        hit_surface(s, tri, h, P, ray.d, &N, &m, &rough, &mid);
        if (st.flags & kOraGbufferQuantize) gbuffer_round_trip(&N, &m, &rough, s.materials[kMatFloats * mid + 7]);
        F3 V = normalize(ray.o - P);
        if (cm && (int)b == cm->bounce) {
            if (!cm->train) {
                F3 c = f3(beta.x * fmax_(cm->rgb.x, 0.0f), beta.y * fmax_(cm->rgb.y, 0.0f),
                          beta.z * fmax_(cm->rgb.z, 0.0f));
                acc_miss = finite3(c) ? c : f3(0, 0, 0);
                break;
            }
            float* r = cm->rec;
            r[0] = P.x; r[1] = P.y; r[2] = P.z;
            r[3] = V.x; r[4] = V.y; r[5] = V.z;
            r[6] = N.x; r[7] = N.y; r[8] = N.z;
            r[9] = rough;
            r[10] = m.base.x; r[11] = m.base.y; r[12] = m.base.z;
            const uint32_t tag = 1u;
            std::memcpy(&r[13], &tag, 4);
            r[14] = r[15] = 0.0f;
            beta = f3(1.0f, 1.0f, 1.0f);
            acc_miss = acc_sky = acc_rect = f3(0, 0, 0);
            for (int l = 0; l < kMaxDirLights; ++l) acc_dir[l] = f3(0, 0, 0);
        }
        Frame fr = make_frame(N);
        const bool suffix = cm && cm->train && (int)b >= cm->bounce;
        Pcg rng = pcg_keyed(suffix ? st.seed ^ cm->suffix_key : st.seed, pixel, sample, b);
        float u0 = pcg_float(rng), u1 = pcg_float(rng), u2 = pcg_float(rng), u3 = pcg_float(rng);
        // directional lights (deferred_lighting_secondary.hlsl:45-60), then point lights
        // (:61-70), with shadow rays: unbounded towards a directional light, up to the
        // light's position for a point light
        for (uint32_t l = 0; l < nl; ++l) {
            F3 ldir;
            float tmax = kShadowTMax, strength;
            int sm_index;
            F3 le;
            if (l < ndir) {
                const float* L = &s.dir_lights[16 * l];
                le = f3(L[0], L[1], L[2]);
                ldir = f3(L[4], L[5], L[6]);
                std::memcpy(&sm_index, &L[3], sizeof sm_index);
                strength = L[7];
            } else {
                const float* L = &s.point_lights[16 * (l - ndir)];
                le = point_light_eval(L, P, &ldir, &tmax);
                std::memcpy(&sm_index, &L[12], sizeof sm_index);
                strength = L[13];
            }
            F3 f = bsdf_eval(fr, m, V, ldir);
            F3 c = (le * f) * beta;
            if (!any_pos(c)) continue;
            if (sm_index < 0) { acc_dir[l] = acc_dir[l] + c; continue; }
            out.shadow_bits[b] |= (uint8_t)(1u << l);
            out.n_light_shadow++;
            Ray sr{P, ldir, kTMin, tmax, st.seed};
            if (!trace_any(s, sr)) {
                acc_dir[l] = acc_dir[l] + c;
            } else {
                float keep = 1.0f - strength;
                if (keep > 0.0f) acc_dir[l] = acc_dir[l] + c * keep;
            }
        }
        // rect lights (deferred_lighting_secondary.hlsl:73-96): unshadowed, in light order
        for (uint32_t r = 0; r < s.num_rect_lights; ++r)
            acc_rect = acc_rect + rect_light_lit(&s.rect_lights[28 * r], s.ltc.data(), P, fr, V, m, nullptr, &s) * beta;
        // Sky next-event estimation with power-heuristic MIS against BSDF sampling, only
        // where the path also samples the next segment: the reference receives sky light
        // only through a traced segment that escapes (deferred_lighting_secondary.hlsl:24-29),
        // so the last vertex and terminated paths get none.
        const bool extend = b + 1 < D && !(max3(beta) < kThroughputMin);
        if (sky_nee && extend) {
            F3 wi, ls;
            float pl;
            if (sky_sample(s, u0, u1, &wi, &pl, &ls)) {
                F3 f = bsdf_eval(fr, m, V, wi);
                if (any_pos(f)) {
                    F3 hl = to_local(fr, normalize(V + wi));
                    float pb = vndf_pdf_l(hl, to_local(fr, V), m.rx, m.ry);
                    float w = power_heuristic(pl, pb);
                    F3 c = ((ls * f) * (w / pl)) * beta;
                    if (any_pos(c) && finite3(c)) {
                        out.shadow_bits[b] |= 0x80u;
                        out.n_sky_shadow++;
                        Ray sr{P, wi, kTMin, kShadowTMax, st.seed};
                        if (!trace_any(s, sr)) acc_sky = acc_sky + c;
                    }
                }
            }
        }
        // extension ray (sample_secondary_ray.hlsl:17-68)
        if (!extend) break;
        F3 vl = to_local(fr, V);
        F3 hh = vndf_sample(vl, m.rx, m.ry, u2, u3);
        float hv = dot(vl, hh);
        F3 ll = hh * (2.0f * hv) - vl;  // reflect(-V, h)
        float pdf = vndf_pdf_l(hh, vl, m.rx, m.ry);
        F3 L = to_world(fr, ll);
        F3 f = bsdf_eval(fr, m, V, L);
        F3 w = f3(f.x / pdf, f.y / pdf, f.z / pdf);
        if (!finite3(w)) w = f3(0, 0, 0);
        F3 nb = beta * w;
        if (!(max3(nb) > 0.0f)) break;
        beta = nb;
        pdf_prev = pdf;
        ray.o = P;
        ray.d = L;
    }
    F3 r = acc_miss;
    for (uint32_t l = 0; l < nl; ++l) r = r + acc_dir[l];
    r = r + acc_sky;
    if (s.num_rect_lights) r = r + acc_rect;
    out.radiance = r;
    return out;
}

}  // namespace

uint32_t pixel_of(uint32_t q, uint32_t W, uint32_t H) {
    if ((W & 7u) == 0 && (H & 7u) == 0) {
        uint32_t tile = q >> 6, within = q & 63u;
        uint32_t tiles_x = W >> 3;
        uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
        return ((ty << 3) + (within >> 3)) * W + (tx << 3) + (within & 7u);
    }
    return q;
}

// float -> IEEE half, round to nearest even (the conversion of an rgba16f store); NaN
// keeps a quiet payload, overflow goes to infinity, small values to half subnormals.
uint16_t half_rne(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   // |f| >= 65520 rounds past 65504
    const uint32_t e = ax >> 23;
    if (e >= 113u) {   // normal half: 2^-14 <= |f|
        uint32_t r = ax >> 13;                               // exponent and top 10 mantissa bits
        const uint32_t rem = ax & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
        return (uint16_t)(sign | (r - (112u << 10)));
    }
    // half subnormal: units of 2^-24; |f| = M * 2^(e - 150) with M the significand
    const uint32_t M = e ? ((ax & 0x7fffffu) | 0x800000u) : (ax & 0x7fffffu);
    const uint32_t sh = (e ? 126u - e : 125u);
    if (sh > 31u) return (uint16_t)sign;
    uint32_t r = M >> sh;
    const uint32_t rem = M & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (r & 1u))) ++r;
    return (uint16_t)(sign | r);
}
uint32_t cvt_u32_sat(float x) { return x > 0.0f ? (x >= 4294967296.0f ? 0xffffffffu : (uint32_t)x) : 0u; }
// pack_color_rg11b10 (pack.hlsl:14-16)
uint32_t pack_rg11b10(F3 c) {
    return cvt_u32_sat(c.x * 2047.0f) | (cvt_u32_sat(c.y * 2047.0f) << 11) | (cvt_u32_sat(c.z * 1023.0f) << 22);
}
// pack_u32_to_unorm16x2's halves h / 65536 (pack.hlsl:6-8) as an rgba16 unorm texel holds
// them: round(h * 65535 / 65536), exactly.
uint16_t unorm16_half_word(uint32_t h) { return (uint16_t)((h * 65535u + 32768u) >> 16); }
uint8_t unorm8(float v) {
    const float c = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;
    return (uint8_t)std::nearbyint(c * 255.0f);   // round to nearest even (default mode)
}
// pack_normal_and_tangent (pack.hlsl:89-110)
void pack_normal_tangent(F3 N, F3 T, float out[3]) {
    const float s = (std::fabs(N.x) + std::fabs(N.y)) + std::fabs(N.z);
    float x = N.x / s, y = N.y / s;
    const float z = N.z / s;
    if (!(z >= 0.0f)) {   // oct_wrap
        const float wx = (1.0f - std::fabs(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        const float wy = (1.0f - std::fabs(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = wx;
        y = wy;
    }
    const Frame f = make_frame(N);
    float px = dot(T, f.t), py = dot(T, f.b);
    const float l = std::fabs(px) + std::fabs(py);
    px = px / l;
    py = py / l;
    const float pxx = px * 0.5f + 0.5f;
    out[0] = x;
    out[1] = y;
    out[2] = py < 0.0f ? -pxx : pxx;
}

}  // namespace ora

// ================================================================== C API
using namespace ora;

extern "C" {

uint32_t ora_pixel_of(uint32_t q, uint32_t width, uint32_t height) { return pixel_of(q, width, height); }

namespace {
int render32_impl(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st,
                  uint32_t sample_base, uint32_t spp, float* sum_rgb, uint8_t* path_segments,
                  uint8_t* shadow_bits, uint64_t* ray_counts, int num_threads, const CacheMode* cm);
}

int ora_render32(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st,
                 uint32_t sample_base, uint32_t spp, float* sum_rgb, uint8_t* path_segments,
                 uint8_t* shadow_bits, uint64_t* ray_counts, int num_threads) {
    return render32_impl(ctx, cam, st, sample_base, spp, sum_rgb, path_segments, shadow_bits, ray_counts,
                         num_threads, nullptr);
}

int ora_render32_cached(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t cache_bounce,
                        const float* cache_rgb, uint32_t sample_base, uint32_t spp, float* sum_rgb, int num_threads) {
    CacheMode cm;
    cm.bounce = (int)cache_bounce;
    cm.rgb = f3(cache_rgb[0], cache_rgb[1], cache_rgb[2]);
    return render32_impl(ctx, cam, st, sample_base, spp, sum_rgb, nullptr, nullptr, nullptr, num_threads, &cm);
}

int ora_train_samples(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t cache_bounce,
                      uint32_t n, uint32_t suffixes, float* rec, float* target) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    if (cam->width == 0 || cam->height == 0) return -1;
    const uint64_t npix = (uint64_t)cam->width * cam->height;
    ora_settings js = *st;
    js.flags |= 1u;   // training paths are jittered
    for (uint32_t q = 0; q < n; ++q) {
        Pcg r = pcg_keyed(st->seed, q, 0u, 254u);
        const uint32_t pixel = (uint32_t)(((uint64_t)pcg_next(r) * npix) >> 32);
        CacheMode cm;
        cm.bounce = (int)cache_bounce;
        cm.train = true;
        cm.rec = &rec[16 * (size_t)q];
        std::memset(cm.rec, 0, 64);
        PathOut po = trace_path(s, *cam, js, pixel, q, &cm);
        F3 t = po.radiance;
        uint32_t tag = 0;
        std::memcpy(&tag, &cm.rec[13], 4);
        if (suffixes > 1 && tag == 1u) {
            // mean over `suffixes` continuations from the recorded vertex (run 0 above)
            double m[3] = {t.x, t.y, t.z};
            float keep[16];
            std::memcpy(keep, cm.rec, 64);
            for (uint32_t k = 1; k < suffixes; ++k) {
                cm.suffix_key = k * 0x9E3779B9u;
                PathOut pk = trace_path(s, *cam, js, pixel, q, &cm);
                m[0] += pk.radiance.x; m[1] += pk.radiance.y; m[2] += pk.radiance.z;
            }
            std::memcpy(cm.rec, keep, 64);
            t = f3((float)(m[0] / suffixes), (float)(m[1] / suffixes), (float)(m[2] / suffixes));
        }
        target[3 * (size_t)q] = t.x;
        target[3 * (size_t)q + 1] = t.y;
        target[3 * (size_t)q + 2] = t.z;
    }
    return 0;
}

namespace {
int render32_impl(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st,
                  uint32_t sample_base, uint32_t spp, float* sum_rgb, uint8_t* path_segments,
                  uint8_t* shadow_bits, uint64_t* ray_counts, int num_threads, const CacheMode* cm) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    const uint32_t W = cam->width, H = cam->height, npix = W * H;
    if (W == 0 || H == 0) return -1;
    if (num_threads <= 0) num_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint64_t> n_ext{0}, n_ls{0}, n_ss{0};
    std::atomic<uint32_t> next{0};
    const uint32_t chunk = 64;
    auto worker = [&]() {
        uint64_t e = 0, ls = 0, ss = 0;
        for (;;) {
            uint32_t q0 = next.fetch_add(chunk);
            if (q0 >= npix) break;
            uint32_t q1 = std::min(npix, q0 + chunk);
            for (uint32_t q = q0; q < q1; ++q) {
                uint32_t p = pixel_of(q, W, H);
                F3 sum = f3(sum_rgb[3 * p], sum_rgb[3 * p + 1], sum_rgb[3 * p + 2]);
                for (uint32_t sl = 0; sl < spp; ++sl) {
                    PathOut po = trace_path(s, *cam, *st, p, sample_base + sl, cm);
                    sum = sum + po.radiance;
                    size_t pi = (size_t)sl * npix + q;
                    if (path_segments) path_segments[pi] = (uint8_t)po.segments;
                    if (shadow_bits) std::memcpy(&shadow_bits[16 * pi], po.shadow_bits, 16);
                    e += po.segments - 1;
                    ls += po.n_light_shadow;
                    ss += po.n_sky_shadow;
                }
                sum_rgb[3 * p] = sum.x; sum_rgb[3 * p + 1] = sum.y; sum_rgb[3 * p + 2] = sum.z;
            }
        }
        n_ext += e; n_ls += ls; n_ss += ss;
    };
    std::vector<std::thread> th;
    for (int i = 1; i < num_threads; ++i) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    if (ray_counts) {
        ray_counts[0] = (uint64_t)npix * spp;
        ray_counts[1] = n_ext.load();
        ray_counts[2] = n_ls.load();
        ray_counts[3] = n_ss.load();
    }
    return 0;
}
}  // namespace

/* Primary-hit GBuffer of sample `sample` of the listed pixels (bi_pt_render_gbuffer),
 * indexed by list position: base_color and normal_roughness 4 half words each, fresnel 4
 * uint16, material_0 4 bytes (each nullable); zeros where the camera ray misses. */
int ora_gbuffer32(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t sample, uint32_t n,
                  const uint32_t* pixels, uint16_t* base, uint16_t* nr, uint16_t* fres, uint8_t* m0) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t pixel = pixels[i];
        float jx = 0.5f, jy = 0.5f;
        if (st->flags & 1u) {
            Pcg jr = pcg_keyed(st->seed, pixel, sample, kJitterBounce);
            jx = pcg_float(jr);
            jy = pcg_float(jr);
        }
        Ray ray;
        camera_ray(*cam, pixel % cam->width, pixel / cam->width, jx, jy, &ray.o, &ray.d);
        ray.tmin = kTMin;
        ray.tmax = st->ray_length;
        ray.seed = st->seed;
        uint16_t b4[4] = {0, 0, 0, 0}, n4[4] = {0, 0, 0, 0}, f4[4] = {0, 0, 0, 0};
        uint8_t m4[4] = {0, 0, 0, 0};
        const Hit h = trace_closest(s, ray);
        if (h.leaf != kMiss) {
            const uint32_t tri = s.leaf_tri[h.leaf];
            const F3 P = ray.o + ray.d * h.t;
            F3 N;
            Mat m;
            float rough;
            uint32_t mid;
            hit_surface(s, tri, h, P, ray.d, &N, &m, &rough, &mid);
            const float* g = &s.tri_tan[12 * (size_t)tri];
            const F3 T = normalize(f3((g[0] + (g[4] - g[0]) * h.u) + (g[8] - g[0]) * h.v,
                                      (g[1] + (g[5] - g[1]) * h.u) + (g[9] - g[1]) * h.v,
                                      (g[2] + (g[6] - g[2]) * h.u) + (g[10] - g[2]) * h.v));
            float pk[3];
            pack_normal_tangent(N, T, pk);
            b4[0] = half_rne(m.base.x); b4[1] = half_rne(m.base.y); b4[2] = half_rne(m.base.z);
            b4[3] = half_rne(1.0f);
            n4[0] = half_rne(pk[0]); n4[1] = half_rne(pk[1]); n4[2] = half_rne(pk[2]); n4[3] = half_rne(rough);
            const uint32_t p0 = pack_rg11b10(m.f0), p1 = pack_rg11b10(m.f90);
            f4[0] = unorm16_half_word(p0 & 0xffffu); f4[1] = unorm16_half_word(p0 >> 16);
            f4[2] = unorm16_half_word(p1 & 0xffffu); f4[3] = unorm16_half_word(p1 >> 16);
            const float* M = &s.materials[kMatFloats * mid];
            m4[0] = unorm8(M[7]);            // anisotropy
            m4[1] = unorm8(1.0f / M[11]);    // 1 / ior
            m4[3] = unorm8(1.0f / 256.0f);   // SurfaceModel::lit / 256
        }
        if (base) std::memcpy(&base[4 * (size_t)i], b4, 8);
        if (nr) std::memcpy(&nr[4 * (size_t)i], n4, 8);
        if (fres) std::memcpy(&fres[4 * (size_t)i], f4, 8);
        if (m0) std::memcpy(&m0[4 * (size_t)i], m4, 4);
    }
    return 0;
}

uint16_t ora_half32(float f) { return half_rne(f); }

int ora_render32_pixels(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t sample_base,
                        uint32_t spp, uint32_t n, const uint32_t* pixels, float* sum_rgb, int num_threads) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    const uint64_t npix = (uint64_t)cam->width * cam->height;
    if (npix == 0) return -1;
    for (uint32_t i = 0; i < n; ++i)
        if (pixels[i] >= npix) return -1;
    if (num_threads <= 0) num_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            const uint32_t p = pixels[i];
            F3 sum = f3(sum_rgb[3 * (size_t)i], sum_rgb[3 * (size_t)i + 1], sum_rgb[3 * (size_t)i + 2]);
            for (uint32_t sl = 0; sl < spp; ++sl) sum = sum + trace_path(s, *cam, *st, p, sample_base + sl).radiance;
            sum_rgb[3 * (size_t)i] = sum.x;
            sum_rgb[3 * (size_t)i + 1] = sum.y;
            sum_rgb[3 * (size_t)i + 2] = sum.z;
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < num_threads; ++k) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    return 0;
}

void ora_trace(const ora_ctx* ctx, uint32_t n, const float* rays, int any_hit, float* hit_t,
               uint32_t* hit_leaf, float* hit_u, float* hit_v) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = &rays[8 * i];
        Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        if (any_hit) {
            hit_leaf[i] = trace_any(s, ray) ? 1u : 0u;
        } else {
            Hit h = trace_closest(s, ray);
            hit_t[i] = h.t; hit_leaf[i] = h.leaf; hit_u[i] = h.u; hit_v[i] = h.v;
        }
    }
}

void ora_tex_sample32(const ora_ctx* ctx, uint32_t tex, uint32_t n, const float* uv, float* out4) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    for (uint32_t i = 0; i < n; ++i) {
        const F4 c = tex_sample(s, tex, uv[2 * i], uv[2 * i + 1]);
        out4[4 * i] = c.x; out4[4 * i + 1] = c.y; out4[4 * i + 2] = c.z; out4[4 * i + 3] = c.w;
    }
}

void ora_rect_light_scene32(const ora_ctx* ctx, uint32_t light, const float* P, const float* N, const float* V,
                            float rx, float ry, const float* base, const float* f0, const float* f90, float* rgb,
                            float* terms, float* dirs) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    Mat m{};
    m.base = f3(base[0], base[1], base[2]);
    m.f0 = f3(f0[0], f0[1], f0[2]);
    m.f90 = f3(f90[0], f90[1], f90[2]);
    m.rx = rx;
    m.ry = ry;
    const Frame fr = make_frame(f3(N[0], N[1], N[2]));
    const F3 c = rect_light_lit(&s.rect_lights[28 * (size_t)light], s.ltc.data(), f3(P[0], P[1], P[2]), fr,
                                f3(V[0], V[1], V[2]), m, terms, &s, dirs);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
void ora_rect_tex_lookup32(const ora_ctx* ctx, uint32_t light, const float* dir, const float* P, float r, float* rgb,
                           float* uvl) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    const F3 c = rect_tex_lookup(s, &s.rect_lights[28 * (size_t)light], f3(dir[0], dir[1], dir[2]),
                                 f3(P[0], P[1], P[2]), r, uvl);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
void ora_rect_tex_sample32(const ora_ctx* ctx, uint32_t tex, float u, float v, float level, float* out4) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    const F4 c = rect_tex_sample(s, s.rect_textures[tex], u, v, level);
    out4[0] = c.x; out4[1] = c.y; out4[2] = c.z; out4[3] = c.w;
}
void ora_log2_32(uint32_t n, const float* x, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = log2_32(x[i]);
}

void ora_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    Pcg r = pcg_seed(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_next(r);
}
void ora_pcg32_keyed(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t n,
                     float* out) {
    Pcg r = pcg_keyed(seed, pixel, sample, bounce);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_float(r);
}
uint32_t ora_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }

void ora_math32(int fn, uint32_t n, const float* x, const float* y, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = sin32(x[i]); break;
            case 1: out[i] = cos32(x[i]); break;
            case 2: out[i] = acos32(x[i]); break;
            default: out[i] = atan2_32(x[i], y[i]); break;
        }
    }
}

void ora_bsdf32(const float* material, const float* n, const float* v, const float* l, float* out3) {
    Mat m = load_mat(material);
    Frame fr = make_frame(f3(n[0], n[1], n[2]));
    F3 r = bsdf_eval(fr, m, f3(v[0], v[1], v[2]), f3(l[0], l[1], l[2]));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

void ora_vndf32(const float* v_local, float alpha, float u0, float u1, float* h_out, float* pdf_out) {
    F3 v = f3(v_local[0], v_local[1], v_local[2]);
    F3 h = vndf_sample(v, alpha, alpha, u0, u1);
    h_out[0] = h.x; h_out[1] = h.y; h_out[2] = h.z;
    *pdf_out = vndf_pdf_l(h, v, alpha, alpha);
}

void ora_sky_eval32(const ora_ctx* ctx, const float* dir, float* rgb) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    F3 c = sky_eval(s, f3(dir[0], dir[1], dir[2]));
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
void ora_sky_sample32(const ora_ctx* ctx, float u0, float u1, float* dir, float* pdf) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    F3 d = f3(0, 0, 0), le = f3(0, 0, 0);
    sky_sample(s, u0, u1, &d, pdf, &le);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}
void ora_sky_sample_le32(const ora_ctx* ctx, float u0, float u1, float* dir, float* pdf, float* le) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    F3 d = f3(0, 0, 0), l = f3(0, 0, 0);
    sky_sample(s, u0, u1, &d, pdf, &l);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    le[0] = l.x; le[1] = l.y; le[2] = l.z;
}
float ora_sky_pdf32(const ora_ctx* ctx, const float* dir) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    return sky_pdf(s, f3(dir[0], dir[1], dir[2]));
}
void ora_rect_light32(const float* rect, const float* ltc, const float* P, const float* N, const float* V, float rx,
                      float ry, const float* base, const float* f0, const float* f90, float* rgb, float* terms) {
    Mat m{};
    m.base = f3(base[0], base[1], base[2]);
    m.f0 = f3(f0[0], f0[1], f0[2]);
    m.f90 = f3(f90[0], f90[1], f90[2]);
    m.rx = rx;
    m.ry = ry;
    const Frame fr = make_frame(f3(N[0], N[1], N[2]));
    const F3 c = rect_light_lit(rect, ltc, f3(P[0], P[1], P[2]), fr, f3(V[0], V[1], V[2]), m, terms);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}

void ora_ltc_fetch32(const float* ltc, float u3, float u2, float u1, float u0, float* m, float* b) {
    M3 mt;
    ltc_fetch(ltc, u3, u2, u1, u0, &mt, &b[0], &b[1]);
    const F3 r[3] = {mt.r0, mt.r1, mt.r2};
    for (int k = 0; k < 3; ++k) { m[3 * k] = r[k].x; m[3 * k + 1] = r[k].y; m[3 * k + 2] = r[k].z; }
}

void ora_camera_ray32(const ora_camera* cam, uint32_t px, uint32_t py, float jx, float jy, float* o,
                      float* d) {
    F3 oo, dd;
    camera_ray(*cam, px, py, jx, jy, &oo, &dd);
    o[0] = oo.x; o[1] = oo.y; o[2] = oo.z;
    d[0] = dd.x; d[1] = dd.y; d[2] = dd.z;
}

}  // extern "C"

extern "C" {

void ora_sky_batch32(const ora_ctx* ctx, uint32_t count, const float* dirs, float* rgb, float* pdf) {
    for (uint32_t i = 0; i < count; ++i) {
        if (rgb) ora_sky_eval32(ctx, dirs + 3 * (size_t)i, rgb + 3 * (size_t)i);
        if (pdf) pdf[i] = ora_sky_pdf32(ctx, dirs + 3 * (size_t)i);
    }
}

}  // extern "C"
