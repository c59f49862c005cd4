// ref64.cpp -- FP64 oracle of the path-tracing integrand.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Same light-transport model as ref32.cpp and the HIP kernels (lit BSDF of
// shaders/core/material/lit.hlsl:5-35, GGX VNDF bounce sampling of
// sample_secondary_ray.hlsl:51-68, directional lights with exact visibility,
// equirect sky on misses plus sky next-event estimation with power-heuristic MIS,
// termination when max(throughput) < 0.001), evaluated independently: double
// precision everywhere, libm transcendentals, and its own top-down median-split BVH
// over the same world-space triangles.  Used for convergence checks only.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"
#include "ref_scene.hpp"

namespace ora {
namespace {

struct D3 { double x, y, z; };
inline D3 d3(double x, double y, double z) { return D3{x, y, z}; }
inline D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline D3 operator*(D3 a, D3 b) { return d3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
inline D3 operator-(D3 a) { return d3(-a.x, -a.y, -a.z); }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline D3 normalize(D3 v) { return v * (1.0 / std::sqrt(dot(v, v))); }
inline double max3(D3 v) { return std::max(std::max(v.x, v.y), v.z); }
inline bool any_pos(D3 v) { return v.x > 0 || v.y > 0 || v.z > 0; }

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------- BVH (median)
struct BNode {
    double lo[3], hi[3];
    int32_t left, right;   // internal: children; leaf: left = first, right = -count
};
struct Bvh64 {
    std::vector<BNode> nodes;
    std::vector<uint32_t> tris;
    std::vector<double> v;  // 9 per triangle (input order)
};

void build64(const Scene& s, Bvh64& b) {
    const uint32_t n = s.num_tris;
    b.v.resize((size_t)9 * n);
    for (size_t i = 0; i < (size_t)9 * n; ++i) b.v[i] = s.tri_v[i];
    b.tris.resize(n);
    for (uint32_t i = 0; i < n; ++i) b.tris[i] = i;
    std::vector<double> cen((size_t)3 * n);
    for (uint32_t t = 0; t < n; ++t)
        for (int k = 0; k < 3; ++k)
            cen[3 * t + k] = (b.v[9 * t + k] + b.v[9 * t + 3 + k] + b.v[9 * t + 6 + k]) / 3.0;
    struct Item { int node; uint32_t first, count; };
    std::vector<Item> work;
    if (n == 0) return;
    b.nodes.push_back(BNode{});
    work.push_back({0, 0, n});
    while (!work.empty()) {
        Item it = work.back();
        work.pop_back();
        BNode nd{};
        double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (int k = 0; k < 3; ++k) { nd.lo[k] = DBL_MAX; nd.hi[k] = -DBL_MAX; }
        for (uint32_t i = it.first; i < it.first + it.count; ++i) {
            uint32_t t = b.tris[i];
            for (int k = 0; k < 3; ++k) {
                for (int c = 0; c < 3; ++c) {
                    nd.lo[k] = std::min(nd.lo[k], b.v[9 * t + 3 * c + k]);
                    nd.hi[k] = std::max(nd.hi[k], b.v[9 * t + 3 * c + k]);
                }
                clo[k] = std::min(clo[k], cen[3 * t + k]);
                chi[k] = std::max(chi[k], cen[3 * t + k]);
            }
        }
        if (it.count <= 4) {
            nd.left = (int32_t)it.first;
            nd.right = -(int32_t)it.count;
            b.nodes[it.node] = nd;
            continue;
        }
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
        uint32_t mid = it.first + it.count / 2;
        std::nth_element(b.tris.begin() + it.first, b.tris.begin() + mid, b.tris.begin() + it.first + it.count,
                         [&](uint32_t a, uint32_t c) {
                             double ca = cen[3 * a + axis], cc = cen[3 * c + axis];
                             return ca < cc || (ca == cc && a < c);
                         });
        int l = (int)b.nodes.size();
        b.nodes.push_back(BNode{});
        b.nodes.push_back(BNode{});
        nd.left = l;
        nd.right = l + 1;
        b.nodes[it.node] = nd;
        work.push_back({l, it.first, mid - it.first});
        work.push_back({l + 1, mid, it.first + it.count - mid});
    }
}

const Bvh64& bvh64(const Scene& s) {
    // built once per scene by the first caller; callers are single-threaded entry points
    if (!s.bvh64) {
        Bvh64* b = new Bvh64();
        build64(s, *b);
        s.bvh64 = b;
    }
    return *static_cast<const Bvh64*>(s.bvh64);
}

inline bool slab(const BNode& nd, D3 o, D3 inv, double tmin, double tmax) {
    double t0 = tmin, t1 = tmax;
    const double oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
    for (int k = 0; k < 3; ++k) {
        double a = (nd.lo[k] - oo[k]) * ii[k], b = (nd.hi[k] - oo[k]) * ii[k];
        if (a > b) std::swap(a, b);
        if (std::isnan(a)) a = -DBL_MAX;
        if (std::isnan(b)) b = DBL_MAX;
        t0 = std::max(t0, a);
        t1 = std::min(t1, b * (1.0 + 1e-12));
        if (t0 > t1) return false;
    }
    return true;
}

inline bool tri64(const double* v, D3 o, D3 d, double* t, double* u, double* w) {
    D3 v0 = d3(v[0], v[1], v[2]);
    D3 e1 = d3(v[3] - v[0], v[4] - v[1], v[5] - v[2]);
    D3 e2 = d3(v[6] - v[0], v[7] - v[1], v[8] - v[2]);
    D3 p = cross(d, e2);
    double det = dot(e1, p);
    if (det == 0.0) return false;
    double inv = 1.0 / det;
    D3 tv = o - v0;
    double uu = dot(tv, p) * inv;
    if (uu < 0.0 || uu > 1.0) return false;
    D3 q = cross(tv, e1);
    double vv = dot(d, q) * inv;
    if (vv < 0.0 || uu + vv > 1.0) return false;
    *t = dot(e2, q) * inv;
    *u = uu; *w = vv;
    return true;
}

struct Hit64 { double t; uint32_t tri; double u, v; };

Hit64 closest64(const Bvh64& b, D3 o, D3 d, double tmin, double tmax) {
    Hit64 h{tmax, 0xffffffffu, 0, 0};
    if (b.nodes.empty()) return h;
    D3 inv = d3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& nd = b.nodes[stack[--sp]];
        if (!slab(nd, o, inv, tmin, h.t)) continue;
        if (nd.right <= 0) {
            for (int i = 0; i < -nd.right; ++i) {
                uint32_t t = b.tris[nd.left + i];
                double tt, uu, vv;
                if (tri64(&b.v[9 * (size_t)t], o, d, &tt, &uu, &vv) && tt >= tmin &&
                    (tt < h.t || (tt == h.t && t < h.tri)))
                    h = Hit64{tt, t, uu, vv};
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return h;
}

bool any64(const Bvh64& b, D3 o, D3 d, double tmin, double tmax) {
    if (b.nodes.empty()) return false;
    D3 inv = d3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& nd = b.nodes[stack[--sp]];
        if (!slab(nd, o, inv, tmin, tmax)) continue;
        if (nd.right <= 0) {
            for (int i = 0; i < -nd.right; ++i) {
                double tt, uu, vv;
                if (tri64(&b.v[9 * (size_t)b.tris[nd.left + i]], o, d, &tt, &uu, &vv) && tt >= tmin && tt <= tmax)
                    return true;
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return false;
}

// ------------------------------------------------------------------ shading
// Checkerboard materials (word 13 = 1): even cells of floor(P.x) ^ floor(P.z) shade
// with the record in word 14, evaluated at the FP64 hit point.
uint32_t resolve_mat64(const std::vector<float>& mats, uint32_t id, D3 P) {
    uint32_t pattern, alt;
    std::memcpy(&pattern, &mats[kMatFloats * (size_t)id + 13], 4);
    std::memcpy(&alt, &mats[kMatFloats * (size_t)id + 14], 4);
    if (pattern != 1u) return id;
    auto cell = [](double x) {
        double f = std::floor(x);
        if (!(f >= -2147483648.0)) f = -2147483648.0;
        if (f > 2147483520.0) f = 2147483520.0;
        return (long long)f;
    };
    return ((cell(P.x) ^ cell(P.z)) & 1) == 0 ? alt : id;
}

struct Mat64 { D3 base, f0, f90; double rx, ry; bool two_sided; };
Mat64 load_mat64(const float* m) {
    Mat64 r;
    r.base = d3(m[0], m[1], m[2]);
    r.f0 = d3(m[4], m[5], m[6]);
    r.f90 = d3(m[8], m[9], m[10]);
    double aniso = std::sqrt(1.0 - (double)m[7] * 0.9);
    double r2 = (double)m[3] * (double)m[3];
    r.rx = std::max(r2 / aniso, 0.001);
    r.ry = std::max(r2 * aniso, 0.001);
    uint32_t ts;
    std::memcpy(&ts, &m[12], sizeof ts);
    r.two_sided = ts != 0u;
    return r;
}
struct Frame64 { D3 t, b, n; };
Frame64 frame64(D3 n) {
    double sign = n.z > 0 ? 1.0 : -1.0;
    double a = -1.0 / (sign + n.z);
    double b = n.x * n.y * a;
    return Frame64{d3(1.0 + sign * n.x * n.x * a, sign * b, -sign * n.x), d3(b, sign + n.y * n.y * a, -n.y), n};
}
inline D3 loc(const Frame64& f, D3 v) { return d3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
inline double sq(double x) { return x * x; }
inline double ndf64(D3 h, double rx, double ry) {
    double a = sq(h.x / rx) + sq(h.y / ry) + sq(h.z);
    return 1.0 / (kPi * rx * ry * a * a);
}
inline double g1_64(D3 v, double rx, double ry) {
    double a = (sq(rx * v.x) + sq(ry * v.y)) / std::max(v.z * v.z, 0.0001);
    return 2.0 / (1.0 + std::sqrt(1.0 + a));
}
D3 bsdf64(const Frame64& fr, const Mat64& m, D3 V, D3 L) {
    D3 H = normalize(V + L);
    D3 lh = loc(fr, H), lv = loc(fr, V), ll = loc(fr, L);
    if (lv.z <= 0 || ll.z <= 0) return d3(0, 0, 0);
    double c = std::max(dot(V, H), 0.0);
    double p5 = std::pow(1.0 - c, 5.0);
    D3 fres = m.f0 + (m.f90 - m.f0) * p5;
    D3 diffuse = (d3(1, 1, 1) - fres) * m.base * (ll.z / kPi);
    double a = ll.z * std::sqrt(sq(m.rx * lv.x) + sq(m.ry * lv.y) + sq(lv.z));
    double b = lv.z * std::sqrt(sq(m.rx * ll.x) + sq(m.ry * ll.y) + sq(ll.z));
    double vis = 0.5 / std::max(a + b, 0.0001);
    D3 spec = fres * (ndf64(lh, m.rx, m.ry) * vis * ll.z);
    return diffuse + spec;
}
D3 vndf_sample64(D3 v, double rx, double ry, double u0, double u1) {
    if (v.z < 0) v = -v;
    D3 vh = normalize(d3(rx * v.x, ry * v.y, v.z));
    double ls = vh.x * vh.x + vh.y * vh.y;
    D3 t1v = ls > 0 ? d3(-vh.y, vh.x, 0) * (1.0 / std::sqrt(ls)) : d3(1, 0, 0);
    D3 t2v = cross(vh, t1v);
    double r = std::sqrt(u0), phi = 2.0 * kPi * u1;
    double t1 = r * std::cos(phi), t2 = r * std::sin(phi);
    double s = 0.5 * (1.0 + vh.z);
    t2 = (1.0 - s) * std::sqrt(std::max(1.0 - t1 * t1, 0.0)) + s * t2;
    D3 nh = t1v * t1 + t2v * t2 + vh * std::sqrt(std::max(1.0 - t1 * t1 - t2 * t2, 0.0));
    return normalize(d3(rx * nh.x, ry * nh.y, std::max(nh.z, 0.0)));
}
double vndf_pdf_l64(D3 h, D3 v, double rx, double ry) {
    double hv = dot(h, v);
    double pdf_h = g1_64(v, rx, ry) * ndf64(h, rx, ry) * std::max(hv, 0.0) / std::max(v.z, 0.0001);
    return pdf_h / (4.0 * std::fabs(hv));
}

// -------------------------------------------------------------------- sky
struct Sky64 {
    std::vector<double> w, ccdf, rows, marg;
    double total = 0;
};
Sky64 build_sky64(const Scene& s) {
    Sky64 k;
    const uint32_t W = s.sky_w, H = s.sky_h;
    if (!W || !H) return k;
    k.w.resize((size_t)W * H);
    k.ccdf.resize((size_t)(W + 1) * H);
    k.rows.resize(H);
    k.marg.assign(H + 1, 0.0);
    for (uint32_t i = 0; i < H; ++i) {
        double st = std::sin((i + 0.5) * kPi / H), run = 0;
        for (uint32_t j = 0; j < W; ++j) {
            const float* p = &s.sky_rgba[4 * ((size_t)i * W + j)];
            double lum = 0.2126 * p[0] * s.sky_color[0] + 0.7152 * p[1] * s.sky_color[1] + 0.0722 * p[2] * s.sky_color[2];
            double w = std::max(lum * st, 0.0);
            k.w[(size_t)i * W + j] = w;
            k.ccdf[(size_t)i * (W + 1) + j] = run;
            run += w;
        }
        k.ccdf[(size_t)i * (W + 1) + W] = run;
        k.rows[i] = run;
        k.marg[i + 1] = k.marg[i] + run;
    }
    k.total = k.marg[H];
    return k;
}
inline D3 sky_local(const Scene& s, D3 d) {
    const float* R = s.sky_rot;
    return d3(R[0] * d.x + R[1] * d.y + R[2] * d.z, R[3] * d.x + R[4] * d.y + R[5] * d.z,
              R[6] * d.x + R[7] * d.y + R[8] * d.z);
}
void sky_uv64(D3 ds, double* u, double* v, double* th) {
    double y = std::min(std::max(ds.y / std::sqrt(dot(ds, ds)), -1.0), 1.0);
    *th = std::acos(y);
    double ph = ds.z == 0 ? 0.0 : std::atan2(ds.x, ds.z);
    if (ph < 0) ph += 2 * kPi;
    *u = ph / (2 * kPi);
    *v = *th / kPi;
}
D3 sky_eval64(const Scene& s, D3 d) {
    if (!s.sky_w) return d3(0, 0, 0);
    double u, v, th;
    sky_uv64(sky_local(s, d), &u, &v, &th);
    const int W = (int)s.sky_w, H = (int)s.sky_h;
    double x = u * W - 0.5, y = v * H - 0.5;
    double fx0 = std::floor(x), fy0 = std::floor(y), fx = x - fx0, fy = y - fy0;
    int ix0 = (int)fx0, iy0 = (int)fy0, ix1 = ix0 + 1, iy1 = iy0 + 1;
    ix0 = (ix0 % W + W) % W;
    ix1 = (ix1 % W + W) % W;
    iy0 = std::min(std::max(iy0, 0), H - 1);
    iy1 = std::min(std::max(iy1, 0), H - 1);
    auto tex = [&](int i, int j) {
        const float* p = &s.sky_rgba[4 * ((size_t)i * W + j)];
        return d3(p[0], p[1], p[2]);
    };
    D3 c = (tex(iy0, ix0) * (1 - fx) + tex(iy0, ix1) * fx) * (1 - fy) + (tex(iy1, ix0) * (1 - fx) + tex(iy1, ix1) * fx) * fy;
    return c * d3(s.sky_color[0], s.sky_color[1], s.sky_color[2]);
}
double sky_pdf64(const Scene& s, const Sky64& k, D3 d) {
    if (!(k.total > 0)) return 0;
    double u, v, th;
    sky_uv64(sky_local(s, d), &u, &v, &th);
    const int W = (int)s.sky_w, H = (int)s.sky_h;
    int j = std::min(std::max((int)(u * W), 0), W - 1), i = std::min(std::max((int)(v * H), 0), H - 1);
    double st = std::sin(th), w = k.w[(size_t)i * W + j];
    if (!(st > 0) || !(w > 0)) return 0;
    return w * W * H / k.total / (2 * kPi * kPi * st);
}
bool sky_sample64(const Scene& s, const Sky64& k, double u0, double u1, D3* dir, double* pdf) {
    *pdf = 0;
    if (!(k.total > 0)) return false;
    const int W = (int)s.sky_w, H = (int)s.sky_h;
    double t0 = u0 * k.total;
    int i = (int)(std::upper_bound(k.marg.begin() + 1, k.marg.end(), t0) - (k.marg.begin() + 1));
    i = std::min(i, H - 1);
    if (!(k.rows[i] > 0)) return false;
    double dv = std::min(std::max((t0 - k.marg[i]) / k.rows[i], 0.0), 1.0);
    const double* cc = &k.ccdf[(size_t)i * (W + 1)];
    double t1 = u1 * k.rows[i];
    int j = (int)(std::upper_bound(cc + 1, cc + W + 1, t1) - (cc + 1));
    j = std::min(j, W - 1);
    double w = k.w[(size_t)i * W + j];
    if (!(w > 0)) return false;
    double du = std::min(std::max((t1 - cc[j]) / w, 0.0), 1.0);
    double u = (j + du) / W, v = (i + dv) / H;
    double th = v * kPi, ph = u * 2 * kPi, st = std::sin(th);
    if (!(st > 0)) return false;
    D3 ds = d3(st * std::sin(ph), std::cos(th), st * std::cos(ph));
    const float* R = s.sky_rot;
    *dir = d3(R[0] * ds.x + R[3] * ds.y + R[6] * ds.z, R[1] * ds.x + R[4] * ds.y + R[7] * ds.z,
              R[2] * ds.x + R[5] * ds.y + R[8] * ds.z);
    *pdf = w * W * H / k.total / (2 * kPi * kPi * st);
    return true;
}

// --------------------------------------------------------------------- RNG
struct Pcg { uint64_t state, inc; };
inline uint32_t pcg_next(Pcg& r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
inline double pcg_double(Pcg& r) {
    uint64_t hi = pcg_next(r) >> 5, lo = pcg_next(r) >> 6;  // 53-bit uniform
    return (double)((hi << 26) | lo) * 0x1p-53;
}

// --------------------------------------------------------------- path loop
D3 path64(const Scene& s, const Bvh64& b, const Sky64& k, const ora_camera& cam, const ora_settings& st,
          uint32_t pixel, uint32_t sample, uint32_t seed64) {
    const uint32_t M = std::min(std::max(st.max_bounces, 2u), 16u), D = M - 1;
    const bool sky_nee = (st.flags & 2u) == 0 && k.total > 0;
    Pcg rng{0, (((uint64_t)seed64 << 1) | 1u)};
    pcg_next(rng);
    rng.state += ((uint64_t)pixel << 32) | sample;
    pcg_next(rng);
    double jx = 0.5, jy = 0.5;
    if (st.flags & 1u) { jx = pcg_double(rng); jy = pcg_double(rng); }
    // camera ray (generate_camera_ray.hlsl:9-15)
    double u = ((pixel % cam.width) + jx) / cam.width, v = ((pixel / cam.width) + jy) / cam.height;
    double nx = u * 2 - 1, ny = 1 - v * 2;
    const float* ip = cam.inv_proj;
    D3 q = d3(ip[0] * nx + ip[1] * ny + ip[2] + ip[3], ip[4] * nx + ip[5] * ny + ip[6] + ip[7],
              ip[8] * nx + ip[9] * ny + ip[10] + ip[11]);
    q = normalize(q);
    const float* iv = cam.inv_view;
    D3 dir = normalize(d3(iv[0] * q.x + iv[1] * q.y + iv[2] * q.z, iv[4] * q.x + iv[5] * q.y + iv[6] * q.z,
                          iv[8] * q.x + iv[9] * q.y + iv[10] * q.z));
    D3 o = d3(iv[3], iv[7], iv[11]);
    D3 beta = d3(1, 1, 1), L = d3(0, 0, 0);
    double pdf_prev = 0;
    for (uint32_t bo = 0; bo < D; ++bo) {
        Hit64 h = closest64(b, o, dir, 0.001, st.ray_length);
        if (h.tri == 0xffffffffu) {
            double w = 1;
            if (bo > 0 && sky_nee) {
                double pl = sky_pdf64(s, k, dir);
                w = pdf_prev * pdf_prev / (pdf_prev * pdf_prev + pl * pl);
            }
            L = L + beta * sky_eval64(s, dir) * w;
            break;
        }
        D3 P = o + dir * h.t;
        const float* tn = &s.tri_n[9 * (size_t)h.tri];
        D3 n0 = d3(tn[0], tn[1], tn[2]), n1 = d3(tn[3], tn[4], tn[5]), n2 = d3(tn[6], tn[7], tn[8]);
        D3 N = normalize(n0 + (n1 - n0) * h.u + (n2 - n0) * h.v);
        Mat64 m = load_mat64(&s.materials[kMatFloats * (size_t)resolve_mat64(s.materials, s.tri_mat[h.tri], P)]);
        if (m.two_sided && dot(dir, N) > 0) N = -N;
        D3 V = normalize(o - P);
        Frame64 fr = frame64(N);
        for (uint32_t l = 0; l < std::min(s.num_dir_lights, 4u); ++l) {
            const float* Ld = &s.dir_lights[16 * l];
            D3 ldir = d3(Ld[4], Ld[5], Ld[6]);
            D3 c = d3(Ld[0], Ld[1], Ld[2]) * bsdf64(fr, m, V, ldir) * beta;
            if (!any_pos(c)) continue;
            int32_t sm;
            std::memcpy(&sm, &Ld[3], 4);
            double vis = 1;
            if (sm >= 0 && any64(b, P, ldir, 0.001, DBL_MAX)) vis = 1.0 - Ld[7];
            L = L + c * vis;
        }
        double u0 = pcg_double(rng), u1 = pcg_double(rng), u2 = pcg_double(rng), u3 = pcg_double(rng);
        const bool extend = bo + 1 < D && !(max3(beta) < 0.001);
        if (sky_nee && extend) {
            D3 wi;
            double pl;
            if (sky_sample64(s, k, u0, u1, &wi, &pl)) {
                D3 f = bsdf64(fr, m, V, wi);
                if (any_pos(f) && !any64(b, P, wi, 0.001, DBL_MAX)) {
                    double pb = vndf_pdf_l64(loc(fr, normalize(V + wi)), loc(fr, V), m.rx, m.ry);
                    double w = pl * pl / (pl * pl + pb * pb);
                    L = L + sky_eval64(s, wi) * f * beta * (w / pl);
                }
            }
        }
        if (!extend) break;
        D3 vl = loc(fr, V);
        D3 hh = vndf_sample64(vl, m.rx, m.ry, u2, u3);
        D3 ll = hh * (2 * dot(vl, hh)) - vl;
        double pdf = vndf_pdf_l64(hh, vl, m.rx, m.ry);
        D3 Lw = fr.t * ll.x + fr.b * ll.y + fr.n * ll.z;
        D3 f = bsdf64(fr, m, V, Lw);
        D3 w = f * (1.0 / pdf);
        if (!std::isfinite(w.x) || !std::isfinite(w.y) || !std::isfinite(w.z)) w = d3(0, 0, 0);
        beta = beta * w;
        if (!(max3(beta) > 0)) break;
        pdf_prev = pdf;
        o = P;
        dir = Lw;
    }
    return L;
}

}  // namespace

void free_bvh64(const Scene& s) {
    delete static_cast<Bvh64*>(s.bvh64);
    s.bvh64 = nullptr;
}

}  // namespace ora

using namespace ora;

extern "C" {

void ora_trace_brute64(const ora_ctx* ctx, uint32_t n, const float* rays, double* hit_t, uint32_t* hit_tri) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = &rays[8 * i];
        D3 o = d3(r[0], r[1], r[2]), d = d3(r[4], r[5], r[6]);
        double best = r[7];
        uint32_t bt = 0xffffffffu;
        for (uint32_t t = 0; t < s.num_tris; ++t) {
            double v[9];
            for (int k = 0; k < 9; ++k) v[k] = s.tri_v[9 * (size_t)t + k];
            double tt, uu, vv;
            if (tri64(v, o, d, &tt, &uu, &vv) && tt >= r[3] && (tt < best || (tt == best && t < bt))) {
                best = tt;
                bt = t;
            }
        }
        hit_t[i] = best;
        hit_tri[i] = bt;
    }
}

int ora_render64(const ora_ctx* ctx, const ora_camera* cam, const ora_settings* st, uint32_t num_pixels,
                 const uint32_t* pixels, uint32_t spp, uint32_t seed64, double* mean_rgb, double* var_rgb,
                 int num_threads) {
    const Scene& s = *reinterpret_cast<const Scene*>(ctx);
    // the FP64 oracle covers opaque parameter materials, directional lights and the
    // equirect sky (textured and non-opaque materials, point and rect lights and the
    // cube-map sky are pinned through the FP32 restatement's closed-form tests)
    if (s.sky_layout != 0u || s.num_point_lights != 0u || s.num_rect_lights != 0u) return -2;
    for (size_t m = 0; m < s.materials.size() / kMatFloats; ++m) {
        uint32_t w[32];
        std::memcpy(w, &s.materials[kMatFloats * m], sizeof w);
        if (w[15] || w[17] || w[23] || w[24] || w[25] || w[26]) return -2;
    }
    const Bvh64& b = bvh64(s);
    Sky64 k = build_sky64(s);
    if (num_threads <= 0) num_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= num_pixels) break;
            double sum[3] = {0, 0, 0}, sum2[3] = {0, 0, 0};
            for (uint32_t sm = 0; sm < spp; ++sm) {
                D3 r = path64(s, b, k, *cam, *st, pixels[i], sm, seed64);
                const double rr[3] = {r.x, r.y, r.z};
                for (int c = 0; c < 3; ++c) { sum[c] += rr[c]; sum2[c] += rr[c] * rr[c]; }
            }
            for (int c = 0; c < 3; ++c) {
                double m = sum[c] / spp;
                mean_rgb[3 * i + c] = m;
                var_rgb[3 * i + c] = spp > 1 ? std::max(sum2[c] / spp - m * m, 0.0) * spp / (spp - 1) : 0.0;
            }
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < num_threads; ++i) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    return 0;
}

void ora_bsdf64(const float* material, const double* n, const double* v, const double* l, double* out3) {
    Mat64 m = load_mat64(material);
    Frame64 fr = frame64(d3(n[0], n[1], n[2]));
    D3 r = bsdf64(fr, m, d3(v[0], v[1], v[2]), d3(l[0], l[1], l[2]));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}

double ora_vndf_pdf64(const double* v_local, const double* l_local, double alpha) {
    D3 v = d3(v_local[0], v_local[1], v_local[2]), l = d3(l_local[0], l_local[1], l_local[2]);
    return vndf_pdf_l64(normalize(v + l), v, alpha, alpha);
}

}  // extern "C"

extern "C" {

void ora_bsdf64_batch(const float* material, const double* n, const double* v, uint32_t count, const double* l,
                      double* out3) {
    for (uint32_t i = 0; i < count; ++i) ora_bsdf64(material, n, v, l + 3 * (size_t)i, out3 + 3 * (size_t)i);
}

void ora_vndf_pdf64_batch(const double* v_local, uint32_t count, const double* l_local, double alpha, double* out) {
    for (uint32_t i = 0; i < count; ++i) out[i] = ora_vndf_pdf64(v_local, l_local + 3 * (size_t)i, alpha);
}

}  // extern "C"
