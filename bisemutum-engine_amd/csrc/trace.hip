// trace.hip -- software BVH traversal on gfx950 (CDNA4 has no ray-tracing units).
//
// Replaces TraceRay of rt_gbuffer.hlsl:25 (closest hit, TMin 0.001, TMax
// ray_length) and the shadow-map lookups of lights.hlsl:69-101 (exact any-hit
// shadow rays).  One lane per ray, rays read from SoA queues (coalesced),
// nodes read whole (4 x 16 B).  Short stack: the top `L` entries live in LDS
// ([slot][lane], bank-conflict free for equal depths), deeper entries spill to
// a per-lane private array; the total depth bound (64) covers any Karras tree
// over fewer than 2^34 triangles.
//
// Closest hit uses the "while-while" loop of Aila & Laine 2009: a lane descends
// internal nodes until it reaches a leaf, then tests leaves until it pops an
// internal node.  Each lane's visiting order is a function of its own ray only
// (nearer child first, ties to child 0; stack LIFO), the same order the CPU
// restatement uses, so hits agree bit for bit.
#include "pt_kernels.hpp"

namespace bipt {
namespace {

constexpr int kDone = (int)0x80000000;

__device__ __forceinline__ V3 safe_inv(V3 d) {
    auto one = [](float x) {
        float ax = x < 0.0f ? -x : x;
        if (ax < 1e-20f) x = x < 0.0f ? -1e-20f : 1e-20f;
        return 1.0f / x;
    };
    return v3(one(d.x), one(d.y), one(d.z));
}

struct NodeBoxes {
    float4 a, b, c;
    int2 ch;
};
__device__ __forceinline__ NodeBoxes load_node(const BvhNode* nodes, int i) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    NodeBoxes n;
    n.a = p[0];
    n.b = p[1];
    n.c = p[2];
    int4 d = reinterpret_cast<const int4*>(p)[3];
    n.ch = make_int2(d.x, d.y);
    return n;
}

// Slab test of both children.  No operand can be NaN ((b - o) is finite and
// inv is finite and nonzero), so fminf/fmaxf select exactly what the CPU's
// comparisons select, up to the sign of zero, which only ever reaches comparisons.
__device__ __forceinline__ void box2(const NodeBoxes& n, V3 o, V3 inv, float tmin, float tmax, bool* h0, float* tn0,
                                     bool* h1, float* tn1) {
    float x0 = (n.a.x - o.x) * inv.x, x1 = (n.a.w - o.x) * inv.x;
    float y0 = (n.a.y - o.y) * inv.y, y1 = (n.b.x - o.y) * inv.y;
    float z0 = (n.a.z - o.z) * inv.z, z1 = (n.b.y - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    *tn0 = tn;
    *h0 = tn <= tf * kRobust;
    x0 = (n.b.z - o.x) * inv.x; x1 = (n.c.y - o.x) * inv.x;
    y0 = (n.b.w - o.y) * inv.y; y1 = (n.c.z - o.y) * inv.y;
    z0 = (n.c.x - o.z) * inv.z; z1 = (n.c.w - o.z) * inv.z;
    tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    *tn1 = tn;
    *h1 = tn <= tf * kRobust;
}

// Moller-Trumbore with division-free barycentric acceptance; t = tn / det.
__device__ __forceinline__ bool tri_test(const LeafGeo* geo, uint32_t leaf, V3 o, V3 d, float* t_out, float* un_out,
                                         float* vn_out, float* det_out) {
    const float4* p = reinterpret_cast<const float4*>(geo + leaf);
    float4 g0 = p[0], g1 = p[1], g2 = p[2];
    V3 v0 = v3(g0.x, g0.y, g0.z);
    V3 e1 = v3(g0.w, g1.x, g1.y);
    V3 e2 = v3(g1.z, g1.w, g2.x);
    V3 pv = cross(d, e2);
    float det = dot(e1, pv);
    if (det == 0.0f) return false;
    V3 tv = o - v0;
    float un = dot(tv, pv);
    V3 qv = cross(tv, e1);
    float vn = dot(d, qv);
    if (det > 0.0f) {
        if (un < 0.0f || vn < 0.0f || un + vn > det) return false;
    } else {
        if (un > 0.0f || vn > 0.0f || un + vn < det) return false;
    }
    float tn = dot(e2, qv);
    *t_out = tn / det;
    *un_out = un;
    *vn_out = vn;
    *det_out = det;
    return true;
}

template <int L>
struct ShortStack {
    int spill[kStackMax - L];
    int sp = 0;
    __device__ __forceinline__ bool push(int (*lds)[256], int x) {
        if (sp >= kStackMax) return false;
        if (sp < L) lds[sp][threadIdx.x] = x;
        else spill[sp - L] = x;
        ++sp;
        return true;
    }
    __device__ __forceinline__ int pop(int (*lds)[256]) {
        --sp;
        return sp < L ? lds[sp][threadIdx.x] : spill[sp - L];
    }
};

template <bool PRIMARY, int L>
__global__ __launch_bounds__(256) void k_trace_closest(TraceArgs a) {
    __shared__ int lds[L > 0 ? L : 1][256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t count = PRIMARY ? a.wave.num_paths : *a.count_in;
    if (i == 0 && a.ticket_clear) *a.ticket_clear = 0u;
    if (a.desc_clear && i < (count + 255u) / 256u) a.desc_clear[i] = 0ull;
    if (i >= count) return;
    V3 o, d;
    if (PRIMARY) {
        const uint32_t q = i % a.wave.npix, sl = i / a.wave.npix;
        const uint32_t pixel = pixel_of(q, a.wave.cam.width, a.wave.cam.height);
        float jx = 0.5f, jy = 0.5f;
        if (a.wave.flags & 1u) {
            Pcg jr = pcg_keyed(a.wave.seed, pixel, a.wave.sample0 + sl, kJitterBounce);
            jx = pcg_float(jr);
            jy = pcg_float(jr);
        }
        camera_ray(a.wave.cam, pixel % a.wave.cam.width, pixel / a.wave.cam.width, jx, jy, &o, &d);
    } else {
        o = v3(a.q.ox[i], a.q.oy[i], a.q.oz[i]);
        d = v3(a.q.dx[i], a.q.dy[i], a.q.dz[i]);
    }
    const float tmin = kTMin;
    float tbest = a.wave.ray_length;
    uint32_t best = kLeafMiss;
    float bun = 0.0f, bvn = 0.0f, bdet = 1.0f;
    if (a.bvh.num_tris > 0) {
        const V3 inv = safe_inv(d);
        ShortStack<L> st;
        int node = 0;
        for (;;) {
            while (node >= 0) {
                NodeBoxes nb = load_node(a.bvh.nodes, node);
                bool h0, h1;
                float tn0, tn1;
                box2(nb, o, inv, tmin, tbest, &h0, &tn0, &h1, &tn1);
                if (h0 && h1) {
                    bool first0 = tn0 <= tn1;
                    int nearc = first0 ? nb.ch.x : nb.ch.y;
                    int farc = first0 ? nb.ch.y : nb.ch.x;
                    if (!st.push(lds, farc)) { atomicOr(a.error, 1u); node = kDone; break; }
                    node = nearc;
                } else if (h0) {
                    node = nb.ch.x;
                } else if (h1) {
                    node = nb.ch.y;
                } else {
                    node = st.sp == 0 ? kDone : st.pop(lds);
                    if (node == kDone) break;
                }
            }
            if (node == kDone) break;
            while (node < 0 && node != kDone) {
                const uint32_t leaf = (uint32_t)~node;
                float t, un, vn, det;
                if (tri_test(a.bvh.leaf_geo, leaf, o, d, &t, &un, &vn, &det) && t >= tmin &&
                    (t < tbest || (t == tbest && leaf < best))) {
                    tbest = t;
                    best = leaf;
                    bun = un;
                    bvn = vn;
                    bdet = det;
                }
                node = st.sp == 0 ? kDone : st.pop(lds);
            }
            if (node == kDone) break;
        }
    }
    a.hits.t[i] = tbest;
    a.hits.leaf[i] = best;
    a.hits.u[i] = best != kLeafMiss ? bun / bdet : 0.0f;
    a.hits.v[i] = best != kLeafMiss ? bvn / bdet : 0.0f;
}

// Any-hit: the answer does not depend on the visiting order (no pruning by a
// running distance), so children are taken in stored order without sorting.
template <int L>
__device__ __forceinline__ bool occluded(const BvhView& bvh, int (*lds)[256], V3 o, V3 d, float tmin, float tmax,
                                         uint32_t* error) {
    if (bvh.num_tris == 0) return false;
    const V3 inv = safe_inv(d);
    ShortStack<L> st;
    int node = 0;
    for (;;) {
        while (node >= 0) {
            NodeBoxes nb = load_node(bvh.nodes, node);
            bool h0, h1;
            float tn0, tn1;
            box2(nb, o, inv, tmin, tmax, &h0, &tn0, &h1, &tn1);
            if (h0 && h1) {
                if (!st.push(lds, nb.ch.y)) { atomicOr(error, 1u); return false; }
                node = nb.ch.x;
            } else if (h0) {
                node = nb.ch.x;
            } else if (h1) {
                node = nb.ch.y;
            } else {
                if (st.sp == 0) return false;
                node = st.pop(lds);
            }
        }
        while (node < 0) {
            float t, un, vn, det;
            if (tri_test(bvh.leaf_geo, (uint32_t)~node, o, d, &t, &un, &vn, &det) && t >= tmin && t <= tmax)
                return true;
            if (st.sp == 0) return false;
            node = st.pop(lds);
        }
    }
}

template <int L>
__global__ __launch_bounds__(256) void k_trace_shadow(ShadowArgs a) {
    __shared__ int lds[L > 0 ? L : 1][256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= *a.count_in) return;
    V3 o = v3(a.q.ox[i], a.q.oy[i], a.q.oz[i]);
    V3 d = v3(a.q.dx[i], a.q.dy[i], a.q.dz[i]);
    const uint32_t pk = a.q.pk[i];
    const uint32_t path = pk >> 3, kind = pk & 7u;
    V3 c = v3(a.q.cx[i], a.q.cy[i], a.q.cz[i]);
    bool occ = occluded<L>(a.bvh, lds, o, d, kTMin, kShadowTMax, a.error);
    if (occ) {
        if (kind == kKindSky) return;
        float keep = a.keep[kind];
        if (!(keep > 0.0f)) return;
        c = c * keep;
    }
    const uint32_t slot = kind == kKindSky ? a.num_lights + 1u : kind + 1u;
    const size_t cap = a.ps.cap;
    float* acc = a.ps.acc + (size_t)slot * 3u * cap;
    acc[path] = acc[path] + c.x;
    acc[cap + path] = acc[cap + path] + c.y;
    acc[2 * cap + path] = acc[2 * cap + path] + c.z;
}

template <int L>
__global__ __launch_bounds__(256) void k_debug_anyhit(BvhView bvh, RayQueue q, uint32_t n, float tmax,
                                                      uint32_t* out, uint32_t* error) {
    __shared__ int lds[L][256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    V3 o = v3(q.ox[i], q.oy[i], q.oz[i]);
    V3 d = v3(q.dx[i], q.dy[i], q.dz[i]);
    out[i] = occluded<L>(bvh, lds, o, d, kTMin, tmax, error) ? 1u : 0u;
}

template <bool P>
void dispatch_closest(const TraceArgs& a, dim3 g, uint32_t lds, hipStream_t s) {
    if (lds <= 8) hipLaunchKernelGGL((k_trace_closest<P, 8>), g, dim3(256), 0, s, a);
    else if (lds <= 16) hipLaunchKernelGGL((k_trace_closest<P, 16>), g, dim3(256), 0, s, a);
    else if (lds <= 24) hipLaunchKernelGGL((k_trace_closest<P, 24>), g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_trace_closest<P, 32>), g, dim3(256), 0, s, a);
}

}  // namespace

void launch_trace_closest(const TraceArgs& a, uint32_t capacity, bool primary, uint32_t lds_stack, hipStream_t s) {
    dim3 g((capacity + 255u) / 256u);
    if (g.x == 0) return;
    if (primary) dispatch_closest<true>(a, g, lds_stack, s);
    else dispatch_closest<false>(a, g, lds_stack, s);
}

void launch_debug_anyhit(const BvhView& bvh, const RayQueue& q, uint32_t n, float tmax, uint32_t* out,
                         uint32_t* error, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_debug_anyhit<16>, dim3((n + 255u) / 256u), dim3(256), 0, s, bvh, q, n, tmax, out, error);
}

void launch_trace_shadow(const ShadowArgs& a, uint32_t capacity, uint32_t lds, hipStream_t s) {
    dim3 g((capacity + 255u) / 256u);
    if (g.x == 0) return;
    if (lds <= 8) hipLaunchKernelGGL(k_trace_shadow<8>, g, dim3(256), 0, s, a);
    else if (lds <= 16) hipLaunchKernelGGL(k_trace_shadow<16>, g, dim3(256), 0, s, a);
    else if (lds <= 24) hipLaunchKernelGGL(k_trace_shadow<24>, g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_trace_shadow<32>, g, dim3(256), 0, s, a);
}

}  // namespace bipt
