// trace.hip -- software BVH traversal on gfx950 (CDNA4 has no ray-tracing units).
//
// Replaces TraceRay of rt_gbuffer.hlsl:25 (closest hit, TMin 0.001, TMax
// ray_length) and the shadow-map lookups of lights.hlsl:69-101 (exact any-hit
// shadow rays).  Rays come from SoA queues (coalesced); nodes are read whole
// (4 x 16 B).  Short stack: the top `L` entries live in LDS ([slot][lane], bank
// conflict free for equal depths), deeper entries spill to a per-lane private
// array; the total depth bound (64) covers any Karras tree over < 2^34 triangles.
//
// Persistent lanes with dynamic ray fetch: a workgroup-resident grid of wavefronts
// keeps every lane busy -- a lane whose ray has finished takes the next ray index
// from a global counter (one atomic per wavefront and refill, lanes ranked by a
// ballot), so secondary rays of very different lengths do not idle the SIMD.
// Inside a ray, the "while-while" loop of Aila & Laine 2009 descends internal nodes
// until a leaf, then tests leaves until an internal node is popped.  Each lane's
// visiting order is a function of its own ray only (nearer child first, ties to
// child 0; stack LIFO) -- the order of the CPU restatement -- so hits agree bit for bit.
#include "pt_kernels.hpp"

namespace bipt {
namespace {

constexpr int kDone = (int)0x80000000;
constexpr uint32_t kNoRay = 0xffffffffu;

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
    int2 pad;
};
__device__ __forceinline__ NodeBoxes load_node(const BvhNode* nodes, int i) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    NodeBoxes n;
    n.a = p[0];
    n.b = p[1];
    n.c = p[2];
    const int4 d = reinterpret_cast<const int4*>(p)[3];
    n.ch = make_int2(d.x, d.y);
    n.pad = make_int2(d.z, d.w);
    return n;
}
// The last 16 B of a node are loaded whole (the compiler widens a partial load to
// it); keeping the two padding words formally live until the node is consumed stops
// the register allocator from reusing them while the load is in flight, which would
// force a wait for the load right after issuing it.
__device__ __forceinline__ void consume_pad(const NodeBoxes& n) { asm volatile("" ::"v"(n.pad.x), "v"(n.pad.y)); }

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

// Leaf intersection record held in registers: v0, e1, e2.
struct LeafRegs {
    float4 g0, g1;
    float g2;
};
__device__ __forceinline__ LeafRegs load_leaf(const LeafGeo* geo, uint32_t leaf) {
    const float4* p = reinterpret_cast<const float4*>(geo + leaf);
    LeafRegs r;
    r.g0 = p[0];
    r.g1 = p[1];
    r.g2 = reinterpret_cast<const float*>(p + 2)[0];
    return r;
}

// Moller-Trumbore with division-free barycentric acceptance; t = tn / det.
__device__ __forceinline__ bool tri_test(const LeafRegs& g, V3 o, V3 d, float* t_out, float* un_out, float* vn_out,
                                         float* det_out) {
    V3 v0 = v3(g.g0.x, g.g0.y, g.g0.z);
    V3 e1 = v3(g.g0.w, g.g1.x, g.g1.y);
    V3 e2 = v3(g.g1.z, g.g1.w, g.g2);
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
__device__ __forceinline__ bool tri_test(const LeafGeo* geo, uint32_t leaf, V3 o, V3 d, float* t_out, float* un_out,
                                         float* vn_out, float* det_out) {
    return tri_test(load_leaf(geo, leaf), o, d, t_out, un_out, vn_out, det_out);
}

// The stack depth lives in a register of the caller (`sp`), apart from the spill
// array: an array indexed at run time is placed in private memory, and a counter
// stored next to it would follow it there.  The spill array is touched only inside
// branches taken when some lane of the wavefront is that deep, and its loaded value is
// consumed inside the branch: the common path is then pure LDS, with no generic
// (flat) access and no wait on the vector-memory counter, which would also wait for
// the node load issued ahead of it.
template <int L>
struct ShortStack {
    int spill[kStackMax - L];
    __device__ __forceinline__ bool push(int (*lds)[256], int& sp, int x) {
        const int s = sp;
        if (s >= kStackMax) return false;
        if (s < L) lds[s][threadIdx.x] = x;
        if (__ballot(s >= L) != 0ull) {
            if (s >= L) spill[s - L] = x;
        }
        sp = s + 1;
        return true;
    }
    __device__ __forceinline__ int pop(int (*lds)[256], int& sp) {
        const int s = sp - 1;
        sp = s;
        int v = lds[s < L ? s : L - 1][threadIdx.x];
        if (__ballot(s >= L) != 0ull) {
            const int w = spill[s >= L ? s - L : 0];
            v = s >= L ? w : v;
        }
        return v;
    }
};

// Refill idle lanes of the wavefront, once `refill` lanes are idle (or no lane is
// busy): one atomic per refill.  The queue is split into `parts` contiguous
// parts, each with its own counter on its own 128-B line; a wavefront starts on the
// part of its workgroup's XCD (workgroups go to the XCDs round-robin) and moves on to
// the following parts when its part runs dry, so no single counter sees every
// refill.  Returns the new ray index for a lane that asked and got one, kNoRay
// otherwise; sets *exhausted when every part is drained.
struct WorkCursor {
    uint32_t part;    // part the wavefront draws from
    uint32_t left;    // parts not yet found empty
    uint32_t parts;
    uint32_t chunk;   // rays per part (the last parts may hold fewer)
    uint32_t refill;
};
__device__ __forceinline__ WorkCursor work_cursor(uint32_t count, uint32_t parts, uint32_t refill) {
    return WorkCursor{blockIdx.x % parts, parts, parts, (count + parts - 1u) / parts, refill};
}
__device__ __forceinline__ uint32_t fetch_ray(uint32_t* work, uint32_t count, bool want, bool* exhausted,
                                              bool any_busy, WorkCursor& wc) {
    const uint64_t m = __ballot(want);
    if (m == 0ull) return kNoRay;
    if (wc.left == 0) {
        if (want) *exhausted = true;
        return kNoRay;
    }
    if (any_busy && (uint32_t)__popcll(m) < wc.refill) return kNoRay;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__builtin_ctzll(m);
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    for (;;) {
        const uint32_t lo = min(wc.part * wc.chunk, count);
        const uint32_t hi = min(lo + wc.chunk, count);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(work + wc.part * kWorkStride, (uint32_t)__popcll(m));
        base = __shfl(base, (int)leader);
        if (base < hi - lo) {
            const uint32_t idx = base + rank;
            return (want && idx < hi - lo) ? lo + idx : kNoRay;
        }
        wc.part = wc.part + 1u == wc.parts ? 0u : wc.part + 1u;
        if (--wc.left == 0) {
            if (want) *exhausted = true;
            return kNoRay;
        }
    }
}

template <bool PRIMARY, int L>
__global__ __launch_bounds__(256) void k_trace_closest(TraceArgs a) {
    __shared__ int lds[L][256];
    const uint32_t count = PRIMARY ? a.wave.num_paths : *a.count_in;
    const float tmin = kTMin;
    if (a.bvh.num_tris == 0) {
        for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
            a.hits.t[i] = a.wave.ray_length;
            a.hits.leaf[i] = kLeafMiss;
            a.hits.u[i] = 0.0f;
            a.hits.v[i] = 0.0f;
        }
        return;
    }
    // Each iteration: (1) the lanes holding a ray test both child boxes of `cur` and
    // choose the next node, issuing the loads of the hit leaf children; (2) lanes whose
    // ray ends here and idle lanes ask for new rays (a new ray starts at the root) and
    // issue the loads of their ray data; (3) the next node is loaded, at this single
    // place, so the loop keeps it in the same registers; (4) the leaf tests of (1) and
    // the ray set-up of (2) run while that load is in flight.
    uint32_t ray = kNoRay;
    bool exhausted = false;
    WorkCursor wc = work_cursor(count, a.parts, a.refill);
    V3 o = v3(0.f, 0.f, 0.f), d = v3(0.f, 0.f, 1.f), inv = v3(0.f, 0.f, 0.f);
    float tbest = 0.0f, bun = 0.0f, bvn = 0.0f, bdet = 1.0f;
    uint32_t best = kLeafMiss;
    ShortStack<L> st;
    int sp = 0;
    NodeBoxes cur = load_node(a.bvh.nodes, 0);
    for (;;) {
        const bool active = ray != kNoRay;
        int next = kDone;
        int cn, cf;              // set, with the leaf records, where ln / lf are set
        bool ln = false, lf = false;
        LeafRegs gn, gf;
        if (active) {
            // node-centric step: both boxes against [tmin, tbest]; hit leaf children are
            // intersected below (nearer first), the nearer internal child is entered
            const NodeBoxes nb = cur;
            consume_pad(nb);
            bool h0, h1;
            float tn0, tn1;
            box2(nb, o, inv, tmin, tbest, &h0, &tn0, &h1, &tn1);
            const int first = (h0 && h1) ? (tn0 <= tn1 ? 0 : 1) : (h0 ? 0 : 1);
            cn = first == 0 ? nb.ch.x : nb.ch.y;
            cf = first == 0 ? nb.ch.y : nb.ch.x;
            const bool hn = first == 0 ? h0 : h1, hf = first == 0 ? h1 : h0;
            ln = hn && cn < 0;
            lf = hf && cf < 0;
            if (ln) gn = load_leaf(a.bvh.leaf_geo, (uint32_t)~cn);
            if (lf) gf = load_leaf(a.bvh.leaf_geo, (uint32_t)~cf);
            const bool i0 = h0 && nb.ch.x >= 0, i1 = h1 && nb.ch.y >= 0;
            if (i0 && i1) {
                if (!st.push(lds, sp, cf)) atomicOr(a.error, 1u);
                else next = cn;
            } else if (i0) {
                next = nb.ch.x;
            } else if (i1) {
                next = nb.ch.y;
            } else if (sp > 0) {
                next = st.pop(lds, sp);
            }
        }
        const bool done = active && next == kDone;
        const uint32_t got = fetch_ray(a.work, count, (!active || done) && !exhausted, &exhausted,
                                       __ballot(active && !done) != 0ull, wc);
        const bool fresh = got != kNoRay;
        V3 no = o, nd = d;
        if (fresh) {
            if (PRIMARY) {
                const uint32_t q = got % a.wave.npix, sl = got / a.wave.npix;
                const uint32_t pixel = pixel_of(q, a.wave.cam.width, a.wave.cam.height);
                float jx = 0.5f, jy = 0.5f;
                if (a.wave.flags & 1u) {
                    Pcg jr = pcg_keyed(a.wave.seed, pixel, a.wave.sample0 + sl, kJitterBounce);
                    jx = pcg_float(jr);
                    jy = pcg_float(jr);
                }
                camera_ray(a.wave.cam, pixel % a.wave.cam.width, pixel / a.wave.cam.width, jx, jy, &no, &nd);
            } else {
                no = v3(a.q.ox[got], a.q.oy[got], a.q.oz[got]);
                nd = v3(a.q.dx[got], a.q.dy[got], a.q.dz[got]);
            }
            next = 0;
        }
        cur = load_node(a.bvh.nodes, next != kDone ? next : 0);
        if (active) {
            // the running best is replaced only by a strictly nearer hit or, at equal
            // distance, by the lower leaf: the result does not depend on the order
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (!(k == 0 ? ln : lf)) continue;
                const uint32_t leaf = (uint32_t)~(k == 0 ? cn : cf);
                float t, un, vn, det;
                if (tri_test(k == 0 ? gn : gf, o, d, &t, &un, &vn, &det) && t >= tmin &&
                    (t < tbest || (t == tbest && leaf < best))) {
                    tbest = t;
                    best = leaf;
                    bun = un;
                    bvn = vn;
                    bdet = det;
                }
            }
        }
        if (done) {
            a.hits.t[ray] = tbest;
            a.hits.leaf[ray] = best;
            a.hits.u[ray] = best != kLeafMiss ? bun / bdet : 0.0f;
            a.hits.v[ray] = best != kLeafMiss ? bvn / bdet : 0.0f;
            ray = kNoRay;
        }
        if (fresh) {
            ray = got;
            o = no;
            d = nd;
            inv = safe_inv(nd);
            tbest = a.wave.ray_length;
            best = kLeafMiss;
            bun = bvn = 0.0f;
            bdet = 1.0f;
            sp = 0;
        }
        if (__ballot(ray != kNoRay) == 0ull && __ballot(!exhausted) == 0ull) break;
    }
}

// Any-hit traversal of one ray, resumable: returns once the ray is decided.
// The answer does not depend on the visiting order (no pruning by a running
// distance), so children are taken in stored order without sorting.
template <int L>
__device__ __forceinline__ bool occluded(const BvhView& bvh, int (*lds)[256], V3 o, V3 d, float tmin, float tmax,
                                         uint32_t* error) {
    if (bvh.num_tris == 0) return false;
    const V3 inv = safe_inv(d);
    ShortStack<L> st;
    int sp = 0;
    int node = 0;
    for (;;) {
        while (node >= 0) {
            NodeBoxes nb = load_node(bvh.nodes, node);
            bool h0, h1;
            float tn0, tn1;
            box2(nb, o, inv, tmin, tmax, &h0, &tn0, &h1, &tn1);
            if (h0 && h1) {
                if (!st.push(lds, sp, nb.ch.y)) { atomicOr(error, 1u); return false; }
                node = nb.ch.x;
            } else if (h0) {
                node = nb.ch.x;
            } else if (h1) {
                node = nb.ch.y;
            } else {
                if (sp == 0) return false;
                node = st.pop(lds, sp);
            }
        }
        while (node < 0) {
            float t, un, vn, det;
            if (tri_test(bvh.leaf_geo, (uint32_t)~node, o, d, &t, &un, &vn, &det) && t >= tmin && t <= tmax)
                return true;
            if (sp == 0) return false;
            node = st.pop(lds, sp);
        }
    }
}

__device__ __forceinline__ void apply_shadow(const ShadowArgs& a, uint32_t i, bool occ) {
    const uint32_t pk = a.q.pk[i];
    const uint32_t path = pk >> 3, kind = pk & 7u;
    V3 c = v3(a.q.cx[i], a.q.cy[i], a.q.cz[i]);
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
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_trace_shadow(ShadowArgs a) {
    __shared__ int lds[L][256];
    const uint32_t count = *a.count_in;
    const float tmin = kTMin, tmax = kShadowTMax;
    if (a.bvh.num_tris == 0) {
        for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) apply_shadow(a, i, false);
        return;
    }
    // Same iteration as the closest-hit kernel: box tests and the next node, refill,
    // the single node load, then the leaf tests and ray set-up while it is in flight.
    // A ray found occluded by its leaf tests ends after the load was issued for it;
    // its lane takes a new ray at the next iteration.
    uint32_t ray = kNoRay;
    bool exhausted = false;
    WorkCursor wc = work_cursor(count, a.parts, a.refill);
    V3 o = v3(0.f, 0.f, 0.f), d = v3(0.f, 0.f, 1.f), inv = v3(0.f, 0.f, 0.f);
    ShortStack<L> st;
    int sp = 0;
    NodeBoxes cur = load_node(a.bvh.nodes, 0);
    for (;;) {
        const bool active = ray != kNoRay;
        int next = kDone;
        int c0, c1;              // set, with the leaf records, where l0 / l1 are set
        bool l0 = false, l1 = false;
        LeafRegs g0, g1;
        if (active) {
            // node-centric any-hit step: the next internal node is chosen from the box
            // tests (child 0 first; the answer is order-independent)
            const NodeBoxes nb = cur;
            consume_pad(nb);
            bool h0, h1;
            float tn0, tn1;
            box2(nb, o, inv, tmin, tmax, &h0, &tn0, &h1, &tn1);
            c0 = nb.ch.x;
            c1 = nb.ch.y;
            l0 = h0 && c0 < 0;
            l1 = h1 && c1 < 0;
            if (l0) g0 = load_leaf(a.bvh.leaf_geo, (uint32_t)~c0);
            if (l1) g1 = load_leaf(a.bvh.leaf_geo, (uint32_t)~c1);
            const bool i0 = h0 && c0 >= 0, i1 = h1 && c1 >= 0;
            if (i0 && i1) {
                if (!st.push(lds, sp, c1)) atomicOr(a.error, 1u);
                else next = c0;
            } else if (i0) {
                next = c0;
            } else if (i1) {
                next = c1;
            } else if (sp > 0) {
                next = st.pop(lds, sp);
            }
        }
        const bool done = active && next == kDone;
        const uint32_t got = fetch_ray(a.work, count, (!active || done) && !exhausted, &exhausted,
                                       __ballot(active && !done) != 0ull, wc);
        const bool fresh = got != kNoRay;
        V3 no = o, nd = d;
        if (fresh) {
            no = v3(a.q.ox[got], a.q.oy[got], a.q.oz[got]);
            nd = v3(a.q.dx[got], a.q.dy[got], a.q.dz[got]);
            next = 0;
        }
        cur = load_node(a.bvh.nodes, next != kDone ? next : 0);
        bool occ = false;
        if (active) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                if (!(c == 0 ? l0 : l1) || occ) continue;
                float t, un, vn, det;
                if (tri_test(c == 0 ? g0 : g1, o, d, &t, &un, &vn, &det) && t >= tmin && t <= tmax) occ = true;
            }
        }
        if (done || occ) {
            apply_shadow(a, ray, occ);
            ray = kNoRay;
        }
        if (fresh) {
            ray = got;
            o = no;
            d = nd;
            inv = safe_inv(nd);
            sp = 0;
        }
        if (__ballot(ray != kNoRay) == 0ull && __ballot(!exhausted) == 0ull) break;
    }
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

// Workgroups that stay resident on the whole chip for a kernel (persistent grid).
template <class K>
uint32_t resident_groups(K kernel) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
    return (uint32_t)(cus * per_cu);
}

template <bool P, int L>
void closest_l(const TraceArgs& a, uint32_t capacity, hipStream_t s) {
    uint32_t groups = (capacity + 255u) / 256u;
    static uint32_t resident = 0;
    if (resident == 0) resident = resident_groups(k_trace_closest<P, L>);
    if (groups > resident) groups = resident;
    if (groups == 0) groups = 1;
    hipLaunchKernelGGL((k_trace_closest<P, L>), dim3(groups), dim3(256), 0, s, a);
}

template <int L>
void shadow_l(const ShadowArgs& a, uint32_t capacity, hipStream_t s) {
    uint32_t groups = (capacity + 255u) / 256u;
    static uint32_t resident = 0;
    if (resident == 0) resident = resident_groups(k_trace_shadow<L>);
    if (groups > resident) groups = resident;
    if (groups == 0) groups = 1;
    hipLaunchKernelGGL(k_trace_shadow<L>, dim3(groups), dim3(256), 0, s, a);
}

template <bool P>
void dispatch_closest(const TraceArgs& a, uint32_t capacity, uint32_t lds, hipStream_t s) {
    if (lds <= 8) closest_l<P, 8>(a, capacity, s);
    else if (lds <= 16) closest_l<P, 16>(a, capacity, s);
    else if (lds <= 24) closest_l<P, 24>(a, capacity, s);
    else closest_l<P, 32>(a, capacity, s);
}

}  // namespace

void launch_trace_closest(const TraceArgs& a, uint32_t capacity, bool primary, uint32_t lds_stack, hipStream_t s) {
    if (capacity == 0) return;
    if (primary) dispatch_closest<true>(a, capacity, lds_stack, s);
    else dispatch_closest<false>(a, capacity, lds_stack, s);
}

void launch_debug_anyhit(const BvhView& bvh, const RayQueue& q, uint32_t n, float tmax, uint32_t* out,
                         uint32_t* error, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_debug_anyhit<16>, dim3((n + 255u) / 256u), dim3(256), 0, s, bvh, q, n, tmax, out, error);
}

void launch_trace_shadow(const ShadowArgs& a, uint32_t capacity, uint32_t lds, hipStream_t s) {
    if (capacity == 0) return;
    if (lds <= 8) shadow_l<8>(a, capacity, s);
    else if (lds <= 16) shadow_l<16>(a, capacity, s);
    else if (lds <= 24) shadow_l<24>(a, capacity, s);
    else shadow_l<32>(a, capacity, s);
}

}  // namespace bipt
