// trace.hip -- software BVH traversal on gfx950 (CDNA4 has no ray-tracing units).
//
// Replaces TraceRay of rt_gbuffer.hlsl:25 (closest hit, TMin 0.001, TMax
// ray_length) and the shadow-map lookups of lights.hlsl:69-101 (exact any-hit
// shadow rays).  Rays come from SoA queues (coalesced); the traversal structure is
// the BVH4 collapsed from the LBVH, one 128-B node per step (7 x 16-B loads).  Short
// stack: the top `L` entries live in LDS ([slot][lane], bank conflict free for equal
// depths), deeper entries spill to a per-lane private array; the total depth bound
// kStackMax covers any tree the build can produce.
//
// Persistent lanes with dynamic ray fetch: a workgroup-resident grid of wavefronts
// keeps lanes busy -- lanes whose rays have finished take the next ray indices from
// per-XCD counters (one atomic per wavefront and refill, lanes ranked by a ballot), so
// secondary rays of very different lengths do not idle the SIMD.  Each step is
// node-centric: the four child boxes of one node are tested, hit leaves intersected,
// the nearest hit internal child entered and the other hit internal children pushed
// (ascending slot order).  Each lane's visiting order is a function of its own ray
// only, the order of the CPU restatement, so hits and error flags agree bit for bit.
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

struct Node4 {
    float4 lx, ly, lz, hx, hy, hz;   // child k's box in component k
    int4 ch;
};
// Node and leaf records are read through global-address-space pointers (the integer
// round trip keeps the compiler from folding the cast away): global loads, not flat ones,
// so the seven 16-B loads of a node issue back to back and the waits count only vector
// memory (flat loads, which might address LDS, made the compiler wait between them).
typedef float VF4 __attribute__((ext_vector_type(4)));
typedef float VF2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const VF4 GlobalF4;
typedef __attribute__((address_space(1))) const VF2 GlobalF2;
__device__ __forceinline__ float4 f4(VF4 v) { return make_float4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ Node4 load_node(const Bvh4Node* nodes, int i) {
    const GlobalF4* p = (const GlobalF4*)(uintptr_t)(nodes + i);
    Node4 n;
    n.lx = f4(p[0]);
    n.ly = f4(p[1]);
    n.lz = f4(p[2]);
    n.hx = f4(p[3]);
    n.hy = f4(p[4]);
    n.hz = f4(p[5]);
    const VF4 c = p[6];
    n.ch = make_int4(__float_as_int(c.x), __float_as_int(c.y), __float_as_int(c.z), __float_as_int(c.w));
    return n;
}

__device__ __forceinline__ float comp(float4 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
__device__ __forceinline__ int comp(int4 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// Slab test of child k (k a compile-time constant after unrolling).  No operand can
// be NaN ((b - o) is finite or +inf and inv is finite and nonzero), so fminf/fmaxf
// select exactly what the CPU's comparisons select, up to the sign of zero, which
// only ever reaches comparisons.
__device__ __forceinline__ bool box_k(const Node4& n, int k, V3 o, V3 inv, float tmin, float tmax, float* tn_out) {
    const float x0 = (comp(n.lx, k) - o.x) * inv.x, x1 = (comp(n.hx, k) - o.x) * inv.x;
    const float y0 = (comp(n.ly, k) - o.y) * inv.y, y1 = (comp(n.hy, k) - o.y) * inv.y;
    const float z0 = (comp(n.lz, k) - o.z) * inv.z, z1 = (comp(n.hz, k) - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    *tn_out = tn;
    return tn <= tf * kRobust;
}

// Leaf intersection record held in registers: v0, e1, e2 and the flag word.
struct LeafRegs {
    float4 g0, g1;
    float2 g2;
};
__device__ __forceinline__ LeafRegs load_leaf(const LeafGeo* geo, uint32_t leaf) {
    const GlobalF4* p = (const GlobalF4*)(uintptr_t)(geo + leaf);
    LeafRegs r;
    r.g0 = f4(p[0]);
    r.g1 = f4(p[1]);
    const VF2 g2 = ((const GlobalF2*)(p + 2))[0];
    r.g2 = make_float2(g2.x, g2.y);
    return r;
}
__device__ __forceinline__ bool leaf_non_opaque(const LeafRegs& r) {
    return (__float_as_uint(r.g2.y) & kLeafNonOpaque) != 0u;
}
// Any-hit test of a non-opaque leaf hit at barycentric numerators (un, vn) / det by the
// ray (o, d) of a job: whether its surface is present for that ray.
__device__ __forceinline__ bool leaf_present(const JobRec& J, uint32_t leaf, float un, float vn, float det, V3 o,
                                             V3 d) {
    const MaterialRec& M = J.materials[J.leaf_shade[leaf].material];
    return surface_present(M, J.tex, J.leaf_uv, leaf, un / det, vn / det, opacity_random(J.seed, o, d));
}

// Moller-Trumbore with division-free barycentric acceptance; t = tn / det.
__device__ __forceinline__ bool tri_test(const LeafRegs& g, V3 o, V3 d, float* t_out, float* un_out, float* vn_out,
                                         float* det_out) {
    V3 v0 = v3(g.g0.x, g.g0.y, g.g0.z);
    V3 e1 = v3(g.g0.w, g.g1.x, g.g1.y);
    V3 e2 = v3(g.g1.z, g.g1.w, g.g2.x);
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
template <int L, int B = 256>
struct ShortStack {
    int spill[kStackMax - L];
    __device__ __forceinline__ bool push(int (*lds)[B], int& sp, int x) {
        const int s = sp;
        if (s >= kStackMax) return false;
        if (s < L) lds[s][threadIdx.x] = x;
        if (__ballot(s >= L) != 0ull) {
            if (s >= L) spill[s - L] = x;
        }
        sp = s + 1;
        return true;
    }
    __device__ __forceinline__ int pop(int (*lds)[B], int& sp) {
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

// Ray fetch for the lanes that ask (the caller decides when a wavefront refills: once
// `refill` lanes are idle, or no lane is busy): one atomic per refill.  The queue is
// split into `parts` contiguous parts, each with its own counter on its own 128-B
// line; a wavefront starts on the part of its workgroup's XCD (workgroups go to the
// XCDs round-robin) and moves on to the following parts when its part runs dry, so
// no single counter sees every refill.  Returns the new ray index for a lane that
// asked and got one, kNoRay otherwise; sets *exhausted when every part is drained.
struct WorkCursor {
    uint32_t part;    // part the wavefront draws from
    uint32_t left;    // parts not yet found empty
    uint32_t parts;
    uint32_t chunk;   // rays per part (the last parts may hold fewer)
};
__device__ __forceinline__ WorkCursor work_cursor(uint32_t count, uint32_t parts) {
    return WorkCursor{blockIdx.x % parts, parts, parts, (count + parts - 1u) / parts};
}
__device__ __forceinline__ uint32_t fetch_ray(uint32_t* work, uint32_t count, bool want, bool* exhausted,
                                              WorkCursor& wc) {
    const uint64_t m = __ballot(want);
    if (m == 0ull) return kNoRay;
    if (wc.left == 0) {
        if (want) *exhausted = true;
        return kNoRay;
    }
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

// Deferred leaf tests.  A box step does not intersect the leaves it reaches: it appends
// (leaf, owner lane) to its wavefront's queue in LDS, and the queue is flushed -- 64
// tests at a time, one per lane, each lane testing a leaf of any owner with the
// owner's ray fetched by a cross-lane read -- once it holds a wavefront's worth of
// tests, before lanes take new rays, or when no lane has a node left to visit.  So the
// triangle tests run with full wavefronts instead of the few lanes that reach a leaf
// in a given step.  A closest-hit candidate is accepted only where its leaf's own box
// passes the slab test up to the candidate's distance t.  Boxes nest and the slab
// arithmetic rounds monotonically, so every ancestor box of an accepted candidate then
// passes for any running distance >= t: each accepted candidate is reached whatever the
// pruning schedule, and the closest hit -- the minimum of (t, leaf) over accepted
// candidates -- does not depend on when or in which order leaves are tested.  An
// occlusion is an OR over leaves tested against a fixed extent.  A deferred hit only
// prunes later box tests a little later.
constexpr uint32_t kLeafQueue = 64u + 4u * 64u;   // < 64 before a step, which adds <= 4 per lane

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Pointer held by another lane.
template <class T>
__device__ __forceinline__ T* shfl_ptr(T* p, int src) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src), hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// Appends the hit leaf children (mask lm over the slots of ch) of every lane, with their
// box entry distances when qtn is given; returns the new queue length (uniform).
__device__ __forceinline__ uint32_t append_leaves(uint32_t* qleaf, uint8_t* qown, float* qtn, uint32_t qn,
                                                  uint32_t lm, int4 ch, float4 tn, uint32_t lane) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool bit = (lm >> k) & 1u;
        const uint64_t m = __ballot(bit);
        if (m == 0ull) continue;
        if (bit) {
            const uint32_t idx = qn + lanes_below(m);
            qleaf[idx] = (uint32_t)~comp(ch, k);
            qown[idx] = (uint8_t)lane;
            if (qtn) qtn[idx] = comp(tn, k);
        }
        qn += (uint32_t)__popcll(m);
    }
    return qn;
}

// AH: the call's scenes hold non-opaque (alpha-tested or translucent) triangles, whose
// hits pass the any-hit test; without them the test and its registers are compiled out.
// ONE: every ray of the launch belongs to job 0, so the scene pointers are uniform
// (scalar registers) instead of per lane.
template <bool PRIMARY, bool AH, int L, int B, bool ONE>
__device__ __forceinline__ void trace_closest_body(const TraceArgs& a) {
    __shared__ int lds[L][B];
    __shared__ uint32_t s_qleaf[B / 64][kLeafQueue];
    __shared__ uint8_t s_qown[B / 64][kLeafQueue];
    __shared__ float s_qtn[B / 64][kLeafQueue];
    __shared__ unsigned long long s_key[B / 64][64];
    const uint32_t count = PRIMARY ? a.jobs.num_paths : *a.count_in;
    const float tmin = kTMin;
    // Each lane keeps its ray's scene (BVH nodes, leaf records, triangle count): a wave
    // may hold rays of several jobs.  A job without triangles points at a one-node tree
    // whose slots are all empty.
    // Each iteration: (1) lanes with a node to visit test its four child boxes, push the
    // hit internal children but the nearest, queue the hit leaves and move to the
    // nearest (a lane with nothing left is parked until its queued leaves are tested);
    // (2) the leaf queue is flushed when full, before a refill, or when no lane is busy,
    // and parked lanes with no queued leaves write their hit; (3) idle lanes take new
    // rays (starting at the root) once enough of them are idle.  The node a lane visits
    // is loaded where its boxes are tested: its 28 registers are then free during the
    // leaf flush, which keeps the kernel at 79 VGPRs (6 wavefronts per SIMD when the LDS
    // stack is 12 deep or less).
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t* qleaf = s_qleaf[w];
    uint8_t* qown = s_qown[w];
    float* qtn = s_qtn[w];
    unsigned long long* key = s_key[w];
    key[lane] = ~0ull;
    uint32_t qn = 0;
    uint32_t ray = kNoRay;
    bool parked = false, exhausted = false;
    WorkCursor wc = work_cursor(count, a.parts);
    V3 o = v3(0.f, 0.f, 0.f), d = v3(0.f, 0.f, 1.f), inv = v3(0.f, 0.f, 0.f);
    float tbest = 0.0f;
    uint32_t best = kLeafMiss;
    const Bvh4Node* nodes = a.jobs.rec[0].nodes;
    const LeafGeo* geo = a.jobs.rec[0].leaf_geo;
    const JobRec* jr = a.jobs.rec;   // the lane's job (materials for any-hit tests)
    ShortStack<L, B> st;
    int sp = 0;
    uint32_t steps_left = 0;   // node visits left before the ray is taken to loop
    int cur_i = 0;             // the node the lane visits next
    for (;;) {
        int next = kDone;
        uint32_t lm = 0;
        int4 ch = make_int4(0, 0, 0, 0);
        float4 tnv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ray != kNoRay && !parked) {
            const Node4 nb = load_node(nodes, cur_i);
            ch = nb.ch;
            float tn[4];
            bool h[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) h[k] = box_k(nb, k, o, inv, tmin, tbest, &tn[k]);
            tnv = make_float4(tn[0], tn[1], tn[2], tn[3]);
            int nearest = -1, nearc = kDone;
            float tnear = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int c = comp(ch, k);
                if (h[k] && c < 0) lm |= 1u << k;
                if (h[k] && c >= 0 && (nearest < 0 || tn[k] < tnear)) {
                    nearest = k;
                    nearc = c;
                    tnear = tn[k];
                }
            }
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int c = comp(ch, k);
                if (ok && k != nearest && h[k] && c >= 0) ok = st.push(lds, sp, c);
            }
            if (!ok) atomicOr(a.error, 1u);
            else if (nearest >= 0) next = nearc;
            else if (sp > 0) next = st.pop(lds, sp);
            // a ray enters each node at most once: more steps than nodes means corrupt
            // links, and the ray is ended instead of looping
            if (steps_left-- == 0u) {
                atomicOr(a.error, 8u);
                next = kDone;
            }
            parked = next == kDone;
        }
        qn = append_leaves(qleaf, qown, qtn, qn, lm, ch, tnv, lane);
        const bool busy = ray != kNoRay && !parked;
        const uint64_t wm = __ballot((ray == kNoRay || parked) && !exhausted);
        const bool any_busy = __ballot(busy) != 0ull;
        const bool refill_now = wm != 0ull && (!any_busy || (uint32_t)__popcll(wm) >= a.refill);
        if (qn != 0u && (qn >= 64u || refill_now || !any_busy)) {
            __builtin_amdgcn_wave_barrier();
            for (uint32_t b = 0; b < qn; b += 64u) {
                const uint32_t e = b + lane;
                const bool valid = e < qn;
                const uint32_t leaf = valid ? qleaf[e] : 0u;
                const float tnl = valid ? qtn[e] : 0.0f;
                const int owner = valid ? (int)qown[e] : (int)lane;
                const V3 oo = v3(__shfl(o.x, owner), __shfl(o.y, owner), __shfl(o.z, owner));
                const V3 dd = v3(__shfl(d.x, owner), __shfl(d.y, owner), __shfl(d.z, owner));
                const float tb = __shfl(tbest, owner);
                const LeafGeo* og = ONE ? geo : shfl_ptr(geo, owner);
                // a hit counts only where the leaf's own box passes the slab test up to the
                // hit distance (the acceptance rule that makes the closest hit independent
                // of the order leaves are found in), and, for a non-opaque leaf, where its
                // surface is present for the ray (a function of the ray and the leaf alone)
                float t = 0.0f, un = 0.0f, vn = 0.0f, det = 1.0f;
                const LeafRegs lr = valid ? load_leaf(og, leaf) : LeafRegs{};
                bool cand = valid && tri_test(lr, oo, dd, &t, &un, &vn, &det) && t >= tmin && t <= tb &&
                            tnl <= t * kRobust;
                if constexpr (AH) {
                    const bool nop = cand && leaf_non_opaque(lr);
                    if (__ballot(nop) != 0ull) {
                        const JobRec* oj = ONE ? jr : shfl_ptr(jr, owner);
                        if (nop) cand = leaf_present(*oj, leaf, un, vn, det, oo, dd);
                    }
                }
                if (cand) atomicMin(&key[owner], ((unsigned long long)__float_as_uint(t) << 32) | leaf);
            }
            __builtin_amdgcn_wave_barrier();
            qn = 0;
            const unsigned long long kb = key[lane];
            if (kb != ~0ull) {
                key[lane] = ~0ull;
                const float t = __uint_as_float((uint32_t)(kb >> 32));
                const uint32_t leaf = (uint32_t)kb;
                if (t < tbest || (t == tbest && leaf < best)) {
                    tbest = t;
                    best = leaf;
                }
            }
        }
        if (parked && qn == 0u) {
            // every leaf of the ray is tested: write the hit; u, v come from repeating
            // the winning test
            float u = 0.0f, v = 0.0f;
            if (best != kLeafMiss) {
                float t, un, vn, det;
                tri_test(geo, best, o, d, &t, &un, &vn, &det);
                u = un / det;
                v = vn / det;
            }
            a.hits.rec[ray] = HitRec{tbest, best, u, v};
            ray = kNoRay;
            parked = false;
        }
        uint32_t got = kNoRay;
        if (refill_now) got = fetch_ray(a.work, count, ray == kNoRay && !exhausted, &exhausted, wc);
        const bool fresh = got != kNoRay;
        V3 no = o, nd = d;
        float tmax = 0.0f;
        uint32_t setup_steps = 0;
        if (fresh) {
            // one job: its record sits at a uniform address (scalar loads); several: the
            // path's job is looked up per lane
            auto setup = [&](const JobRec& J, uint32_t local) {
                if constexpr (!ONE) {
                    nodes = J.nodes;
                    geo = J.leaf_geo;
                    jr = &J;
                }
                setup_steps = J.num_tris > 1u ? J.num_tris : 1u;
                tmax = J.ray_length;
                if (PRIMARY) {
                    const PathKey k = path_key(J, local);
                    float jx = 0.5f, jy = 0.5f;
                    if (J.flags & 1u) {
                        Pcg jr = pcg_keyed(J.seed, k.pixel, k.sample, kJitterBounce);
                        jx = pcg_float(jr);
                        jy = pcg_float(jr);
                    }
                    camera_ray(J.cam, k.pixel % J.cam.width, k.pixel / J.cam.width, jx, jy, &no, &nd);
                }
            };
            if (ONE || a.jobs.num_jobs == 1u) {
                setup(a.jobs.rec[0], got);
            } else {
                const uint32_t path = PRIMARY ? got : a.q.path[got];
                const uint32_t jid = job_of(a.jobs.path_base, a.jobs.num_jobs, path);
                setup(a.jobs.rec[jid], path - a.jobs.path_base[jid]);
            }
            if (!PRIMARY) {
                no = v3(a.q.ox[got], a.q.oy[got], a.q.oz[got]);
                nd = v3(a.q.dx[got], a.q.dy[got], a.q.dz[got]);
            }
            next = 0;
        }
        cur_i = next != kDone ? next : 0;
        if (fresh) {
            ray = got;
            o = no;
            d = nd;
            inv = safe_inv(nd);
            tbest = tmax;
            best = kLeafMiss;
            sp = 0;
            steps_left = setup_steps;
        }
        if (__ballot(ray != kNoRay) == 0ull && __ballot(!exhausted) == 0ull) break;
    }
}

template <bool PRIMARY, bool AH, int L, int B>
__global__ __launch_bounds__(B) void k_trace_closest(TraceArgs a) {
    trace_closest_body<PRIMARY, AH, L, B, false>(a);
}
// One job, opaque scene, 256-thread blocks and a 12-deep LDS stack (the defaults): the
// scene pointers sit in scalar registers, which frees six vector registers per lane.
template <bool PRIMARY>
__global__ __launch_bounds__(256) void k_trace_closest_one(TraceArgs a) {
    trace_closest_body<PRIMARY, false, 12, 256, true>(a);
}

// Any-hit traversal of one ray (inspection kernel).  The answer does not depend on the
// visiting order (no pruning by a running distance); the order is the shadow kernel's:
// the hit internal child entered farthest along the ray first, the others pushed in slot
// order.
template <int L>
__device__ __forceinline__ bool occluded(const BvhView& bvh, int (*lds)[256], V3 o, V3 d, float tmin, float tmax,
                                         uint32_t* error) {
    if (bvh.num_tris == 0) return false;
    const V3 inv = safe_inv(d);
    ShortStack<L> st;
    int sp = 0;
    int node = 0;
    for (uint32_t steps = 1;; ++steps) {
        if (steps > bvh.num_tris) {   // corrupt links: end the ray instead of looping
            atomicOr(error, 8u);
            return false;
        }
        const Node4 nb = load_node(bvh.nodes, node);
        int next = kDone;
        bool ok = true;
        bool h[4];
        int far = -1;   // the shadow kernel's order: the farthest-entered hit internal child first
        float tfar = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float tn;
            h[k] = box_k(nb, k, o, inv, tmin, tmax, &tn);
            if (h[k] && comp(nb.ch, k) >= 0 && (far < 0 || tn > tfar)) {
                far = k;
                tfar = tn;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!h[k]) continue;
            const int c = comp(nb.ch, k);
            if (c < 0) {
                float t, un, vn, det;
                const uint32_t leaf = (uint32_t)~c;
                const LeafRegs lr = load_leaf(bvh.leaf_geo, leaf);
                if (tri_test(lr, o, d, &t, &un, &vn, &det) && t >= tmin && t <= tmax &&
                    (!leaf_non_opaque(lr) ||
                     surface_present(bvh.materials[bvh.leaf_shade[leaf].material], bvh.tex, bvh.leaf_uv, leaf,
                                     un / det, vn / det, opacity_random(bvh.seed, o, d))))
                    return true;
            } else if (k == far) {
                next = c;
            } else if (ok) {
                ok = st.push(lds, sp, c);
            }
        }
        if (!ok) {
            atomicOr(error, 1u);
            return false;
        }
        if (next == kDone) {
            if (sp == 0) return false;
            next = st.pop(lds, sp);
        }
        node = next;
    }
}

// Adds shadow ray i's contribution unless it is occluded; returns whether it is a sky ray.
__device__ __forceinline__ bool apply_shadow(const ShadowArgs& a, uint32_t i, bool occ) {
    const uint32_t pk = a.q.pk[i];
    const uint32_t path = pk >> 3, kind = pk & 7u;
    float keep = 1.0f;
    uint32_t nl;
    if (a.jobs.num_jobs == 1u) {   // uniform record address: scalar loads
        nl = a.jobs.rec[0].num_lights;
        if (occ && kind != kKindSky) keep = a.jobs.rec[0].keep[kind];
    } else {
        const JobRec& J = a.jobs.rec[job_of(a.jobs.path_base, a.jobs.num_jobs, path)];
        nl = J.num_lights;
        if (occ && kind != kKindSky) keep = J.keep[kind];
    }
    V3 c = v3(a.q.cx[i], a.q.cy[i], a.q.cz[i]);
    if (occ) {
        if (kind == kKindSky) return true;
        if (!(keep > 0.0f)) return false;
        c = c * keep;
    }
    const uint32_t slot = kind == kKindSky ? nl + 1u : kind + 1u;
    const size_t cap = a.ps.cap;
    float* acc = a.ps.acc + (size_t)slot * 3u * cap;
    acc[path] = acc[path] + c.x;
    acc[cap + path] = acc[cap + path] + c.y;
    acc[2 * cap + path] = acc[2 * cap + path] + c.z;
    return kind == kKindSky;
}

// AH (the general variant): any-hit tests of non-opaque leaves and rays that end at a
// point light; launched when a scene of the call has either.
template <bool AH, int L, int B>
__global__ __launch_bounds__(B) void k_trace_shadow(ShadowArgs a) {
    __shared__ int lds[L][B];
    __shared__ uint32_t s_qleaf[B / 64][kLeafQueue];
    __shared__ uint8_t s_qown[B / 64][kLeafQueue];
    __shared__ uint32_t s_occ[B / 64][64];
    const uint32_t count = *a.count_in;
    const float tmin = kTMin;
    // a ray's extent: unbounded towards the sky and directional lights, the distance to a
    // point light (the general variant only; the other has no point lights)
    float tmax = kShadowTMax;
    // Same iteration as the closest-hit kernel: box steps queue their leaves, the queue
    // is flushed when full, before a refill or when no lane is busy, and parked lanes
    // finish once their leaves are tested.  The hit internal child whose box the ray enters
    // farthest away is entered (occluders are found sooner there than near the ray's own
    // surface: C1 shadow rays visit 8.9 nodes instead of 11.3 in slot order) and the others
    // pushed in slot order (the answer is order-independent).  A flush that finds a ray
    // occluded ends it at once, busy or parked.
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t* qleaf = s_qleaf[w];
    uint8_t* qown = s_qown[w];
    uint32_t* occf = s_occ[w];
    occf[lane] = 0u;
    uint32_t qn = 0;
    uint32_t ray = kNoRay;
    bool parked = false, exhausted = false;
    WorkCursor wc = work_cursor(count, a.parts);
    V3 o = v3(0.f, 0.f, 0.f), d = v3(0.f, 0.f, 1.f), inv = v3(0.f, 0.f, 0.f);
    const Bvh4Node* nodes = a.jobs.rec[0].nodes;   // the lane's ray's scene (see k_trace_closest)
    const LeafGeo* geo = a.jobs.rec[0].leaf_geo;
    const JobRec* jr = a.jobs.rec;
    ShortStack<L, B> st;
    int sp = 0;
    uint32_t steps_left = 0;   // node visits left before the ray is taken to loop
    uint32_t sky_rays = 0;   // ray statistics: sky shadow rays this lane finished
    int cur_i = 0;             // the node the lane visits next
    for (;;) {
        int next = kDone;
        uint32_t lm = 0;
        int4 ch = make_int4(0, 0, 0, 0);
        if (ray != kNoRay && !parked) {
            const Node4 nb = load_node(nodes, cur_i);
            ch = nb.ch;
            bool h[4];
            int far = -1;   // the hit internal child entered farthest along the ray
            float tfar = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float tn;
                h[k] = box_k(nb, k, o, inv, tmin, tmax, &tn);
                if (h[k] && comp(ch, k) >= 0 && (far < 0 || tn > tfar)) {
                    far = k;
                    tfar = tn;
                }
            }
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int c = comp(ch, k);
                if (!h[k]) continue;
                if (c < 0) lm |= 1u << k;
                else if (k == far) next = c;
                else if (ok) ok = st.push(lds, sp, c);
            }
            if (!ok) {
                atomicOr(a.error, 1u);
                next = kDone;
            } else if (next == kDone && sp > 0) {
                next = st.pop(lds, sp);
            }
            if (steps_left-- == 0u) {   // corrupt links: end the ray instead of looping
                atomicOr(a.error, 8u);
                next = kDone;
            }
            parked = next == kDone;
        }
        qn = append_leaves(qleaf, qown, nullptr, qn, lm, ch, make_float4(0.f, 0.f, 0.f, 0.f), lane);
        const bool busy = ray != kNoRay && !parked;
        const uint64_t wm = __ballot((ray == kNoRay || parked) && !exhausted);
        const bool any_busy = __ballot(busy) != 0ull;
        const bool refill_now = wm != 0ull && (!any_busy || (uint32_t)__popcll(wm) >= a.refill);
        bool occ = false;
        if (qn != 0u && (qn >= 64u || refill_now || !any_busy)) {
            __builtin_amdgcn_wave_barrier();
            for (uint32_t b = 0; b < qn; b += 64u) {
                const uint32_t e = b + lane;
                const bool valid = e < qn;
                const uint32_t leaf = valid ? qleaf[e] : 0u;
                const int owner = valid ? (int)qown[e] : (int)lane;
                const V3 oo = v3(__shfl(o.x, owner), __shfl(o.y, owner), __shfl(o.z, owner));
                const V3 dd = v3(__shfl(d.x, owner), __shfl(d.y, owner), __shfl(d.z, owner));
                const LeafGeo* og = shfl_ptr(geo, owner);
                float t = 0.0f, un = 0.0f, vn = 0.0f, det = 1.0f;
                const LeafRegs lr = valid ? load_leaf(og, leaf) : LeafRegs{};
                const float tm = AH ? __shfl(tmax, owner) : kShadowTMax;
                bool cand = valid && tri_test(lr, oo, dd, &t, &un, &vn, &det) && t >= tmin && t <= tm;
                if constexpr (AH) {
                    const bool nop = cand && leaf_non_opaque(lr);
                    if (__ballot(nop) != 0ull) {
                        const JobRec* oj = shfl_ptr(jr, owner);
                        if (nop) cand = leaf_present(*oj, leaf, un, vn, det, oo, dd);
                    }
                }
                if (cand) occf[owner] = 1u;
            }
            __builtin_amdgcn_wave_barrier();
            qn = 0;
            occ = occf[lane] != 0u;
            if (occ) occf[lane] = 0u;
        }
        if (ray != kNoRay && (occ || (parked && qn == 0u))) {
            sky_rays += apply_shadow(a, ray, occ) ? 1u : 0u;
            ray = kNoRay;
            parked = false;
        }
        uint32_t got = kNoRay;
        if (refill_now) got = fetch_ray(a.work, count, ray == kNoRay && !exhausted, &exhausted, wc);
        const bool fresh = got != kNoRay;
        V3 no = o, nd = d;
        uint32_t setup_steps = 0;
        if (fresh) {
            auto setup = [&](const JobRec& J) {
                nodes = J.nodes;
                geo = J.leaf_geo;
                jr = &J;
                setup_steps = J.num_tris > 1u ? J.num_tris : 1u;
            };
            const uint32_t pk = a.q.pk[got];
            const JobRec& J = a.jobs.num_jobs == 1u ? a.jobs.rec[0]
                                                    : a.jobs.rec[job_of(a.jobs.path_base, a.jobs.num_jobs, pk >> 3)];
            setup(J);
            no = v3(a.q.ox[got], a.q.oy[got], a.q.oz[got]);
            // a sky ray's direction is in the queue; a light's ray is rebuilt from the
            // light as the shading built it (without point lights, every light is
            // directional and every extent unbounded)
            const uint32_t kind = pk & 7u;
            if (kind == kKindSky) {
                nd = v3(a.q.dx[got], a.q.dy[got], a.q.dz[got]);
                if constexpr (AH) tmax = kShadowTMax;
            } else if constexpr (AH) {
                light_shadow_ray(J.lights[kind], no, &nd, &tmax);
            } else {
                const LightRec& Lk = J.lights[kind];
                nd = v3(Lk.direction[0], Lk.direction[1], Lk.direction[2]);
            }
            next = 0;
        }
        cur_i = next != kDone ? next : 0;
        if (fresh) {
            ray = got;
            o = no;
            d = nd;
            inv = safe_inv(nd);
            sp = 0;
            steps_left = setup_steps;
        }
        if (__ballot(ray != kNoRay) == 0ull && __ballot(!exhausted) == 0ull) break;
    }
    // the shade launch counted every shadow ray as a light shadow ray: move the sky
    // rays over, one atomic pair per workgroup
    for (int off = 32; off >= 1; off >>= 1) sky_rays += __shfl_xor(sky_rays, off);
    __shared__ uint32_t s_sky[B / 64];
    if (lane == 0) s_sky[w] = sky_rays;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int k = 0; k < B / 64; ++k) tot += s_sky[k];
        if (tot) {
            atomicAdd(reinterpret_cast<unsigned long long*>(a.ray_counters + 2), (unsigned long long)(0ull - tot));
            atomicAdd(reinterpret_cast<unsigned long long*>(a.ray_counters + 3), (unsigned long long)tot);
        }
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
uint32_t resident_groups(K kernel, int block) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1024 / block;
    return (uint32_t)(cus * per_cu);
}

template <bool P, bool AH, int L, int B>
void closest_lb(const TraceArgs& a, uint32_t capacity, hipStream_t s) {
    uint32_t groups = (capacity + B - 1u) / B;
    static uint32_t resident = 0;
    if (resident == 0) resident = resident_groups(k_trace_closest<P, AH, L, B>, B);
    if (groups > resident) groups = resident;
    if (groups == 0) groups = 1;
    hipLaunchKernelGGL((k_trace_closest<P, AH, L, B>), dim3(groups), dim3(B), 0, s, a);
}

template <bool P>
void closest_one(const TraceArgs& a, uint32_t capacity, hipStream_t s) {
    uint32_t groups = (capacity + 255u) / 256u;
    static uint32_t resident = 0;
    if (resident == 0) resident = resident_groups(k_trace_closest_one<P>, 256);
    if (groups > resident) groups = resident;
    if (groups == 0) groups = 1;
    hipLaunchKernelGGL((k_trace_closest_one<P>), dim3(groups), dim3(256), 0, s, a);
}

template <bool AH, int L, int B>
void shadow_lb(const ShadowArgs& a, uint32_t capacity, hipStream_t s) {
    uint32_t groups = (capacity + B - 1u) / B;
    static uint32_t resident = 0;
    if (resident == 0) resident = resident_groups(k_trace_shadow<AH, L, B>, B);
    if (groups > resident) groups = resident;
    if (groups == 0) groups = 1;
    hipLaunchKernelGGL((k_trace_shadow<AH, L, B>), dim3(groups), dim3(B), 0, s, a);
}

// Block shapes: 128, 256 (default) or 512 threads; LDS stack depths 8, 12, 16 (default), 24, 32.
template <bool P, bool AH, int B>
void closest_b(const TraceArgs& a, uint32_t capacity, uint32_t lds, hipStream_t s) {
    if (lds <= 8) closest_lb<P, AH, 8, B>(a, capacity, s);
    else if (lds <= 12) closest_lb<P, AH, 12, B>(a, capacity, s);
    else if (lds <= 16) closest_lb<P, AH, 16, B>(a, capacity, s);
    else if (lds <= 24) closest_lb<P, AH, 24, B>(a, capacity, s);
    else closest_lb<P, AH, 32, B>(a, capacity, s);
}

template <bool P, bool AH>
void dispatch_closest(const TraceArgs& a, uint32_t capacity, uint32_t lds, uint32_t block, hipStream_t s) {
    if (block == 128) closest_b<P, AH, 128>(a, capacity, lds, s);
    else if (block == 512) closest_b<P, AH, 512>(a, capacity, lds, s);
    else closest_b<P, AH, 256>(a, capacity, lds, s);
}

template <bool AH, int B>
void shadow_b(const ShadowArgs& a, uint32_t capacity, uint32_t lds, hipStream_t s) {
    if (lds <= 8) shadow_lb<AH, 8, B>(a, capacity, s);
    else if (lds <= 12) shadow_lb<AH, 12, B>(a, capacity, s);
    else if (lds <= 16) shadow_lb<AH, 16, B>(a, capacity, s);
    else if (lds <= 24) shadow_lb<AH, 24, B>(a, capacity, s);
    else shadow_lb<AH, 32, B>(a, capacity, s);
}

template <bool AH>
void dispatch_shadow(const ShadowArgs& a, uint32_t capacity, uint32_t lds, uint32_t block, hipStream_t s) {
    if (block == 128) shadow_b<AH, 128>(a, capacity, lds, s);
    else if (block == 512) shadow_b<AH, 512>(a, capacity, lds, s);
    else shadow_b<AH, 256>(a, capacity, lds, s);
}

}  // namespace

void launch_trace_closest(const TraceArgs& a, uint32_t capacity, bool primary, uint32_t lds_stack, uint32_t block,
                          hipStream_t s) {
    if (capacity == 0) return;
    if (a.jobs.num_jobs == 1u && !a.any_hit && block == 256 && lds_stack > 8 && lds_stack <= 12) {
        if (primary) closest_one<true>(a, capacity, s);
        else closest_one<false>(a, capacity, s);
        return;
    }
    if (primary) {
        if (a.any_hit) dispatch_closest<true, true>(a, capacity, lds_stack, block, s);
        else dispatch_closest<true, false>(a, capacity, lds_stack, block, s);
    } else {
        if (a.any_hit) dispatch_closest<false, true>(a, capacity, lds_stack, block, s);
        else dispatch_closest<false, false>(a, capacity, lds_stack, block, s);
    }
}

void launch_debug_anyhit(const BvhView& bvh, const RayQueue& q, uint32_t n, float tmax, uint32_t* out,
                         uint32_t* error, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_debug_anyhit<16>, dim3((n + 255u) / 256u), dim3(256), 0, s, bvh, q, n, tmax, out, error);
}

void launch_trace_shadow(const ShadowArgs& a, uint32_t capacity, uint32_t lds, uint32_t block, hipStream_t s) {
    if (capacity == 0) return;
    if (a.any_hit) dispatch_shadow<true>(a, capacity, lds, block, s);
    else dispatch_shadow<false>(a, capacity, lds, block, s);
}

}  // namespace bipt
