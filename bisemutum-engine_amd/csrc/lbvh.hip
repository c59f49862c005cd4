// lbvh.hip -- per-frame LBVH build on gfx950.
//
// Replaces the driver-side BLAS/TLAS build the reference requests every frame
// (src/graphics/accel.cpp:11-159 -> vkCmdBuildAccelerationStructuresKHR,
// src/rhi/backend_vulkan/command.cpp:425,483).  The reference has no BVH code of
// its own; this is the build's design:
//   1. flatten    instances -> world-space triangles (+ cofactor normals), centroid bounds
//   2. morton     30-bit Morton code of each triangle's bounding-box centre
//   3. sort       stable LSD radix sort of (code, triangle) with 8-bit digits:
//                 LDS histogram -> device scan -> LDS-ranked scatter, the rank of
//                 equal digits inside a wavefront taken from 8 ballots
//   4. karras     binary radix tree over the sorted codes (Karras 2012), index
//                 tie-break for equal codes
//   5. leaves     leaf intersection/shading records in sorted order, then the
//                 bottom-up AABB refit (second arrival at each node continues)
//   6. treelets   rounds of treelet restructuring of the refit tree (surface-area
//                 cost; bi_pt_set_build_options, default 3 rounds)
//   7. collapse   BVH4 traversal nodes
// All steps are integer or exact min/max except the Morton quantisation and the
// surface-area costs, whose float operations are written in a fixed order.
#include "pt_kernels.hpp"

namespace bipt {
namespace {

constexpr uint32_t kBoxPending = 0x7fc00000u;   // quiet NaN: never a valid box coordinate

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__global__ void k_init_bounds(uint32_t* cb) {
    int t = threadIdx.x;
    if (t < 3) cb[t] = 0xffffffffu;       // min slots
    else if (t < 6) cb[t] = 0u;           // max slots
}

__global__ __launch_bounds__(256) void k_flatten(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                 const uint32_t* __restrict__ idx,
                                                 const InstanceRec* __restrict__ inst, uint32_t ninst,
                                                 uint32_t n, float* __restrict__ tri_v,
                                                 float* __restrict__ tri_n, uint32_t* __restrict__ tri_mat,
                                                 uint32_t* __restrict__ cb) {
    uint32_t t = blockIdx.x * 256u + threadIdx.x;
    float cmin[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    float cmax[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    if (t < n) {
        uint32_t lo = 0, hi = ninst - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (inst[mid].first_tri <= t) lo = mid; else hi = mid - 1;
        }
        const InstanceRec& I = inst[lo];
        const float* m = I.xform;
        float c00 = m[5] * m[10] - m[6] * m[9], c01 = m[6] * m[8] - m[4] * m[10], c02 = m[4] * m[9] - m[5] * m[8];
        float c10 = m[2] * m[9] - m[1] * m[10], c11 = m[0] * m[10] - m[2] * m[8], c12 = m[1] * m[8] - m[0] * m[9];
        float c20 = m[1] * m[6] - m[2] * m[5], c21 = m[2] * m[4] - m[0] * m[6], c22 = m[0] * m[5] - m[1] * m[4];
        float det = (m[0] * c00 + m[1] * c01) + m[2] * c02;
        float sgn = det < 0.0f ? -1.0f : 1.0f;
        uint32_t k = t - I.first_tri;
        float w[9];
        for (int c = 0; c < 3; ++c) {
            uint32_t vi = I.vtx_off + idx[I.idx_off + 3 * k + c];
            float x = pos[3 * (size_t)vi], y = pos[3 * (size_t)vi + 1], z = pos[3 * (size_t)vi + 2];
            w[3 * c + 0] = ((m[0] * x + m[1] * y) + m[2] * z) + m[3];
            w[3 * c + 1] = ((m[4] * x + m[5] * y) + m[6] * z) + m[7];
            w[3 * c + 2] = ((m[8] * x + m[9] * y) + m[10] * z) + m[11];
            float nx = nrm[3 * (size_t)vi], ny = nrm[3 * (size_t)vi + 1], nz = nrm[3 * (size_t)vi + 2];
            tri_n[9 * (size_t)t + 3 * c + 0] = ((c00 * nx + c01 * ny) + c02 * nz) * sgn;
            tri_n[9 * (size_t)t + 3 * c + 1] = ((c10 * nx + c11 * ny) + c12 * nz) * sgn;
            tri_n[9 * (size_t)t + 3 * c + 2] = ((c20 * nx + c21 * ny) + c22 * nz) * sgn;
        }
        for (int a = 0; a < 9; ++a) tri_v[9 * (size_t)t + a] = w[a];
        tri_mat[t] = I.material;
        for (int a = 0; a < 3; ++a) {
            float l = sel_min(sel_min(w[a], w[3 + a]), w[6 + a]);
            float h = sel_max(sel_max(w[a], w[3 + a]), w[6 + a]);
            float ce = (l + h) * 0.5f;
            cmin[a] = ce;
            cmax[a] = ce;
        }
    }
    // wavefront reduction of the centroid bounds, then one atomic per wave and axis
    for (int a = 0; a < 3; ++a) {
        float mn = cmin[a], mx = cmax[a];
        for (int off = 32; off >= 1; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off));
            mx = fmaxf(mx, __shfl_xor(mx, off));
        }
        if ((threadIdx.x & 63u) == 0) {
            atomicMin(&cb[a], f2ord(mn));
            atomicMax(&cb[3 + a], f2ord(mx));
        }
    }
}

// Texture coordinates and world tangents of each leaf's vertices (scenes whose materials
// sample textures or normal maps): leaf k is triangle order[k]; its instance is found by
// binary search over the instances' first triangles; tangents take the instance's linear
// part (ObjectToWorld3x4 * (t.xyz, 0), hit.hlsl:155-156), their w (the bitangent sign) as is.
__global__ __launch_bounds__(256) void k_leaf_attrs(const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ idx,
                                                    const InstanceRec* __restrict__ inst, uint32_t ninst,
                                                    uint32_t n, const float* __restrict__ uvs,
                                                    const float* __restrict__ tans, LeafUv* __restrict__ leaf_uv,
                                                    LeafTan* __restrict__ leaf_tan) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = order[k];
    uint32_t lo = 0, hi = ninst - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (inst[mid].first_tri <= t) lo = mid; else hi = mid - 1;
    }
    const InstanceRec& I = inst[lo];
    const float* m = I.xform;
    const uint32_t kk = t - I.first_tri;
    LeafUv u;
    LeafTan g;
    for (int c = 0; c < 3; ++c) {
        const uint32_t vi = I.vtx_off + idx[I.idx_off + 3 * kk + c];
        if (leaf_uv) {
            u.uv[2 * c] = uvs[2 * (size_t)vi];
            u.uv[2 * c + 1] = uvs[2 * (size_t)vi + 1];
        }
        if (leaf_tan) {
            const float x = tans[4 * (size_t)vi], y = tans[4 * (size_t)vi + 1], z = tans[4 * (size_t)vi + 2];
            g.t[4 * c + 0] = (m[0] * x + m[1] * y) + m[2] * z;
            g.t[4 * c + 1] = (m[4] * x + m[5] * y) + m[6] * z;
            g.t[4 * c + 2] = (m[8] * x + m[9] * y) + m[10] * z;
            g.t[4 * c + 3] = tans[4 * (size_t)vi + 3];
        }
    }
    if (leaf_uv) {
        u.uv[6] = u.uv[7] = 0.0f;
        leaf_uv[k] = u;
    }
    if (leaf_tan) leaf_tan[k] = g;
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ tri_v, uint32_t n,
                                                const uint32_t* __restrict__ cb, uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals, float* __restrict__ cb_out) {
    uint32_t t = blockIdx.x * 256u + threadIdx.x;
    float cmin[3], inv_ext[3];
    for (int a = 0; a < 3; ++a) {
        cmin[a] = ord2f(cb[a]);
        float e = ord2f(cb[3 + a]) - cmin[a];
        inv_ext[a] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    if (t == 0 && cb_out) {
        for (int a = 0; a < 6; ++a) cb_out[a] = ord2f(cb[a]);
    }
    if (t >= n) return;
    const float* w = tri_v + 9 * (size_t)t;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        float l = sel_min(sel_min(w[a], w[3 + a]), w[6 + a]);
        float h = sel_max(sel_max(w[a], w[3 + a]), w[6 + a]);
        float ce = (l + h) * 0.5f;
        float f = (ce - cmin[a]) * inv_ext[a] * 1024.0f;
        int qi = (int)f;
        q[a] = (uint32_t)(qi < 0 ? 0 : (qi > 1023 ? 1023 : qi));
    }
    keys[t] = morton3(q[0], q[1], q[2]);
    vals[t] = t;
}

constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortItems = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;

__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* __restrict__ kin, uint32_t n, uint32_t shift,
                                                    uint32_t nblocks, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    for (uint32_t r = 0; r < kSortItems; ++r) {
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(kin[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m entries in place, one workgroup of 1024 lanes.
__global__ __launch_bounds__(1024) void k_scan_single(uint32_t* __restrict__ a, uint32_t m) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t start = 0; start < m; start += 1024u) {
        uint32_t i = start + threadIdx.x;
        uint32_t v = i < m ? a[i] : 0u;
        uint32_t inc = v;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(inc, off);
            if (lane >= (uint32_t)off) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t wbase = 0, total = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            uint32_t s = wsum[w];
            if (w < wave) wbase += s;
            total += s;
        }
        uint32_t carry = carry_s;
        if (i < m) a[i] = carry + wbase + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                       const uint32_t* __restrict__ hist, uint32_t n,
                                                       uint32_t nblocks, uint32_t shift) {
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_wcnt[4][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    s_off[tid] = hist[tid * nblocks + blockIdx.x];
    for (int w = 0; w < 4; ++w) s_wcnt[w][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t r = 0; r < kSortItems; ++r) {
        uint32_t i = base + r * kSortThreads + tid;
        bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t d = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            m &= bit ? bal : ~bal;
        }
        uint32_t rank = __popcll(m & lt);
        if (valid && rank == 0) s_wcnt[wave][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t dest = s_off[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) dest += s_wcnt[w][d];
            kout[dest] = key;
            vout[dest] = val;
        }
        __syncthreads();
        s_off[tid] += s_wcnt[0][tid] + s_wcnt[1][tid] + s_wcnt[2][tid] + s_wcnt[3][tid];
        for (int w = 0; w < 4; ++w) s_wcnt[w][tid] = 0;
        __syncthreads();
    }
}

__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ codes, int n,
                                                BvhNode* __restrict__ nodes, int32_t* __restrict__ parent,
                                                uint32_t* __restrict__ refit_cnt, float* __restrict__ sub_cost,
                                                int32_t* __restrict__ sub_height) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (n == 1) {
        if (i == 0) {
            nodes[0].child[0] = ~0;
            nodes[0].child[1] = ~0;
            nodes[0].pad[0] = nodes[0].pad[1] = 0;
            parent[0] = -1;
            parent[1] = 0;
            refit_cnt[0] = 0;
        }
        return;
    }
    if (i >= n - 1) return;
    int d = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(codes, n, i, i - d);
    int lmax = 2;
    while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(codes, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) / 2;
        if (kdelta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int left = (min(i, j) == gamma) ? ~gamma : gamma;
    int right = (max(i, j) == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    nodes[i].child[0] = left;
    nodes[i].child[1] = right;
    nodes[i].pad[0] = nodes[i].pad[1] = 0;
    // child boxes start as NaN: the refit reads a sibling's box once it is no longer NaN
    for (int a = 0; a < 12; ++a) nodes[i].box[a] = __uint_as_float(kBoxPending);
    parent[left < 0 ? (n - 1) + ~left : left] = i;
    parent[right < 0 ? (n - 1) + ~right : right] = i;
    if (i == 0) parent[0] = -1;
    refit_cnt[i] = 0;
    // subtree cost and height, published by the refit when the node is complete
    sub_cost[i] = __uint_as_float(kBoxPending);
    sub_height[i] = -1;
}

// Device-coherent (agent scope) relaxed accesses of box words: no cache write-back or
// invalidation is needed because a reader waits for the word itself to change.
__device__ __forceinline__ uint32_t ld_box(const float* p) {
    return __hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_box(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
// Half surface of a box (lo.xyz, hi.xyz), the surface-area cost unit.
__device__ __forceinline__ float half_area_box(const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return (dx * dy + dy * dz) + dz * dx;
}

__global__ __launch_bounds__(256) void k_leaves(const uint32_t* __restrict__ order, const float* __restrict__ tri_v,
                                                const float* __restrict__ tri_n,
                                                const uint32_t* __restrict__ tri_mat,
                                                const MaterialRec* __restrict__ mats, int n,
                                                LeafGeo* __restrict__ geo, LeafShade* __restrict__ shade,
                                                BvhNode* nodes, const int32_t* __restrict__ parent,
                                                uint32_t* refit_cnt, float* sub_cost, int32_t* sub_height,
                                                uint32_t* error) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    uint32_t t = order[k];
    const float* w = tri_v + 9 * (size_t)t;
    float v[9];
    for (int a = 0; a < 9; ++a) v[a] = w[a];
    LeafGeo g;
    for (int a = 0; a < 3; ++a) {
        g.v[a] = v[a];
        g.v[3 + a] = v[3 + a] - v[a];
        g.v[6 + a] = v[6 + a] - v[a];
    }
    // alpha-tested and translucent triangles are marked for the traversal's any-hit test
    // (the instance flags force_opaque / force_non_opaque of accel.cpp:108-116, per material)
    g.v[9] = __uint_as_float(mats[tri_mat[t]].blend_mode != kBlendOpaque ? kLeafNonOpaque : 0u);
    g.v[10] = g.v[11] = 0.0f;
    geo[k] = g;
    LeafShade sh;
    for (int a = 0; a < 9; ++a) sh.n[a] = tri_n[9 * (size_t)t + a];
    sh.material = tri_mat[t];
    sh.pad[0] = sh.pad[1] = 0;
    shade[k] = sh;
    float box[6];
    for (int a = 0; a < 3; ++a) {
        box[a] = sel_min(sel_min(v[a], v[3 + a]), v[6 + a]) + 0.0f;
        box[3 + a] = sel_max(sel_max(v[a], v[3 + a]), v[6 + a]) + 0.0f;
    }
    if (n == 1) {
        for (int a = 0; a < 6; ++a) { nodes[0].box[a] = box[a]; nodes[0].box[6 + a] = box[a]; }
        return;
    }
    int p = parent[(n - 1) + k];
    int slot = nodes[p].child[0] == ~k ? 0 : 1;
    // surface-area cost and height of the subtree this thread carries upward (a
    // triangle: its box's surface, height 0)
    float cost = half_area_box(box);
    int height = 0;
    for (;;) {
        for (int a = 0; a < 6; ++a) st_box(&nodes[p].box[6 * slot + a], box[a]);
        uint32_t old = __hip_atomic_fetch_add(&refit_cnt[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;   // the sibling subtree is still being built; it continues upward
        // second arrival: the sibling has issued its box; wait until every word is visible
        const float* ob = nodes[p].box + 6 * (1 - slot);
        float other[6];
        for (int a = 0; a < 6; ++a) {
            uint32_t bits = ld_box(ob + a);
            uint32_t spins = 0;
            while (bits == kBoxPending) {
                if (++spins > (1u << 20)) { atomicOr(error, 4u); bits = 0u; break; }
                __builtin_amdgcn_s_sleep(1);
                bits = ld_box(ob + a);
            }
            other[a] = __uint_as_float(bits);
        }
        for (int a = 0; a < 3; ++a) {
            box[a] = sel_min(box[a], other[a]) + 0.0f;
            box[3 + a] = sel_max(box[3 + a], other[3 + a]) + 0.0f;
        }
        // the sibling subtree's cost and height: a triangle's from its box, an internal
        // node's once its thread has published them
        const int sib = nodes[p].child[1 - slot];
        float sib_cost;
        int sib_height;
        if (sib < 0) {
            sib_cost = half_area_box(other);
            sib_height = 0;
        } else {
            uint32_t bits = ld_box(sub_cost + sib);
            uint32_t spins = 0;
            while (bits == kBoxPending) {
                if (++spins > (1u << 20)) { atomicOr(error, 4u); bits = 0u; break; }
                __builtin_amdgcn_s_sleep(1);
                bits = ld_box(sub_cost + sib);
            }
            sib_cost = __uint_as_float(bits);
            sib_height = __hip_atomic_load(sub_height + sib, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            spins = 0;
            while (sib_height < 0) {
                if (++spins > (1u << 20)) { atomicOr(error, 4u); sib_height = 0; break; }
                __builtin_amdgcn_s_sleep(1);
                sib_height = __hip_atomic_load(sub_height + sib, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const float c0 = slot == 0 ? cost : sib_cost, c1 = slot == 0 ? sib_cost : cost;
        cost = (half_area_box(box) + c0) + c1;
        height = 1 + max(height, sib_height);
        st_box(sub_cost + p, cost);
        __hip_atomic_store(sub_height + p, height, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == 0) return;
        int q = parent[p];
        slot = nodes[q].child[0] == p ? 0 : 1;
        p = q;
    }
}

// BVH4 collapse.  Every internal BVH2 node k is turned into a BVH4 node stored at
// index k, independently of the others: starting from k's two children, the internal
// child with the largest box surface (ties: the lower slot) is replaced in its slot by
// its first child and its second child is appended, until four slots are filled or
// only leaves are left.  Traversal starts at node 0 and follows internal links, which
// are BVH2 node indices and so name the BVH4 node built from that node; nodes not
// reached that way are never read.  Runs after the refit: every box is final.
__device__ __forceinline__ float half_area(const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return (dx * dy + dy * dz) + dz * dx;
}

// ---------------------------------------------------------------- treelet restructuring
// Rounds of treelet restructuring (Karras & Aila 2013, "Fast Parallel Construction of
// High-Quality Bounding Volume Hierarchies") lower the tree's surface-area cost after
// the refit.  Node n's treelet starts as its two children; the treelet leaf that is an
// internal node with the largest box surface (ties: the lower position) is replaced in
// place by its first child and its second child appended, until 7 treelet leaves (or
// only triangles) remain.  Over the subsets of those leaves, cost(s) = area(s) + the
// cheapest split of s into two parts (the part holding the lowest leaf first, parts
// enumerated as p = (p - 1) & s, first minimum kept), a single leaf costing its
// subtree's cost (a triangle: its box's surface).  If the best topology is cheaper than
// the current one, and does not raise the subtree's height or keeps it <= 48 (so the
// tree stays within 63 levels and the traversal stack bound holds), the treelet is
// rebuilt from the same internal nodes, assigned in pre-order (left part first).
// Only nodes whose subtree is at least kTreeletMinHeight levels tall at the start of a
// round are restructured, children first: restructuring a node changes nothing outside
// its subtree and subtrees of nodes where neither contains the other are disjoint, so
// every children-first order gives the same tree -- the CPU restatement's recursion.
// One launch per round schedules them bottom-up like the refit: a wavefront takes each
// node of height kTreeletMinHeight..kTreeletSerialHeight whose parent stands higher,
// restructures that node's subtree alone (children first), then climbs to the parent --
// directly if it is the parent's only restructured child, else it counts itself in and
// the second arrival carries on, the first ends.  No wavefront waits for another.  The
// refit publishes every subtree's cost and height for the first round; a round leaves
// them current for the next.
// [host-wave begin] From here to the end marker: the treelet code of one wavefront,
// which tests/test_treelet_host.py also compiles for the CPU (64 threads as the lanes).
constexpr int kTreeletLeaves = 7;
constexpr int kTreeletMaxHeight = 48;
constexpr int kTreeletMinHeight = 5;   // lower subtrees are left as built (their share of the cost is small)
constexpr uint32_t kTreeletSoleChild = 0x100u;   // arrival count of a parent with one restructured child
// s_waitcnt operand (gfx9 encoding) that waits for vmcnt 0 only: expcnt and lgkmcnt at
// their maxima (bits 6:4 and 11:8), vmcnt 0 (bits 3:0 and 15:14)
constexpr unsigned kWaitVmcnt0 = 0x0F70u;
// Subtrees whose root stands at most this high (and whose parent stands higher) are
// restructured by one wavefront alone, children first: at most 2^(8 - 4) - 1 = 15 nodes
// of height >= kTreeletMinHeight each.
constexpr int kTreeletSerialHeight = 8;
constexpr int kTreeletSerialMax = 1 << (kTreeletSerialHeight - 4);

__device__ __forceinline__ float half_area6(const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return (dx * dy + dy * dz) + dz * dx;
}

struct TreeletArgs {
    BvhNode* nodes;
    int32_t* parent;
    int ni, n;
    float* cost;        // per internal node: surface-area cost of its subtree
    int32_t* height;    // per internal node: levels below it
    int32_t* snap;      // per internal node: its height at the start of the round
    uint32_t* arrivals; // per internal node: kTreeletSoleChild if one child is restructured
                        // this round, else 0 + arrivals
    uint32_t* starts;   // [0]: count, then the nodes a climb starts from (at most n / 6)
    uint32_t start_cap;
    uint32_t* error;
};

// One treelet per wavefront.  Every lane forms the same treelet (uniform values, leaf
// boxes in registers with compile-time indices, leaf and internal node indices in LDS);
// lane l computes the unions of subsets l + 1 and l + 65; the dynamic programme runs
// over subset sizes (one lane per subset of up to 5 leaves, 8 lanes per subset of 6, the
// wavefront for the full set), the table in LDS; an accepted treelet is rebuilt with
// every link written by its own lane.
struct TreeletLds {
    float box[1 << kTreeletLeaves][6];   // union of every subset's boxes
    float area[1 << kTreeletLeaves];     // its half surface
    float cost[1 << kTreeletLeaves];
    uint8_t split[1 << kTreeletLeaves];
    uint8_t hh[1 << kTreeletLeaves];
    int8_t slot[1 << kTreeletLeaves];   // rebuild: pre-order slot of an internal subset, else -1
    int8_t up[1 << kTreeletLeaves];     // rebuild: slot of the parent of a subset in the tree, else -1
    int8_t side[1 << kTreeletLeaves];   // rebuild: 0 left part, 1 right part
    int ref[8];     // treelet leaves: internal node index, or ~triangle
    int inner[8];   // the treelet's internal nodes, root first
    int list[kTreeletSerialMax];   // the start node's subtree: its nodes to restructure, level by level
};

// Subsets of 7 leaves by size (1..5): entry [s][l] is the l-th subset of s leaves in
// increasing order, 0 past the last; lane l of the dynamic programme owns [s][l].
struct TreeletBySize {
    uint8_t v[6][64];
};
constexpr TreeletBySize make_treelet_by_size() {
    TreeletBySize t{};
    int cnt[6] = {};
    for (int q = 1; q < (1 << kTreeletLeaves); ++q) {
        const int pc = __builtin_popcount(q);
        if (pc <= 5) t.v[pc][cnt[pc]++] = (uint8_t)q;
    }
    return t;
}
__constant__ TreeletBySize kTreeletBySize = make_treelet_by_size();

__device__ __forceinline__ void tl_union(const float (&bx)[kTreeletLeaves][6], int k, int s, float* u) {
    bool first = true;
#pragma unroll
    for (int i = 0; i < kTreeletLeaves; ++i) {
        if (i >= k || !((s >> i) & 1)) continue;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            u[e] = first ? bx[i][e] : sel_min(u[e], bx[i][e]);
            u[3 + e] = first ? bx[i][3 + e] : sel_max(u[3 + e], bx[i][3 + e]);
        }
        first = false;
    }
}

// Cheapest split of a subset: the minimum of cost[p] + cost[sub ^ p] over the parts p
// holding the lowest leaf, the largest p among equal minima (the one the serial
// enumeration p = (p - 1) & sub meets first), 3e38 and p = 0 if no sum is below 3e38.
// Candidate c of a subset deposits its bits, lowest first, at the subset's set bits
// above the lowest; the candidates are evaluated unrolled so their LDS reads overlap.
__device__ __forceinline__ void tl_take(float v, int p, float& best, int& bp) {
    if (v < best || (v == best && bp != 0 && p > bp)) {
        best = v;
        bp = p;
    }
}
__device__ __forceinline__ int tl_deposit(int c, int low, const int (&b)[kTreeletLeaves - 1]) {
    int p = low;
#pragma unroll
    for (int j = 0; j < kTreeletLeaves - 1; ++j) p |= ((c >> j) & 1) ? b[j] : 0;
    return p;
}
__device__ __forceinline__ void tl_bits(int sub, int& low, int (&b)[kTreeletLeaves - 1]) {
    low = sub & -sub;
    int m = sub ^ low;
#pragma unroll
    for (int j = 0; j < kTreeletLeaves - 1; ++j) {
        b[j] = m & -m;
        m ^= b[j];
    }
}

// All candidates of one subset of S leaves in one lane.
template <int S>
__device__ __forceinline__ void tl_split_lane(const TreeletLds& L, int sub, float& best, int& bp) {
    constexpr int kCand = (1 << (S - 1)) - 1;
    int low, b[kTreeletLeaves - 1];
    tl_bits(sub, low, b);
    float v[kCand];
    int pp[kCand];
#pragma unroll
    for (int c = 0; c < kCand; ++c) {
        pp[c] = tl_deposit(c, low, b);
        v[c] = L.cost[pp[c]] + L.cost[sub ^ pp[c]];
    }
    best = 3.0e38f;
    bp = 0;
#pragma unroll
    for (int c = 0; c < kCand; ++c) tl_take(v[c], pp[c], best, bp);
}

// The candidates of `sub` (0: none) spread over a group of `width` lanes (a power of two,
// lane j of the group takes candidates j, j + width, ...; at most kPer each), the
// result shuffled to every lane of the group.
template <int kPer>
__device__ __forceinline__ void tl_split_group(const TreeletLds& L, int sub, int j, int width, float& best, int& bp) {
    best = 3.0e38f;
    bp = 0;
    if (sub != 0) {
        const int ncand = (1 << (__popc(sub) - 1)) - 1;
        int low, b[kTreeletLeaves - 1];
        tl_bits(sub, low, b);
        float v[kPer];
        int pp[kPer];
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int c = j + width * t;
            pp[t] = c < ncand ? tl_deposit(c, low, b) : 0;
            v[t] = c < ncand ? L.cost[pp[t]] + L.cost[sub ^ pp[t]] : 3.0e38f;
        }
#pragma unroll
        for (int t = 0; t < kPer; ++t)
            if (pp[t] != 0) tl_take(v[t], pp[t], best, bp);
    }
    for (int o = 1; o < width; o <<= 1) {
        const float ob = __shfl_xor(best, o);
        const int op = __shfl_xor(bp, o);
        if (op != 0 && (bp == 0 || ob < best || (ob == best && op > bp))) {
            best = ob;
            bp = op;
        }
    }
}

// One subset size of the dynamic programme: `sub` is the lane's subset of S leaves.
template <int S>
__device__ __forceinline__ void tl_dp_size(const TreeletArgs& a, TreeletLds& L, int sub, int full) {
    if (sub != 0 && sub <= full) {
        if constexpr (S == 1) {
            const int r = L.ref[__builtin_ctz(sub)];
            L.cost[sub] = r >= 0 ? a.cost[r] : L.area[sub];
            L.hh[sub] = (uint8_t)(r >= 0 ? a.height[r] : 0);
            L.split[sub] = 0;
        } else {
            float best;
            int bp;
            tl_split_lane<S>(L, sub, best, bp);
            L.cost[sub] = L.area[sub] + best;
            L.split[sub] = (uint8_t)bp;
            L.hh[sub] = (uint8_t)(1 + max((int)L.hh[bp], (int)L.hh[sub ^ bp]));
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// subs[s]: the lane's subset of s leaves (kTreeletBySize), s = 1..5.
__device__ void tl_process_wave(const TreeletArgs& a, int n, uint32_t lane, TreeletLds& L, const int (&subs)[6]) {
    BvhNode* nodes = a.nodes;
    // the treelet's leaf boxes stay in registers (compile-time indices only); leaf and
    // internal node indices sit in LDS, identical in every lane; bit i of `internal`
    // marks a leaf that is an internal node
    float bx[kTreeletLeaves][6];
#pragma unroll
    for (int i = 0; i < kTreeletLeaves; ++i)
#pragma unroll
        for (int e = 0; e < 6; ++e) bx[i][e] = 0.0f;
    uint32_t internal = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int ch = nodes[n].child[c];
        L.ref[c] = ch;
        internal |= ch >= 0 ? 1u << c : 0u;
#pragma unroll
        for (int e = 0; e < 6; ++e) bx[c][e] = nodes[n].box[6 * c + e];
    }
    L.inner[0] = n;
    int k = 2;
    for (int step = 0; step < kTreeletLeaves - 2; ++step) {
        int best = -1;
        float ba = -1.0f;
#pragma unroll
        for (int i = 0; i < kTreeletLeaves; ++i) {
            if (!((internal >> i) & 1u)) continue;
            const float ar = half_area6(bx[i]);
            if (ar > ba) { ba = ar; best = i; }
        }
        if (best < 0) break;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int m = L.ref[best];
        const int c0 = nodes[m].child[0], c1 = nodes[m].child[1];
        float b0[6], b1[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            b0[e] = nodes[m].box[e];
            b1[e] = nodes[m].box[6 + e];
        }
        // every lane has read ref[best] before any lane overwrites it
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < kTreeletLeaves; ++i) {
            if (i == best) {
#pragma unroll
                for (int e = 0; e < 6; ++e) bx[i][e] = b0[e];
            }
            if (i == k) {
#pragma unroll
                for (int e = 0; e < 6; ++e) bx[i][e] = b1[e];
            }
        }
        internal = (internal & ~(1u << best)) | (c0 >= 0 ? 1u << best : 0u) | (c1 >= 0 ? 1u << k : 0u);
        L.ref[best] = c0;
        L.ref[k] = c1;
        L.inner[step + 1] = m;
        ++k;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float all[6];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        all[e] = sel_min(nodes[n].box[e], nodes[n].box[6 + e]);
        all[3 + e] = sel_max(nodes[n].box[3 + e], nodes[n].box[9 + e]);
    }
    const int r0 = nodes[n].child[0], r1 = nodes[n].child[1];
    const float cc0 = r0 >= 0 ? a.cost[r0] : half_area6(nodes[n].box);
    const float cc1 = r1 >= 0 ? a.cost[r1] : half_area6(nodes[n].box + 6);
    const float cur = (half_area6(all) + cc0) + cc1;
    const int curh = 1 + max(r0 >= 0 ? a.height[r0] : 0, r1 >= 0 ? a.height[r1] : 0);
    bool rebuilt = false;
    if (k >= 3) {
        const int full = (1 << k) - 1;
        // every subset's union and its half surface once, lane l owning subsets l + 1
        // and l + 65
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int sub = (int)lane + 1 + 64 * h;
            float u[6];
            tl_union(bx, k, sub, u);
            if (sub <= full) {
#pragma unroll
                for (int e = 0; e < 6; ++e) L.box[sub][e] = u[e];
                L.area[sub] = half_area6(u);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // subsets of up to 5 leaves by size, one lane per subset
        tl_dp_size<1>(a, L, subs[1], full);
        tl_dp_size<2>(a, L, subs[2], full);
        tl_dp_size<3>(a, L, subs[3], full);
        if (k >= 4) tl_dp_size<4>(a, L, subs[4], full);
        if (k >= 5) tl_dp_size<5>(a, L, subs[5], full);
        // subsets of 6 and 7 leaves: 8 lanes for each subset of 6 out of 7 leaves, the
        // whole wavefront for the full set
        if (k == 7) {
            const int g = (int)(lane >> 3);
            const int sub = g < kTreeletLeaves ? full ^ (1 << g) : 0;
            float best;
            int bp;
            tl_split_group<4>(L, sub, (int)(lane & 7u), 8, best, bp);
            if (sub != 0 && (lane & 7u) == 0) {
                L.cost[sub] = L.area[sub] + best;
                L.split[sub] = (uint8_t)bp;
                L.hh[sub] = (uint8_t)(1 + max((int)L.hh[bp], (int)L.hh[sub ^ bp]));
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        if (k >= 6) {
            float best;
            int bp;
            tl_split_group<1>(L, full, (int)lane, 64, best, bp);
            if (lane == 0) {
                L.cost[full] = L.area[full] + best;
                L.split[full] = (uint8_t)bp;
                L.hh[full] = (uint8_t)(1 + max((int)L.hh[bp], (int)L.hh[full ^ bp]));
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        const float best_cost = L.cost[full];
        const int best_h = L.hh[full];
        if (best_cost < cur && (best_h <= curh || best_h <= kTreeletMaxHeight)) {
            rebuilt = true;
            // rebuild in pre-order (left part first) from the treelet's internal nodes:
            // the root takes slot 0; the parts of an internal subset at slot t take
            // t + 1 (left) and t + popc(left) (right: the left part's subtree holds
            // popc(left) - 1 internal nodes), leaves no slot.  Filled top-down, one tree
            // level per pass (at most 6 levels below the root); then every subset in
            // the tree writes its link, box and parent in parallel.
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = (int)lane + 64 * h;
                L.slot[q] = q == full ? 0 : -1;
                L.up[q] = -1;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int pass = 0; pass < kTreeletLeaves - 1; ++pass) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int q = (int)lane + 64 * h;
                    const int t = q <= full ? L.slot[q] : -1;
                    if (t < 0) continue;
                    const int lp = L.split[q], rp = q ^ lp;
                    L.up[lp] = (int8_t)t;
                    L.side[lp] = 0;
                    L.up[rp] = (int8_t)t;
                    L.side[rp] = 1;
                    if (lp & (lp - 1)) L.slot[lp] = (int8_t)(t + 1);
                    if (rp & (rp - 1)) L.slot[rp] = (int8_t)(t + __popc(lp));
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = (int)lane + 64 * h;
                if (q == 0 || q > full) continue;
                const int t = L.slot[q], u = L.up[q];
                if (t < 0 && u < 0) continue;   // not in the rebuilt treelet
                const int me = t >= 0 ? L.inner[t] : L.ref[__builtin_ctz(q)];
                if (t >= 0) {
                    a.cost[me] = L.cost[q];
                    a.height[me] = L.hh[q];
                }
                if (u < 0) continue;   // the root
                const int par = L.inner[u], sd = L.side[q];
                nodes[par].child[sd] = me;
#pragma unroll
                for (int e = 0; e < 6; ++e) nodes[par].box[6 * sd + e] = L.box[q][e];
                a.parent[me >= 0 ? me : (a.n - 1) + ~me] = par;
            }
        }
    }
    if (!rebuilt && lane == 0) {
        a.cost[n] = cur;
        a.height[n] = curh;
    }
}

// [host-wave end]

__global__ __launch_bounds__(256) void k_tl_prep(TreeletArgs a) {
    const int k = (int)(blockIdx.x * 256u + threadIdx.x);
    if (k >= a.ni) return;
    const int h = a.height[k];
    a.snap[k] = h;
    uint32_t t = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int ch = a.nodes[k].child[c];
        t += (ch >= 0 && a.height[ch] >= kTreeletMinHeight) ? 1u : 0u;
    }
    a.arrivals[k] = t == 1u ? kTreeletSoleChild : 0u;
    // a wavefront starts at every node of height kTreeletMinHeight..kTreeletSerialHeight
    // whose parent stands higher (or that is the root): these subtrees are disjoint and
    // hold at least 6 triangles each.  One atomic per wavefront.
    const bool start = h >= kTreeletMinHeight && h <= kTreeletSerialHeight &&
                       (k == 0 || a.height[a.parent[k]] > kTreeletSerialHeight);
    const uint64_t m = __ballot(start);
    if (!start) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t first = (uint32_t)__builtin_ctzll(m);
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(a.starts, (uint32_t)__popcll(m));
    base = __shfl(base, (int)first);
    const uint32_t idx = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (idx < a.start_cap) a.starts[1 + idx] = (uint32_t)k;
    else atomicOr(a.error, 16u);
}

// Workgroup w (one wavefront) restructures the subtree of the w-th start node, then
// climbs from there.  A parent whose only
// restructured child is this one is taken over directly (a workgroup-scope fence makes
// this wavefront's own stores visible to all its lanes).  Otherwise: release fence, one
// relaxed counter increment by lane 0, acquire fence on the arrival that carries on, so
// the next restructuring sees every write below it whichever XCD made them.
__global__ __launch_bounds__(64) void k_tl_climb(TreeletArgs a) {
    __shared__ TreeletLds s_tl;
    const uint32_t lane = threadIdx.x;
    const uint32_t i = blockIdx.x;
    const uint32_t count = __builtin_amdgcn_readfirstlane(a.starts[0]);
    if (i >= min(count, a.start_cap)) return;
    int subs[6];
#pragma unroll
    for (int sz = 0; sz < 6; ++sz) subs[sz] = kTreeletBySize.v[sz][lane];
    int n = __builtin_amdgcn_readfirstlane((int)a.starts[1 + i]);
    // the start node's subtree, level by level (lane j expands entry j of a level; the
    // nodes it adds keep lane order), then restructured in reverse: children first
    if (lane == 0) s_tl.list[0] = n;
    int lo = 0, hi = 1;
    while (lo < hi) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        int c0 = -1, c1 = -1;
        bool t0 = false, t1 = false;
        if (lo + (int)lane < hi) {
            const int x = s_tl.list[lo + lane];
            c0 = a.nodes[x].child[0];
            c1 = a.nodes[x].child[1];
            t0 = c0 >= 0 && a.snap[c0] >= kTreeletMinHeight;
            t1 = c1 >= 0 && a.snap[c1] >= kTreeletMinHeight;
        }
        const uint64_t b0 = __ballot(t0), b1 = __ballot(t1);
        const int added = __popcll(b0) + __popcll(b1);
        if (hi + added > kTreeletSerialMax) {   // cannot happen for a start node (see above)
            if (lane == 0) atomicOr(a.error, 16u);
            return;
        }
        const uint64_t below = (1ull << lane) - 1ull;
        const int pos = hi + __popcll(b0 & below) + __popcll(b1 & below);
        if (t0) s_tl.list[pos] = c0;
        if (t1) s_tl.list[pos + (t0 ? 1 : 0)] = c1;
        lo = hi;
        hi += added;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int j = hi - 1; j > 0; --j) {
        tl_process_wave(a, __builtin_amdgcn_readfirstlane(s_tl.list[j]), lane, s_tl, subs);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    for (;;) {
        tl_process_wave(a, n, lane, s_tl, subs);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (n == 0) break;
        const int p = __builtin_amdgcn_readfirstlane(a.parent[n]);
        const uint32_t pre = __builtin_amdgcn_readfirstlane(a.arrivals[p]);
        if (pre == kTreeletSoleChild) {   // the only restructured child of p (no count is kept)
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            n = p;
            continue;
        }
        // release: the fence writes this XCD's L2 back (buffer_wbl2), and the write-back
        // must be complete before the counter moves; hipcc of ROCm 7.2 drops the
        // vmcnt wait after buffer_wbl2 when no other vector-memory operation is
        // outstanding, so the wait is stated here
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        uint32_t old = 0;
        if (lane == 0) old = __hip_atomic_fetch_add(&a.arrivals[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        old = __builtin_amdgcn_readfirstlane(old);
        if (old != 1u) break;   // the other restructured child of p is still to come; its wavefront carries on
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // buffer_inv: later loads miss stale lines
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        n = p;
    }
}

__global__ __launch_bounds__(256) void k_collapse4(const BvhNode* __restrict__ nodes, int ni,
                                                   Bvh4Node* __restrict__ out) {
    const int k = (int)(blockIdx.x * 256u + threadIdx.x);
    if (k >= ni) return;
    int id[4];
    float box[4][6];
    int cnt = 2;
    for (int c = 0; c < 2; ++c) {
        id[c] = nodes[k].child[c];
        for (int a = 0; a < 6; ++a) box[c][a] = nodes[k].box[6 * c + a];
    }
    while (cnt < 4) {
        int pick = -1;
        float best = -1.0f;
        for (int j = 0; j < cnt; ++j) {
            if (id[j] < 0) continue;
            const float ar = half_area(box[j]);
            if (ar > best) { best = ar; pick = j; }
        }
        if (pick < 0) break;
        const BvhNode& m = nodes[id[pick]];
        id[cnt] = m.child[1];
        for (int a = 0; a < 6; ++a) box[cnt][a] = m.box[6 + a];
        id[pick] = m.child[0];
        for (int a = 0; a < 6; ++a) box[pick][a] = m.box[a];
        ++cnt;
    }
    Bvh4Node r;
    const float inf = __builtin_inff();
    for (int j = 0; j < 4; ++j) {
        const bool used = j < cnt;
        r.lox[j] = used ? box[j][0] : inf;
        r.loy[j] = used ? box[j][1] : inf;
        r.loz[j] = used ? box[j][2] : inf;
        r.hix[j] = used ? box[j][3] : inf;
        r.hiy[j] = used ? box[j][4] : inf;
        r.hiz[j] = used ? box[j][5] : inf;
        r.child[j] = used ? id[j] : 0;
        r.pad[j] = 0;
    }
    out[k] = r;
}

}  // namespace

void launch_build_lbvh(const BuildArgs& a, hipStream_t s) {
    const uint32_t n = a.num_tris;
    if (n == 0) return;
    const uint32_t g256 = (n + 255) / 256;
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, a.cbounds_ord);
    hipLaunchKernelGGL(k_flatten, dim3(g256), dim3(256), 0, s, a.positions, a.normals, a.indices, a.instances,
                       a.num_instances, n, a.tri_v, a.tri_n, a.tri_mat, a.cbounds_ord);
    hipLaunchKernelGGL(k_morton, dim3(g256), dim3(256), 0, s, a.tri_v, n, a.cbounds_ord, a.keys[0], a.vals[0],
                       a.cbounds_out);
    const uint32_t nblocks = (n + kSortTile - 1) / kSortTile;
    int cur = 0;
    for (uint32_t shift = 0; shift < 30; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(256), 0, s, a.keys[cur], n, shift, nblocks, a.hist);
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(1024), 0, s, a.hist, 256u * nblocks);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(256), 0, s, a.keys[cur], a.vals[cur],
                           a.keys[cur ^ 1], a.vals[cur ^ 1], a.hist, n, nblocks, shift);
        cur ^= 1;
    }
    // 4 passes: the sorted keys/values end in buffer 0
    hipLaunchKernelGGL(k_karras, dim3(g256), dim3(256), 0, s, a.keys[cur], (int)n, a.nodes, a.parent, a.refit_cnt,
                       a.tl_cost, a.tl_height);
    hipLaunchKernelGGL(k_leaves, dim3(g256), dim3(256), 0, s, a.vals[cur], a.tri_v, a.tri_n, a.tri_mat,
                       a.materials, (int)n, a.leaf_geo, a.leaf_shade, a.nodes, a.parent, a.refit_cnt, a.tl_cost,
                       a.tl_height, a.error);
    if (a.leaf_uv || a.leaf_tan)
        hipLaunchKernelGGL(k_leaf_attrs, dim3(g256), dim3(256), 0, s, a.vals[cur], a.indices, a.instances,
                           a.num_instances, n, a.uvs, a.tangents, a.leaf_uv, a.leaf_tan);
    const int ni = n > 1 ? (int)n - 1 : 1;
    if (a.treelet_rounds > 0 && n >= 3) {
        // the refit's arrival counters are free again: they count arrivals of the climb
        const uint32_t cap = treelet_start_cap(n);
        TreeletArgs t{a.nodes, a.parent, ni, (int)n, a.tl_cost, a.tl_height, a.tl_snap, a.refit_cnt,
                      a.tl_starts, cap, a.error};
        const uint32_t gi = (uint32_t)(ni + 255) / 256u;
        for (uint32_t r = 0; r < a.treelet_rounds; ++r) {
            launch_fill_u32(a.tl_starts, 0u, 1u, s);
            hipLaunchKernelGGL(k_tl_prep, dim3(gi), dim3(256), 0, s, t);
            hipLaunchKernelGGL(k_tl_climb, dim3(cap), dim3(64), 0, s, t);
        }
    }
    hipLaunchKernelGGL(k_collapse4, dim3((ni + 255) / 256), dim3(256), 0, s, a.nodes, ni, a.nodes4);
}

uint32_t treelet_start_cap(uint32_t n) { return n / 6u + 1u; }

uint32_t sort_hist_entries(uint32_t n) { return 256u * ((n + kSortTile - 1) / kSortTile); }

}  // namespace bipt
