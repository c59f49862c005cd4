// shade.hip -- surface shading, next-event estimation and queue compaction (gfx950).
//
// One segment's shading replaces the reference's per-bounce passes
// sample_secondary_ray_cs (sample_secondary_ray.hlsl:11-69) and
// deferred_lighting_secondary_cs (deferred_lighting_secondary.hlsl:11-111):
//   miss   -> sky radiance * throughput (MIS-weighted against sky NEE after bounce 0)
//   hit    -> interpolated normal, lit BSDF; one shadow ray per directional light
//             (exact visibility replaces the shadow map); one sky NEE shadow ray;
//             VNDF-sampled extension ray while max(throughput) >= 0.001
// Instead of full-screen passes with dead lanes, live rays are compacted into the
// next queues by the shading kernel itself, in entry order (deterministic contents):
// one lane per queue entry keeps its outputs in registers; workgroup t shades tile t
// (256 entries), scans its entries' emissions (extension ray, shadow rays) and
// publishes the tile total, then finds its offset by decoupled look-back over the
// tiles before it (one wavefront reads 64 tile words at a time) and publishes its
// inclusive prefix; the lanes then write their rays at their final places.  A tile
// waits only for lower tiles, and publishes its total before it waits.  Workgroups
// reach each XCD in index order, so the lowest unfinished tile is always running and
// every wait ends; the wait is bounded all the same (error flag 32).  (A ticket taken
// with one atomic per workgroup instead of the index costs more than the compaction
// saves: ~130k same-address atomics per launch.)
#include "pt_kernels.hpp"

namespace bipt {
namespace {

constexpr uint32_t kTile = 256;
constexpr uint32_t kMaskExt = 1u;   // bit 0: extension ray; bit 1 + k: shadow slot k (lights 0..3, sky 4)

// Look-back word of a tile: status in bits 62-63 (1 tile total, 2 inclusive prefix),
// shadow rays in bits 31-61, extension rays in bits 0-30 (at most 5 * 2^28 and 2^28:
// fields never carry into each other when words are added).
constexpr uint64_t kLbTotal = 1ull << 62, kLbPrefix = 2ull << 62, kLbFields = (1ull << 62) - 1ull;
constexpr uint32_t kLbSpinMax = 1u << 22;
__device__ __forceinline__ uint64_t lb_word(uint32_t ext, uint32_t sh) { return ((uint64_t)sh << 31) | ext; }
__device__ __forceinline__ uint32_t lb_ext(uint64_t w) { return (uint32_t)(w & 0x7fffffffull); }
__device__ __forceinline__ uint32_t lb_sh(uint64_t w) { return (uint32_t)((w >> 31) & 0x7fffffffull); }

// Outputs of one queue entry, held until the entry's place in the queues is known.  NL:
// the most lights of any job in the launch (1, 2 or 4), so only that many contributions
// take registers.
template <uint32_t NL>
struct ShadeOut {
    uint32_t mask = 0, path = 0;
    V3 P, ext, sky;
    V3 c[NL + 1];   // shadow contribution of light l at [l], of the sky at [NL]
};

// Texture-driven surface fields at a hit (material_function statements that sample the
// material's textures at vertex.texcoord, hits/rt_gbuffer_hit.hlsl:8-11): base colour,
// f0, roughness, and the normal map applied in the interpolated tangent frame,
// N = normalize(n.x T + n.y B + n.z N) with n = 2 texel - 1, T = normalize(tangent.xyz),
// B = normalize(cross(N, T)) * tangent.w (hit.hlsl:155-157).
__device__ __forceinline__ void apply_textures(const JobRec& J, const MaterialRec& M, uint32_t leaf, float hu,
                                               float hv, Mat& m, V3& N, float& rough) {
    float s = 0.0f, t = 0.0f;
    if (J.leaf_uv) leaf_texcoord(J.leaf_uv[leaf], hu, hv, &s, &t);
    if (M.base_color_tex != kNoTexture) {
        const V4 c = tex_sample(J.tex, M.base_color_tex - 1u, s, t);
        m.base = v3(c.x, c.y, c.z);
    }
    if (M.f0_tex != kNoTexture) {
        const V4 c = tex_sample(J.tex, M.f0_tex - 1u, s, t);
        m.f0 = v3(c.x, c.y, c.z);
    }
    if (M.roughness_tex != kNoTexture) {
        rough = comp4(tex_sample(J.tex, M.roughness_tex - 1u, s, t), M.roughness_channel);
        mat_alpha(m, rough, M.anisotropy);
    }
    if (M.normal_tex != kNoTexture && J.leaf_tan) {
        const V4 c = tex_sample(J.tex, M.normal_tex - 1u, s, t);
        const V3 nt = v3(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f);
        const float* g = J.leaf_tan[leaf].t;
        const V3 T = normalize(v3((g[0] + (g[4] - g[0]) * hu) + (g[8] - g[0]) * hv,
                                  (g[1] + (g[5] - g[1]) * hu) + (g[9] - g[1]) * hv,
                                  (g[2] + (g[6] - g[2]) * hu) + (g[10] - g[2]) * hv));
        const float tw = (g[3] + (g[7] - g[3]) * hu) + (g[11] - g[3]) * hv;
        const V3 B = normalize(cross(N, T)) * tw;
        N = normalize((T * nt.x + B * nt.y) + N * nt.z);
    }
}

// The surface a ray (direction d) sees at a hit of leaf `leaf` at (hu, hv), point P: the
// interpolated normal, the material (a checkerboard's cell by P), textures and normal
// map, then the two-sided flip against d.
__device__ __forceinline__ void hit_surface(const JobRec& J, uint32_t leaf, float hu, float hv, V3 P, V3 d, Mat& m,
                                            V3& N, float& rough, uint32_t& mat_out) {
    const float4* lsp = reinterpret_cast<const float4*>(J.leaf_shade + leaf);
    const float4 s0 = lsp[0], s1 = lsp[1], s2 = lsp[2];
    V3 n0 = v3(s0.x, s0.y, s0.z), n1 = v3(s0.w, s1.x, s1.y), n2 = v3(s1.z, s1.w, s2.x);
    uint32_t mat_id = __float_as_uint(s2.y);
    N = normalize((n0 + (n1 - n0) * hu) + (n2 - n0) * hv);
    // the triangle's record is loaded at once (its pattern word with it); a checkerboard's
    // even cells then reload from the pattern material, so plain materials wait for one load
    m = load_mat(J.materials[mat_id]);
    if (const uint32_t cell_id = resolve_material(J.materials, mat_id, P); cell_id != mat_id) {
        mat_id = cell_id;
        m = load_mat(J.materials[mat_id]);
    }
    rough = J.materials[mat_id].roughness;
    if (mat_textured(J.materials[mat_id])) apply_textures(J, J.materials[mat_id], leaf, hu, hv, m, N, rough);
    if (m.two_sided && dot(d, N) > 0.0f) N = -N;
    mat_out = mat_id;
}

// Shades queue entry i (path `path`, the job's local path `local`) of job J: its rays
// go to `out` (mask: which of them exist).
template <uint32_t NL, bool RECT>
__device__ __forceinline__ void shade_entry(const ShadeArgs& a, const JobRec& J, uint32_t i, uint32_t path,
                                            uint32_t local, bool primary, ShadeOut<NL>& out) {
    const size_t cap = a.ps.cap;
    uint32_t mask = 0;
    const PathKey key = path_key(J, local);
    const uint32_t pixel = key.pixel;
    const uint32_t sample = key.sample;
    const uint32_t nl = J.num_lights;
    if (a.nrc_query) a.nrc_query[i].tag = kNoPath;   // entries without a cache query
    const uint32_t nslots = nl + 2u + (RECT && J.num_rects ? 1u : 0u);
    V3 o, d, P = v3(0.0f, 0.0f, 0.0f);
    V3 beta;
    float pdf_prev;
    if (primary) {
        float jx = 0.5f, jy = 0.5f;
        if (J.flags & 1u) {
            Pcg jr = pcg_keyed(J.seed, pixel, sample, kJitterBounce);
            jx = pcg_float(jr);
            jy = pcg_float(jr);
        }
        camera_ray(J.cam, pixel % J.cam.width, pixel / J.cam.width, jx, jy, &o, &d);
        beta = v3(1.0f, 1.0f, 1.0f);
        pdf_prev = 0.0f;
        for (uint32_t k = 0; k < nslots; ++k) {
            float* acc = a.ps.acc + (size_t)k * 3u * cap;
            acc[path] = 0.0f;
            acc[cap + path] = 0.0f;
            acc[2 * cap + path] = 0.0f;
        }
    } else {
        o = v3(a.qin.ox[i], a.qin.oy[i], a.qin.oz[i]);
        d = v3(a.qin.dx[i], a.qin.dy[i], a.qin.dz[i]);
        beta = v3(a.ps.bx[path], a.ps.by[path], a.ps.bz[path]);
        pdf_prev = a.ps.pdf[path];
    }
    const HitRec hit = a.hits.rec[i];
    const uint32_t leaf = hit.leaf;
    const bool want_depth = primary && J.depth_out != nullptr && sample == J.depth_sample;
    if (leaf == kLeafMiss) {
        V3 ls = sky_eval(J.sky, d);
        float w = 1.0f;
        if (!primary && J.sky_nee) w = power_heuristic(pdf_prev, sky_pdf(J.sky, d));
        V3 c = (beta * ls) * w;
        a.ps.acc[path] = c.x;
        a.ps.acc[cap + path] = c.y;
        a.ps.acc[2 * cap + path] = c.z;
        if (want_depth) J.depth_out[pixel] = 0.0f;
    } else {
        const float t = hit.t, hu = hit.u, hv = hit.v;
        P = o + d * t;
        Mat m;
        V3 N;
        float rough;
        uint32_t mat_id;
        hit_surface(J, leaf, hu, hv, P, d, m, N, rough, mat_id);
        V3 V = normalize(o - P);
        if (a.bounce == J.cache_bounce && (J.flags & (kJobCache | kJobTrain))) {
            NrcRecord r;
            r.p[0] = P.x; r.p[1] = P.y; r.p[2] = P.z;
            r.wo[0] = V.x; r.wo[1] = V.y; r.wo[2] = V.z;
            r.n[0] = N.x; r.n[1] = N.y; r.n[2] = N.z;
            r.roughness = rough;
            r.albedo[0] = m.base.x; r.albedo[1] = m.base.y; r.albedo[2] = m.base.z;
            r.pad[0] = 0.0f; r.pad[1] = 0.0f;
            if (J.flags & kJobCache) {
                // the path ends here: the cache's estimate of the radiance leaving P
                // towards V (added by the cache kernel, times the throughput) stands
                // for this vertex's lighting and everything after it
                r.tag = path;
                a.nrc_query[i] = r;
                out.mask = 0u;
                return;
            }
            // training path: the radiance gathered from here on, per unit throughput,
            // is the target of this sample
            r.tag = 1u;
            J.cache_rec[key.q] = r;
            beta = v3(1.0f, 1.0f, 1.0f);
            for (uint32_t k = 0; k < nslots; ++k) {
                float* acc = a.ps.acc + (size_t)k * 3u * cap;
                acc[path] = 0.0f;
                acc[cap + path] = 0.0f;
                acc[2 * cap + path] = 0.0f;
            }
        }
        Frame fr = make_frame(N);
        Pcg rng = pcg_keyed(J.seed, pixel, sample, a.bounce);
        float u0 = pcg_float(rng), u1 = pcg_float(rng), u2 = pcg_float(rng), u3 = pcg_float(rng);
#pragma unroll
        for (uint32_t l = 0; l < NL; ++l) {
            if (l >= nl) break;
            const LightRec& L = J.lights[l];
            V3 ldir;
            const V3 le = light_eval(L, P, &ldir);
            V3 f = bsdf_eval(fr, m, V, ldir);
            V3 c = (le * f) * beta;
            if (!any_pos(c)) continue;
            if (L.sm_index < 0) {
                float* acc = a.ps.acc + (size_t)(l + 1u) * 3u * cap;
                acc[path] = acc[path] + c.x;
                acc[cap + path] = acc[cap + path] + c.y;
                acc[2 * cap + path] = acc[2 * cap + path] + c.z;
                continue;
            }
            out.c[l] = c;
            mask |= 2u << l;
        }
        // rect lights (deferred_lighting_secondary.hlsl:73-96): no shadow rays, each light's
        // lighting times the throughput added in light order to the slot after the sky's
        if constexpr (RECT) {
            if (J.num_rects) {
                float* acc = a.ps.acc + (size_t)(nl + 2u) * 3u * cap;
                V3 sum = v3(acc[path], acc[cap + path], acc[2 * cap + path]);
                for (uint32_t r = 0; r < J.num_rects; ++r)
                    sum = sum + rect_light_lit(J.rects[r], J.ltc, J.rect_tex, J.rect_tex_levels, P, fr, V, m) * beta;
                acc[path] = sum.x;
                acc[cap + path] = sum.y;
                acc[2 * cap + path] = sum.z;
            }
        }
        // sky NEE only where the next segment is sampled too (the reference sees the
        // sky only through a traced segment that escapes)
        const bool extend = a.bounce < J.last_bounce && !(max3(beta) < kThroughputMin);
        if (J.sky_nee && extend) {
            V3 wi, ls;
            float pl;
            if (sky_sample(J.sky, u0, u1, &wi, &pl, &ls)) {
                V3 f = bsdf_eval(fr, m, V, wi);
                if (any_pos(f)) {
                    V3 hl = to_local(fr, normalize(V + wi));
                    float pb = vndf_pdf_l(hl, to_local(fr, V), m.rx, m.ry);
                    float w = power_heuristic(pl, pb);
                    V3 c = ((ls * f) * (w / pl)) * beta;
                    if (any_pos(c) && finite3(c)) {
                        out.c[NL] = c;
                        out.sky = wi;
                        mask |= 2u << kMaxDirLights;
                    }
                }
            }
        }
        if (extend) {
            V3 vl = to_local(fr, V);
            V3 hh = vndf_sample(vl, m.rx, m.ry, u2, u3);
            float hvd = dot(vl, hh);
            V3 ll = hh * (2.0f * hvd) - vl;
            float pdf = vndf_pdf_l(hh, vl, m.rx, m.ry);
            V3 Lext = to_world(fr, ll);
            V3 f = bsdf_eval(fr, m, V, Lext);
            V3 w = v3(f.x / pdf, f.y / pdf, f.z / pdf);
            if (!finite3(w)) w = v3(0.0f, 0.0f, 0.0f);
            V3 nb = beta * w;
            if (max3(nb) > 0.0f) {
                out.ext = Lext;
                a.ps.bx[path] = nb.x; a.ps.by[path] = nb.y; a.ps.bz[path] = nb.z;
                a.ps.pdf[path] = pdf;
                mask |= kMaskExt;
            }
        }
        if (want_depth) {
            const float* pv = J.cam.proj_view;
            float z = ((pv[8] * P.x + pv[9] * P.y) + pv[10] * P.z) + pv[11];
            float w = ((pv[12] * P.x + pv[13] * P.y) + pv[14] * P.z) + pv[15];
            J.depth_out[pixel] = z / w;
        }
    }
    out.mask = mask;
    out.P = P;
    out.path = path;
}

// Writes an entry's rays: the extension ray at queue place e, its shadow rays from
// place o on in kind order (lights 0..3, then the sky).
template <uint32_t NL>
__device__ __forceinline__ void emit_entry(const ShadeArgs& a, const ShadeOut<NL>& out, uint32_t e, uint32_t o) {
    const uint32_t mask = out.mask;
    if (mask & kMaskExt) {
        a.qout.ox[e] = out.P.x; a.qout.oy[e] = out.P.y; a.qout.oz[e] = out.P.z;
        a.qout.dx[e] = out.ext.x; a.qout.dy[e] = out.ext.y; a.qout.dz[e] = out.ext.z;
        a.qout.path[e] = out.path;
    }
#pragma unroll
    for (uint32_t k = 0; k <= NL; ++k) {
        const bool sky = k == NL;
        if (!(mask & (2u << (sky ? (uint32_t)kMaxDirLights : k)))) continue;
        a.sq.ox[o] = out.P.x; a.sq.oy[o] = out.P.y; a.sq.oz[o] = out.P.z;
        if (sky) {   // a light's ray takes its direction from the light in the shadow kernel
            a.sq.dx[o] = out.sky.x; a.sq.dy[o] = out.sky.y; a.sq.dz[o] = out.sky.z;
        }
        a.sq.cx[o] = out.c[k].x; a.sq.cy[o] = out.c[k].y; a.sq.cz[o] = out.c[k].z;
        a.sq.pk[o] = (out.path << 3) | (sky ? kKindSky : k);
        ++o;
    }
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Offset of tile t (> 0) in the queues: the sum of the words of tiles t-1, t-2, ...
// down to the nearest published inclusive prefix, 64 tiles per step (lane j reads
// tile p - j; tiles before 0 count as an empty prefix).  Runs in one wavefront.
__device__ __forceinline__ uint64_t look_back(const uint64_t* state, uint32_t t, uint32_t lane, uint32_t* error) {
    uint64_t sum = 0;
    int p = (int)t - 1;
    for (;;) {
        const int q = p - (int)lane;
        uint64_t w = q >= 0 ? __hip_atomic_load(state + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kLbPrefix;
        uint32_t spins = 0;
        while (__ballot((w >> 62) == 0ull) != 0ull) {
            if (++spins > kLbSpinMax) {   // a tile never published: report, count it as empty
                if (lane == 0) atomicOr(error, 32u);
                if ((w >> 62) == 0ull) w = kLbPrefix;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            if ((w >> 62) == 0ull) w = __hip_atomic_load(state + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const uint64_t pm = __ballot((w >> 62) == 2ull);
        const uint32_t j = pm ? (uint32_t)__builtin_ctzll(pm) : 64u;
        sum += wave_sum64(lane <= j ? (w & kLbFields) : 0ull);
        if (j < 64u) return sum;
        p -= 64;
    }
}

template <uint32_t NL, bool RECT>
__device__ __forceinline__ void shade_tile(const ShadeArgs& a) {
    __shared__ uint32_t s_wsum[4];
    __shared__ uint64_t s_base;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const bool primary = a.bounce == 0;
    const uint32_t count = primary ? a.jobs.num_paths : *a.count_in;
    if (blockIdx.x == 0 && tid < kWorkPartsMax) {
        // the closest-hit launch of this segment has finished and this segment's shadow
        // launch has not started: reset their persistent-lane part counters
        a.trace_work_reset[tid * kWorkStride] = 0u;
        a.shadow_work_reset[tid * kWorkStride] = 0u;
    }
    const uint32_t ntiles = (count + kTile - 1u) / kTile;
    const uint32_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    const uint32_t i = tile * kTile + tid;
    ShadeOut<NL> out;
    const JobRec* J = a.jobs.rec;
    if (i < count) {
        const uint32_t path = primary ? i : a.qin.path[i];
        if (a.jobs.num_jobs == 1u) {
            shade_entry<NL, RECT>(a, a.jobs.rec[0], i, path, path, primary, out);   // uniform record: scalar loads
        } else {
            const uint32_t jid = job_of(a.jobs.path_base, a.jobs.num_jobs, path);
            J = a.jobs.rec + jid;
            shade_entry<NL, RECT>(a, *J, i, path, path - a.jobs.path_base[jid], primary, out);
        }
    }
    // the tile's emissions: (extension << 16 | shadow) per entry, scanned in entry order
    const uint32_t packed = ((out.mask & kMaskExt) << 16) | (uint32_t)__popc(out.mask >> 1);
    uint32_t inc = packed;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if (lane >= (uint32_t)off) inc += y;
    }
    if (lane == 63) s_wsum[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) wbase += w < wave ? s_wsum[w] : 0u;
    const uint32_t excl = wbase + inc - packed;
    if (wave == 0) {
        const uint32_t agg = (s_wsum[0] + s_wsum[1]) + (s_wsum[2] + s_wsum[3]);
        const uint64_t own = lb_word(agg >> 16, agg & 0xffffu);
        uint64_t base = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(a.tile_state, kLbPrefix | own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(a.tile_state + tile, kLbTotal | own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            base = look_back(a.tile_state, tile, lane, a.error);
            if (lane == 0)
                __hip_atomic_store(a.tile_state + tile, kLbPrefix | (base + own), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_base = base;
            if (tile == ntiles - 1u) {   // the last tile: totals of the segment
                const uint64_t tot = base + own;
                *a.count_ext = lb_ext(tot);
                *a.count_shadow = lb_sh(tot);
                // atomic: a shadow launch of the previous segment may be moving its
                // sky rays between these counters meanwhile
                atomicAdd(reinterpret_cast<unsigned long long*>(a.ray_counters + 1), (unsigned long long)lb_ext(tot));
                atomicAdd(reinterpret_cast<unsigned long long*>(a.ray_counters + 2), (unsigned long long)lb_sh(tot));
            }
        }
    }
    __syncthreads();
    if (out.mask) {
        const uint64_t base = s_base;
        emit_entry(a, out, lb_ext(base) + (excl >> 16), lb_sh(base) + (excl & 0xffffu));
    }
}

// One kernel per light bound; the one-light form (the common case) fits 80 registers,
// 6 wavefronts per SIMD, without spilling.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_shade_l1(ShadeArgs a) {
    shade_tile<1, false>(a);
}
__global__ __launch_bounds__(256) void k_shade_l2(ShadeArgs a) { shade_tile<2, false>(a); }
__global__ __launch_bounds__(256) void k_shade_l4(ShadeArgs a) { shade_tile<kMaxDirLights, false>(a); }
// With rect lights (launches where some job has them): the LTC integrals inline.
__global__ __launch_bounds__(256) void k_shade_rect_l1(ShadeArgs a) { shade_tile<1, true>(a); }
__global__ __launch_bounds__(256) void k_shade_rect_l2(ShadeArgs a) { shade_tile<2, true>(a); }
__global__ __launch_bounds__(256) void k_shade_rect_l4(ShadeArgs a) { shade_tile<kMaxDirLights, true>(a); }

// Output counts and look-back words of the next shade launch.
__global__ __launch_bounds__(256) void k_shade_reset(uint64_t* state, const uint32_t* count_in, uint32_t num_entries,
                                                     uint32_t* count_ext, uint32_t* count_shadow) {
    const uint32_t count = count_in ? *count_in : num_entries;
    const uint32_t ntiles = (count + kTile - 1u) / kTile;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < ntiles) state[t] = 0ull;
    if (t == 0) {
        *count_ext = 0u;
        *count_shadow = 0u;
    }
}

// Per pixel slot (job j's slot q in path order): running sum += radiance of each of the
// job's samples in this wave, in sample order; the radiance of a path is the sum of its
// kinds in kind order (miss, lights, sky).
__global__ __launch_bounds__(256) void k_resolve(ResolveArgs a) {
    const uint32_t slot = blockIdx.x * 256u + threadIdx.x;
    if (slot >= a.num_pix) return;
    const uint32_t jid = job_of(a.jobs.pix_base, a.jobs.num_jobs, slot);
    const JobRec& J = a.jobs.rec[jid];
    const uint32_t q = slot - J.pix_base;
    const uint32_t npix = J.npix;
    // a training job's slot q sums into target q
    const uint32_t pixel = (J.flags & kJobTrain) ? q : pixel_of(q, J.cam.width, J.cam.height);
    const uint32_t nslots = J.num_lights + 2u + (J.num_rects ? 1u : 0u);
    const uint32_t samples = J.samples;
    const uint32_t pbase = a.jobs.path_base[jid];
    float* sums = J.sums;
    const size_t cap = a.ps.cap;
    V3 sum = v3(sums[3 * (size_t)pixel], sums[3 * (size_t)pixel + 1], sums[3 * (size_t)pixel + 2]);
    for (uint32_t s = 0; s < samples; ++s) {
        const size_t p = (size_t)pbase + (size_t)s * npix + q;
        V3 r = v3(0.f, 0.f, 0.f);
        for (uint32_t k = 0; k < nslots; ++k) {
            const float* acc = a.ps.acc + (size_t)k * 3u * cap;
            V3 c = v3(acc[p], acc[cap + p], acc[2 * cap + p]);
            r = k == 0 ? c : r + c;
        }
        sum = sum + r;
    }
    sums[3 * (size_t)pixel] = sum.x;
    sums[3 * (size_t)pixel + 1] = sum.y;
    sums[3 * (size_t)pixel + 2] = sum.z;
}

__global__ __launch_bounds__(256) void k_output(WaveJobs jobs, uint32_t num_pix) {
    const uint32_t slot = blockIdx.x * 256u + threadIdx.x;
    if (slot >= num_pix) return;
    const uint32_t jid = job_of(jobs.pix_base, jobs.num_jobs, slot);
    const JobRec& J = jobs.rec[jid];
    if (!J.out_rgba) return;   // training targets have no image
    const uint32_t p = slot - J.pix_base;
    const float den = J.out_den;
    const float* sums = J.sums;
    float4 o;
    o.x = sums[3 * (size_t)p] / den;
    o.y = sums[3 * (size_t)p + 1] / den;
    o.z = sums[3 * (size_t)p + 2] / den;
    o.w = 1.0f;
    if (J.flags & kJobHalfOut) {   // RGBA16F, round to nearest even (v_cvt_f16_f32)
        const _Float16 h[4] = {(_Float16)o.x, (_Float16)o.y, (_Float16)o.z, (_Float16)o.w};
        ushort4 u;
        __builtin_memcpy(&u, h, sizeof u);
        reinterpret_cast<ushort4*>(J.out_rgba)[p] = u;
    } else {
        reinterpret_cast<float4*>(J.out_rgba)[p] = o;
    }
}

// Primary-hit GBuffer (hits/rt_gbuffer_hit.hlsl:7-17, gbuffer.hlsl:18-34) of job J's
// first sample in this call: thread q reads the segment-0 hit record of the job's path q
// (the sample's pixel q in path order), rebuilds the camera ray and the surface as the
// shading does, and stores the four packed texels (zeros on a miss).
__global__ __launch_bounds__(256) void k_gbuffer(const JobRec* jrec, const HitRec* hits, uint32_t path0) {
    const JobRec& J = *jrec;
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= J.npix) return;
    const PathKey key = path_key(J, q);
    const uint32_t pixel = key.pixel;
    float jx = 0.5f, jy = 0.5f;
    if (J.flags & 1u) {
        Pcg jr = pcg_keyed(J.seed, pixel, key.sample, kJitterBounce);
        jx = pcg_float(jr);
        jy = pcg_float(jr);
    }
    V3 o, d;
    camera_ray(J.cam, pixel % J.cam.width, pixel / J.cam.width, jx, jy, &o, &d);
    const HitRec hit = hits[path0 + q];
    ushort4 base = make_ushort4(0, 0, 0, 0), nr = make_ushort4(0, 0, 0, 0), fr = make_ushort4(0, 0, 0, 0);
    uchar4 m0 = make_uchar4(0, 0, 0, 0);
    if (hit.leaf != kLeafMiss) {
        const V3 P = o + d * hit.t;
        Mat m;
        V3 N;
        float rough;
        uint32_t mat_id;
        hit_surface(J, hit.leaf, hit.u, hit.v, P, d, m, N, rough, mat_id);
        V3 T = v3(1.0f, 0.0f, 0.0f);
        if (J.leaf_tan) {
            const float* g = J.leaf_tan[hit.leaf].t;
            T = normalize(v3((g[0] + (g[4] - g[0]) * hit.u) + (g[8] - g[0]) * hit.v,
                             (g[1] + (g[5] - g[1]) * hit.u) + (g[9] - g[1]) * hit.v,
                             (g[2] + (g[6] - g[2]) * hit.u) + (g[10] - g[2]) * hit.v));
        }
        float pk[3];
        pack_normal_tangent(N, T, pk);
        auto h = [](float x) {
            const _Float16 v = (_Float16)x;   // round to nearest even (v_cvt_f16_f32)
            unsigned short u;
            __builtin_memcpy(&u, &v, sizeof u);
            return u;
        };
        base = make_ushort4(h(m.base.x), h(m.base.y), h(m.base.z), h(1.0f));
        nr = make_ushort4(h(pk[0]), h(pk[1]), h(pk[2]), h(rough));
        const uint32_t f0 = pack_rg11b10(m.f0), f90 = pack_rg11b10(m.f90);
        fr = make_ushort4((unsigned short)unorm16_of_half_word(f0 & 0xffffu), (unsigned short)unorm16_of_half_word(f0 >> 16),
                          (unsigned short)unorm16_of_half_word(f90 & 0xffffu),
                          (unsigned short)unorm16_of_half_word(f90 >> 16));
        const MaterialRec& M = J.materials[mat_id];
        m0 = make_uchar4((unsigned char)unorm8_of(M.anisotropy), (unsigned char)unorm8_of(1.0f / M.ior), 0,
                         (unsigned char)unorm8_of(kSurfaceModelLit / 256.0f));
    }
    if (J.gb_base) reinterpret_cast<ushort4*>(J.gb_base)[pixel] = base;
    if (J.gb_nr) reinterpret_cast<ushort4*>(J.gb_nr)[pixel] = nr;
    if (J.gb_fres) reinterpret_cast<ushort4*>(J.gb_fres)[pixel] = fr;
    if (J.gb_m0) reinterpret_cast<uchar4*>(J.gb_m0)[pixel] = m0;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void launch_shade_reset(uint64_t* tile_state, const uint32_t* count_in, uint32_t num_entries, uint32_t* count_ext,
                        uint32_t* count_shadow, uint32_t capacity, hipStream_t s) {
    const uint32_t tiles = (capacity + kTile - 1u) / kTile;
    hipLaunchKernelGGL(k_shade_reset, dim3((tiles + 255u) / 256u + 1u), dim3(256), 0, s, tile_state, count_in,
                       num_entries, count_ext, count_shadow);
}

void launch_shade(const ShadeArgs& a, uint32_t capacity, uint32_t max_lights, bool rect, hipStream_t s) {
    const uint32_t tiles = (capacity + kTile - 1) / kTile;
    const dim3 grid(tiles ? tiles : 1), block(256);
    if (rect) {
        if (max_lights <= 1) hipLaunchKernelGGL(k_shade_rect_l1, grid, block, 0, s, a);
        else if (max_lights <= 2) hipLaunchKernelGGL(k_shade_rect_l2, grid, block, 0, s, a);
        else hipLaunchKernelGGL(k_shade_rect_l4, grid, block, 0, s, a);
    } else if (max_lights <= 1) {
        hipLaunchKernelGGL(k_shade_l1, grid, block, 0, s, a);
    } else if (max_lights <= 2) {
        hipLaunchKernelGGL(k_shade_l2, grid, block, 0, s, a);
    } else {
        hipLaunchKernelGGL(k_shade_l4, grid, block, 0, s, a);
    }
}

void launch_resolve(const ResolveArgs& a, hipStream_t s) {
    if (a.num_pix == 0) return;
    hipLaunchKernelGGL(k_resolve, dim3((a.num_pix + 255u) / 256u), dim3(256), 0, s, a);
}

void launch_output(const WaveJobs& jobs, uint32_t num_pix, hipStream_t s) {
    if (num_pix == 0) return;
    hipLaunchKernelGGL(k_output, dim3((num_pix + 255u) / 256u), dim3(256), 0, s, jobs, num_pix);
}

void launch_gbuffer(const JobRec* job, const HitRec* hits, uint32_t path0, uint32_t npix, hipStream_t s) {
    if (npix == 0) return;
    hipLaunchKernelGGL(k_gbuffer, dim3((npix + 255u) / 256u), dim3(256), 0, s, job, hits, path0);
}

void launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255u) / 256u), dim3(256), 0, s, p, value, n);
}

}  // namespace bipt
