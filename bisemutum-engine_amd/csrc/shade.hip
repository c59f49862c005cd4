// shade.hip -- surface shading, next-event estimation and queue compaction (gfx950).
//
// One kernel per segment replaces the reference's per-bounce passes
// sample_secondary_ray_cs (sample_secondary_ray.hlsl:11-69) and
// deferred_lighting_secondary_cs (deferred_lighting_secondary.hlsl:11-111):
//   miss   -> sky radiance * throughput (MIS-weighted against sky NEE after bounce 0)
//   hit    -> interpolated normal, lit BSDF; one shadow ray per directional light
//             (exact visibility replaces the shadow map); one sky NEE shadow ray;
//             VNDF-sampled extension ray while max(throughput) >= 0.001
// Instead of full-screen passes with dead lanes, live rays are compacted into
// the next queue: a wavefront prefix scan inside each 256-lane tile, and a
// single-pass decoupled look-back (Merrill & Garland 2016) across tiles taken in
// ticket order, walked 64 predecessors at a time by one wavefront, so queue
// order equals input order (deterministic contents).
#include "pt_kernels.hpp"

namespace bipt {
namespace {

constexpr uint32_t kTile = 256;
constexpr uint64_t kStatusAgg = 1ull << 62;
constexpr uint64_t kStatusPrefix = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 31) - 1ull;
constexpr uint32_t kSpinLimit = 1u << 20;   // bounded wait: a stuck predecessor becomes an error, not a hang
constexpr uint32_t kShadeGridMax = 2048;

__device__ __forceinline__ uint64_t pack_desc(uint64_t status, uint32_t e, uint32_t s) {
    return status | ((uint64_t)e << 31) | (uint64_t)s;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Exclusive prefix (extension, shadow) of all tiles before `tile`, computed by one
// wavefront: lane l inspects descriptor tile-1-l of the current window.
__device__ __forceinline__ void look_back(uint64_t* desc, uint32_t tile, uint32_t lane, uint32_t* error,
                                          uint32_t* pe_out, uint32_t* ps_out) {
    uint32_t pe = 0, ps = 0;
    int base = (int)tile - 1;
    uint32_t spins = 0;
    while (base >= 0) {
        const int j = base - (int)lane;
        uint64_t d = 0;
        uint64_t st = 2;   // lanes before tile 0 read as an inclusive prefix of zero
        if (j >= 0) {
            // relaxed: the descriptor word is the only datum exchanged between tiles, so no
            // acquire (and no cache invalidation) is needed; device scope keeps it coherent
            d = __hip_atomic_load(&desc[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st = d >> 62;
        }
        if (__ballot(st == 0) != 0ull) {
            if (++spins > kSpinLimit) {
                if (lane == 0) atomicOr(error, 2u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        const uint64_t inclusive = __ballot(st == 2);
        uint32_t e = (uint32_t)((d >> 31) & kValMask), s = (uint32_t)(d & kValMask);
        if (inclusive != 0ull) {
            const uint32_t stop = (uint32_t)__builtin_ctzll(inclusive);   // nearest inclusive predecessor
            if (lane > stop) { e = 0; s = 0; }
            pe += wave_sum(e);
            ps += wave_sum(s);
            break;
        }
        pe += wave_sum(e);
        ps += wave_sum(s);
        base -= 64;
    }
    *pe_out = pe;
    *ps_out = ps;
}

__global__ __launch_bounds__(256) void k_shade(ShadeArgs a) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_excl[2];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const bool primary = a.bounce == 0;
    const uint32_t count = primary ? a.wave.num_paths : *a.count_in;
    const uint32_t ntiles = (count + kTile - 1) / kTile;
    const uint32_t nl = a.num_lights;
    const uint32_t nslots = nl + 2u;
    const size_t cap = a.ps.cap;
    if (blockIdx.x == 0 && tid == 0) {
        // the closest-hit launch of this segment has finished and this segment's shadow
        // launch has not started: reset their persistent-lane ray counters
        *a.trace_work_reset = 0u;
        *a.shadow_work_reset = 0u;
    }
    if (count == 0) {
        if (blockIdx.x == 0 && tid == 0) { *a.count_out = 0u; *a.count_shadow = 0u; }
        return;
    }
    for (;;) {
        if (tid == 0) s_tile = atomicAdd(a.ticket, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        __syncthreads();
        if (tile >= ntiles) break;
        const uint32_t i = tile * kTile + tid;
        const bool active = i < count;

        uint32_t n_ext = 0;
        uint32_t sh_mask = 0;   // bit l: shadow ray toward light l; bit kMaxDirLights: sky ray
        uint32_t path = 0;
        V3 P = v3(0.f, 0.f, 0.f), Lext = v3(0.f, 0.f, 0.f), beta_next = v3(0.f, 0.f, 0.f);
        float pdf_next = 0.0f;
        V3 c_slot[kMaxDirLights + 1];
        V3 wi_sky = v3(0.f, 0.f, 0.f);
        for (int k = 0; k <= kMaxDirLights; ++k) c_slot[k] = v3(0.f, 0.f, 0.f);
        if (active) {
            path = primary ? i : a.qin.path[i];
            const uint32_t q = path % a.wave.npix, sl = path / a.wave.npix;
            const uint32_t pixel = pixel_of(q, a.wave.cam.width, a.wave.cam.height);
            const uint32_t sample = a.wave.sample0 + sl;
            V3 o, d;
            V3 beta;
            float pdf_prev;
            if (primary) {
                float jx = 0.5f, jy = 0.5f;
                if (a.wave.flags & 1u) {
                    Pcg jr = pcg_keyed(a.wave.seed, pixel, sample, kJitterBounce);
                    jx = pcg_float(jr);
                    jy = pcg_float(jr);
                }
                camera_ray(a.wave.cam, pixel % a.wave.cam.width, pixel / a.wave.cam.width, jx, jy, &o, &d);
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
            const uint32_t leaf = a.hits.leaf[i];
            const bool want_depth = primary && a.depth_out != nullptr && sample == a.depth_sample;
            if (leaf == kLeafMiss) {
                V3 ls = sky_eval(a.sky, d);
                float w = 1.0f;
                if (!primary && a.sky_nee) w = power_heuristic(pdf_prev, sky_pdf(a.sky, d));
                V3 c = (beta * ls) * w;
                a.ps.acc[path] = c.x;
                a.ps.acc[cap + path] = c.y;
                a.ps.acc[2 * cap + path] = c.z;
                if (want_depth) a.depth_out[pixel] = 0.0f;
            } else {
                const float t = a.hits.t[i], hu = a.hits.u[i], hv = a.hits.v[i];
                P = o + d * t;
                const float4* lsp = reinterpret_cast<const float4*>(a.leaf_shade + leaf);
                const float4 s0 = lsp[0], s1 = lsp[1], s2 = lsp[2];
                V3 n0 = v3(s0.x, s0.y, s0.z), n1 = v3(s0.w, s1.x, s1.y), n2 = v3(s1.z, s1.w, s2.x);
                const uint32_t mat_id = __float_as_uint(s2.y);
                V3 N = normalize((n0 + (n1 - n0) * hu) + (n2 - n0) * hv);
                Mat m = load_mat(a.materials[mat_id]);
                if (m.two_sided && dot(d, N) > 0.0f) N = -N;
                V3 V = normalize(o - P);
                Frame fr = make_frame(N);
                Pcg rng = pcg_keyed(a.wave.seed, pixel, sample, a.bounce);
                float u0 = pcg_float(rng), u1 = pcg_float(rng), u2 = pcg_float(rng), u3 = pcg_float(rng);
#pragma unroll
                for (int l = 0; l < kMaxDirLights; ++l) {
                    if ((uint32_t)l >= nl) break;
                    const LightRec& L = a.lights[l];
                    V3 le = v3(L.emission[0], L.emission[1], L.emission[2]);
                    V3 ldir = v3(L.direction[0], L.direction[1], L.direction[2]);
                    V3 f = bsdf_eval(fr, m, V, ldir);
                    V3 c = (le * f) * beta;
                    if (!any_pos(c)) continue;
                    if (L.sm_index < 0) {
                        float* acc = a.ps.acc + (size_t)(l + 1) * 3u * cap;
                        acc[path] = acc[path] + c.x;
                        acc[cap + path] = acc[cap + path] + c.y;
                        acc[2 * cap + path] = acc[2 * cap + path] + c.z;
                        continue;
                    }
                    c_slot[l] = c;
                    sh_mask |= 1u << l;
                }
                // sky NEE only where the next segment is sampled too (the reference sees the
                // sky only through a traced segment that escapes)
                const bool extend = a.bounce < a.last_bounce && !(max3(beta) < kThroughputMin);
                if (a.sky_nee && extend) {
                    V3 wi;
                    float pl;
                    if (sky_sample(a.sky, u0, u1, &wi, &pl)) {
                        V3 f = bsdf_eval(fr, m, V, wi);
                        if (any_pos(f)) {
                            V3 hl = to_local(fr, normalize(V + wi));
                            float pb = vndf_pdf_l(hl, to_local(fr, V), m.rx, m.ry);
                            float w = power_heuristic(pl, pb);
                            V3 c = ((sky_eval(a.sky, wi) * f) * (w / pl)) * beta;
                            if (any_pos(c) && finite3(c)) {
                                c_slot[kMaxDirLights] = c;
                                wi_sky = wi;
                                sh_mask |= 1u << kMaxDirLights;
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
                    Lext = to_world(fr, ll);
                    V3 f = bsdf_eval(fr, m, V, Lext);
                    V3 w = v3(f.x / pdf, f.y / pdf, f.z / pdf);
                    if (!finite3(w)) w = v3(0.0f, 0.0f, 0.0f);
                    V3 nb = beta * w;
                    if (max3(nb) > 0.0f) {
                        n_ext = 1;
                        beta_next = nb;
                        pdf_next = pdf;
                    }
                }
                if (want_depth) {
                    const float* pv = a.wave.cam.proj_view;
                    float z = ((pv[8] * P.x + pv[9] * P.y) + pv[10] * P.z) + pv[11];
                    float w = ((pv[12] * P.x + pv[13] * P.y) + pv[14] * P.z) + pv[15];
                    a.depth_out[pixel] = z / w;
                }
            }
        }
        const uint32_t n_sh = __popc(sh_mask);

        // ---- tile scan of (extension, shadow) emissions: wavefront prefix, then LDS
        const uint32_t packed = (n_ext << 16) | n_sh;
        uint32_t inc = packed;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(inc, off);
            if (lane >= (uint32_t)off) inc += y;
        }
        if (lane == 63) s_wsum[wave] = inc;
        __syncthreads();
        uint32_t wbase = 0, agg = 0;
        for (uint32_t w = 0; w < 4; ++w) {
            uint32_t s = s_wsum[w];
            if (w < wave) wbase += s;
            agg += s;
        }
        const uint32_t excl_local = wbase + inc - packed;
        const uint32_t agg_e = agg >> 16, agg_s = agg & 0xffffu;

        // ---- decoupled look-back across tiles (wavefront 0)
        if (wave == 0) {
            if (lane == 0)
                __hip_atomic_store(&a.desc[tile], pack_desc(tile == 0 ? kStatusPrefix : kStatusAgg, agg_e, agg_s),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t pe = 0, ps = 0;
            if (tile > 0) look_back(a.desc, tile, lane, a.error, &pe, &ps);
            if (lane == 0) {
                if (tile > 0)
                    __hip_atomic_store(&a.desc[tile], pack_desc(kStatusPrefix, pe + agg_e, ps + agg_s),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_excl[0] = pe;
                s_excl[1] = ps;
                if (tile == ntiles - 1) {
                    *a.count_out = pe + agg_e;
                    *a.count_shadow = ps + agg_s;
                    atomicAdd((unsigned long long*)&a.ray_counters[1], (unsigned long long)(pe + agg_e));
                }
            }
        }
        __syncthreads();
        const uint32_t base_e = s_excl[0] + (excl_local >> 16);
        uint32_t o_sh = s_excl[1] + (excl_local & 0xffffu);
        if (n_ext) {
            a.qout.ox[base_e] = P.x; a.qout.oy[base_e] = P.y; a.qout.oz[base_e] = P.z;
            a.qout.dx[base_e] = Lext.x; a.qout.dy[base_e] = Lext.y; a.qout.dz[base_e] = Lext.z;
            a.qout.path[base_e] = path;
            a.ps.bx[path] = beta_next.x; a.ps.by[path] = beta_next.y; a.ps.bz[path] = beta_next.z;
            a.ps.pdf[path] = pdf_next;
        }
        // shadow rays in kind order: lights 0..3, then the sky
#pragma unroll
        for (int k = 0; k <= kMaxDirLights; ++k) {
            if (!(sh_mask & (1u << k))) continue;
            const bool sky = k == kMaxDirLights;
            const V3 dir = sky ? wi_sky : v3(a.lights[k < kMaxDirLights ? k : 0].direction[0],
                                            a.lights[k < kMaxDirLights ? k : 0].direction[1],
                                            a.lights[k < kMaxDirLights ? k : 0].direction[2]);
            a.sq.ox[o_sh] = P.x; a.sq.oy[o_sh] = P.y; a.sq.oz[o_sh] = P.z;
            a.sq.dx[o_sh] = dir.x; a.sq.dy[o_sh] = dir.y; a.sq.dz[o_sh] = dir.z;
            a.sq.cx[o_sh] = c_slot[k].x; a.sq.cy[o_sh] = c_slot[k].y; a.sq.cz[o_sh] = c_slot[k].z;
            a.sq.pk[o_sh] = (path << 3) | (sky ? kKindSky : (uint32_t)k);
            ++o_sh;
        }
        // ray statistics: one atomic per wavefront and kind
        const uint32_t cl = wave_sum(__popc(sh_mask & ((1u << kMaxDirLights) - 1u)));
        const uint32_t cs = wave_sum((sh_mask >> kMaxDirLights) & 1u);
        if (lane == 0) {
            if (cl) atomicAdd((unsigned long long*)&a.ray_counters[2], (unsigned long long)cl);
            if (cs) atomicAdd((unsigned long long*)&a.ray_counters[3], (unsigned long long)cs);
        }
    }
}

// Per pixel (path order q): running sum += radiance of each sample in sample order.
__global__ __launch_bounds__(256) void k_resolve(ResolveArgs a) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= a.npix) return;
    const uint32_t pixel = pixel_of(q, a.width, a.height);
    const size_t cap = a.ps.cap;
    V3 sum = v3(a.sums[3 * (size_t)pixel], a.sums[3 * (size_t)pixel + 1], a.sums[3 * (size_t)pixel + 2]);
    for (uint32_t s = 0; s < a.samples; ++s) {
        const size_t p = (size_t)s * a.npix + q;
        V3 r = v3(0.f, 0.f, 0.f);
        for (uint32_t k = 0; k < a.num_slots; ++k) {
            const float* acc = a.ps.acc + (size_t)k * 3u * cap;
            V3 c = v3(acc[p], acc[cap + p], acc[2 * cap + p]);
            r = k == 0 ? c : r + c;
        }
        sum = sum + r;
    }
    a.sums[3 * (size_t)pixel] = sum.x;
    a.sums[3 * (size_t)pixel + 1] = sum.y;
    a.sums[3 * (size_t)pixel + 2] = sum.z;
}

__global__ __launch_bounds__(256) void k_output(const float* __restrict__ sums, uint32_t npix, float den,
                                                float* __restrict__ out) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= npix) return;
    float4 o;
    o.x = sums[3 * (size_t)p] / den;
    o.y = sums[3 * (size_t)p + 1] / den;
    o.z = sums[3 * (size_t)p + 2] / den;
    o.w = 1.0f;
    reinterpret_cast<float4*>(out)[p] = o;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void launch_shade(const ShadeArgs& a, uint32_t capacity, hipStream_t s) {
    uint32_t tiles = (capacity + kTile - 1) / kTile;
    uint32_t grid = tiles < kShadeGridMax ? tiles : kShadeGridMax;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(256), 0, s, a);
}

void launch_resolve(const ResolveArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_resolve, dim3((a.npix + 255u) / 256u), dim3(256), 0, s, a);
}

void launch_output(const float* sums, uint32_t npix, float den, uint32_t, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_output, dim3((npix + 255u) / 256u), dim3(256), 0, s, sums, npix, den, out);
}

void launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255u) / 256u), dim3(256), 0, s, p, value, n);
}

}  // namespace bipt
