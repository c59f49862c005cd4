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
// next queues in three steps:
//   k_shade        one lane per queue entry; outputs go to fixed staging slots (entry
//                  i, shadow kind k at k*cap + i), an emission mask per entry and the
//                  (extension, shadow) totals of each 256-entry tile
//   k_scan_tiles   exclusive scan of the tile totals (one workgroup)
//   k_compact      wavefront prefix of the masks inside each tile plus the tile's
//                  offset: copies survivors, in entry order, into the compact queues
// so queue order equals input order (deterministic contents), and no lane holds its
// outputs while waiting for other tiles.
#include "pt_kernels.hpp"

namespace bipt {
namespace {

constexpr uint32_t kTile = 256;
constexpr uint32_t kMaskExt = 1u;   // bit 0: extension ray; bit 1 + k: shadow slot k (lights 0..3, sky 4)

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Shades queue entry i (path `path`, the job's local path `local`) of job J: writes the
// entry's staging slots and returns its emission mask.
__device__ __forceinline__ uint32_t shade_entry(const ShadeArgs& a, const JobRec& J, uint32_t i, uint32_t path,
                                                uint32_t local, bool primary) {
    const size_t cap = a.ps.cap;
    uint32_t mask = 0;
    const PathKey key = path_key(J, local);
    const uint32_t pixel = key.pixel;
    const uint32_t sample = key.sample;
    const uint32_t nl = J.num_lights;
    if (a.nrc_query) a.nrc_query[i].tag = kNoPath;   // entries without a cache query
    const uint32_t nslots = nl + 2u;
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
        const float4* lsp = reinterpret_cast<const float4*>(J.leaf_shade + leaf);
        const float4 s0 = lsp[0], s1 = lsp[1], s2 = lsp[2];
        V3 n0 = v3(s0.x, s0.y, s0.z), n1 = v3(s0.w, s1.x, s1.y), n2 = v3(s1.z, s1.w, s2.x);
        const uint32_t mat_id = __float_as_uint(s2.y);
        V3 N = normalize((n0 + (n1 - n0) * hu) + (n2 - n0) * hv);
        Mat m = load_mat(J.materials[mat_id]);
        if (m.two_sided && dot(d, N) > 0.0f) N = -N;
        V3 V = normalize(o - P);
        if (a.bounce == J.cache_bounce && (J.flags & (kJobCache | kJobTrain))) {
            NrcRecord r;
            r.p[0] = P.x; r.p[1] = P.y; r.p[2] = P.z;
            r.wo[0] = V.x; r.wo[1] = V.y; r.wo[2] = V.z;
            r.n[0] = N.x; r.n[1] = N.y; r.n[2] = N.z;
            r.roughness = J.materials[mat_id].roughness;
            r.albedo[0] = m.base.x; r.albedo[1] = m.base.y; r.albedo[2] = m.base.z;
            r.pad[0] = 0.0f; r.pad[1] = 0.0f;
            if (J.flags & kJobCache) {
                // the path ends here: the cache's estimate of the radiance leaving P
                // towards V (added by the cache kernel, times the throughput) stands
                // for this vertex's lighting and everything after it
                r.tag = path;
                a.nrc_query[i] = r;
                return 0u;
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
        for (uint32_t l = 0; l < nl; ++l) {
            const LightRec& L = J.lights[l];
            V3 le = v3(L.emission[0], L.emission[1], L.emission[2]);
            V3 ldir = v3(L.direction[0], L.direction[1], L.direction[2]);
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
            const size_t s = (size_t)l * cap + i;
            a.st.cx[s] = c.x; a.st.cy[s] = c.y; a.st.cz[s] = c.z;
            mask |= 2u << l;
        }
        // sky NEE only where the next segment is sampled too (the reference sees the
        // sky only through a traced segment that escapes)
        const bool extend = a.bounce < J.last_bounce && !(max3(beta) < kThroughputMin);
        if (J.sky_nee && extend) {
            V3 wi;
            float pl;
            if (sky_sample(J.sky, u0, u1, &wi, &pl)) {
                V3 f = bsdf_eval(fr, m, V, wi);
                if (any_pos(f)) {
                    V3 hl = to_local(fr, normalize(V + wi));
                    float pb = vndf_pdf_l(hl, to_local(fr, V), m.rx, m.ry);
                    float w = power_heuristic(pl, pb);
                    V3 c = ((sky_eval(J.sky, wi) * f) * (w / pl)) * beta;
                    if (any_pos(c) && finite3(c)) {
                        const size_t s = (size_t)nl * cap + i;
                        a.st.cx[s] = c.x; a.st.cy[s] = c.y; a.st.cz[s] = c.z;
                        a.st.sx[i] = wi.x; a.st.sy[i] = wi.y; a.st.sz[i] = wi.z;
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
                a.st.ex[i] = Lext.x; a.st.ey[i] = Lext.y; a.st.ez[i] = Lext.z;
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
    if (mask) {
        a.st.px[i] = P.x; a.st.py[i] = P.y; a.st.pz[i] = P.z;
        a.st.path[i] = path;
    }
    return mask;
}

__global__ __launch_bounds__(256) void k_shade(ShadeArgs a) {
    __shared__ uint32_t s_wsum[4], s_wsky[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const bool primary = a.bounce == 0;
    const uint32_t count = primary ? a.jobs.num_paths : *a.count_in;
    if (blockIdx.x == 0 && tid < kWorkPartsMax) {
        // the closest-hit launch of this segment has finished and this segment's shadow
        // launch has not started: reset their persistent-lane part counters
        a.trace_work_reset[tid * kWorkStride] = 0u;
        a.shadow_work_reset[tid * kWorkStride] = 0u;
    }
    const uint32_t tile = blockIdx.x;
    if (tile * kTile >= count) return;
    const uint32_t i = tile * kTile + tid;
    uint32_t mask = 0;
    if (i < count) {
        const uint32_t path = primary ? i : a.qin.path[i];
        if (a.jobs.num_jobs == 1u) {
            mask = shade_entry(a, a.jobs.rec[0], i, path, path, primary);   // uniform record: scalar loads
        } else {
            const uint32_t jid = job_of(a.jobs.path_base, a.jobs.num_jobs, path);
            mask = shade_entry(a, a.jobs.rec[jid], i, path, path - a.jobs.path_base[jid], primary);
        }
        a.st.mask[i] = mask;
    }
    // tile totals (extension, shadow) and the tile's sky shadow rays (ray statistics)
    const uint32_t packed = ((mask & kMaskExt) << 16) | (uint32_t)__popc(mask >> 1);
    const uint32_t sum = wave_sum(packed);
    const uint32_t sky = wave_sum((mask >> (1 + kMaxDirLights)) & 1u);
    if (lane == 0) {
        s_wsum[wave] = sum;
        s_wsky[wave] = sky;
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t agg = (s_wsum[0] + s_wsum[1]) + (s_wsum[2] + s_wsum[3]);
        a.st.tile_sums[tile] = ((uint64_t)(agg >> 16) << 32) | (agg & 0xffffu);
        a.st.tile_sky[tile] = (s_wsky[0] + s_wsky[1]) + (s_wsky[2] + s_wsky[3]);
    }
}

// Exclusive scan of the per-tile (extension << 32 | shadow) totals, one workgroup.
__global__ __launch_bounds__(1024) void k_scan_tiles(ScanArgs a) {
    __shared__ uint64_t wsum[16];
    __shared__ uint64_t carry_s;
    __shared__ uint32_t wsky[16];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t count = a.count_in ? *a.count_in : a.num_entries;
    const uint32_t ntiles = (count + kTile - 1) / kTile;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    uint32_t sky = 0;
    for (uint32_t t = threadIdx.x; t < ntiles; t += 1024u) sky += a.tile_sky[t];
    for (uint32_t start = 0; start < ntiles; start += 1024u) {
        const uint32_t t = start + threadIdx.x;
        const uint64_t v = t < ntiles ? a.tile_sums[t] : 0ull;
        uint64_t inc = v;
        for (int off = 1; off < 64; off <<= 1) {
            uint64_t y = __shfl_up(inc, off);
            if (lane >= (uint32_t)off) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint64_t wbase = 0, total = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            const uint64_t s = wsum[w];
            if (w < wave) wbase += s;
            total += s;
        }
        const uint64_t carry = carry_s;
        if (t < ntiles) a.tile_sums[t] = carry + wbase + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + total;
        __syncthreads();
    }
    sky = (uint32_t)wave_sum(sky);
    if (lane == 0) wsky[wave] = sky;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t nsky = 0;
        for (uint32_t w = 0; w < 16; ++w) nsky += wsky[w];
        const uint64_t tot = carry_s;
        const uint32_t ne = (uint32_t)(tot >> 32), ns = (uint32_t)(tot & 0xffffffffull);
        *a.count_ext = ne;
        *a.count_shadow = ns;
        // ray statistics: extension, light shadow and sky shadow rays of this segment
        a.ray_counters[1] += ne;
        a.ray_counters[2] += ns - nsky;
        a.ray_counters[3] += nsky;
    }
}

// Copies the staged outputs of one tile into the compact queues, in entry order.
__global__ __launch_bounds__(256) void k_compact(CompactArgs a) {
    __shared__ uint32_t s_wsum[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t count = a.count_in ? *a.count_in : a.num_entries;
    const uint32_t tile = blockIdx.x;
    if (tile * kTile >= count) return;
    const uint32_t i = tile * kTile + tid;
    const uint32_t mask = i < count ? a.st.mask[i] : 0u;
    const uint32_t packed = ((mask & kMaskExt) << 16) | (uint32_t)__popc(mask >> 1);
    uint32_t inc = packed;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(inc, off);
        if (lane >= (uint32_t)off) inc += y;
    }
    if (lane == 63) s_wsum[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0;
    for (uint32_t w = 0; w < wave; ++w) wbase += s_wsum[w];
    const uint32_t excl = wbase + inc - packed;
    const uint64_t base = a.st.tile_sums[tile];
    if (mask) {
        const float px = a.st.px[i], py = a.st.py[i], pz = a.st.pz[i];
        const uint32_t path = a.st.path[i];
        if (mask & kMaskExt) {
            const uint32_t e = (uint32_t)(base >> 32) + (excl >> 16);
            a.qout.ox[e] = px; a.qout.oy[e] = py; a.qout.oz[e] = pz;
            a.qout.dx[e] = a.st.ex[i]; a.qout.dy[e] = a.st.ey[i]; a.qout.dz[e] = a.st.ez[i];
            a.qout.path[e] = path;
        }
        // shadow rays in kind order: lights 0..3, then the sky
        const JobRec& J = a.jobs.rec[job_of(a.jobs.path_base, a.jobs.num_jobs, path)];
        const uint32_t nl = J.num_lights;
        uint32_t o = (uint32_t)(base & 0xffffffffull) + (excl & 0xffffu);
        for (uint32_t k = 0; k <= nl; ++k) {
            const bool sky = k == nl;
            if (!(mask & (2u << (sky ? (uint32_t)kMaxDirLights : k)))) continue;
            const size_t s = (size_t)k * a.cap + i;
            a.sq.ox[o] = px; a.sq.oy[o] = py; a.sq.oz[o] = pz;
            if (sky) {
                a.sq.dx[o] = a.st.sx[i]; a.sq.dy[o] = a.st.sy[i]; a.sq.dz[o] = a.st.sz[i];
            } else {
                const float* ld = J.lights[k].direction;
                a.sq.dx[o] = ld[0]; a.sq.dy[o] = ld[1]; a.sq.dz[o] = ld[2];
            }
            a.sq.cx[o] = a.st.cx[s]; a.sq.cy[o] = a.st.cy[s]; a.sq.cz[o] = a.st.cz[s];
            a.sq.pk[o] = (path << 3) | (sky ? kKindSky : k);
            ++o;
        }
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
    const uint32_t nslots = J.num_lights + 2u;
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
    reinterpret_cast<float4*>(J.out_rgba)[p] = o;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void launch_shade(const ShadeArgs& a, uint32_t capacity, hipStream_t s) {
    const uint32_t tiles = (capacity + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_shade, dim3(tiles ? tiles : 1), dim3(256), 0, s, a);
}

void launch_scan_tiles(const ScanArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, a);
}

void launch_compact(const CompactArgs& a, uint32_t capacity, hipStream_t s) {
    const uint32_t tiles = (capacity + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_compact, dim3(tiles ? tiles : 1), dim3(256), 0, s, a);
}

void launch_resolve(const ResolveArgs& a, hipStream_t s) {
    if (a.num_pix == 0) return;
    hipLaunchKernelGGL(k_resolve, dim3((a.num_pix + 255u) / 256u), dim3(256), 0, s, a);
}

void launch_output(const WaveJobs& jobs, uint32_t num_pix, hipStream_t s) {
    if (num_pix == 0) return;
    hipLaunchKernelGGL(k_output, dim3((num_pix + 255u) / 256u), dim3(256), 0, s, jobs, num_pix);
}

void launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255u) / 256u), dim3(256), 0, s, p, value, n);
}

}  // namespace bipt
