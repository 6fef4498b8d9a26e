// Photon mapping on the device.
//
// Particle tracing restates photon_tracer::trace_particle (photon_tracer.h:
// 23-96) with emitter rays chosen by luminance(power) (scene.h:113-142,
// area_emitter.h:93-122); one lane per particle, photons appended with a wave
// ballot and later ordered by (particle, depth) with a hand-written LSD radix
// sort, so the stored set does not depend on scheduling.  The photon map is
// the reference's median kd tree (photon_map.h:83-114, built on the host with
// the same partitions), with photons stored in tree order; the k-nearest
// search (photon_map.h:117-165) walks it with a per-lane stack and a max-heap
// of (distance, tree index) pairs in LDS, and the irradiance sum follows
// estimate_irradiance (photon_map.h:38-70), including its heap pop order.
#pragma once

#include "bsdf.h"
#include "emit.h"

namespace lrt {

struct PhotonBuf {  // SoA, capacity entries
    float *px, *py, *pz, *dx, *dy, *dz, *nx, *ny, *nz, *wr, *wg, *wb;
    uint32_t* key;     // particle << 8 | depth
    uint32_t* count;   // [0]: photons appended (may exceed capacity)
    uint32_t capacity;
};

struct EmitCdf {
    int32_t n;
    int32_t index[kMaxLights];
    float cdf[kMaxLights];
    float total;
};

// scene::sample_emitter_ray + area_emitter::sample_ray (area_emitter.h:93-122)
LRT_D bool emitter_ray(const DevScene& S, const DevLights& Ls, const EmitCdf& E, float sx, float sy, float dx, float dy,
                       Ray& r, f3& weight) {
    if (E.total <= 0.0f || E.n <= 0) return false;
    const float target = sx * E.total;
    int bin = 0;
    while (bin < E.n && !(target < E.cdf[bin])) ++bin;  // upper_bound
    if (bin >= E.n) bin = E.n - 1;
    const float prev = bin == 0 ? 0.0f : E.cdf[bin - 1];
    const float span = E.cdf[bin] - prev;
    if (span <= 0.0f) return false;
    const float sel = span / E.total;
    const float u = (target - prev) / span;
    const DevLight& L = Ls.l[E.index[bin]];
    const DevObject& ob = S.obj[L.object];
    if (ob.area <= 0.0f) return false;
    f3 p, n;
    float pdf_area;
    if (!shape_sample(S, Ls, L, u, sy, p, n, pdf_area) || !(pdf_area > 0.0f)) return false;
    const float ct = sqrtf(dx), st = sqrtf(1.0f - dx);
    const float phi = 2.0f * kPiF * dy;
    float sp, cp;
    sincos_rn(phi, &sp, &cp);
    const Frame f = frame_from(n);
    r = Ray{p, to_world(f, f3{cp * st, sp * st, ct})};
    const f3 power = ld3(L.radiance) * (kPiF * ob.area);
    weight = f3{power.x / sel, power.y / sel, power.z / sel};
    return true;
}

struct PhotonTraceParams {
    uint32_t first, count, max_depth, rr_depth, kind;
    uint2 key;
};

LRT_D float particle_u(uint32_t index, uint32_t tag, uint32_t d, uint2 key) {
    const uint4 v = philox(make_uint4(index, tag, d >> 2, 0u), key);
    return u24(pick(v, d & 3u));
}

__global__ __launch_bounds__(256) void k_photon_trace(const DevScene S, const DevLights Ls, const Mats M, const EmitCdf E,
                                                      PhotonBuf B, uint32_t* __restrict__ per_particle,
                                                      PhotonTraceParams prm) {
    __shared__ uint2 stack[kLdsStack * 256];
    uint2* lds = stack + threadIdx.x;
    const uint32_t tag = 0xF0000000u + prm.kind;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n_round = (prm.count + stride - 1) / stride * stride;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_round; t += stride) {
        const bool valid = t < prm.count;
        const uint32_t index = prm.first + t;
        uint32_t d = 0, stored = 0;
        Ray r;
        f3 power;
        bool alive = false;
        if (valid) {
            const float sx = particle_u(index, tag, 0, prm.key), sy = particle_u(index, tag, 1, prm.key);
            const float ex = particle_u(index, tag, 2, prm.key), ey = particle_u(index, tag, 3, prm.key);
            d = 4;
            alive = emitter_ray(S, Ls, E, sx, sy, ex, ey, r, power);
        }
        f3 T = splat(1.0f);
        bool delta = false;
        for (uint32_t depth = 1; depth <= prm.max_depth;) {
            bool store = false;
            SurfaceHit sh;
            if (alive) {
                HitInfo h;
                Counters c{0, 0, 0};
                if (!trace_closest<false>(S, r, 0.001f, kInfF, h, lds, c)) {
                    alive = false;
                } else {
                    sh = surface(S, r, h.obj, h.prim, h.t, h.u, h.v);
                    if (sh.material < 0) alive = false;
                }
                if (alive) {
                    const uint32_t flags = bsdf_flags(M, sh.material);
                    const bool smooth = (flags & F_DIFFUSE) != 0 || (flags & F_GLOSSY) != 0;
                    store = smooth && depth > 1 && (prm.kind == 1 ? delta : !delta);
                }
            }
            // photon append (store_if_accepted, photon_tracer.h:80-96)
            const uint32_t slot = wave_append(B.count, store);
            if (store && slot < B.capacity) {
                const f3 dir = normalize(r.d);
                const f3 pw = T * power;
                B.px[slot] = sh.p.x; B.py[slot] = sh.p.y; B.pz[slot] = sh.p.z;
                B.dx[slot] = dir.x; B.dy[slot] = dir.y; B.dz[slot] = dir.z;
                B.nx[slot] = sh.n.x; B.ny[slot] = sh.n.y; B.nz[slot] = sh.n.z;
                B.wr[slot] = pw.x; B.wg[slot] = pw.y; B.wb[slot] = pw.z;
                B.key[slot] = (index << 8) | depth;
            }
            stored += store ? 1u : 0u;
            if (__ballot(alive) == 0) break;
            if (alive) {
                const Frame fr = frame_from(sh.n);
                BsdfCtx ctx;
                ctx.mode = LRT_MODE_IMPORTANCE;
                ctx.lobe = particle_u(index, tag, d, prm.key);
                SRec sr;
                sr.wi = to_local(fr, -normalize(r.d));
                sr.front = sh.front;
                const float ux = particle_u(index, tag, d + 1, prm.key), uy = particle_u(index, tag, d + 2, prm.key);
                d += 3;
                const f3 w = bsdf_sample(M, sh.material, ctx, sr, ux, uy);
                if (is_zero(w)) {
                    alive = false;
                } else {
                    const f3 outgoing = to_world(fr, sr.wo);
                    const f3 incoming = to_world(fr, sr.wi);
                    if (dot(sh.n, incoming) * sr.wi.z <= 0.0f || dot(sh.n, outgoing) * sr.wo.z <= 0.0f) {
                        alive = false;
                    } else {
                        T = T * w;
                        delta = (sr.type & F_SPECULAR) != 0;
                        r = Ray{sh.p, outgoing};
                        if (depth >= prm.rr_depth) {
                            const float surv = fminf(fmaxf(fmaxf(T.x, T.y), T.z), 0.95f);
                            if (particle_u(index, tag, d, prm.key) >= surv) alive = false;
                            else T = f3{T.x / surv, T.y / surv, T.z / surv};
                            d += 1;
                        }
                    }
                }
            }
            ++depth;
            if (is_zero(T)) alive = false;
        }
        if (valid) per_particle[t] = stored;
    }
}

// ---- stable LSD radix sort of (u32 key, u32 value), 4-bit digits -------------
constexpr int kRsBlock = 256, kRsTile = 4096;

__global__ __launch_bounds__(kRsBlock) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                      uint32_t* __restrict__ hist, uint32_t nblocks) {
    __shared__ uint32_t cnt[16];
    if (threadIdx.x < 16) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t b0 = blockIdx.x * kRsTile, b1 = min(n, b0 + kRsTile);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += kRsBlock) atomicAdd(&cnt[(keys[i] >> shift) & 15u], 1u);
    __syncthreads();
    if (threadIdx.x < 16) hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// exclusive scan of m counters in one block of 1024 threads
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ a, uint32_t m) {
    __shared__ uint32_t part[1024];
    const uint32_t seg = (m + 1023) / 1024, lo = threadIdx.x * seg, hi = min(m, lo + seg);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kRsBlock) void k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                         uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n,
                                                         int shift, const uint32_t* __restrict__ hist, uint32_t nblocks) {
    __shared__ uint32_t base[16];
    __shared__ uint32_t wcnt[kRsBlock / 64][16];
    if (threadIdx.x < 16) base[threadIdx.x] = hist[threadIdx.x * nblocks + blockIdx.x];
    __syncthreads();
    const uint32_t b0 = blockIdx.x * kRsTile, b1 = min(n, b0 + kRsTile);
    const uint32_t w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t off = b0; off < b1; off += kRsBlock) {
        const uint32_t i = off + threadIdx.x;
        const bool valid = i < b1;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t dg = (key >> shift) & 15u;
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t dd = 0; dd < 16; ++dd) {
            const uint64_t m = __ballot(valid && dg == dd);
            if (valid && dg == dd) rank = (uint32_t)__popcll(m & lt);
            if (lane == 0) wcnt[w][dd] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = base[dg] + rank;
            for (uint32_t ww = 0; ww < w; ++ww) pos += wcnt[ww][dg];
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
        if (threadIdx.x < 16) {
            uint32_t tot = 0;
            for (uint32_t ww = 0; ww < kRsBlock / 64; ++ww) tot += wcnt[ww][threadIdx.x];
            base[threadIdx.x] += tot;
        }
        __syncthreads();
    }
}

// ---- photon map: photons in kd-tree order -------------------------------------
struct PhotonMapDev {
    const float4* pos;   // xyz + depth (as float), tree order
    const float* pay;    // 9 floats: direction, normal, power
    const uint4* nodes;  // interior {split bits, axis, left, right}; leaf {0, 3, first, count}
    uint32_t n;
    float scale, radius;
    int32_t k;
};

// tree position i <- raw slot perm[order[i]] (perm: particle order)
__global__ void k_photon_pack(const PhotonBuf B, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ order,
                              uint32_t n, float4* __restrict__ pos, float* __restrict__ pay) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = perm[order[i]];
        pos[i] = make_float4(B.px[s], B.py[s], B.pz[s], (float)(B.key[s] & 0xffu));
        float* o = pay + 9 * (size_t)i;
        o[0] = B.dx[s]; o[1] = B.dy[s]; o[2] = B.dz[s];
        o[3] = B.nx[s]; o[4] = B.ny[s]; o[5] = B.nz[s];
        o[6] = B.wr[s]; o[7] = B.wg[s]; o[8] = B.wb[s];
    }
}

// kept photons in particle order -> packed positions for the host tree build
__global__ void k_photon_positions(const PhotonBuf B, const uint32_t* __restrict__ perm, uint32_t n,
                                   float* __restrict__ xyz) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = perm[i];
        xyz[3 * (size_t)i] = B.px[s];
        xyz[3 * (size_t)i + 1] = B.py[s];
        xyz[3 * (size_t)i + 2] = B.pz[s];
    }
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

constexpr int kMaxK = 128;
constexpr int kPkStack = 32;       // > photon kd depth (checked at build)
constexpr int kGroup = 8;          // lanes per photon query (a leaf holds <= 8 photons)
constexpr int kGatherBlock = 256;  // 32 queries per block
// waves per SIMD the eye-pass gather kernel is compiled to fit (0: the compiler's choice)
#ifndef LRT_GATHER_WAVES
#define LRT_GATHER_WAVES 6
#endif
#if LRT_GATHER_WAVES > 0
#define LRT_GATHER_ATTR __attribute__((amdgpu_waves_per_eu(LRT_GATHER_WAVES)))
#else
#define LRT_GATHER_ATTR
#endif

// lanes of the wave: value of v in lane src (0..63)
LRT_D uint32_t lane_read(uint32_t v, int src) { return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v); }
LRT_D float lane_read(float v, int src) { return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v))); }

// maximum of k over the 8 lanes of a query's group (every lane gets it)
LRT_D uint64_t group_max(uint64_t k, int lane) {
#pragma unroll
    for (int m = 1; m < kGroup; m <<= 1) {
        const int src = lane ^ m;
        const uint64_t o = ((uint64_t)lane_read((uint32_t)(k >> 32), src) << 32) | lane_read((uint32_t)k, src);
        k = o > k ? o : k;
    }
    return k;
}

template <int S>
LRT_D uint64_t slots_max(const uint64_t (&key)[S]) {
    uint64_t m = key[0];
#pragma unroll
    for (int s = 1; s < S; ++s) m = key[s] > m ? key[s] : m;
    return m;
}

// The 8 S keys of a query's group sorted in descending order, in place:
// afterwards slot s of group lane l holds the (8 s + l)-th largest key (slot
// major, so the 8 lanes of a group hold the 8 entries of a summing batch).
// A bitonic network over the 8 S elements, element e = 8 s + l: partners 8
// or more apart sit in the same lane (two slots), nearer ones in two lanes
// of the group (one exchange per slot).
template <int S>
LRT_D void group_sort_desc(uint64_t (&key)[S], int lane) {
    const int gl = lane & (kGroup - 1);
    constexpr int N = kGroup * S;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kGroup) {
                const int js = j / kGroup;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s & js) continue;
                    const bool desc = ((s * kGroup + gl) & k) == 0;
                    const uint64_t a = key[s], b = key[s | js];
                    const bool sw = desc ? (b > a) : (a > b);
                    key[s] = sw ? b : a;
                    key[s | js] = sw ? a : b;
                }
            } else {
                const int src = lane ^ j;
                const bool lo = (gl & j) == 0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const bool desc = ((s * kGroup + gl) & k) == 0;
                    const uint64_t o = ((uint64_t)lane_read((uint32_t)(key[s] >> 32), src) << 32) |
                                       lane_read((uint32_t)key[s], src);
                    const bool take_max = desc == lo;
                    const bool other = take_max ? (o > key[s]) : (o < key[s]);
                    key[s] = other ? o : key[s];
                }
            }
        }
    }
}

// k nearest photons within r of p (photon_map.h:117-165) and the irradiance
// estimate over them (photon_map.h:38-70), computed by the 8 lanes of a group
// together; every lane returns the estimate.
//
// The search keeps the reference's order: the far child is pushed when the
// node is entered if delta^2 <= the limit at that moment (r^2 until k photons
// are held, then the largest held distance, capped at r^2); a leaf's photons
// are tested in index order, each against the limit after the previous
// insertion.  The k-nearest set lives in registers as 64-bit keys
// (distance^2 bits << 32 | index + 1; 0 = empty), S slots in each of the 8
// lanes: the reference's max-heap pops the largest (distance^2, index) pair and
// pushes the new one, which leaves the same set as replacing the largest key.
// The estimate sums the set in the heap's pop order (largest key first).
// COUNT: also count the work into c (nodes visited, photon positions tested,
// photons summed) in lane 0 of the group.
template <int S, bool COUNT>
LRT_D f3 photon_estimate_group(const PhotonMapDev& Mp, f3 p, f3 n, int max_depth, uint32_t* stack, Counters& c) {
    if (Mp.n == 0 || Mp.radius <= 0.0f || Mp.k <= 0) return splat(0.0f);
    const int lane = (int)(threadIdx.x & 63u), gl = lane & (kGroup - 1), gb = lane - gl;
    const bool counter = COUNT && gl == 0;
    const float r2 = Mp.radius * Mp.radius;
    const int K = Mp.k;
    uint64_t key[S];
#pragma unroll
    for (int s = 0; s < S; ++s) key[s] = 0;
    int size = 0, sp = 0;
    float maxd = 0.0f;
    uint64_t maxkey = 0;
    uint32_t node = 0;
    for (;;) {
        const uint4 nd = Mp.nodes[node];
        if (counter) ++c.nodes;
        float limit = r2;
        if (size == K) limit = fminf(limit, maxd);
        if (nd.y == 3u) {
            if (counter) c.refs += nd.w;
            float d = 0.0f;
            bool cand = false;
            if ((uint32_t)gl < nd.w) {
                const float4 q = Mp.pos[nd.z + (uint32_t)gl];
                d = len2(f3{q.x, q.y, q.z} - p);
                cand = !(d > limit);  // the limit only shrinks: a photon beyond it now stays out
            }
            uint32_t m = (uint32_t)(__ballot(cand) >> gb) & 0xffu;
            while (m) {
                const int ci = __builtin_ctz(m);
                m &= m - 1u;
                const float dc = lane_read(d, gb + ci);
                if (dc > limit) continue;
                const uint64_t nk = ((uint64_t)__float_as_uint(dc) << 32) | (uint64_t)(nd.z + (uint32_t)ci + 1u);
                if (size < K) {
                    const int t = size++;
#pragma unroll
                    for (int s = 0; s < S; ++s)
                        if (s == (t >> 3) && gl == (t & 7)) key[s] = nk;
                    if (size < K) continue;
                } else {
#pragma unroll
                    for (int s = 0; s < S; ++s)
                        if (key[s] == maxkey) key[s] = nk;
                }
                maxkey = group_max(slots_max(key), lane);
                maxd = __uint_as_float((uint32_t)(maxkey >> 32));
                limit = fminf(r2, maxd);
            }
            if (sp == 0) break;
            node = stack[--sp];
            continue;
        }
        const float pv = nd.y == 0u ? p.x : (nd.y == 1u ? p.y : p.z);
        const float delta = pv - __uint_as_float(nd.x);
        const bool left_near = delta <= 0.0f;
        if (delta * delta <= limit) stack[sp++] = left_near ? nd.w : nd.z;
        node = left_near ? nd.z : nd.w;
    }
    if (counter) c.tris += (uint32_t)size;
    // the set in descending key order (the heap's pop order), 8 entries at a
    // time: after the sort entry 8 s + j is slot s of lane j, the 8 lanes load
    // and weigh their photons together, and the terms are added in order
    const float inv_r2 = 1.0f / r2;
    f3 sum = splat(0.0f);
    group_sort_desc(key, lane);
#pragma unroll
    for (int sb = 0; sb < S; ++sb) {
        const int b = sb * kGroup;
        if (b >= size) break;
        const int cnt = min(kGroup, size - b);
        const uint64_t mine = key[sb];
        f3 term = splat(0.0f);
        uint32_t use = 0;
        if (gl < cnt) {
            const uint32_t idx = (uint32_t)mine - 1u;
            const float d = __uint_as_float((uint32_t)(mine >> 32));
            const float4 q = Mp.pos[idx];
            if (!((int)q.w > max_depth)) {
                const float* pay = Mp.pay + 9 * (size_t)idx;
                const f3 in = -ld3(pay);
                const f3 pn = ld3(pay + 3);
                const float ip = dot(pn, in), is = dot(n, in);
                if (!(is <= 0.0f || dot(pn, n) <= 1e-1f || ip <= 1e-2f)) {
                    const float sq = 1.0f - d * inv_r2;
                    const float w = (is / ip) * (sq * sq);
                    term = ld3(pay + 6) * fabsf(w);
                    use = 1u;
                }
            }
        }
        for (int j = 0; j < cnt; ++j) {
            const uint32_t u = lane_read(use, gb + j);
            const f3 v{lane_read(term.x, gb + j), lane_read(term.y, gb + j), lane_read(term.z, gb + j)};
            if (u) sum = sum + v;
        }
    }
    return sum * (Mp.scale * 3.0f * (1.0f / kPiF) * inv_r2);
}

// queries handled by groups of 8 lanes; S = key slots per lane (k <= 8 S)
template <int S>
__global__ __launch_bounds__(kGatherBlock) void k_photon_estimate(const PhotonMapDev Mp, const float* __restrict__ q6,
                                                                  uint32_t n, int max_depth, float* __restrict__ out) {
    __shared__ uint32_t stacks[(kGatherBlock / kGroup) * kPkStack];
    uint32_t* stack = stacks + (threadIdx.x / kGroup) * kPkStack;
    const uint32_t groups = gridDim.x * (kGatherBlock / kGroup);
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; i < n; i += groups) {
        Counters c{0, 0, 0};
        const f3 e = photon_estimate_group<S, false>(Mp, ld3(q6 + 6 * (size_t)i), ld3(q6 + 6 * (size_t)i + 3), max_depth,
                                                     stack, c);
        if ((threadIdx.x & (kGroup - 1)) == 0) {
            out[3 * (size_t)i] = e.x;
            out[3 * (size_t)i + 1] = e.y;
            out[3 * (size_t)i + 2] = e.z;
        }
    }
}

// key slots per lane for k photons per lookup
inline int gather_slots(int k) { return k <= 64 ? 8 : 16; }

}  // namespace lrt
