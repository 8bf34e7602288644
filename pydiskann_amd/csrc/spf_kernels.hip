// gfx950 kernels of the SPFresh batch-search path.
//
// Stages (one batch of queries, all device-resident):
//   k_prep      per-query values the metric consumes (f16 rounding, int8 codes, norms)
//   k_scan      distances: dense mode = centroid probe (every query x every
//               centroid, Inner::nearest src/spfresh.rs:358-372); listed mode =
//               posting scan over work items (posting, <=G queries, <=1024
//               entries) grouped by posting so one posting is streamed once
//               for all queries that probe it (Inner::search loop
//               src/spfresh.rs:703-717)
//   k_select    per-segment top-K of 64-bit keys (fkey(dist) << 32 | rank):
//               radix select (11/11/10-bit digits) + LDS bitonic sort
//   k_plan_*    probe_ratio cut (src/spfresh.rs:683-687), candidate offsets,
//               posting-major inversion of the (query, probe) pairs
//   k_rerank    exact f32 distance of the pool (src/spfresh.rs:719-723)
//   k_finalize  rank -> id, key -> distance
//
// Distances follow one canonical order shared with the CPU oracle: each f32
// reduction is a sequential fused multiply-add chain from +0 over dimensions
// 0..d-1; int8 L2 is an exact integer sum.  Build with -ffp-contract=off so no
// other multiply-add is fused.
#include "spf_internal.h"

namespace spf {

__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_to_float(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

// Angular distance from (a.b, |a|^2, |b|^2); zero-norm cases as the oracle.
__device__ __forceinline__ float cosine_from(float ab, float a2, float b2) {
    if (a2 == 0.0f && b2 == 0.0f) return 0.0f;
    if (ab == 0.0f) return 1.0f;
    float r = 1.0f - ab / sqrtf(a2 * b2);
    return r > 0.0f ? r : 0.0f;
}

template <int MET>
__device__ __forceinline__ float finish_f(float acc, float qn, float xn) {
    if (MET == M_L2SQ) return acc;
    if (MET == M_L2) return sqrtf(acc);
    if (MET == M_DOT) return 1.0f - acc;
    return cosine_from(acc, qn, xn);
}

template <int MET>
__device__ __forceinline__ void step(float &acc, float q, float x) {
    if (MET == M_L2 || MET == M_L2SQ) {
        float t = q - x;
        acc = __builtin_fmaf(t, t, acc);
    } else {
        acc = __builtin_fmaf(q, x, acc);
    }
}

__device__ __forceinline__ uint32_t next_pow2(uint32_t v) {
    uint32_t m = 1;
    while (m < v) m <<= 1;
    return m;
}

// ---------------------------------------------------------------- k_prep
__global__ __launch_bounds__(64) void k_prep(const PrepParams p) {
    extern __shared__ float s_v[];   // dpad floats
    const uint32_t q = blockIdx.x, lane = threadIdx.x;
    if (p.zero)
        for (uint32_t i = q * 64 + lane; i < p.zero_n; i += gridDim.x * 64) p.zero[i] = 0u;
    const float *src = p.q + (size_t)q * p.dim;
    for (uint32_t i = lane; i < p.dpad; i += 64) {
        float v = i < p.dim ? src[i] : 0.0f;
        s_v[i] = v;
        p.qv_probe[(size_t)q * p.dpad + i] = v;
        if (p.round_f16) p.qv_scan[(size_t)q * p.dpad + i] = (float)(_Float16)v;
    }
    if (p.q16) {
        for (uint32_t i = lane; i < p.dpad32; i += 64)
            p.q16[(size_t)q * p.dpad32 + i] = __builtin_bit_cast(uint16_t, (_Float16)(i < p.dim ? src[i] : 0.0f));
    }
    if (p.qbf) {
        for (uint32_t i = lane; i < p.dpad32; i += 64) {
            const float v = i < p.dim ? src[i] - p.mu[i] : 0.0f;
            p.qbf[(size_t)q * p.dpad32 + i] = __builtin_bit_cast(uint16_t, (__bf16)v);
        }
    }
    __syncthreads();
    if (p.q8m) {
        // symmetric int8 query: q ~ sq * q8 with sq = max|q| / 127; then
        // q.decode(c) = scale * sq * (sum q8 (c - 128) + 128 sum q8) + offset * sum q
        float m = 0.0f;
        for (uint32_t i = lane; i < p.dim; i += 64) m = fmaxf(m, fabsf(s_v[i]));
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        const float inv = m > 0.0f ? 127.0f / m : 0.0f;
        int q8sum = 0;
        float qsum = 0.0f;
        for (uint32_t i = lane; i < p.dpad64; i += 64) {
            int v = 0;
            if (i < p.dim) {
                v = (int)rintf(s_v[i] * inv);
                v = v > 127 ? 127 : (v < -127 ? -127 : v);
                qsum += s_v[i];
            }
            q8sum += v;
            p.q8m[(size_t)q * p.dpad64 + i] = (int8_t)v;
        }
        for (int off = 32; off > 0; off >>= 1) {
            q8sum += __shfl_xor(q8sum, off, 64);
            qsum += __shfl_xor(qsum, off, 64);
        }
        if (lane == 0) {
            const float a = p.scale * (m / 127.0f);
            p.q8a[q] = a;
            p.q8b[q] = a * 128.0f * (float)q8sum + p.offset * qsum;
        }
    }
    if (p.want_q8) {
        const float alpha = 1.0f / p.scale;
        const float beta = -p.offset / p.scale;
        for (uint32_t w = lane; w < p.dpad / 4; w += 64) {
            uint32_t word = 0;
            for (uint32_t j = 0; j < 4; j++) {
                uint32_t i = 4 * w + j;
                uint32_t c = 0;
                if (i < p.dim) {
                    float s = s_v[i] * alpha;
                    s = s + beta;
                    if (s > 0.0f) c = s >= 255.0f ? 255u : (uint32_t)rintf(s);
                }
                word |= c << (8 * j);
            }
            p.q8[(size_t)q * (p.dpad / 4) + w] = word;
        }
    }
    __syncthreads();
    if (lane == 0) {
        float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t i = 0; i < p.dim; i++) c[(i >> 2) & 3] = __builtin_fmaf(s_v[i], s_v[i], c[(i >> 2) & 3]);
        p.qnorm[q] = (c[0] + c[1]) + (c[2] + c[3]);
        if (p.q16) {
            float a16 = 0.0f, e16 = 0.0f;
            for (uint32_t i = 0; i < p.dim; i++) {
                const float h = (float)(_Float16)s_v[i];
                a16 = __builtin_fmaf(h, h, a16);
                const float t = s_v[i] - h;   // exact: h is s_v[i] rounded to 11 bits
                e16 = __builtin_fmaf(t, t, e16);
            }
            p.qn16[q] = a16;
            if (p.eq16) p.eq16[q] = sqrtf(e16);
        }
        if (p.qbf) {
            float ab = 0.0f;
            for (uint32_t i = 0; i < p.dim; i++) {
                const float h = (float)(__bf16)(s_v[i] - p.mu[i]);
                ab = __builtin_fmaf(h, h, ab);
            }
            p.qnbf[q] = ab;
        }
        if (p.want_q8) {
            uint32_t s = 0;
            for (uint32_t w = 0; w < p.dpad / 4; w++) {
                uint32_t word = p.q8[(size_t)q * (p.dpad / 4) + w];
                for (uint32_t j = 0; j < 4; j++) {
                    uint32_t c = (word >> (8 * j)) & 0xffu;
                    s += c * c;
                }
            }
            p.qsq[q] = s;
        }
    }
}

// ---------------------------------------------------------------- k_scan
template <int CK, int MET, int G, bool DENSE>
__global__ __launch_bounds__(256) void k_scan(const ScanParams p) {
    extern __shared__ __align__(16) float s_q[];   // G * dpad f32 (or packed int8 words)
    __shared__ uint64_t s_slot[G];
    __shared__ uint32_t s_qi[G];
    __shared__ float s_qn[G];
    __shared__ uint32_t s_qs[G];
    __shared__ uint64_t s_mask[G];
    constexpr bool INT_L2 = (CK == CK_U8) && (MET == M_L2 || MET == M_L2SQ);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t dq4 = p.dpad >> 2;   // float4 (or u32 words) per staged query row
    const uint32_t n_items = DENSE ? ((p.nq + G - 1) / G) * p.n_echunks
                             : (p.n_items_dev ? *p.n_items_dev : p.n_items);

    for (XcdWalk it(n_items); it.more(); it.next()) {
        const uint32_t item = it.item;
        uint32_t e_begin, e_end, gcnt, base_e;
        if (DENSE) {
            const uint32_t qc = item / p.n_echunks, ec = item - qc * p.n_echunks;
            const uint32_t q0 = qc * G;
            gcnt = min((uint32_t)G, p.nq - q0);
            e_begin = ec * kEChunk;
            e_end = min(e_begin + kEChunk, p.n_entries);
            base_e = 0;
            if (tid < gcnt) {
                s_qi[tid] = q0 + tid;
                s_slot[tid] = (uint64_t)(q0 + tid) * p.out_stride;
            }
        } else {
            const uint32_t pp = p.item_post[item];
            const uint32_t local = item - p.item_off[pp];
            const uint32_t pb = p.post_off[pp], pe = p.post_off[pp + 1];
            const uint32_t nec = (pe - pb + p.echunk - 1) / p.echunk;
            const uint32_t qc = local / nec, ec = local - qc * nec;
            const uint32_t pr0 = p.pair_off[pp] + qc * G;
            gcnt = min((uint32_t)G, p.pair_off[pp + 1] - pr0);
            e_begin = pb + ec * p.echunk;
            e_end = min(e_begin + p.echunk, pe);
            base_e = pb;
            if (tid < gcnt) {
                s_qi[tid] = p.pair_q[pr0 + tid];
                s_slot[tid] = p.pair_slot[pr0 + tid];
                if (p.qmask) s_mask[tid] = p.qmask[s_qi[tid]];
            }
        }
        __syncthreads();
        if (INT_L2) {
            uint32_t *sw = reinterpret_cast<uint32_t *>(s_q);
            for (uint32_t i = tid; i < gcnt * dq4; i += kWG) {
                const uint32_t g = i / dq4;
                sw[i] = p.q8[(size_t)s_qi[g] * dq4 + (i - g * dq4)];
            }
            if (tid < gcnt) s_qs[tid] = p.qsq[s_qi[tid]];
        } else {
            float4 *s4 = reinterpret_cast<float4 *>(s_q);
            const float4 *q4 = reinterpret_cast<const float4 *>(p.qv);
            for (uint32_t i = tid; i < gcnt * dq4; i += kWG) {
                const uint32_t g = i / dq4;
                s4[i] = q4[(size_t)s_qi[g] * dq4 + (i - g * dq4)];
            }
            if (MET == M_COS && tid < gcnt) s_qn[tid] = p.qnorm[s_qi[tid]];
        }
        __syncthreads();

        for (uint32_t e0 = e_begin + wave * 64; e0 < e_end; e0 += kWG) {
            const uint32_t e = e0 + lane;
            const bool valid = e < e_end;
            const uint4 *src = p.data + (size_t)(e >> 6) * p.units * 64 + (e & 63);
            // four FMA chains per query: chain r takes dims i with (i/4)%4 == r
            float accf[G][4];
            uint32_t acci[G];
#pragma unroll
            for (int g = 0; g < G; g++) {
                accf[g][0] = accf[g][1] = accf[g][2] = accf[g][3] = 0.0f;
                acci[g] = 0u;
            }

            if (CK == CK_F32) {   // unit u = dims 4u..4u+3 -> chain u%4 (units % 4 == 0)
                const float4 *s4 = reinterpret_cast<const float4 *>(s_q);
                for (uint32_t v = 0; v < p.units; v += 4) {
                    uint4 w[4];
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        w[r] = valid ? src[(size_t)(v + r) * 64] : make_uint4(0, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const float x0 = __uint_as_float(w[r].x), x1 = __uint_as_float(w[r].y);
                        const float x2 = __uint_as_float(w[r].z), x3 = __uint_as_float(w[r].w);
#pragma unroll
                        for (int g = 0; g < G; g++) {
                            const float4 qq = s4[g * dq4 + v + r];
                            step<MET>(accf[g][r], qq.x, x0);
                            step<MET>(accf[g][r], qq.y, x1);
                            step<MET>(accf[g][r], qq.z, x2);
                            step<MET>(accf[g][r], qq.w, x3);
                        }
                    }
                }
            } else if (CK == CK_F16) {   // unit u = dims 8u..8u+7 -> chains 2(u%2), 2(u%2)+1
                const float4 *s4 = reinterpret_cast<const float4 *>(s_q);
                for (uint32_t v = 0; v < p.units; v += 2) {
                    union { uint4 v; _Float16 h[8]; } cv[2];
#pragma unroll
                    for (int r = 0; r < 2; r++)
                        cv[r].v = valid ? src[(size_t)(v + r) * 64] : make_uint4(0, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        float x[8];
#pragma unroll
                        for (int j = 0; j < 8; j++) x[j] = (float)cv[r].h[j];
#pragma unroll
                        for (int g = 0; g < G; g++) {
                            const float4 qa = s4[g * dq4 + 2 * (v + r)];
                            const float4 qb = s4[g * dq4 + 2 * (v + r) + 1];
                            step<MET>(accf[g][2 * r], qa.x, x[0]);
                            step<MET>(accf[g][2 * r], qa.y, x[1]);
                            step<MET>(accf[g][2 * r], qa.z, x[2]);
                            step<MET>(accf[g][2 * r], qa.w, x[3]);
                            step<MET>(accf[g][2 * r + 1], qb.x, x[4]);
                            step<MET>(accf[g][2 * r + 1], qb.y, x[5]);
                            step<MET>(accf[g][2 * r + 1], qb.z, x[6]);
                            step<MET>(accf[g][2 * r + 1], qb.w, x[7]);
                        }
                    }
                }
            } else if (INT_L2) {
                const uint32_t *sw = reinterpret_cast<const uint32_t *>(s_q);
#pragma unroll 2
                for (uint32_t u = 0; u < p.units; u++) {
                    const uint4 w = valid ? src[(size_t)u * 64] : make_uint4(0, 0, 0, 0);
#pragma unroll
                    for (int g = 0; g < G; g++) {
                        const uint4 qq = *reinterpret_cast<const uint4 *>(sw + g * dq4 + 4 * u);
                        uint32_t a = acci[g];
                        a = __builtin_amdgcn_udot4(w.x, qq.x, a, false);
                        a = __builtin_amdgcn_udot4(w.y, qq.y, a, false);
                        a = __builtin_amdgcn_udot4(w.z, qq.z, a, false);
                        a = __builtin_amdgcn_udot4(w.w, qq.w, a, false);
                        acci[g] = a;
                    }
                }
            } else {   // CK_U8 decode path (cosine / dot), src/quantized.rs:28-36
                const float4 *s4 = reinterpret_cast<const float4 *>(s_q);
                const float sc = p.scale, of = p.offset;
                // unit u = dims 16u.. -> chains 0..3, in increasing u; four
                // units' loads in flight at a time
#pragma unroll 1
                for (uint32_t u0 = 0; u0 < p.units; u0 += 4) {
                    uint4 w4[4];
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        w4[r] = (valid && u0 + r < p.units) ? src[(size_t)(u0 + r) * 64]
                                                            : make_uint4(0, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t u = u0 + r;
                        if (u >= p.units) break;   // (uniform)
                        const uint32_t ws[4] = {w4[r].x, w4[r].y, w4[r].z, w4[r].w};
                        float x[16];
#pragma unroll
                        for (int j = 0; j < 16; j++) {
                            const float c = (float)((ws[j >> 2] >> (8 * (j & 3))) & 0xffu);
                            const float t = c * sc;
                            x[j] = t + of;
                        }
#pragma unroll
                        for (int g = 0; g < G; g++) {
#pragma unroll
                            for (int j4 = 0; j4 < 4; j4++) {
                                const float4 qq = s4[g * dq4 + 4 * u + j4];
                                step<MET>(accf[g][j4], qq.x, x[4 * j4 + 0]);
                                step<MET>(accf[g][j4], qq.y, x[4 * j4 + 1]);
                                step<MET>(accf[g][j4], qq.z, x[4 * j4 + 2]);
                                step<MET>(accf[g][j4], qq.w, x[4 * j4 + 3]);
                            }
                        }
                    }
                }
            }

            if (valid) {
                const uint64_t low = p.low ? (uint64_t)p.low[e] : (uint64_t)e;
                const bool dead = p.del && ((p.del[e >> 5] >> (e & 31)) & 1u);
                const float xn = (MET == M_COS) ? p.xnorm[e] : 0.0f;
                const uint32_t xs = INT_L2 ? p.xsq[e] : 0u;
                const uint32_t lab = (!DENSE && p.qmask) ? p.label_e[e] : 0u;
#pragma unroll
                for (int g = 0; g < G; g++) {
                    if (g < (int)gcnt) {
                        const bool skip = dead || (!DENSE && p.qmask && !label_ok(s_mask[g], lab));
                        float dist;
                        if (INT_L2) {
                            const uint32_t s = xs + s_qs[g] - 2u * acci[g];
                            float v = (float)s;
                            v = v * p.scale;
                            v = v * p.scale;
                            dist = (MET == M_L2) ? sqrtf(v) : v;
                        } else {
                            const float acc = (accf[g][0] + accf[g][1]) + (accf[g][2] + accf[g][3]);
                            dist = finish_f<MET>(acc, MET == M_COS ? s_qn[g] : 0.0f, xn);
                        }
                        const uint64_t key = skip ? kNoKey : (((uint64_t)fkey(dist) << 32) | low);
                        p.out[s_slot[g] + (e - base_e)] = key;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- selection
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *part) {
    const uint32_t tid = threadIdx.x;
    part[tid] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kWG; off <<= 1) {
        const uint32_t t = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += t;
        __syncthreads();
    }
    const uint32_t incl = part[tid];
    __syncthreads();
    return incl - v;
}

__device__ void bitonic_sort_lds(uint64_t *buf, uint32_t m) {
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < (m >> 1); i += kWG) {
                const uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                const uint32_t hi = lo + j;
                const uint64_t a = buf[lo], b = buf[hi];
                const bool up = (lo & k) == 0;
                if ((a > b) == up) {
                    buf[lo] = b;
                    buf[hi] = a;
                }
            }
            __syncthreads();
        }
}

constexpr uint32_t kSelSlack = 256;     // extra keys tolerated in the final sort

// Top-K of one segment.  R > 0: the segment (at most R * 256 keys) is held in
// registers, key i = tid + 256 j in kr[j]; R == 0: every pass re-reads L2.
template <int R>
__device__ __forceinline__ void select_body(const uint64_t *kp, uint64_t n, uint32_t K,
                                            uint32_t pad_to, uint64_t *o, uint32_t *cnt_out,
                                            uint64_t *buf,
                                            uint32_t *hist, uint32_t *part, uint32_t *s_u,
                                            uint64_t (*s_lh)[4]) {
    const uint32_t tid = threadIdx.x;
    uint64_t kr[R > 0 ? R : 1];
    if (R > 0) {
#pragma unroll
        for (int j = 0; j < R; j++) {
            const uint64_t i = tid + 256ull * j;
            kr[j] = i < n ? kp[i] : kNoKey;
        }
    }
    // visit every key of the segment (padding keys read as kNoKey)
#define SPF_FOR_KEYS(BODY)                                                        \
    if (R > 0) {                                                                  \
        _Pragma("unroll") for (int j = 0; j < (R > 0 ? R : 1); j++) {             \
            const uint64_t k = kr[j];                                             \
            BODY                                                                  \
        }                                                                         \
    } else {                                                                      \
        for (uint64_t i = tid; i < n; i += kWG) {                                 \
            const uint64_t k = kp[i];                                             \
            BODY                                                                  \
        }                                                                         \
    }
    if (tid == 0) s_u[0] = 0;
    __syncthreads();
    // invalid count, min and max of the valid keys
    uint32_t inval = 0;
    uint64_t lo = kNoKey, hi = 0;
    SPF_FOR_KEYS({
        if (k == kNoKey) {
            inval++;
        } else {
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    })
    for (int off = 32; off > 0; off >>= 1) {
        inval += __shfl_xor(inval, off, 64);
        const uint64_t a = __shfl_xor(lo, off, 64), b2 = __shfl_xor(hi, off, 64);
        lo = a < lo ? a : lo;
        hi = b2 > hi ? b2 : hi;
    }
    if ((tid & 63) == 0) {
        s_lh[0][tid >> 6] = lo;
        s_lh[1][tid >> 6] = hi;
        atomicAdd(&s_u[0], inval);
    }
    __syncthreads();
    lo = s_lh[0][0];
    hi = s_lh[1][0];
    for (int w = 1; w < 4; w++) {
        lo = s_lh[0][w] < lo ? s_lh[0][w] : lo;
        hi = s_lh[1][w] > hi ? s_lh[1][w] : hi;
    }
    // with registers, padding beyond n was counted as invalid
    const uint64_t pad = R > 0 ? (uint64_t)R * 256 - n : 0;
    const uint64_t nvalid = n + pad - s_u[0];
    const uint32_t Kv = (uint32_t)min((uint64_t)K, nvalid);
    __syncthreads();
    if (Kv == 0) {
        for (uint32_t i = tid; i < pad_to; i += kWG) o[i] = kNoKey;
        if (tid == 0 && cnt_out) *cnt_out = 0;
        return;
    }
    // digits start at the highest bit where the valid keys differ, so the first
    // histogram spreads over bins instead of piling into one
    const int hb = lo == hi ? -1 : 63 - __clzll(lo ^ hi);
    uint64_t mask = hb < 0 ? ~0ull : (hb >= 63 ? 0ull : ~((2ull << hb) - 1ull));
    uint64_t prefix = lo & mask;
    uint32_t rem = Kv;
    int top = hb;
    while (top >= 0) {
        const int wdt = top + 1 < 11 ? top + 1 : 11;
        const int sh = top + 1 - wdt;
        const uint32_t dm = (1u << wdt) - 1u;
        for (uint32_t i = tid; i < 2048; i += kWG) hist[i] = 0;
        __syncthreads();
        SPF_FOR_KEYS({
            if ((k & mask) == prefix) atomicAdd(&hist[(uint32_t)(k >> sh) & dm], 1u);
        })
        __syncthreads();
        uint32_t loc[8], sm = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            loc[j] = hist[tid * 8 + j];
            sm += loc[j];
        }
        const uint32_t excl = block_excl_scan256(sm, part);
        if (excl < rem && rem <= excl + sm) {
            uint32_t c = excl;
            for (int j = 0; j < 8; j++) {
                if (rem <= c + loc[j]) {
                    s_u[1] = tid * 8 + j;
                    s_u[2] = c;
                    s_u[3] = loc[j];
                    break;
                }
                c += loc[j];
            }
        }
        __syncthreads();
        const uint32_t bin = s_u[1], below = s_u[2], cnt = s_u[3];
        __syncthreads();
        rem -= below;
        prefix |= (uint64_t)bin << sh;
        mask |= (uint64_t)dm << sh;
        top = sh - 1;
        if ((Kv - rem) + cnt <= Kv + kSelSlack) break;
    }
    if (tid == 0) s_u[0] = 0;
    __syncthreads();
    SPF_FOR_KEYS({
        if (k != kNoKey && (k & mask) <= prefix) {
            const uint32_t pos = atomicAdd(&s_u[0], 1u);
            if (pos < kSelCap) buf[pos] = k;
        }
    })
#undef SPF_FOR_KEYS
    __syncthreads();
    const uint32_t m = min(s_u[0], kSelCap);
    const uint32_t m2 = next_pow2(m < 2 ? 2 : m);
    for (uint32_t i = m + tid; i < m2; i += kWG) buf[i] = kNoKey;
    __syncthreads();
    bitonic_sort_lds(buf, m2);
    for (uint32_t i = tid; i < Kv; i += kWG) o[i] = buf[i];
    for (uint32_t i = Kv + tid; i < pad_to; i += kWG) o[i] = kNoKey;
    if (tid == 0 && cnt_out) *cnt_out = Kv;
}

// Segments with n_lo <= n <= n_hi keys; R > 0: held in registers (n <= R * 256),
// one launch per size band so the register-heavy band does not lower the
// occupancy of the others.
template <int R>
__global__ __launch_bounds__(256) void k_select(const uint64_t *keys, const uint64_t *seg_off,
                                                const uint32_t *seg_cnt, uint64_t seg_stride,
                                                uint32_t K, uint64_t *out, uint32_t *out_cnt,
                                                uint64_t n_lo, uint64_t n_hi) {
    __shared__ uint64_t buf[kSelCap];
    __shared__ uint32_t hist[2048];
    __shared__ uint32_t part[kWG];
    __shared__ uint32_t s_u[4];
    __shared__ uint64_t s_lh[2][4];
    const uint32_t seg = blockIdx.x;
    const uint64_t b = seg_off ? seg_off[seg] : (uint64_t)seg * seg_stride;
    const uint64_t e = seg_off ? seg_off[seg + 1] : b + seg_stride;
    const uint64_t n = seg_cnt ? (uint64_t)seg_cnt[seg] : e - b;
    if (n < n_lo || n > n_hi) return;   // another band's (or, long, its chunks')
    select_body<R>(keys + b, n, K, 0, out + (size_t)seg * K, out_cnt + seg, buf, hist, part, s_u,
                   s_lh);
}

constexpr uint64_t kSelRegSmall = 16 * 256, kSelRegBig = 96 * 256;   // keys held in registers

// every segment of at most n_max keys, by size band
static hipError_t select_bands(const uint64_t *keys, const uint64_t *seg_off,
                               const uint32_t *seg_cnt, uint64_t seg_stride, uint32_t nseg,
                               uint32_t K, uint64_t *out, uint32_t *out_cnt, uint64_t n_max,
                               hipStream_t s) {
    // fixed-length rows (no seg_cnt) need only the band of their length;
    // segments above n_max are done in chunks by the caller
    const uint64_t lo = seg_cnt ? 0 : seg_stride;
    if (lo > n_max) return hipSuccess;
    const uint64_t hi = seg_cnt ? n_max : seg_stride;
    if (lo <= kSelRegSmall)
        hipLaunchKernelGGL(k_select<16>, dim3(nseg), dim3(kWG), 0, s, keys, seg_off, seg_cnt,
                           seg_stride, K, out, out_cnt, (uint64_t)0, std::min(n_max, kSelRegSmall));
    if (hi > kSelRegSmall && lo <= kSelRegBig)
        hipLaunchKernelGGL(k_select<96>, dim3(nseg), dim3(kWG), 0, s, keys, seg_off, seg_cnt,
                           seg_stride, K, out, out_cnt, kSelRegSmall + 1, std::min(n_max, kSelRegBig));
    if (hi > kSelRegBig)
        hipLaunchKernelGGL(k_select<0>, dim3(nseg), dim3(kWG), 0, s, keys, seg_off, seg_cnt,
                           seg_stride, K, out, out_cnt, kSelRegBig + 1, n_max);
    return hipGetLastError();
}

// Long segments (more than kSelChunk keys) in two rounds, so one query's
// Zipf-sized candidate list does not leave a single workgroup streaming it:
// chunk counts per segment, then the top K of every chunk (padded to K with
// "no key"), then the top K of each long segment's chunk results.
__global__ void k_sel_chunk_count(const uint64_t *seg_off, const uint32_t *seg_cnt,
                                  uint64_t seg_stride, uint32_t nseg, uint32_t chunk,
                                  uint32_t *chunk_cnt) {
    const uint32_t seg = blockIdx.x * blockDim.x + threadIdx.x;
    if (seg >= nseg) return;
    const uint64_t n = seg_cnt ? (uint64_t)seg_cnt[seg]
                       : seg_off ? seg_off[seg + 1] - seg_off[seg] : seg_stride;
    chunk_cnt[seg] = n > chunk ? (uint32_t)((n + chunk - 1) / chunk) : 0u;
}

// R: keys per thread held in registers for chunks up to R * 256 keys (0: none)
template <int R>
__global__ __launch_bounds__(256) void k_select_chunk(const uint64_t *keys, const uint64_t *seg_off,
                                                      const uint32_t *seg_cnt, uint64_t seg_stride,
                                                      const uint32_t *chunk_off, uint32_t nseg,
                                                      uint32_t chunk, uint32_t K, uint64_t *tmp) {
    __shared__ uint64_t buf[kSelCap];
    __shared__ uint32_t hist[2048];
    __shared__ uint32_t part[kWG];
    __shared__ uint32_t s_u[4];
    __shared__ uint64_t s_lh[2][4];
    const uint32_t c = blockIdx.x;
    if (c >= chunk_off[nseg]) return;
    uint32_t lo = 0, hi = nseg;   // the segment whose chunk range holds c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (chunk_off[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t seg = lo;
    const uint64_t j = c - chunk_off[seg];
    const uint64_t s0 = seg_off ? seg_off[seg] : (uint64_t)seg * seg_stride;
    const uint64_t n = seg_cnt ? (uint64_t)seg_cnt[seg] : seg_off ? seg_off[seg + 1] - s0 : seg_stride;
    const uint64_t b = s0 + j * chunk;
    const uint64_t m = min((uint64_t)chunk, n - j * chunk);
    if (R > 0 && m <= (uint64_t)R * 256)
        select_body<R>(keys + b, m, K, K, tmp + (size_t)c * K, nullptr, buf, hist, part, s_u, s_lh);
    else
        select_body<0>(keys + b, m, K, K, tmp + (size_t)c * K, nullptr, buf, hist, part, s_u, s_lh);
}

__global__ __launch_bounds__(256) void k_select_merge(const uint64_t *tmp, const uint32_t *chunk_off,
                                                      uint32_t K, uint64_t *out, uint32_t *out_cnt) {
    __shared__ uint64_t buf[kSelCap];
    __shared__ uint32_t hist[2048];
    __shared__ uint32_t part[kWG];
    __shared__ uint32_t s_u[4];
    __shared__ uint64_t s_lh[2][4];
    const uint32_t seg = blockIdx.x;
    const uint32_t c0 = chunk_off[seg], c1 = chunk_off[seg + 1];
    if (c1 == c0) return;   // done whole
    const uint64_t n = (uint64_t)(c1 - c0) * K;
    if (n <= 16 * 256)
        select_body<16>(tmp + (size_t)c0 * K, n, K, 0, out + (size_t)seg * K, out_cnt + seg, buf,
                        hist, part, s_u, s_lh);
    else
        select_body<0>(tmp + (size_t)c0 * K, n, K, 0, out + (size_t)seg * K, out_cnt + seg, buf,
                       hist, part, s_u, s_lh);
}

// ---------------------------------------------------------------- planning
// One wavefront per query (four per workgroup): the probe_ratio cut (the first
// probe beyond d0 * max(ratio, 1) ends the list), the kept postings, the
// query's candidate count, each probe's slot relative to the query's segment
// (wave prefix sums in probe order), and the pair's rank within its posting
// (the posting-major counting sort's count, taken with the returned value).
__global__ __launch_bounds__(256) void k_plan_q(const PlanParams p) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= p.nq) return;   // whole wave
    const uint64_t *sel = p.probe_sel + (size_t)q * p.npr;
    const uint32_t c = p.probe_cnt[q];
    uint32_t kept = c;
    if (c > 0) {
        const float d0 = key_to_float((uint32_t)(sel[0] >> 32));
        const float ratio = p.probe_ratio;
        const bool cut = isfinite(ratio) && d0 >= 0.0f;
        const float lim = d0 * fmaxf(ratio, 1.0f);
        if (cut)
            for (uint32_t j0 = 0; j0 < c; j0 += 64) {
                const uint32_t j = j0 + lane;
                const bool beyond = j < c && key_to_float((uint32_t)(sel[j] >> 32)) > lim;
                const uint64_t m = __ballot(beyond);
                if (m) {
                    kept = j0 + (uint32_t)(__ffsll((unsigned long long)m) - 1);
                    break;
                }
            }
    }
    // slots and entries of the probes before this round; the first probe of
    // the second phase (two-phase plans) and the slots before it
    uint32_t run = 0, real = 0, m1 = kept, s1 = 0;
    for (uint32_t j0 = 0; j0 < kept; j0 += 64) {
        const uint32_t j = j0 + lane;
        uint32_t len = 0, rl = 0, post = 0;
        if (j < kept) {
            post = (uint32_t)sel[j];
            rl = p.post_len[post];
            len = (rl + p.align - 1) / p.align * p.align;   // slots reserved
            p.probe_post[(size_t)q * p.npr + j] = post;
        }
        uint32_t incl = len, rin = rl;
#pragma unroll
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, 64), r = __shfl_up(rin, off, 64);
            if (lane >= off) {
                incl += t;
                rin += r;
            }
        }
        // second phase: the probes before j already hold `want` entries
        const bool second = p.split && j < kept && real + rin - rl >= p.want;
        if (j < kept) {
            const size_t i = (size_t)q * p.npr + j;
            p.qp_rel[i] = run + incl - len;
            if (!p.qb) p.prank[i] = atomicAdd(&p.post_cnt[post + (second ? p.split : 0u)], 1u);
        }
        const uint64_t sm = __ballot(second);
        if (m1 == kept && sm) {
            const uint32_t f = (uint32_t)(__ffsll((unsigned long long)sm) - 1);
            m1 = j0 + f;
            s1 = run + __shfl(incl - len, f, 64);
        }
        run += __shfl(incl, 63, 64);
        real += __shfl(rin, 63, 64);
    }
    if (m1 == kept) s1 = run;
    if (lane == 0) {
        if (p.cand_real) p.cand_real[q] = real;
        if (p.split) {
            p.m1[q] = m1;
            p.s1[q] = s1;
        }
        p.np_q[q] = kept;
        p.cand_cnt[q] = run;
        p.cand_len[q] = (run + 31) & ~31ull;   // 32-aligned segments: a query owns whole bitmap words
    }
}

// posting-major pairs, the absolute candidate slot of each (query, probe)
// pair (one thread per pair)
__global__ void k_scatter_pairs(const PlanParams p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nq * p.npr) return;
    const uint32_t q = i / p.npr, j = i - q * p.npr;
    if (j >= p.np_q[q]) return;
    const uint64_t slot = p.cand_off[q] + p.qp_rel[i];
    // (two phases: second-phase probes rank under post + split)
    const uint32_t post = p.probe_post[i] + (p.split && j >= p.m1[q] ? p.split : 0u);
    const uint32_t pos = p.pair_off[post] + p.prank[i] +
                         (p.qb ? p.blk_cnt[(size_t)(q / p.qb) * p.n_post + post] : 0u);
    p.qp_slot[i] = slot;
    p.pair_q[pos] = q;
    p.pair_slot[pos] = slot;
    if (p.pair_src) p.pair_src[pos] = i;
}

// Pair ranks within their posting, by query block: the block's pairs take
// ranks from LDS counters (one per posting), and the block's counts go to
// blk_cnt[block][posting].  Hot postings cost LDS atomics, not serialised
// device-wide atomics.
__global__ __launch_bounds__(1024) void k_pair_rank(const PlanParams p) {
    extern __shared__ uint32_t s_h[];   // n_post counters
    const uint32_t tid = threadIdx.x, b = blockIdx.x;
    for (uint32_t i = tid; i < p.n_post; i += 1024) s_h[i] = 0u;
    __syncthreads();
    const uint32_t q0 = b * p.qb, q1 = min(q0 + p.qb, p.nq);
    for (uint32_t i = q0 * p.npr + tid; i < q1 * p.npr; i += 1024) {
        const uint32_t q = i / p.npr, j = i - q * p.npr;
        if (j < p.np_q[q])
            p.prank[i] = atomicAdd(&s_h[p.probe_post[i] + (p.split && j >= p.m1[q] ? p.split : 0u)], 1u);
    }
    __syncthreads();
    uint32_t *row = p.blk_cnt + (size_t)b * p.n_post;
    for (uint32_t i = tid; i < p.n_post; i += 1024) row[i] = s_h[i];
}

// Per posting (one thread each): blk_cnt's column becomes exclusive over the
// blocks and post_cnt its total.
__global__ void k_blk_scan(const PlanParams p, uint32_t nblk) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.n_post) return;
    uint32_t *col = p.blk_cnt + b;
    const size_t st = p.n_post;
    uint32_t run = 0, k = 0;
    for (; k + 8 <= nblk; k += 8) {   // eight loads in flight
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = col[(k + i) * st];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            col[(k + i) * st] = run;
            run += v[i];
        }
    }
    for (; k < nblk; k++) {
        const uint32_t v = col[k * st];
        col[k * st] = run;
        run += v;
    }
    p.post_cnt[b] = run;
}

__global__ void k_items(const PlanParams p) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= p.n_post) return;
    const uint32_t cnt = p.pair_off[b + 1] - p.pair_off[b];
    const uint32_t len = p.post_len[p.split && b >= p.split ? b - p.split : b];
    p.item_cnt[b] = (cnt && len) ? ((cnt + p.G - 1) / p.G) * ((len + p.echunk - 1) / p.echunk) : 0u;
}

// work item -> posting, one thread per item: the posting b with
// item_off[b] <= i < item_off[b + 1] (a long posting's items spread over threads)
__global__ void k_item_table(const PlanParams p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (p.n_items_dev ? *p.n_items_dev : p.n_items)) return;
    uint32_t lo = 0, hi = p.n_post;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (p.item_off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    p.item_post[i] = lo;
}

// Exclusive scan by one 1024-thread workgroup: thread t takes E consecutive
// values per pass (all loads in flight), so up to 1024 * E values take one
// pass; out[n] = total.
template <typename T, int E>
__global__ __launch_bounds__(1024) void k_excl_scan(const T *in, T *out, uint32_t n) {
    __shared__ T s_w[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T carry = 0;
    for (uint32_t base = 0; base < n; base += 1024 * E) {
        const uint32_t i0 = base + tid * E;
        T v[E], sum = 0;
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = i0 + e < n ? in[i0 + e] : (T)0;
#pragma unroll
        for (int e = 0; e < E; e++) sum += v[e];
        T x = sum;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const T y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        __syncthreads();
        T before = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < 16; w++) {
            const T t = s_w[w];
            before += w < wave ? t : (T)0;
            total += t;
        }
        T run = carry + before + x - sum;
#pragma unroll
        for (int e = 0; e < E; e++) {
            if (i0 + e < n) out[i0 + e] = run;
            run += v[e];
        }
        carry += total;
        __syncthreads();   // s_w is rewritten by the next pass
    }
    if (tid == 0) out[n] = carry;
}

// ---------------------------------------------------------------- rerank
// Exact f32 distance with four lanes per candidate: lane r of the quad runs
// chain r over units u = 4t + r (dims 16t+4r..16t+4r+3), so the quad reads 64
// contiguous bytes per step; chains combine as (c0 + c1) + (c2 + c3).
__device__ __forceinline__ float quad_combine(float c) {
    const float c01 = c + __shfl_xor(c, 1, 64);       // lanes 0,1: c0+c1; lanes 2,3: c2+c3
    return c01 + __shfl_xor(c01, 2, 64);               // (c0+c1) + (c2+c3) on every lane
}

__device__ __forceinline__ float quad_exact(uint32_t metric, const float4 *q4, const float4 *x4,
                                            uint32_t units, uint32_t r, float qn) {
    float acc = 0.0f, b2 = 0.0f;
    const bool l2 = metric == M_L2 || metric == M_L2SQ;
    for (uint32_t u = r; u < units; u += 4) {
        const float4 a = q4[u], b = x4[u];
        if (l2) {
            float t = a.x - b.x; acc = __builtin_fmaf(t, t, acc);
            t = a.y - b.y; acc = __builtin_fmaf(t, t, acc);
            t = a.z - b.z; acc = __builtin_fmaf(t, t, acc);
            t = a.w - b.w; acc = __builtin_fmaf(t, t, acc);
        } else {
            acc = __builtin_fmaf(a.x, b.x, acc);
            acc = __builtin_fmaf(a.y, b.y, acc);
            acc = __builtin_fmaf(a.z, b.z, acc);
            acc = __builtin_fmaf(a.w, b.w, acc);
            b2 = __builtin_fmaf(b.x, b.x, b2);
            b2 = __builtin_fmaf(b.y, b.y, b2);
            b2 = __builtin_fmaf(b.z, b.z, b2);
            b2 = __builtin_fmaf(b.w, b.w, b2);
        }
    }
    acc = quad_combine(acc);
    switch (metric) {
    case M_L2SQ: return acc;
    case M_L2: return sqrtf(acc);
    case M_DOT: return 1.0f - acc;
    default:
        b2 = quad_combine(b2);
        return cosine_from(acc, qn, b2);
    }
}

__global__ __launch_bounds__(256) void k_rerank(const RerankParams p) {
    __shared__ uint64_t buf[kMaxK];
    extern __shared__ __align__(16) float s_qv[];   // dpad floats
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint32_t c = p.sel_cnt[q];
    const uint64_t *sel = p.sel + (size_t)q * p.want;
    for (uint32_t i = tid; i < p.dpad; i += kWG) s_qv[i] = p.qv[(size_t)q * p.dpad + i];
    const float qn = p.qnorm[q];
    __syncthreads();
    const uint32_t r = tid & 3, units = p.dpad / 4;
    const uint32_t nround = (c + 63) / 64;          // 64 quads per workgroup
    for (uint32_t rd = 0; rd < nround; rd++) {
        const uint32_t i = rd * 64 + (tid >> 2);
        const bool ok = i < c;
        const uint32_t rank = ok ? (uint32_t)sel[i] : 0u;
        const uint32_t e = ok ? p.rank_entry[rank] : 0u;
        const float d = quad_exact(p.metric, reinterpret_cast<const float4 *>(s_qv),
                                   reinterpret_cast<const float4 *>(p.raw + (size_t)e * p.dpad),
                                   units, r, qn);
        if (ok && r == 0) buf[i] = ((uint64_t)fkey(d) << 32) | rank;
    }
    const uint32_t m2 = next_pow2(c < 2 ? 2 : c);
    for (uint32_t i = c + tid; i < m2; i += kWG) buf[i] = kNoKey;
    __syncthreads();
    bitonic_sort_lds(buf, m2);
    const uint32_t kk = min(p.k, c);
    for (uint32_t i = tid; i < kk; i += kWG) p.out[(size_t)q * p.k + i] = buf[i];
    if (tid == 0) p.out_cnt[q] = kk;
}

__global__ void k_finalize(const uint64_t *keys, uint32_t stride, const uint32_t *cnt,
                           uint32_t nq, uint32_t k, const uint64_t *rank_id, uint64_t *out_ids,
                           float *out_dists, uint32_t *out_counts) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)nq * k) return;
    const uint32_t q = (uint32_t)(i / k), j = (uint32_t)(i - (uint64_t)q * k);
    const uint32_t c = min(cnt[q], k);
    if (j < c) {
        const uint64_t key = keys[(size_t)q * stride + j];
        out_ids[i] = rank_id[(uint32_t)key];
        out_dists[i] = key_to_float((uint32_t)(key >> 32));
    } else {
        out_ids[i] = ~0ull;
        out_dists[i] = __builtin_inff();
    }
    if (j == 0) out_counts[q] = c;
}

// |x|^2 of every stored row in its decoded form (canonical order), and the
// exact sum of squared u8 codes
__global__ void k_entry_norms(uint32_t ck, const uint4 *data, uint32_t units, uint32_t n,
                              uint32_t dim, float scale, float offset, float *xnorm,
                              uint32_t *xsq) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint4 *src = data + (size_t)(e >> 6) * units * 64 + (e & 63);
    float c4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t si = 0, d = 0;
    for (uint32_t u = 0; u < units; u++) {
        const uint4 w = src[(size_t)u * 64];
        if (ck == CK_F32) {
            const float x[4] = {__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z),
                                __uint_as_float(w.w)};
            for (int j = 0; j < 4 && d < dim; j++, d++)
                c4[(d >> 2) & 3] = __builtin_fmaf(x[j], x[j], c4[(d >> 2) & 3]);
        } else if (ck == CK_F16) {
            union { uint4 v; _Float16 h[8]; } cv;
            cv.v = w;
            for (int j = 0; j < 8 && d < dim; j++, d++) {
                const float x = (float)cv.h[j];
                c4[(d >> 2) & 3] = __builtin_fmaf(x, x, c4[(d >> 2) & 3]);
            }
        } else {
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
            for (int j = 0; j < 16 && d < dim; j++, d++) {
                const uint32_t c = (ws[j >> 2] >> (8 * (j & 3))) & 0xffu;
                si += c * c;
                float x = (float)c * scale;
                x = x + offset;
                c4[(d >> 2) & 3] = __builtin_fmaf(x, x, c4[(d >> 2) & 3]);
            }
        }
    }
    if (xnorm) xnorm[e] = (c4[0] + c4[1]) + (c4[2] + c4[3]);
    if (xsq) xsq[e] = si;
}

// ---------------------------------------------------------------- launchers
hipError_t launch_prep(const PrepParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prep, dim3(p.nq), dim3(64), p.dpad * sizeof(float), s, p);
    return hipGetLastError();
}

template <int CK, int MET, int G>
static hipError_t scan_mode(bool dense, const ScanParams &p, uint32_t grid, hipStream_t s) {
    const size_t lds = (size_t)G * p.dpad * sizeof(float);
    if (dense)
        hipLaunchKernelGGL((k_scan<CK, MET, G, true>), dim3(grid), dim3(kWG), lds, s, p);
    else
        hipLaunchKernelGGL((k_scan<CK, MET, G, false>), dim3(grid), dim3(kWG), lds, s, p);
    return hipGetLastError();
}

template <int CK, int G>
static hipError_t scan_metric(uint32_t metric, bool dense, const ScanParams &p, uint32_t grid,
                              hipStream_t s) {
    switch (metric) {
    case M_L2: return scan_mode<CK, M_L2, G>(dense, p, grid, s);
    case M_L2SQ: return scan_mode<CK, M_L2SQ, G>(dense, p, grid, s);
    case M_COS: return scan_mode<CK, M_COS, G>(dense, p, grid, s);
    case M_DOT: return scan_mode<CK, M_DOT, G>(dense, p, grid, s);
    }
    return hipErrorInvalidValue;
}

template <int G>
static hipError_t scan_kind(uint32_t ck, uint32_t metric, bool dense, const ScanParams &p,
                            uint32_t grid, hipStream_t s) {
    switch (ck) {
    case CK_F32: return scan_metric<CK_F32, G>(metric, dense, p, grid, s);
    case CK_F16: return scan_metric<CK_F16, G>(metric, dense, p, grid, s);
    case CK_U8: return scan_metric<CK_U8, G>(metric, dense, p, grid, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_scan(uint32_t ck, uint32_t metric, bool dense, uint32_t G, const ScanParams &p,
                       uint32_t grid, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    if (G == 16) return scan_kind<16>(ck, metric, dense, p, grid, s);
    if (G == 8) return scan_kind<8>(ck, metric, dense, p, grid, s);
    if (G == 4) return scan_kind<4>(ck, metric, dense, p, grid, s);
    return hipErrorInvalidValue;
}

hipError_t launch_select(const uint64_t *keys, const uint64_t *seg_off, const uint32_t *seg_cnt,
                         uint64_t seg_stride, uint32_t nseg, uint32_t K, uint64_t *out,
                         uint32_t *out_cnt, hipStream_t s) {
    if (nseg == 0) return hipSuccess;
    return select_bands(keys, seg_off, seg_cnt, seg_stride, nseg, K, out, out_cnt, ~0ull, s);
}

hipError_t launch_select_long(const uint64_t *keys, const uint64_t *seg_off,
                              const uint32_t *seg_cnt, uint64_t seg_stride, uint32_t nseg,
                              uint32_t chunk, uint32_t K, uint64_t *out, uint32_t *out_cnt,
                              uint32_t *chunk_cnt, uint32_t *chunk_off, uint64_t *tmp,
                              uint64_t max_chunks, hipStream_t s) {
    if (nseg == 0) return hipSuccess;
    hipError_t e = select_bands(keys, seg_off, seg_cnt, seg_stride, nseg, K, out, out_cnt, chunk, s);
    if (e != hipSuccess || max_chunks == 0) return e;
    hipLaunchKernelGGL(k_sel_chunk_count, dim3((nseg + 255) / 256), dim3(256), 0, s, seg_off,
                       seg_cnt, seg_stride, nseg, chunk, chunk_cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_excl_scan_u32(chunk_cnt, chunk_off, nseg, s)) != hipSuccess) return e;
    if (chunk <= 16 * 256)
        hipLaunchKernelGGL(k_select_chunk<16>, dim3((uint32_t)max_chunks), dim3(kWG), 0, s, keys,
                           seg_off, seg_cnt, seg_stride, chunk_off, nseg, chunk, K, tmp);
    else if (chunk <= 96 * 256)
        hipLaunchKernelGGL(k_select_chunk<96>, dim3((uint32_t)max_chunks), dim3(kWG), 0, s, keys,
                           seg_off, seg_cnt, seg_stride, chunk_off, nseg, chunk, K, tmp);
    else
        hipLaunchKernelGGL(k_select_chunk<0>, dim3((uint32_t)max_chunks), dim3(kWG), 0, s, keys,
                           seg_off, seg_cnt, seg_stride, chunk_off, nseg, chunk, K, tmp);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_select_merge, dim3(nseg), dim3(kWG), 0, s, tmp, chunk_off, K, out, out_cnt);
    return hipGetLastError();
}

// nearest posting of each vector: probe key -> (block index, distance)
__global__ void k_nearest_out(const uint64_t *probe_sel, const uint32_t *probe_cnt, uint32_t n,
                              const uint32_t *block_sorted, uint32_t *out_block, float *out_dist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (probe_cnt[i] == 0) {
        out_block[i] = 0xffffffffu;
        out_dist[i] = __builtin_inff();
        return;
    }
    const uint64_t key = probe_sel[i];
    out_block[i] = block_sorted[(uint32_t)key];
    out_dist[i] = key_to_float((uint32_t)(key >> 32));
}

hipError_t launch_nearest_out(const uint64_t *probe_sel, const uint32_t *probe_cnt, uint32_t n,
                              const uint32_t *block_sorted, uint32_t *out_block, float *out_dist,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nearest_out, dim3((n + 255) / 256), dim3(256), 0, s, probe_sel, probe_cnt,
                       n, block_sorted, out_block, out_dist);
    return hipGetLastError();
}

hipError_t launch_plan_q(const PlanParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_q, dim3((p.nq + 3) / 4), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_scatter_pairs(const PlanParams &p, hipStream_t s) {
    const uint32_t n = p.nq * p.npr;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_pairs, dim3((n + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_items(const PlanParams &p, hipStream_t s) {
    hipLaunchKernelGGL(k_items, dim3((p.n_post + 2 + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_item_table(const PlanParams &p, hipStream_t s) {
    if (p.n_items == 0) return hipSuccess;
    hipLaunchKernelGGL(k_item_table, dim3((p.n_items + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_excl_scan_u64(const uint64_t *in, uint64_t *out, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_excl_scan<unsigned long long, 16>), dim3(1), dim3(1024), 0, s,
                       (const unsigned long long *)in, (unsigned long long *)out, n);
    return hipGetLastError();
}

hipError_t launch_excl_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_excl_scan<uint32_t, 16>), dim3(1), dim3(1024), 0, s, in, out, n);
    return hipGetLastError();
}

// k_pair_rank + k_blk_scan (p.qb > 0; k_plan_q has written probe_post / np_q)
hipError_t launch_pair_rank(const PlanParams &p, hipStream_t s) {
    if (!p.qb || p.nq == 0) return hipSuccess;
    if (p.n_post > kRankMaxPost) return hipErrorInvalidValue;
    const uint32_t nblk = (p.nq + p.qb - 1) / p.qb;
    hipLaunchKernelGGL(k_pair_rank, dim3(nblk), dim3(1024), p.n_post * sizeof(uint32_t), s, p);
    hipLaunchKernelGGL(k_blk_scan, dim3((p.n_post + 255) / 256), dim3(256), 0, s, p, nblk);
    return hipGetLastError();
}

hipError_t launch_rerank(const RerankParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rerank, dim3(p.nq), dim3(kWG), p.dpad * sizeof(float), s, p);
    return hipGetLastError();
}

hipError_t launch_finalize(const uint64_t *keys, uint32_t stride, const uint32_t *cnt, uint32_t nq,
                           uint32_t k, const uint64_t *rank_id, uint64_t *out_ids,
                           float *out_dists, uint32_t *out_counts, hipStream_t s) {
    const uint64_t n = (uint64_t)nq * k;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, keys,
                       stride, cnt, nq, k, rank_id, out_ids, out_dists, out_counts);
    return hipGetLastError();
}

hipError_t launch_entry_norms(uint32_t ck, const uint4 *data, uint32_t units, uint32_t n,
                              uint32_t dim, float scale, float offset, float *xnorm,
                              uint32_t *xsq, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_entry_norms, dim3((n + 255) / 256), dim3(256), 0, s, ck, data, units, n,
                       dim, scale, offset, xnorm, xsq);
    return hipGetLastError();
}

}  // namespace spf
