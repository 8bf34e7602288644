// RaBitQ postings on gfx950 (src/rabitq.rs, src/spfresh.rs:731-770).
//
//   k_rq_prep         query_with(q, centroid_b) for every (query, probed posting)
//                     pair: residual, norm (sequential for the exact path),
//                     3 x (sign flip + FHT) in registers and lane shuffles with
//                     the reference's butterfly pairing (so every element is the
//                     same IEEE sum), 4-bit quantisation, bit planes by wave
//                     ballot, estimator constants
//   k_rq_search_exact search_rabitq semantics: one wave per query walks its
//                     probes and entries in scan order; estimates, bounds and
//                     exact distances run in parallel per 64-entry chunk, heap
//                     updates are applied in entry order (the bound test of an
//                     entry sees every earlier update, as in the reference)
//   k_rq_decode       decode_with for every entry (cosine / dot postings are
//                     searched on the decoded f32 vectors, src/spfresh.rs:346-348)
//   k_rq_scan         mixed precision: posting-major estimator scan writing the
//                     16-bit key of each estimate, for pool selection + exact rerank
#include "spf_internal.h"

namespace spf {

__device__ __forceinline__ uint32_t fkey_r(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float_r(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// unnormalised fast Hadamard transform of v[0..n) in LDS, one wave;
// pair (j, j+h) of stage h is the reference's (src/rabitq.rs:31-43)
__device__ __forceinline__ void wave_fht(float *v, uint32_t n) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t h = 1; h < n; h *= 2) {
        for (uint32_t idx = lane; idx < n / 2; idx += 64) {
            const uint32_t j = (idx / h) * 2 * h + (idx % h);
            const float x = v[j], y = v[j + h];
            v[j] = x + y;
            v[j + h] = x - y;
        }
        wave_sync_lds();
    }
}

// Lane l's partner l ^ H: DPP within a quad (H = 1, 2) or a 16-lane row
// rotated by 8 (H = 8); the other stages through the LDS crossbar (measured:
// two rotations and a select for H = 4, or the row-swap permutes for 16 and
// 32, were slower than the shuffle)
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <uint32_t H>
__device__ __forceinline__ float xor_lane(float v) {
    if constexpr (H == 1) return dpp_f<0xB1>(v);    // quad_perm [1,0,3,2]
    else if constexpr (H == 2) return dpp_f<0x4E>(v);    // quad_perm [2,3,0,1]
    else if constexpr (H == 8) return dpp_f<0x128>(v);   // row_ror:8
    else return __shfl_xor(v, (int)H, 64);
}
// FHT stage h = H < 64 over lanes: o - r (the lane with bit H) or r + o, as
// one fma: the product by -1 or 1 is exact, so the single rounding is the sum's
template <uint32_t H, uint32_t W>
__device__ __forceinline__ void fht_lanes(float (&r)[W], uint32_t lane) {
    const float sg = (lane & H) ? -1.0f : 1.0f;
#pragma unroll
    for (uint32_t m = 0; m < W; m++) r[m] = __builtin_fmaf(sg, r[m], xor_lane<H>(r[m]));
}

// ---------------------------------------------------------------- query prep
// One wave per (query, probed posting).  Lane l holds elements i = 64 m + l
// (m < W) in registers.  The FHT's butterfly pairs and their IEEE operations
// are the reference's (src/rabitq.rs:31-43): stage h pairs j and j + h with
// v[j] = x + y, v[j + h] = x - y, stages in increasing h.  Stages h < 64 pair
// lane l with lane l ^ h (one cross-lane move per register, DPP or shuffle:
// the lane without bit h adds, the other subtracts its own value from its
// partner's); stages
// h >= 64 pair registers m and m + h / 64 of the same lane.  Which lane
// computes a butterfly does not change its result, so the rotated residual
// is the same as a sequential transform's, bit for bit.
template <uint32_t W>
__global__ __launch_bounds__(64) void k_rq_prep(const RqParams p) {
    constexpr uint32_t P = 64 * W;
    extern __shared__ float s_seq[];   // [P]: the exact path's sequential norm only
    __shared__ float s_norm;
    const uint32_t pair = blockIdx.x, lane = threadIdx.x;
    const uint32_t q = pair / p.npr, j = pair - q * p.npr;
    if (j >= p.np_q[q]) return;
    const uint32_t post = p.probe_post[pair];
    const float *qv = p.qv + (size_t)q * p.dpad;
    const float *c = p.cent + (size_t)post * p.dpad;
    // mixed precision with cosine: the residual of the unit query
    const float qs = p.mixed && p.metric == M_COS
                         ? (p.qnorm[q] > 0.0f ? 1.0f / sqrtf(p.qnorm[q]) : 0.0f) : 1.0f;
    float r[W];
#pragma unroll
    for (uint32_t m = 0; m < W; m++) {
        const uint32_t i = 64 * m + lane;
        r[m] = i < p.dim ? qv[i] * qs - c[i] : 0.0f;
    }
    float norm;
    if (p.mixed) {   // a wave sum (the estimate only orders candidates)
        float acc = 0.0f;
#pragma unroll
        for (uint32_t m = 0; m < W; m++) acc = __builtin_fmaf(r[m], r[m], acc);
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        norm = sqrtf(acc);   // (every lane holds the same sum)
    } else {   // r.iter().map(|x| x * x).sum::<f32>(): sequential, unfused
#pragma unroll
        for (uint32_t m = 0; m < W; m++) s_seq[64 * m + lane] = r[m];
        wave_sync_lds();
        if (lane == 0) {
            float acc = 0.0f;
            for (uint32_t i = 0; i < P; i++) {
                const float sq = s_seq[i] * s_seq[i];
                acc = acc + sq;
            }
            s_norm = sqrtf(acc);
        }
        wave_sync_lds();
        norm = s_norm;
        wave_sync_lds();   // (the next pair rewrites s_seq / s_norm)
    }
    if (norm > 0.0f)
#pragma unroll
        for (uint32_t m = 0; m < W; m++) r[m] = r[m] / norm;
    for (uint32_t rd = 0; rd < 3; rd++) {
#pragma unroll
        for (uint32_t m = 0; m < W; m++) r[m] = r[m] * p.signs[rd * P + 64 * m + lane];
        fht_lanes<1, W>(r, lane);
        fht_lanes<2, W>(r, lane);
        fht_lanes<4, W>(r, lane);
        fht_lanes<8, W>(r, lane);
        fht_lanes<16, W>(r, lane);
        fht_lanes<32, W>(r, lane);
#pragma unroll
        for (uint32_t g = 1; g < W; g *= 2)
#pragma unroll
            for (uint32_t m = 0; m < W; m++)
                if ((m & g) == 0) {
                    const float x = r[m], y = r[m + g];
                    r[m] = x + y;
                    r[m + g] = x - y;
                }
    }
    float lo = 3.402823466e+38f, hi = -3.402823466e+38f;
#pragma unroll
    for (uint32_t m = 0; m < W; m++) {
        lo = fminf(lo, r[m]);
        hi = fmaxf(hi, r[m]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
    }
    const float levels = 15.0f;
    const float delta = (hi - lo) / levels;
    const float inv = delta > 0.0f ? 1.0f / delta : 0.0f;
    uint64_t *planes = p.planes + (size_t)pair * 4 * p.words;
    uint32_t sum_u = 0;
#pragma unroll
    for (uint32_t m = 0; m < W; m++) {   // element i = 64 m + lane -> bit `lane` of word m
        const float t = roundf((r[m] - lo) * inv);
        uint32_t u = t > 0.0f ? (t >= 4294967295.0f ? 0xffffffffu : (uint32_t)t) : 0u;
        u = u > 15u ? 15u : u;
        sum_u += u;
        if (p.levels) p.levels[(size_t)pair * P + 64 * m + lane] = (uint8_t)u;
#pragma unroll
        for (uint32_t b = 0; b < 4; b++) {
            const uint64_t bits = __ballot((u >> b) & 1u);
            if (lane == 0) planes[b * W + m] = bits;
        }
    }
    for (int off = 32; off > 0; off >>= 1) sum_u += __shfl_xor(sum_u, off, 64);
    if (lane == 0) {
        const float sd = sqrtf((float)P);
        const float a = delta * (float)sum_u;
        const float b = (float)P * lo;
        float *cs = p.consts + (size_t)pair * 8;
        cs[0] = norm * norm;
        cs[1] = 2.0f * norm;
        cs[2] = 2.0f * delta / sd;
        cs[3] = 2.0f * lo / sd;
        cs[4] = -(a + b) / sd;
    }
}

// estimate and error radius of one code against prepared planes
// (distance_with_bound, src/rabitq.rs:202-215); bits in 64-bit words
template <uint32_t W>
__device__ __forceinline__ void rq_est(const uint64_t *code_words, float n2, float fac,
                                       const uint64_t *planes, const float *cs, uint32_t P,
                                       float &est, float &err) {
    uint32_t S = 0, C = 0;
#pragma unroll
    for (uint32_t i = 0; i < W; i++) {
        const uint64_t b = code_words[i];
        C += __popcll(b);
        S += __popcll(b & planes[i]) + (__popcll(b & planes[W + i]) << 1) +
             (__popcll(b & planes[2 * W + i]) << 2) + (__popcll(b & planes[3 * W + i]) << 3);
    }
    const float nq2 = cs[0], sq = cs[1], k1 = cs[2], k2 = cs[3], k3 = cs[4];
    float inner = k1 * (float)S + k2 * (float)C;
    inner = inner + k3;
    const float sf = sq * fac;
    const float e = (n2 + nq2) - sf * inner;
    est = fmaxf(e, 0.0f);
    const float ip = fac > 0.0f ? sqrtf(n2) / fac : 0.0f;
    if (ip > 0.0f) {
        const float one = fmaxf(1.0f - ip * ip, 0.0f);
        const float rad = sqrtf(one / ((float)P - 1.0f));
        err = sf * 1.9f * rad;
    } else {
        err = __builtin_inff();
    }
}

// code of entry e from the tiled layout: bit words and (n2, fac)
template <uint32_t W>
__device__ __forceinline__ void rq_load_code(const uint4 *codes, uint32_t units, uint32_t e,
                                             uint64_t *words, float &n2, float &fac) {
    const uint4 *src = codes + (size_t)(e >> 6) * units * 64 + (e & 63);
#pragma unroll
    for (uint32_t u = 0; u < W / 2; u++) {
        const uint4 v = src[(size_t)u * 64];
        words[2 * u] = ((uint64_t)v.y << 32) | v.x;
        words[2 * u + 1] = ((uint64_t)v.w << 32) | v.z;
    }
    if (W & 1) {   // padded = 64: one word, then n2/fac in the same unit
        const uint4 v = src[0];
        words[0] = ((uint64_t)v.y << 32) | v.x;
        n2 = __uint_as_float(v.z);
        fac = __uint_as_float(v.w);
        return;
    }
    const uint4 v = src[(size_t)(W / 2) * 64];
    n2 = __uint_as_float(v.x);
    fac = __uint_as_float(v.y);
}

// exact squared L2 of a raw row, canonical 4-chain order
__device__ __forceinline__ float l2sq_row(const float *q, const float *x, uint32_t dpad) {
    float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const float4 *q4 = reinterpret_cast<const float4 *>(q);
    const float4 *x4 = reinterpret_cast<const float4 *>(x);
    for (uint32_t u = 0; u < dpad / 4; u += 4) {
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) {
            const float4 a = q4[u + r], b = x4[u + r];
            float t = a.x - b.x; c[r] = __builtin_fmaf(t, t, c[r]);
            t = a.y - b.y; c[r] = __builtin_fmaf(t, t, c[r]);
            t = a.z - b.z; c[r] = __builtin_fmaf(t, t, c[r]);
            t = a.w - b.w; c[r] = __builtin_fmaf(t, t, c[r]);
        }
    }
    return (c[0] + c[1]) + (c[2] + c[3]);
}

// ---------------------------------------------------------------- exact search
constexpr uint32_t kRqMaxK = 1024;

template <uint32_t W>
__global__ __launch_bounds__(64) void k_rq_search_exact(const RqParams p) {
    __shared__ uint64_t s_heap[kRqMaxK];
    __shared__ uint64_t s_pl[4 * W];
    __shared__ float s_cs[8];
    __shared__ uint32_t s_cnt;
    const uint32_t q = blockIdx.x, lane = threadIdx.x;
    const uint32_t P = p.padded, k = p.k;
    const float *qv = p.qv + (size_t)q * p.dpad;
    if (lane == 0) s_cnt = 0;
    wave_sync_lds();
    float kth = __builtin_inff();
    const uint32_t np = p.np_q[q];
    const uint64_t qm = p.qmask ? p.qmask[q] : ~0ull;
    for (uint32_t jp = 0; jp < np; jp++) {
        const uint32_t pair = q * p.npr + jp;
        const uint32_t post = p.probe_post[pair];
        for (uint32_t i = lane; i < 4 * W; i += 64) s_pl[i] = p.planes[(size_t)pair * 4 * W + i];
        if (lane < 5) s_cs[lane] = p.consts[(size_t)pair * 8 + lane];
        wave_sync_lds();
        const uint32_t eb = p.post_off[post], ee = p.post_off[post + 1];
        for (uint32_t e0 = eb; e0 < ee; e0 += 64) {
            const uint32_t e = e0 + lane;
            const bool valid = e < ee && !(p.del && ((p.del[e >> 5] >> (e & 31)) & 1u)) &&
                               (!p.qmask || label_ok(qm, p.label_e[e]));
            float lb = __builtin_inff(), exact = 0.0f;
            if (valid) {
                uint64_t words[W];
                float n2, fac, est, err;
                rq_load_code<W>(p.codes, p.units, e, words, n2, fac);
                rq_est<W>(words, n2, fac, s_pl, s_cs, P, est, err);
                lb = est - err;
                if (lb < kth) exact = l2sq_row(qv, p.raw_e + (size_t)e * p.dpad, p.dpad);
            }
            uint64_t mask = __ballot(valid && lb < kth);
            const uint32_t rank = valid ? p.rank[e] : 0u;
            while (mask) {   // heap updates in entry order
                const int src = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                const float lb_i = __shfl(lb, src, 64);
                const float ex_i = __shfl(exact, src, 64);
                const uint32_t rk_i = __shfl(rank, src, 64);
                if (!(lb_i < kth)) continue;
                const uint64_t key = ((uint64_t)fkey_r(ex_i) << 32) | rk_i;
                if (lane == 0) {
                    uint32_t cnt = s_cnt;
                    bool ins = false;
                    if (cnt < k) {
                        s_heap[cnt] = key;
                        cnt++;
                        ins = true;
                    } else if (ex_i < kth) {
                        s_heap[cnt - 1] = key;
                        ins = true;
                    }
                    if (ins)
                        for (uint32_t i = cnt - 1; i > 0 && s_heap[i] < s_heap[i - 1]; i--) {
                            const uint64_t t = s_heap[i];
                            s_heap[i] = s_heap[i - 1];
                            s_heap[i - 1] = t;
                        }
                    s_cnt = cnt;
                }
                wave_sync_lds();
                const uint32_t cnt = s_cnt;
                kth = cnt < k ? __builtin_inff() : key_float_r((uint32_t)(s_heap[cnt - 1] >> 32));
            }
        }
        wave_sync_lds();
    }
    const uint32_t cnt = s_cnt;
    // DistL2 returns the root (src/spfresh.rs:766); re-key and re-sort by (dist, id)
    if (p.metric == M_L2) {
        for (uint32_t i = lane; i < cnt; i += 64) {
            const float v = sqrtf(key_float_r((uint32_t)(s_heap[i] >> 32)));
            s_heap[i] = ((uint64_t)fkey_r(v) << 32) | (uint32_t)s_heap[i];
        }
        wave_sync_lds();
        if (lane == 0)
            for (uint32_t i = 1; i < cnt; i++)
                for (uint32_t t = i; t > 0 && s_heap[t] < s_heap[t - 1]; t--) {
                    const uint64_t x = s_heap[t];
                    s_heap[t] = s_heap[t - 1];
                    s_heap[t - 1] = x;
                }
        wave_sync_lds();
    }
    for (uint32_t i = lane; i < cnt; i += 64) p.out[(size_t)q * k + i] = s_heap[i];
    if (lane == 0) p.out_cnt[q] = cnt;
}

// ---------------------------------------------------------------- decode
// decode_with (src/rabitq.rs:224-238) of entry e into tiled f32 codes
template <uint32_t W>
__global__ __launch_bounds__(64) void k_rq_decode(const RqDecodeParams p) {
    extern __shared__ float s_r[];
    const uint32_t e = blockIdx.x, lane = threadIdx.x;
    const uint32_t P = p.padded;
    uint64_t words[W];
    float n2, fac;
    rq_load_code<W>(p.codes, p.units, e, words, n2, fac);
    const float sd = sqrtf((float)P);
    for (uint32_t i = lane; i < P; i += 64) {
        const uint64_t w = words[i >> 6];
        s_r[i] = ((w >> (i & 63)) & 1u) ? 1.0f / sd : -1.0f / sd;
    }
    wave_sync_lds();
    for (int rd = 2; rd >= 0; rd--) {
        wave_fht(s_r, P);
        for (uint32_t i = lane; i < P; i += 64) s_r[i] = s_r[i] * p.signs[(uint32_t)rd * P + i];
        wave_sync_lds();
    }
    const float norm = sqrtf(n2);
    const float *c = p.cent + (size_t)p.entry_post[e] * p.dpad;
    uint32_t *dst = reinterpret_cast<uint32_t *>(p.out);
    for (uint32_t i = lane; i < p.out_units * 4; i += 64) {
        float v = 0.0f;
        if (i < p.dim) {
            const float t = norm * s_r[i];
            v = c[i] + t;
        }
        // tiled f32 layout: unit i/4 of entry e, word i%4
        dst[(((size_t)(e >> 6) * p.out_units + i / 4) * 64 + (e & 63)) * 4 + (i & 3)] = __float_as_uint(v);
    }
}

// ---------------------------------------------------------------- mixed scan
// Posting-major items of <= 8 queries: lane = entry, popcount estimator per
// (entry, query), 16-bit key of the estimate out.
constexpr uint32_t kRqG = 8;

template <uint32_t W>
__global__ __launch_bounds__(256) void k_rq_scan(const RqParams p) {
    __shared__ uint64_t s_pl[kRqG][4 * W];
    __shared__ float s_cs[kRqG][8];
    __shared__ uint64_t s_slot[kRqG];
    __shared__ uint64_t s_mask[kRqG];
    __shared__ float s_qn[kRqG];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t P = p.padded;
    const bool l2 = p.metric == M_L2 || p.metric == M_L2SQ;
    const uint32_t n_items = p.n_items_dev ? *p.n_items_dev : p.n_items;
    for (XcdWalk it(n_items); it.more(); it.next()) {
        const uint32_t item = it.item;
        const uint32_t pp = p.item_post[item];
        const uint32_t local = item - p.item_off[pp];
        const uint32_t pb = p.post_off[pp], pe = p.post_off[pp + 1];
        const uint32_t nec = (pe - pb + kEChunk - 1) / kEChunk;
        const uint32_t qc = local / nec, ec = local - qc * nec;
        const uint32_t pr0 = p.pair_off[pp] + qc * kRqG;
        const uint32_t gcnt = min(kRqG, p.pair_off[pp + 1] - pr0);
        const uint32_t e_begin = pb + ec * kEChunk, e_end = min(e_begin + kEChunk, pe);
        __syncthreads();
        for (uint32_t i = tid; i < gcnt * 4 * W; i += 256) {
            const uint32_t g = i / (4 * W), w = i - g * 4 * W;
            s_pl[g][w] = p.planes[(size_t)p.pair_src[pr0 + g] * 4 * W + w];
        }
        if (tid < gcnt * 8) {
            const uint32_t g = tid / 8, c = tid - g * 8;
            s_cs[g][c] = p.consts[(size_t)p.pair_src[pr0 + g] * 8 + c];
        }
        if (tid < gcnt) {
            s_slot[tid] = p.pair_slot[pr0 + tid];
            const uint32_t qi = p.pair_src[pr0 + tid] / p.npr;
            if (p.qmask) s_mask[tid] = p.qmask[qi];
            s_qn[tid] = p.metric == M_DOT ? p.qnorm[qi] : 0.0f;
        }
        __syncthreads();
        for (uint32_t e0 = e_begin + 64 * wave; e0 < e_end; e0 += 256) {
            const uint32_t e = e0 + lane;
            if (e >= e_end) continue;
            uint64_t words[W];
            float n2, fac;
            rq_load_code<W>(p.codes, p.units, e, words, n2, fac);
            const bool dead = p.del && ((p.del[e >> 5] >> (e & 31)) & 1u);
            const uint32_t lab = p.qmask ? p.label_e[e] : 0u;
            const float xn = l2 ? 0.0f : p.xn[e];
            const float rx = xn > 0.0f ? 0.5f / sqrtf(xn) : 0.0f;
            for (uint32_t g = 0; g < gcnt; g++) {
                float est, err;
                rq_est<W>(words, n2, fac, s_pl[g], s_cs[g], P, est, err);
                // cosine (unit query): 1 - q.x / |x| with q.x = (1 + |x|^2 - |q - x|^2) / 2;
                // dot: 1 - q.x with q.x = (|q|^2 + |x|^2 - |q - x|^2) / 2
                float d = est;
                if (p.metric == M_COS) d = 1.0f - (1.0f + xn - est) * rx;
                else if (p.metric == M_DOT) d = 1.0f - 0.5f * (s_qn[g] + xn - est);
                const bool skip = dead || (p.qmask && !label_ok(s_mask[g], lab));
                p.out16[s_slot[g] + (e - pb)] = skip ? (uint16_t)0xffff : (uint16_t)(fkey_r(d) >> 16);
            }
        }
    }
}

// ---------------------------------------------------------------- matrix-core estimator
// Items of up to kRqMfmaQ (query, posting) pairs over 64 T entries of one
// posting (T = rq_mfma_tiles(W) tiles of 16 entries per wave).  The
// estimator's integer part S = sum_i bit_i * level_i, which the popcount scan
// sums over four bit planes (src/rabitq.rs:46-58), is an int8 product here:
// code bits unpacked to {0, 1} (B: each lane's 16-bit slice of a 64-dim K
// step) against the pair's 4-bit levels (A), v_mfma_i32_16x16x64_i8 with
// int32 sums.  S is the same integer, so estimates and keys equal the
// popcount scan's bit for bit.
typedef int i32x4r_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t nib_bytes(uint32_t n) {   // bit i of n (i < 4) -> byte i
    return (n * 0x00204081u) & 0x01010101u;
}

// The item's code bits are unpacked once per 64-dim K step and used by every
// 16-pair group of the item: the groups' accumulators (G4 x T) stay live
// across the K loop, which streams one 16-byte code unit (two words) per tile
// and the pairs' level rows per step, the next unit's loads in flight.  The
// keys go through an LDS tile and leave as aligned 16-byte row segments (each
// pair's slot range starts at a multiple of kScanAlign; the posting's last
// item fills the range's tail with "no key").  MET: the metric family of the
// key (0: L2, 1: cosine, 2: dot).
template <uint32_t W, int MET>
__global__ __launch_bounds__(256, 2) void k_rq_scan_mfma(const RqParams p) {
    constexpr uint32_t T = W <= 16 ? 4 : 2;   // rq_mfma_tiles
    constexpr uint32_t EW = 64 * T;
    constexpr uint32_t NU = (W + 1) / 2;      // code units holding bit words
    constexpr uint32_t G4 = kRqMfmaQ / 16;    // 16-pair groups per item
    constexpr uint32_t RS = EW + 8;           // tile row stride (16-byte rows, 4-bank shift per row)
    __shared__ __align__(16) uint16_t s_tile[kRqMfmaQ][RS];
    __shared__ float s_cs[kRqMfmaQ][5];
    __shared__ uint64_t s_slot[kRqMfmaQ];
    __shared__ uint64_t s_mask[kRqMfmaQ];
    __shared__ float s_qn[kRqMfmaQ];
    __shared__ uint32_t s_src[kRqMfmaQ];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t col = lane & 15, kg = lane >> 4;
    const uint32_t P = p.padded;
    const uint32_t n_items = p.n_items_dev ? *p.n_items_dev : p.n_items;
    for (XcdWalk it(n_items); it.more(); it.next()) {
        const uint32_t item = it.item;
        const uint32_t pp = p.item_post[item];
        const uint32_t local = item - p.item_off[pp];
        const uint32_t pb = p.post_off[pp], pe = p.post_off[pp + 1];
        const uint32_t nec = (pe - pb + EW - 1) / EW;
        const uint32_t qc = local / nec, ec = local - qc * nec;
        const uint32_t pr0 = p.pair_off[pp] + qc * kRqMfmaQ;
        const uint32_t gcnt = min(kRqMfmaQ, p.pair_off[pp + 1] - pr0);
        const uint32_t e_begin = pb + ec * EW, e_end = min(e_begin + EW, pe);
        __syncthreads();   // the previous item is done with the pair table
        if (tid < kRqMfmaQ) {
            const bool ok = tid < gcnt;
            const uint32_t src = ok ? p.pair_src[pr0 + tid] : 0u;
            const uint32_t qi = src / p.npr;
            s_src[tid] = src;
            s_slot[tid] = ok ? p.pair_slot[pr0 + tid] : 0ull;
            s_mask[tid] = (ok && p.qmask) ? p.qmask[qi] : 0ull;
            s_qn[tid] = (ok && MET == 2) ? p.qnorm[qi] : 0.0f;
        }
        for (uint32_t i = tid; i < kRqMfmaQ * 5; i += 256) {
            const uint32_t g = i / 5, c = i - g * 5;
            s_cs[g][c] = g < gcnt ? p.consts[(size_t)p.pair_src[pr0 + g] * 8 + c] : 0.0f;
        }
        // the wave's tiles: code unit pointers, |r|^2 and factor, liveness,
        // label and |x|^2 of its column's entry (entries past the posting
        // read e_begin)
        const uint4 *csrc[T];
        float n2[T], fac[T], xn[T];
        uint32_t lab[T];
        bool live[T];
#pragma unroll
        for (uint32_t t = 0; t < T; t++) {
            const uint32_t e_raw = e_begin + 16 * (wave + 4 * t) + col;
            const uint32_t e = e_raw < e_end ? e_raw : e_begin;
            csrc[t] = p.codes + (size_t)(e >> 6) * p.units * 64 + (e & 63);
            if (W & 1) {   // padded = 64: one word, then n2 / fac in the same unit
                const uint4 v = csrc[t][0];
                n2[t] = __uint_as_float(v.z);
                fac[t] = __uint_as_float(v.w);
            } else {
                const uint4 v = csrc[t][(size_t)(W / 2) * 64];
                n2[t] = __uint_as_float(v.x);
                fac[t] = __uint_as_float(v.y);
            }
            const bool dead = p.del && ((p.del[e >> 5] >> (e & 31)) & 1u);
            live[t] = e_raw < e_end && !dead;
            lab[t] = p.qmask ? p.label_e[e] : 0u;
            xn[t] = MET == 0 ? 0.0f : p.xn[e];
        }
        __syncthreads();
        const uint32_t ng = (gcnt + 15) / 16;   // groups in use (uniform)
        // the level row of this lane's pair in every group (A: pair 16 g + col)
        const uint4 *arow[G4];
#pragma unroll
        for (uint32_t g = 0; g < G4; g++) {
            const uint32_t pr = 16 * g + col;
            arow[g] = reinterpret_cast<const uint4 *>(p.levels + (size_t)s_src[pr < gcnt ? pr : 0] * P);
        }
        i32x4r_t acc[G4][T];
#pragma unroll
        for (uint32_t g = 0; g < G4; g++)
#pragma unroll
            for (uint32_t t = 0; t < T; t++) acc[g][t] = (i32x4r_t){0, 0, 0, 0};
        uint32_t cnt[T] = {};
        // unit u: words 2u, 2u + 1 of every tile's entry; levels of steps 2u, 2u + 1
        uint4 cu[T], a[G4][2];
        auto load_unit = [&](uint32_t u, uint4 (&c)[T], uint4 (&l)[G4][2]) {
#pragma unroll
            for (uint32_t t = 0; t < T; t++) c[t] = csrc[t][(size_t)u * 64];
#pragma unroll
            for (uint32_t g = 0; g < G4; g++)
#pragma unroll
                for (uint32_t h = 0; h < 2; h++)
                    if (g < ng && 2 * u + h < W) l[g][h] = arow[g][4 * (2 * u + h) + kg];
        };
        load_unit(0, cu, a);
#pragma unroll 1
        for (uint32_t u = 0; u < NU; u++) {
            uint4 cn[T], an[G4][2];
            if (u + 1 < NU) load_unit(u + 1, cn, an);
#pragma unroll
            for (uint32_t h = 0; h < 2; h++) {
                if (2 * u + h >= W) break;
                i32x4r_t bt[T];
#pragma unroll
                for (uint32_t t = 0; t < T; t++) {
                    const uint32_t lo = h ? cu[t].z : cu[t].x, hi = h ? cu[t].w : cu[t].y;
                    cnt[t] += __popc(lo) + __popc(hi);
                    // the lane's 16-bit slice of the word: bits 16 kg .. 16 kg + 15
                    const uint32_t v = ((kg < 2 ? lo : hi) >> (16 * (kg & 1))) & 0xffffu;
                    bt[t] = (i32x4r_t){(int)nib_bytes(v & 0xfu), (int)nib_bytes((v >> 4) & 0xfu),
                                       (int)nib_bytes((v >> 8) & 0xfu), (int)nib_bytes(v >> 12)};
                }
#pragma unroll
                for (uint32_t g = 0; g < G4; g++) {
                    if (g >= ng) break;
                    union { uint4 u; i32x4r_t i; } av;
                    av.u = a[g][h];   // levels 64 s + 16 kg .. + 15
#pragma unroll
                    for (uint32_t t = 0; t < T; t++)
                        acc[g][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av.i, bt[t], acc[g][t], 0, 0, 0);
                }
            }
            if (u + 1 < NU) {
#pragma unroll
                for (uint32_t t = 0; t < T; t++) cu[t] = cn[t];
#pragma unroll
                for (uint32_t g = 0; g < G4; g++) {
                    a[g][0] = an[g][0];
                    a[g][1] = an[g][1];
                }
            }
        }
        // D[4 kg + j][col]: pair 16 g + 4 kg + j against entry col of tile t;
        // each pair's constants are read once for the lane's tiles
        float rx[T];
#pragma unroll
        for (uint32_t t = 0; t < T; t++)
            rx[t] = MET == 1 ? (xn[t] > 0.0f ? 0.5f / sqrtf(xn[t]) : 0.0f) : 0.0f;
#pragma unroll
        for (uint32_t g = 0; g < G4; g++) {
            if (g >= ng) break;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t g2 = 16 * g + 4 * kg + j;
                const float c0 = s_cs[g2][0], c1 = s_cs[g2][1], c2 = s_cs[g2][2], c3 = s_cs[g2][3],
                            c4 = s_cs[g2][4];
                const float qn2 = MET == 2 ? s_qn[g2] : 0.0f;
                const uint64_t msk = p.qmask ? s_mask[g2] : 0ull;
#pragma unroll
                for (uint32_t t = 0; t < T; t++) {
                    const uint32_t el = 16 * (wave + 4 * t) + col;   // entry within the item
                    // distance_with_bound's estimate (src/rabitq.rs:202-215), as in rq_est
                    float inner = c2 * (float)acc[g][t][j] + c3 * (float)cnt[t];
                    inner = inner + c4;
                    const float sf = c1 * fac[t];
                    const float est = fmaxf((n2[t] + c0) - sf * inner, 0.0f);
                    float d = est;
                    if (MET == 1) d = 1.0f - (1.0f + xn[t] - est) * rx[t];
                    else if (MET == 2) d = 1.0f - 0.5f * (qn2 + xn[t] - est);
                    // ("no key" past the posting: the range's tail; rows past
                    // the item's pairs are written and never stored)
                    const bool ok = e_begin + el < e_end && live[t] && (!p.qmask || label_ok(msk, lab[t]));
                    s_tile[g2][el] = ok ? (uint16_t)(fkey_r(d) >> 16) : (uint16_t)0xffff;
                }
            }
        }
        __syncthreads();
        // the pairs' rows as aligned 16-byte stores, up to the reserved end of
        // the range on the posting's last item
        const uint32_t len = e_end - e_begin;
        const uint32_t lp = e_end == pe ? (len + kScanAlign - 1) / kScanAlign * kScanAlign : len;
        const uint32_t cpr = lp / 8;
        for (uint32_t c = tid; c < gcnt * cpr; c += 256) {
            const uint32_t r = c / cpr, ch = c - r * cpr;
            const uint4 v = *reinterpret_cast<const uint4 *>(&s_tile[r][8 * ch]);
            *reinterpret_cast<uint4 *>(p.out16 + s_slot[r] + (e_begin - pb) + 8 * ch) = v;
        }
    }
}

// ---------------------------------------------------------------- rows
// 64 entries per workgroup: their postings found by 64 threads at once, then
// each wave takes 16 of them one after another, lane l over dims l, l + 64, ...
// (the row read coalesced, once): f16 copy of the raw row, |x|^2, and the
// posting's max |x - f16(x)| in double, rounded up (non-negative floats order
// as their bits, so atomicMax on the bits is a max).
__global__ __launch_bounds__(256) void k_rq_rows(const RqRowsParams p) {
    __shared__ uint32_t s_post[64];
    const uint32_t e0 = blockIdx.x * 64, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < 64 && e0 + threadIdx.x < p.n) {
        const uint32_t e = e0 + threadIdx.x;
        uint32_t lo = 0, hi = p.n_post;   // posting: post_off[lo] <= e < post_off[lo + 1]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (p.post_off[mid] <= e) lo = mid;
            else hi = mid;
        }
        s_post[threadIdx.x] = lo;
    }
    __syncthreads();
    for (uint32_t k = 16 * wave; k < 16 * wave + 16; k++) {
        const uint32_t e = e0 + k;
        if (e >= p.n) break;   // whole wave
        const float *x = p.raw_e + (size_t)e * p.dpad;
        uint16_t *h = p.raw16 + (size_t)e * p.dpad;
        double e2 = 0.0, n2 = 0.0;
        for (uint32_t i = lane; i < p.dpad; i += 64) {
            const float v = i < p.dim ? x[i] : 0.0f;
            const _Float16 f = (_Float16)v;
            h[i] = __builtin_bit_cast(uint16_t, f);
            const double t = (double)v - (double)(float)f;
            e2 += t * t;
            n2 += (double)v * v;
        }
        for (int off = 32; off > 0; off >>= 1) {
            e2 += __shfl_xor(e2, off, 64);
            n2 += __shfl_xor(n2, off, 64);
        }
        if (lane == 0) {
            p.xn[e] = (float)n2;
            const float ex = (float)(sqrt(e2) * (1.0 + 1e-12));
            const uint32_t up = __float_as_uint(ex) + 1u;   // the next float up
            // a long posting's entries would queue on one address: the atomic
            // only when the value seen is below (a stale read just costs the atomic)
            uint32_t *a = p.post_ex + s_post[k];
            if (__atomic_load_n(a, __ATOMIC_RELAXED) < up) atomicMax(a, up);
        }
    }
}

// ---------------------------------------------------------------- launchers
#define SPF_RQ_DISPATCH(W_, LAUNCH)                                               \
    switch (W_) {                                                                 \
    case 1: LAUNCH(1); break;                                                     \
    case 2: LAUNCH(2); break;                                                     \
    case 4: LAUNCH(4); break;                                                     \
    case 8: LAUNCH(8); break;                                                     \
    case 16: LAUNCH(16); break;                                                   \
    case 32: LAUNCH(32); break;                                                   \
    case 64: LAUNCH(64); break;                                                   \
    default: return hipErrorInvalidValue;                                         \
    }

hipError_t launch_rq_prep(const RqParams &p, uint32_t n_pairs, hipStream_t s) {
    if (n_pairs == 0) return hipSuccess;
    if (p.words * 64 != p.padded) return hipErrorInvalidValue;
#define L(W) hipLaunchKernelGGL(k_rq_prep<W>, dim3(n_pairs), dim3(64), p.mixed ? 0 : 64 * W * sizeof(float), s, p)
    SPF_RQ_DISPATCH(p.words, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_rq_search_exact(const RqParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
#define L(W) hipLaunchKernelGGL(k_rq_search_exact<W>, dim3(p.nq), dim3(64), 0, s, p)
    SPF_RQ_DISPATCH(p.words, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_rq_decode(const RqDecodeParams &p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
#define L(W) hipLaunchKernelGGL(k_rq_decode<W>, dim3(n), dim3(64), p.padded * sizeof(float), s, p)
    SPF_RQ_DISPATCH(p.padded / 64, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_rq_scan_mfma(const RqParams &p, uint32_t grid, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    const int met = (p.metric == M_L2 || p.metric == M_L2SQ) ? 0 : p.metric == M_COS ? 1 : 2;
#define L(W)                                                                                       \
    if (met == 0) hipLaunchKernelGGL((k_rq_scan_mfma<W, 0>), dim3(grid), dim3(256), 0, s, p);      \
    else if (met == 1) hipLaunchKernelGGL((k_rq_scan_mfma<W, 1>), dim3(grid), dim3(256), 0, s, p); \
    else hipLaunchKernelGGL((k_rq_scan_mfma<W, 2>), dim3(grid), dim3(256), 0, s, p)
    SPF_RQ_DISPATCH(p.words, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_rq_rows(const RqRowsParams &p, hipStream_t s) {
    if (p.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rq_rows, dim3((p.n + 63) / 64), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_rq_scan(const RqParams &p, uint32_t grid, hipStream_t s) {
    if (grid == 0) return hipSuccess;
#define L(W) hipLaunchKernelGGL(k_rq_scan<W>, dim3(grid), dim3(256), 0, s, p)
    SPF_RQ_DISPATCH(p.words, L)
#undef L
    return hipGetLastError();
}

}  // namespace spf
