// Mixed-precision batch search kernels (gfx950 matrix cores).
//
//   k_scan_mfma     posting scan as a (queries x entries) f16 GEMM on
//                   v_mfma_f32_16x16x32_f16 with f32 accumulate:
//                   dist = |x|^2 + |q|^2 - 2 q.x, stored as a 16-bit key
//                   (top half of the order-preserving f32 key) per candidate
//   k_probe_mfma    centroid probe as a bf16 GEMM (centroids centred on
//                   their mean), 16-bit keys per (query, centroid)
//   k_pool_select   per query: the `want` smallest 16-bit keys (ties by
//                   candidate slot), marked in a bitmap over the candidate
//                   buffer (candidates) or listed (probe band)
//   k_rerank_pool   exact f32 recheck of a query's probe band, sorted
//   k_rerank_pruned per query: exact f32 rerank of the pool's `ca` smallest,
//                   a bound that excludes the other members it can, exact
//                   rerank of the rest, top k written as ids and distances
//
// MFMA fragment layout (v_mfma_f32_16x16x32_{f16,bf16}, wave64), checked on the
// device by spf_selftest: lane l holds A[l%16][8*(l/16) .. +7],
// B[8*(l/16) .. +7][l%16], and D[4*(l/16) + j][l%16] for j = 0..3.
#include <cstdlib>

#include "spf_internal.h"

namespace spf {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef int int32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int32x4_t as_i32x4(uint4 v) {
    union { uint4 u; int32x4_t i; } c;
    c.u = v;
    return c.i;
}

__device__ __forceinline__ uint32_t fkey_m(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float_m(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ half8_t as_half8(uint4 v) {
    union { uint4 u; half8_t h; } c;
    c.u = v;
    return c.h;
}

// ---------------------------------------------------------------- self-test
// A[m][k] = ((3m + 5k) % 7) - 3, B[k][n] = ((2k + 7n) % 5) - 2, D = A B (exact
// in f16/bf16 and f32).  Each lane builds its fragments by the documented layout
// and writes D back by the documented output layout.
__global__ void k_selftest_mfma(float *out_f16, float *out_bf16, float *out_i8) {
    const int l = threadIdx.x;
    half8_t a, b;
    bf16x8_t a2, b2;
    for (int i = 0; i < 8; i++) {
        const int m = l % 16, k = 8 * (l / 16) + i, n = l % 16;
        const float av = (float)(((3 * m + 5 * k) % 7) - 3);
        const float bv = (float)(((2 * k + 7 * n) % 5) - 2);
        a[i] = (_Float16)av;
        b[i] = (_Float16)bv;
        a2[i] = (__bf16)av;
        b2[i] = (__bf16)bv;
    }
    floatx4_t acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
    floatx4_t acc2 = {0, 0, 0, 0};
    acc2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b2, acc2, 0, 0, 0);
    for (int j = 0; j < 4; j++) {
        const int m = 4 * (l / 16) + j, n = l % 16;
        out_f16[m * 16 + n] = acc[j];
        out_bf16[m * 16 + n] = acc2[j];
    }
    // 16x16x64 int8: lane l holds A[l%16][16*(l/16) .. +15] and B[16*(l/16) .. +15][l%16]
    union { int8_t b[16]; int32x4_t v; } ai, bi;
    for (int i = 0; i < 16; i++) {
        const int m = l % 16, k = 16 * (l / 16) + i, n = l % 16;
        ai.b[i] = (int8_t)(((3 * m + 5 * k) % 7) - 3);
        bi.b[i] = (int8_t)(((2 * k + 7 * n) % 5) - 2);
    }
    int32x4_t acc3 = {0, 0, 0, 0};
    acc3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(ai.v, bi.v, acc3, 0, 0, 0);
    for (int j = 0; j < 4; j++) out_i8[(4 * (l / 16) + j) * 16 + l % 16] = (float)acc3[j];
}

hipError_t launch_selftest_mfma(float *d_out_f16, float *d_out_bf16, hipStream_t s) {
    hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, s, d_out_f16, d_out_bf16,
                       d_out_f16 + 512);
    return hipGetLastError();
}

// ---------------------------------------------------------------- MFMA scan
// Work item = (posting, chunk of 64 T entries, up to kMixQ queries that probe
// it).  Wave w owns the chunk's 16-entry tiles w, w+4, ...: their B fragments
// (f16 codes) are loaded once into registers, then reused by every 32-query
// group of the item (two 16-row A tiles per group, 16x16x32 f16 MFMA, f32
// accumulate).  Each (query, entry) writes the 16-bit key of its distance.
// Four workgroups per CU for the unfiltered scan up to 128 dims (the label
// masks, or 256-dim query fragments, would spill there).
//
// I8: int8 postings.  A K step is 64 dims (16x16x64 int8 MFMA, int32 sums),
// so a lane's fragment is again one 16-byte unit: u8 codes shifted to signed
// (c - 128, by flipping the top bit of every byte) against the symmetric int8
// query; q.decode(c) = a * sum + b with the query's affine terms (a, b).
template <int KS, bool FILT, bool I8, bool L2M>
__global__ __launch_bounds__(256, (FILT || KS > 4) ? 3 : 4) void k_scan_mfma(const MixedParams p) {
    constexpr int T = KS <= 4 ? 4 : 2;   // tiles per wave, mfma_tiles()
    constexpr uint32_t EW = 64 * T;      // entries per item
    __shared__ __align__(16) uint16_t s_tile[32][EW];   // one query group's keys
    // one query group's rows (the A fragments), loaded by the whole workgroup
    // (a row padded by one unit: rows start in different banks)
    __shared__ __align__(16) uint4 s_a[32][4 * KS + 1];
    __shared__ uint64_t s_slot[kMixQ];
    __shared__ uint32_t s_qi[kMixQ];
    __shared__ float s_qn[kMixQ];
    __shared__ float s_qa[I8 ? kMixQ : 1], s_qb[I8 ? kMixQ : 1];
    __shared__ uint64_t s_mask[kMixQ];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t col = lane & 15, kg = lane >> 4;
    const uint32_t lo = p.item_lo_dev ? *p.item_lo_dev : 0u;
    const uint32_t n_items = (p.n_items_dev ? *p.n_items_dev : p.n_items) - lo;
    for (XcdWalk it(n_items); it.more(); it.next()) {
        const uint32_t item = lo + it.item;
        const uint32_t vp = p.item_post[item];   // (virtual) posting of the pair table
        const uint32_t pp = vp >= p.n_post ? vp - p.n_post : vp;
        const uint32_t local = item - p.item_off[vp];
        const uint32_t pb = p.post_off[pp], pe = p.post_off[pp + 1];
        const uint32_t nec = (pe - pb + p.echunk - 1) / p.echunk;
        const uint32_t qc = local / nec, ec = local - qc * nec;
        const uint32_t pr0 = p.pair_off[vp] + qc * p.qg;
        const uint32_t gcnt = min(p.qg, p.pair_off[vp + 1] - pr0);
        const uint32_t e_begin = pb + ec * p.echunk, e_end = min(e_begin + p.echunk, pe);
        __syncthreads();   // the previous item is done with the query table
        if (tid < kMixQ) {
            const bool ok = tid < gcnt;
            const uint32_t q = ok ? p.pair_q[pr0 + tid] : 0u;
            s_qi[tid] = q;
            s_slot[tid] = ok ? p.pair_slot[pr0 + tid] : 0ull;
            float qn = ok ? (I8 ? p.qnorm[q] : p.qn16[q]) : 0.0f;
            if (!L2M) qn = p.metric == M_DOT ? 1.0f : (qn > 0.0f ? 1.0f / sqrtf(qn) : 0.0f);
            s_qn[tid] = qn;
            if (I8) {
                s_qa[tid] = ok ? p.q8a[q] : 0.0f;
                s_qb[tid] = ok ? p.q8b[q] : 0.0f;
            }
            if (FILT) s_mask[tid] = ok ? p.qmask[q] : 0ull;
        }
        // the wave's B fragments, and per-entry norm / tombstone word / label
        // of its column: every load of every tile is issued before any value
        // is used (entries past the posting read entry e_begin and are masked)
        uint4 bf[T][KS];
        float xn[T], rx[T];
        bool live[T];
        uint32_t lab[T], dw[T];
#pragma unroll
        for (int t = 0; t < T; t++) {
            const uint32_t e_raw = e_begin + 16 * (wave + 4 * t) + col;
            const uint32_t e = e_raw < e_end ? e_raw : e_begin;
            const uint4 *src = p.codes16 + (size_t)(e >> 6) * p.units * 64 + (e & 63);
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const uint32_t u = 4 * s + kg;
                bf[t][s] = u < p.units ? src[(size_t)u * 64] : make_uint4(0, 0, 0, 0);
                if (I8 && u < p.units) {   // u8 c -> signed c - 128
                    bf[t][s].x ^= 0x80808080u;
                    bf[t][s].y ^= 0x80808080u;
                    bf[t][s].z ^= 0x80808080u;
                    bf[t][s].w ^= 0x80808080u;
                }
            }
            xn[t] = p.xn16[e];
            dw[t] = p.del ? p.del[e >> 5] : 0u;
            lab[t] = FILT ? p.label_e[e] : 0u;
        }
#pragma unroll
        for (int t = 0; t < T; t++) {
            const uint32_t e = e_begin + 16 * (wave + 4 * t) + col;
            live[t] = e < e_end && !((dw[t] >> (e & 31)) & 1u);   // (a column's products
            // depend on its own B values only, so the stand-in entry's are harmless)
            rx[t] = !L2M && p.metric == M_COS ? (xn[t] > 0.0f ? 1.0f / sqrtf(xn[t]) : 0.0f) : 1.0f;
        }
        __syncthreads();
        for (uint32_t g0 = 0; g0 < gcnt; g0 += 32) {
            // the group's query rows into LDS, coalesced (rows hold 4*ks units;
            // rows past the group's queries read zero)
            for (uint32_t i = tid; i < 32 * 4 * KS; i += 256) {
                const uint32_t r = i / (4 * KS), u = i - r * (4 * KS), g = g0 + r;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (g < gcnt && u < 4 * p.ks) v = p.q16[(size_t)s_qi[g] * p.q16_units + u];
                s_a[r][u] = v;
            }
            __syncthreads();
            // two 16-query halves, one after the other (fewer live registers)
#pragma unroll 1
            for (uint32_t h = 0; h < 2; h++) {
                const uint32_t gb = g0 + 16 * h;
                if (gb >= gcnt) break;   // (uniform; the half's tile rows are not stored)
                // A fragments: 16 queries x dims
                uint4 a[KS];
#pragma unroll
                for (int s = 0; s < KS; s++) a[s] = s_a[16 * h + col][4 * s + kg];
                // the lane's four rows 4*kg + j of the half: norm factor and
                // label mask, read once (gb + 15 < kMixQ)
                float qv[4], qa[4], qb[4];
                uint64_t mv[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    qv[j] = s_qn[gb + 4 * kg + j];
                    mv[j] = FILT ? s_mask[gb + 4 * kg + j] : 0ull;
                    qa[j] = I8 ? s_qa[gb + 4 * kg + j] : 0.0f;
                    qb[j] = I8 ? s_qb[gb + 4 * kg + j] : 0.0f;
                }
#pragma unroll
                for (int t = 0; t < T; t++) {
                    const uint32_t e0 = e_begin + 16 * (wave + 4 * t);
                    if (e0 >= e_end) break;   // later tiles are further out
                    floatx4_t acc = {0, 0, 0, 0};
                    if (I8) {
                        int32x4_t ai = {0, 0, 0, 0};
#pragma unroll
                        for (int s = 0; s < KS; s++)
                            ai = __builtin_amdgcn_mfma_i32_16x16x64_i8(as_i32x4(a[s]),
                                                                      as_i32x4(bf[t][s]), ai, 0, 0, 0);
#pragma unroll
                        for (int j = 0; j < 4; j++) acc[j] = qa[j] * (float)ai[j] + qb[j];   // q.x
                    } else {
#pragma unroll
                        for (int s = 0; s < KS; s++)
                            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_half8(a[s]),
                                                                         as_half8(bf[t][s]), acc,
                                                                         0, 0, 0);
                    }
                    // keys into the group tile ("no key" for dead, filtered or
                    // past-the-posting entries: the aligned range's tail).  Rows
                    // past the group's queries are written too and never stored.
                    const uint32_t el = e0 + col - e_begin;
                    uint32_t kv[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        // (2 acc is exact: the fma rounds as (xn + qn) - (acc + acc) does)
                        const float d = L2M ? __builtin_fmaf(-2.0f, acc[j], xn[t] + qv[j])
                                            : 1.0f - acc[j] * qv[j] * rx[t];
                        const bool ok = live[t] && (!FILT || label_ok(mv[j], lab[t]));
                        kv[j] = ok ? fkey_m(d) >> 16 : 0xffffu;
                        s_tile[16 * h + 4 * kg + j][el] = (uint16_t)kv[j];
                    }
                }
            }
            {
                const uint32_t len = e_end - e_begin, l16 = (len + 15) & ~15u;
                // the tiles written above end at the 16-entry multiple past the
                // posting's end; "no key" from there to the reserved end
                const uint32_t lp = e_end == pe ? (len + kScanAlign - 1) / kScanAlign * kScanAlign : len;
                if (lp > l16)
                    for (uint32_t c = tid; c < 32 * (lp - l16); c += 256)
                        s_tile[c / (lp - l16)][l16 + c % (lp - l16)] = 0xffff;
                __syncthreads();
                // the tile's rows as aligned 16-byte stores: each probe's slot
                // range starts at a multiple of kScanAlign and is reserved up to
                // the next multiple past the posting's end
                const uint32_t cpr = lp / 8, rows = min(32u, gcnt - g0);
                for (uint32_t c = tid; c < rows * cpr; c += 256) {
                    const uint32_t r = c / cpr, ch = c - r * cpr;
                    const uint4 v = *reinterpret_cast<const uint4 *>(&s_tile[r][8 * ch]);
                    *reinterpret_cast<uint4 *>(p.out16 + s_slot[g0 + r] + (e_begin - pb) + 8 * ch) = v;
                }
                __syncthreads();   // the next group reuses the tile
            }
        }
    }
}

// ---------------------------------------------------------------- wave scan
// Work item = (posting, echunk entries, up to 16 * NH queries), one wave per
// item.  The wave keeps its queries' fragments (the B operand, 16 queries per
// half) in registers for the whole item and streams the chunk's 16-entry code
// tiles (the A operand) with the next tile's loads in flight, so the only
// dependent load chain of an item is its set-up.  With the codes as A,
// D[4 kg + j][col] is entry 4 kg + j of the tile against query col of the
// half: a lane holds four consecutive keys of one query row, packs their top
// halves into two words (v_perm) and stores them as one 8-byte word at the
// row's slots, so keys never pass through LDS and the wave never waits for
// another.  Keys are bit-identical to k_scan_mfma's: the same f32 expression
// per (query, entry) (2 acc is exact, so fma(-2, acc, xn + qn) rounds as
// (xn + qn) - (acc + acc) does).
__device__ __forceinline__ uint32_t key_pair(float d0, float d1) {
    // top halves of fkey(d0), fkey(d1) as d1 << 16 | d0: the raw top halves,
    // xor'd with 0xffff (negative: every bit flips) or 0x8000 (sign set)
    const uint32_t u0 = __float_as_uint(d0), u1 = __float_as_uint(d1);
    const uint32_t hi = __builtin_amdgcn_perm(u1, u0, 0x07060302u);
    const uint32_t sg = __builtin_amdgcn_perm(u1, u0, 0x0b0b0909u);   // sign bytes of u0, u1
    return hi ^ (sg | 0x80008000u);
}

//
// ADM: second phase of a two-phase scan.  A key at most its query's bound
// tau[q] is appended to the query's list (adm_list from seg_off[q], as
// key << 16 | slot within the segment; one returned atomic on adm_cnt[q] per
// lane with admitted keys); nothing else is stored.  The list order depends
// on timing; the pool taken from it does not.
template <int KS, int NH, bool FILT, bool I8, bool L2M, bool ADM>
__global__ __launch_bounds__(64) void k_scan_wave(const MixedParams p) {
    // keys of GT tiles per query row (8 KiB in all), stored as 16-byte units of
    // a row (up to 32 GT contiguous bytes a row) once the GT tiles are done
    constexpr uint32_t GT = NH == 2 ? 8 : 4;
    __shared__ __align__(16) uint16_t s_k[ADM ? 1 : NH][16][16 * GT];
    const uint32_t lane = threadIdx.x, col = lane & 15, kg = lane >> 4;
    const uint32_t lo = p.item_lo_dev ? *p.item_lo_dev : 0u;
    const uint32_t n_items = (p.n_items_dev ? *p.n_items_dev : p.n_items) - lo;
    for (XcdWalk it(n_items); it.more(); it.next()) {
        const uint32_t item = lo + it.item;
        const uint32_t vp = p.item_post[item];   // (virtual) posting of the pair table
        const uint32_t pp = vp >= p.n_post ? vp - p.n_post : vp;
        const uint32_t local = item - p.item_off[vp];
        const uint32_t pb = p.post_off[pp], pe = p.post_off[pp + 1];
        const uint32_t nec = (pe - pb + p.echunk - 1) / p.echunk;
        const uint32_t qc = local / nec, ec = local - qc * nec;
        const uint32_t pr0 = p.pair_off[vp] + qc * p.qg;
        const uint32_t gcnt = min(p.qg, p.pair_off[vp + 1] - pr0);
        const uint32_t e_begin = pb + ec * p.echunk, e_end = min(e_begin + p.echunk, pe);
        // a probe's slot range is reserved up to the next multiple of
        // kScanAlign past the posting's end (slots relative to pb)
        const uint32_t lim = (pe - pb + kScanAlign - 1) / kScanAlign * kScanAlign;
        // this lane's query (column col) of every half, its fragments and terms
        uint4 bq[NH][KS];
        uint64_t slot[NH], qm[NH];
        float qv[NH], qa[NH], qb[NH];
        uint32_t aq[NH], atau[NH], asl[NH];   // admission: query, bound, slot of the range in its segment
#pragma unroll
        for (int h = 0; h < NH; h++) {
            const uint32_t g = 16 * h + col;
            const bool ok = g < gcnt;
            const uint32_t q = ok ? p.pair_q[pr0 + g] : 0u;
            slot[h] = ok ? p.pair_slot[pr0 + g] : ~0ull;
            if (ADM) {   // (rows past the group's queries admit nothing: bound 0 < every key's 0xffff)
                aq[h] = q;
                atau[h] = ok ? p.tau[q] : 0u;
                asl[h] = ok ? (uint32_t)(slot[h] - p.seg_off[q]) : 0u;
            }
            const uint4 *qrow = p.q16 + (size_t)q * p.q16_units;
#pragma unroll
            for (int s = 0; s < KS; s++)
                bq[h][s] = (ok && s < (int)p.ks) ? qrow[4 * s + kg] : make_uint4(0, 0, 0, 0);
            float v = ok ? (I8 ? p.qnorm[q] : p.qn16[q]) : 0.0f;
            if (!L2M) v = p.metric == M_DOT ? 1.0f : (v > 0.0f ? 1.0f / sqrtf(v) : 0.0f);
            qv[h] = v;
            qa[h] = I8 && ok ? p.q8a[q] : 0.0f;
            qb[h] = I8 && ok ? p.q8b[q] : 0.0f;
            qm[h] = FILT && ok ? p.qmask[q] : 0ull;
        }
        const uint32_t ntile = (e_end - e_begin + 15) / 16;
        // tile t: entry e_begin + 16 t + col (past the chunk: entry e_begin,
        // masked), its code units 4 s + kg, |x|^2, tombstone word and label
        uint4 a[KS];
        float xv;
        uint32_t dw, lb;
        auto load = [&](uint32_t t, uint4 (&fa)[KS], float &fx, uint32_t &fd, uint32_t &fl) {
            const uint32_t e_raw = e_begin + 16 * t + col;
            const uint32_t e = e_raw < e_end ? e_raw : e_begin;
            const uint4 *src = p.codes16 + (size_t)(e >> 6) * p.units * 64 + (e & 63);
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const uint32_t u = 4 * s + kg;
                uint4 v = u < p.units ? src[(size_t)u * 64] : make_uint4(0, 0, 0, 0);
                if (I8 && u < p.units) {   // u8 c -> signed c - 128
                    v.x ^= 0x80808080u;
                    v.y ^= 0x80808080u;
                    v.z ^= 0x80808080u;
                    v.w ^= 0x80808080u;
                }
                fa[s] = v;
            }
            fx = p.xn16[e];
            fd = p.del ? p.del[e >> 5] : 0u;
            fl = FILT ? p.label_e[e] : 0u;
        };
        load(0, a, xv, dw, lb);
        for (uint32_t t = 0; t < ntile; t++) {
            // the next tile's loads in flight (the last tile reloads itself)
            uint4 an[KS];
            float xvn;
            uint32_t dwn, lbn;
            load(min(t + 1, ntile - 1), an, xvn, dwn, lbn);
            const uint32_t e0 = e_begin + 16 * t;
            const uint32_t e_raw = e0 + col;
            const bool lv = e_raw < e_end && !((dw >> (e_raw & 31)) & 1u);
            // rows 4 kg + j of the tile: liveness bits, |x|^2 (or 1/|x|), label
            const uint32_t m4 = (uint32_t)(__ballot(lv) >> (4 * kg)) & 0xfu;
            const float xr = L2M ? xv
                             : p.metric == M_COS ? (xv > 0.0f ? 1.0f / sqrtf(xv) : 0.0f) : 1.0f;
            float xj[4];
            uint32_t lj[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                xj[j] = __shfl(xr, 4 * kg + j);
                lj[j] = FILT ? (uint32_t)__shfl((int)lb, 4 * kg + j) : 0u;
            }
            const uint32_t rel = e0 - pb + 4 * kg;   // slot offset of the lane's four keys
            uint32_t dm01 = ((m4 & 1u) ? 0u : 0xffffu) | ((m4 & 2u) ? 0u : 0xffff0000u);
            uint32_t dm23 = ((m4 & 4u) ? 0u : 0xffffu) | ((m4 & 8u) ? 0u : 0xffff0000u);
#pragma unroll
            for (int h = 0; h < NH; h++) {
                if (16 * h >= (int)gcnt) break;   // (uniform)
                floatx4_t acc = {0, 0, 0, 0};
                if (I8) {
                    int32x4_t ai = {0, 0, 0, 0};
#pragma unroll
                    for (int s = 0; s < KS; s++)
                        ai = __builtin_amdgcn_mfma_i32_16x16x64_i8(as_i32x4(a[s]), as_i32x4(bq[h][s]),
                                                                  ai, 0, 0, 0);
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[j] = qa[h] * (float)ai[j] + qb[h];   // q.x
                } else {
#pragma unroll
                    for (int s = 0; s < KS; s++)
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_half8(a[s]), as_half8(bq[h][s]),
                                                                     acc, 0, 0, 0);
                }
                float d[4];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    d[j] = L2M ? __builtin_fmaf(-2.0f, acc[j], xj[j] + qv[h])
                               : 1.0f - acc[j] * qv[h] * xj[j];
                uint32_t k01 = key_pair(d[0], d[1]), k23 = key_pair(d[2], d[3]);
                if (FILT) {
                    const uint32_t f0 = label_ok(qm[h], lj[0]) ? 0u : 0xffffu;
                    const uint32_t f1 = label_ok(qm[h], lj[1]) ? 0u : 0xffff0000u;
                    const uint32_t f2 = label_ok(qm[h], lj[2]) ? 0u : 0xffffu;
                    const uint32_t f3 = label_ok(qm[h], lj[3]) ? 0u : 0xffff0000u;
                    k01 |= dm01 | f0 | f1;
                    k23 |= dm23 | f2 | f3;
                } else {
                    k01 |= dm01;
                    k23 |= dm23;
                }
                if constexpr (ADM) {
                    // (dead and past-the-range keys read 0xffff, above every bound)
                    const uint32_t kk[4] = {k01 & 0xffffu, k01 >> 16, k23 & 0xffffu, k23 >> 16};
                    const bool row = slot[h] != ~0ull;
                    uint32_t n = 0;
#pragma unroll
                    for (int j = 0; j < 4; j++) n += row && kk[j] <= atau[h] ? 1u : 0u;
                    if (n) {
                        uint32_t pos = atomicAdd(p.adm_cnt + aq[h], n);
                        uint32_t *dst = p.adm_list + (slot[h] - asl[h]);   // the segment's list
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            if (kk[j] <= atau[h]) dst[pos++] = (kk[j] << 16) | (asl[h] + rel + j);
                    }
                } else {
                    *reinterpret_cast<uint2 *>(&s_k[h][col][(t % GT) * 16 + 4 * kg]) = make_uint2(k01, k23);
                }
            }
            if constexpr (!ADM) {
                if (t % GT == GT - 1 || t + 1 == ntile) {   // the group's rows leave as 16-byte units
                    __syncthreads();
                    const uint32_t g0t = t - t % GT, nu = 2 * (t - g0t + 1);   // units a row
                    const uint32_t relg = e_begin + 16 * g0t - pb;        // (a multiple of 16)
                    const uint32_t r8 = lane >> 3, c = lane & 7;
#pragma unroll
                    for (int h = 0; h < NH; h++) {
                        if (16 * h >= (int)gcnt) break;   // (uniform)
#pragma unroll
                        for (uint32_t rr = 0; rr < 16; rr += 8) {
                            const uint32_t r = rr + r8;   // lane r holds row r's slot
                            const uint32_t slo = __shfl((uint32_t)slot[h], (int)r, 64);
                            const uint32_t shi = __shfl((uint32_t)(slot[h] >> 32), (int)r, 64);
                            const uint64_t sl = ((uint64_t)shi << 32) | slo;
#pragma unroll
                            for (uint32_t cu = c; cu < 2 * GT; cu += 8)
                                if (cu < nu && sl != ~0ull && relg + 8 * cu < lim)
                                    *reinterpret_cast<uint4 *>(p.out16 + sl + relg + 8 * cu) =
                                        *reinterpret_cast<const uint4 *>(&s_k[h][r][8 * cu]);
                        }
                    }
                    __syncthreads();   // (the next group rewrites the tile rows)
                }
            }
#pragma unroll
            for (int s = 0; s < KS; s++) a[s] = an[s];
            xv = xvn;
            dw = dwn;
            lb = lbn;
        }
    }
}

// ---------------------------------------------------------------- probe
constexpr int kMixKS = 8;      // max K-steps of 32 dims held in registers (dim <= 256)

// One workgroup = 16 queries x 1024 centroids; each wave takes 16-centroid
// tiles.  bf16 operands, f32 accumulate, 16-bit keys out.
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8v as_bf8(uint4 v) {
    union { uint4 u; bf16x8v h; } c;
    c.u = v;
    return c.h;
}

// Each wave walks its 16-centroid tiles with the next tile's loads in flight;
// keys collect in an LDS tile and leave as 16-byte row segments, so the walk
// itself issues no global stores.  Rows are reserved up to a multiple of 32
// keys: the last segment's keys past the centroids read "no key".
template <int KS>
__global__ __launch_bounds__(256) void k_probe_mfma(const ProbeMfmaParams p) {
    __shared__ __align__(16) uint16_t s_out[16][kEChunk];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t col = lane & 15, kg = lane >> 4;
    const uint32_t nch = (p.n_post + kEChunk - 1) / kEChunk;
    const uint32_t qt = blockIdx.x / nch, ch = blockIdx.x - qt * nch;
    const uint32_t q0 = qt * 16;
    const uint32_t c_begin = ch * kEChunk, c_end = min(c_begin + kEChunk, p.n_post);
    bf16x8v a[KS];
    const uint32_t qa = q0 + col;
#pragma unroll
    for (int s = 0; s < KS; s++) {
        uint4 v = make_uint4(0, 0, 0, 0);
        if (qa < p.nq && s < (int)p.ks) v = p.qbf[(size_t)qa * p.q_units + 4 * s + kg];
        a[s] = as_bf8(v);
    }
    // per row: L2 |bf16(q - mu)|^2, cosine 1 / |q|
    float qn[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t q = q0 + 4 * kg + j;
        float v = 0.0f;
        if (q < p.nq) {
            if (p.metric == M_COS) {
                const float n2 = p.qnorm[q];
                v = n2 > 0.0f ? 1.0f / sqrtf(n2) : 0.0f;
            } else {
                v = p.qnbf[q];
            }
        }
        qn[j] = v;
    }
    auto key = [&](float acc, float cn, float qnj) -> float {
        if (p.metric == M_DOT) return 1.0f - acc;
        if (p.metric == M_COS) return 1.0f - acc * qnj * cn;
        return cn + qnj - 2.0f * acc;
    };
    // a tile's B fragment and the norm of the lane's centroid (columns past
    // the end read the last centroid; their keys are never stored)
    auto load_b = [&](uint32_t c0, uint4 (&b)[KS], float &cn) {
        const uint32_t c = min(c0 + col, c_end - 1);
        const uint4 *src = p.cbf + (size_t)(c >> 6) * p.c_units * 64 + (c & 63);
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const uint32_t u = 4 * s + kg;
            b[s] = (s < (int)p.ks && u < p.c_units) ? src[(size_t)u * 64] : make_uint4(0, 0, 0, 0);
        }
        cn = p.cnbf[c];
    };
    const uint32_t start = c_begin + 16 * wave;
    if constexpr (KS <= kMixKS) {
        auto body = [&](uint32_t c0, const uint4 (&b)[KS], float cn) {
            floatx4_t acc = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < KS; s++)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], as_bf8(b[s]), acc, 0, 0, 0);
            const uint32_t cl = c0 + col - c_begin;
#pragma unroll
            for (int j = 0; j < 4; j++)
                s_out[4 * kg + j][cl] = (uint16_t)(fkey_m(key(acc[j], cn, qn[j])) >> 16);
        };
        uint4 b0[KS], b1[KS];
        float cn0 = 0.0f, cn1 = 0.0f;
        if (start < c_end) load_b(start, b0, cn0);
        for (uint32_t c0 = start; c0 < c_end; c0 += 128) {
            if (c0 + 64 < c_end) load_b(c0 + 64, b1, cn1);
            body(c0, b0, cn0);
            if (c0 + 64 >= c_end) break;
            if (c0 + 128 < c_end) load_b(c0 + 128, b0, cn0);
            body(c0 + 64, b1, cn1);
        }
    } else {
        // long rows: the queries' fragments stay in registers, each tile's
        // centroid fragments stream in chunks of kMixKS K steps
        for (uint32_t c0 = start; c0 < c_end; c0 += 64) {
            const uint32_t c = min(c0 + col, c_end - 1);
            const uint4 *src = p.cbf + (size_t)(c >> 6) * p.c_units * 64 + (c & 63);
            const float cn = p.cnbf[c];
            floatx4_t acc = {0, 0, 0, 0};
#pragma unroll
            for (int s0 = 0; s0 < KS; s0 += kMixKS) {
                uint4 b[kMixKS];
#pragma unroll
                for (int i = 0; i < kMixKS; i++) {
                    const uint32_t u = 4 * (s0 + i) + kg;
                    b[i] = (s0 + i < (int)p.ks && u < p.c_units) ? src[(size_t)u * 64]
                                                                 : make_uint4(0, 0, 0, 0);
                }
#pragma unroll
                for (int i = 0; i < kMixKS; i++)
                    if (s0 + i < (int)p.ks)
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s0 + i], as_bf8(b[i]), acc,
                                                                      0, 0, 0);
            }
            const uint32_t cl = c0 + col - c_begin;
#pragma unroll
            for (int j = 0; j < 4; j++)
                s_out[4 * kg + j][cl] = (uint16_t)(fkey_m(key(acc[j], cn, qn[j])) >> 16);
        }
    }
    // keys past the last centroid up to the next multiple of 8
    const uint32_t n = c_end - c_begin, n8 = (n + 7) & ~7u;
    for (uint32_t i = tid; i < 16 * (n8 - n); i += 256)
        s_out[i / (n8 - n)][n + i % (n8 - n)] = 0xffff;
    __syncthreads();
    const uint32_t cpr = n8 / 8;
    for (uint32_t i = tid; i < 16 * cpr; i += 256) {
        const uint32_t r = i / cpr, c8 = i - r * cpr;
        if (q0 + r < p.nq)
            *reinterpret_cast<uint4 *>(p.out + (size_t)(q0 + r) * p.out_stride + c_begin + 8 * c8) =
                *reinterpret_cast<const uint4 *>(&s_out[r][8 * c8]);
    }
}

// ---------------------------------------------------------------- pool select
// Per query: mark the `want` smallest 16-bit keys (0xffff = no candidate).
// Exactly-tied keys at the boundary are taken in slot order.

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// exclusive scan over a 256-thread block with two barriers; *tot = block sum
__device__ __forceinline__ uint32_t block_scan_fast(uint32_t v, uint32_t *s_w, uint32_t *tot) {
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if ((threadIdx.x & 63) == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t base = 0, sum = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; w++) {
        const uint32_t x = s_w[w];
        base += w < wave ? x : 0u;
        sum += x;
    }
    __syncthreads();
    *tot = sum;
    return base + incl - v;
}

// 32 keys of one bitmap word (64 bytes; segments are 32-slot aligned)
__device__ __forceinline__ void load_word_raw(const uint16_t *kp, uint32_t w, uint4 r[4]) {
    const uint4 *src = reinterpret_cast<const uint4 *>(kp + 32 * w);
#pragma unroll
    for (int j = 0; j < 4; j++) r[j] = src[j];
}

// The query's probe table: segment-relative first slot and first entry of each
// probed posting, in LDS for short probe lists, else read from the plan buffers.
struct ProbeTab {
    uint32_t np;
    const uint32_t *pbase, *ebase;      // LDS form (null when global)
    const uint64_t *qp_row;             // global form
    const uint32_t *probe_row, *post_off;
    uint64_t seg;
    __device__ __forceinline__ uint32_t rel(uint32_t j) const {
        return pbase ? pbase[j] : (uint32_t)(qp_row[j] - seg);
    }
    // index of the probed posting holding `slot` (relative to the query's segment)
    __device__ __forceinline__ uint32_t find(uint32_t slot) const {
        uint32_t lo = 0, hi = np;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (rel(mid) <= slot) lo = mid; else hi = mid;
        }
        return lo;
    }
    __device__ __forceinline__ uint32_t first_entry(uint32_t j) const {
        return pbase ? ebase[j] : post_off[probe_row[j]];
    }
    // slot -> entry, walking forward from probe j (slots visited in increasing order)
    __device__ __forceinline__ uint32_t entry_from(uint32_t &j, uint32_t slot) const {
        while (j + 1 < np && rel(j + 1) <= slot) j++;
        return first_entry(j) + (slot - rel(j));
    }
};

// Wave-level scans and bin search for the pool select below.
__device__ __forceinline__ uint32_t wave_excl(uint32_t v, uint32_t *tot) {
    const uint32_t incl = wave_incl_scan(v);
    *tot = __shfl(incl, 63, 64);
    return incl - v;
}

// bit masks of one 32-key word: live keys, keys below T, keys equal to T
__device__ __forceinline__ void word_masks(const uint4 (&r)[4], uint32_t w, uint32_t n, uint32_t T,
                                           uint32_t &live, uint32_t &lt, uint32_t &eq) {
    uint32_t lv = 0, l = 0, e = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint32_t v4[4] = {r[c].x, r[c].y, r[c].z, r[c].w};
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const uint32_t j = 8 * c + t;
            const uint32_t kk = (v4[t >> 1] >> (16 * (t & 1))) & 0xffffu;
            const uint32_t bit = (uint32_t)(kk != 0xffffu && 32 * w + j < n) << j;
            lv |= bit;
            l |= kk < T ? bit : 0u;
            e |= kk == T ? bit : 0u;
        }
    }
    live = lv;
    lt = l;
    eq = e;
}

// bit masks of one 8-key unit for two thresholds T and TA: live keys, keys
// below T, keys equal to T, keys below TA, keys equal to TA
__device__ __forceinline__ void unit_masks2(const uint4 v, uint32_t u, uint32_t n, uint32_t T,
                                            uint32_t TA, uint32_t &live, uint32_t &lt,
                                            uint32_t &eq, uint32_t &ltA, uint32_t &eqA) {
    uint32_t lv = 0, l = 0, e = 0, la = 0, ea = 0;
    const uint32_t v4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const uint32_t kk = (v4[t >> 1] >> (16 * (t & 1))) & 0xffffu;
        const uint32_t bit = (uint32_t)(kk != 0xffffu && 8 * u + t < n) << t;
        lv |= bit;
        l |= kk < T ? bit : 0u;
        e |= kk == T ? bit : 0u;
        la |= kk < TA ? bit : 0u;
        ea |= kk == TA ? bit : 0u;
    }
    live = lv;
    lt = l;
    eq = e;
    ltA = la;
    eqA = ea;
}


// the lowest `c` set bits of m
__device__ __forceinline__ uint32_t lowest_bits(uint32_t m, uint32_t c) {
    uint32_t out = 0;
    while (c && m) {
        const uint32_t lo = m & (~m + 1u);
        out |= lo;
        m ^= lo;
        c--;
    }
    return out;
}

// The threads that select one query's pool: NW wavefronts (NW = 1: four
// queries per workgroup and no workgroup barrier; NW = 4: one query per
// workgroup).  Thread t of the group owns the segment's 16-byte units (8 keys)
// t, t + 64*NW, t + 2*64*NW, ...; sx is the group's LDS for cross-wave sums.
template <int NW>
struct PoolGroup {
    static constexpr uint32_t T = 64 * NW;
    uint32_t *sx;        // 2 * NW scan totals (double-buffered) + 3 * NW reductions
    uint32_t parity = 0;
    __device__ static __forceinline__ uint32_t tid() { return NW == 1 ? (threadIdx.x & 63) : threadIdx.x; }
    __device__ static __forceinline__ void sync() {
        if (NW == 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        } else {
            __syncthreads();
        }
    }
    // exclusive scan over the group in thread order; *tot = group sum.  A
    // buffer is rewritten two scans later, after a barrier that every reader
    // of it has passed.
    __device__ __forceinline__ uint32_t excl(uint32_t v, uint32_t *tot) {
        uint32_t wt;
        const uint32_t e = wave_excl(v, &wt);
        if (NW == 1) {
            *tot = wt;
            return e;
        }
        const uint32_t wv = threadIdx.x >> 6;
        uint32_t *buf = sx + parity * NW;
        parity ^= 1u;
        if ((threadIdx.x & 63) == 0) buf[wv] = wt;
        __syncthreads();
        uint32_t before = 0, all = 0;
#pragma unroll
        for (int i = 0; i < NW; i++) {
            const uint32_t x = buf[i];
            before += (uint32_t)i < wv ? x : 0u;
            all += x;
        }
        *tot = all;
        return e + before;
    }
    // group-wide sum, min and max (one call per kernel)
    __device__ __forceinline__ void reduce3(uint32_t &sum, uint32_t &mn, uint32_t &mx) {
        for (int off = 32; off > 0; off >>= 1) {
            sum += __shfl_xor(sum, off, 64);
            mn = min(mn, (uint32_t)__shfl_xor(mn, off, 64));
            mx = max(mx, (uint32_t)__shfl_xor(mx, off, 64));
        }
        if (NW == 1) return;
        uint32_t *r = sx + 2 * NW;
        const uint32_t wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            r[3 * wv] = sum;
            r[3 * wv + 1] = mn;
            r[3 * wv + 2] = mx;
        }
        __syncthreads();
        sum = 0;
#pragma unroll
        for (int i = 0; i < NW; i++) {
            sum += r[3 * i];
            mn = min(mn, r[3 * i + 1]);
            mx = max(mx, r[3 * i + 2]);
        }
    }
};

// Two 16-bit keys per 32-bit register, handled with packed 16-bit integer ops.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// keys of unit u at slots >= n (the segment's last unit) read as "no key"
__device__ __forceinline__ void unit_tail(uint4 &v, uint32_t u, uint32_t n) {
    const uint32_t r = n - 8 * u;   // keys of this unit inside the segment
    if (r >= 8) return;
    uint32_t c[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (uint32_t i = 0; i < 4; i++)
        c[i] |= (2 * i >= r ? 0xffffu : 0u) | (2 * i + 1 >= r ? 0xffff0000u : 0u);
    v = make_uint4(c[0], c[1], c[2], c[3]);
}

// smallest key of a unit
__device__ __forceinline__ uint32_t unit_min(const uint4 v) {
    const u16x2 m = __builtin_elementwise_min(
        __builtin_elementwise_min(as_u16x2(v.x), as_u16x2(v.y)),
        __builtin_elementwise_min(as_u16x2(v.z), as_u16x2(v.w)));
    return min((uint32_t)m.x, (uint32_t)m.y);
}

// A wave's view of a histogram of 64 * NB bins in LDS: lane l covers bins
// NB*l .. NB*l + NB-1 (their sum and the count below them); the histogram
// stays in place while find() runs.
template <int NB>
struct BinTab {
    const uint32_t *hist;
    uint32_t sm, excl, tot;
    __device__ __forceinline__ void load(const uint32_t *h) {
        const uint32_t lane = threadIdx.x & 63;
        hist = h;
        sm = 0;
#pragma unroll
        for (int i = 0; i < NB; i++) sm += h[NB * lane + i];
        excl = wave_excl(sm, &tot);
    }
    // bin where the running count reaches rem (1 <= rem <= tot): (bin, below)
    __device__ __forceinline__ void find(uint32_t rem, uint32_t &bin, uint32_t &below) const {
        const uint32_t lane = threadIdx.x & 63;
        uint32_t b = 0, bl = 0;
        const bool mine = excl < rem && rem <= excl + sm;
        if (mine) {
            uint32_t c = excl;
            for (int i = 0; i < NB; i++) {
                const uint32_t h = hist[NB * lane + i];
                if (rem <= c + h) {
                    b = NB * lane + i;
                    bl = c;
                    break;
                }
                c += h;
            }
        }
        const uint64_t who = __ballot(mine);
        const int src = who ? __ffsll((unsigned long long)who) - 1 : 0;
        bin = __shfl(b, src, 64);
        below = __shfl(bl, src, 64);
    }
};

// keys of word w (32 slots, four units) at slots >= n read as "no key"
__device__ __forceinline__ void word_tail(uint4 (&r)[4], uint32_t w, uint32_t n) {
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        const uint32_t u = 4 * w + k;
        if (8 * u >= n) r[k] = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
        else unit_tail(r[k], u, n);
    }
}

// The `want` smallest 16-bit keys of one query's segment, exactly-tied keys
// at the threshold taken in slot order, written as one selection mask per
// 32-slot word (sel_bits, at the segment's word offset; a_bits marks the `ca`
// smallest of them the same way), or else as entries (map) or slots, with
// their keys (pool_key, optional).  hist: 1,152 words of LDS.
// Thread t of the group owns the segment's 32-slot words t, t + 64*NW, ...;
// the first RW stay in registers after the first pass, so a segment of up to
// RW * 32 * 64 * NW keys is read from memory once; the rest is re-read by
// each later pass.  Passes: min / max (packed 16-bit); a 1,024-bin histogram
// of (key - min) >> sh, whose total is the live count; only when the keys
// span more than 1,024 values, histograms of the low sh bits inside the two
// threshold bins; collect in slot order (rounds of 64 * NW words, one group
// scan each), where a unit whose smallest key is above the pool threshold is
// known to hold no member.
template <bool BITS, int NW, int RW>
__device__ __forceinline__ void pool_group(const uint16_t *kp, uint32_t q, uint32_t n,
                                           uint32_t want, uint32_t *pool_entry,
                                           uint16_t *pool_key, uint32_t *pool_want, bool map,
                                           const ProbeTab &tab, uint32_t *hist, PoolGroup<NW> &g,
                                           uint32_t *sel_bits, uint32_t *a_bits, uint32_t ca,
                                           uint32_t *a_min) {
    constexpr uint32_t T = PoolGroup<NW>::T;
    const uint32_t t = PoolGroup<NW>::tid();
    const uint32_t nwords = (n + 31) / 32;
    uint4 reg[RW][4];
#pragma unroll
    for (int i = 0; i < RW; i++)
        if (t + T * i < nwords) load_word_raw(kp, t + T * i, reg[i]);
#pragma unroll
    for (int i = 0; i < RW; i++)
        if (t + T * i < nwords) word_tail(reg[i], t + T * i, n);
    // every owned word once, registers first: body(word, index)
    auto sweep = [&](auto &&body) {
#pragma unroll
        for (int i = 0; i < RW; i++)
            if (t + T * i < nwords) body(reg[i], t + T * i);
        for (uint32_t w0 = t + T * RW; w0 < nwords; w0 += 2 * T) {
            uint4 v[2][4];
#pragma unroll
            for (int i = 0; i < 2; i++)
                if (w0 + T * i < nwords) load_word_raw(kp, w0 + T * i, v[i]);
#pragma unroll
            for (int i = 0; i < 2; i++)
                if (w0 + T * i < nwords) {
                    word_tail(v[i], w0 + T * i, n);
                    body(v[i], w0 + T * i);
                }
        }
    };
    for (uint32_t i = t; i < 1152; i += T) hist[i] = 0;
    // min over raw keys ("no key" = 0xffff never wins); max over key + 1
    // (0xffff wraps to 0 and never wins)
    u16x2 mn2 = {0xffff, 0xffff}, mx2 = {0, 0};
    const u16x2 one2 = {1, 1};
    sweep([&](const uint4 (&r)[4], uint32_t) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t c[4] = {r[k].x, r[k].y, r[k].z, r[k].w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                mn2 = __builtin_elementwise_min(mn2, as_u16x2(c[i]));
                mx2 = __builtin_elementwise_max(mx2, as_u16x2(c[i]) + one2);
            }
        }
    });
    uint32_t zero = 0, kmn = min((uint32_t)mn2.x, (uint32_t)mn2.y);
    uint32_t kmx1 = max((uint32_t)mx2.x, (uint32_t)mx2.y);   // largest live key + 1; 0: none
    g.reduce3(zero, kmn, kmx1);   // (its barrier also orders the hist clear)
    if (NW == 1) g.sync();
    uint32_t *h1 = hist, *h2 = hist + 1024, *h3 = hist + 1088;
    uint32_t nvalid = 0, T1 = 0xffffu, need = 0, TA = 0xffffu, needA = 0, W = 0;
    bool all = true, allA = true;
    if (kmx1 > 0) {
        const uint32_t span = kmx1 - 1 - kmn;
        const uint32_t sh = span < 1024 ? 0 : 32 - __clz(span) - 10;
        const u16x2 kmn2 = {(unsigned short)kmn, (unsigned short)kmn};
        sweep([&](const uint4 (&r)[4], uint32_t) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t c[4] = {r[k].x, r[k].y, r[k].z, r[k].w};
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t d = as_u32(as_u16x2(c[i]) - kmn2);
                    if ((c[i] & 0xffffu) != 0xffffu) atomicAdd(&h1[(d & 0xffffu) >> sh], 1u);
                    if (c[i] < 0xffff0000u) atomicAdd(&h1[d >> (16 + sh)], 1u);
                }
            }
        });
        g.sync();
        BinTab<16> b1;   // every wave finds the same bins
        b1.load(h1);
        nvalid = b1.tot;
        W = min(want, nvalid);
        all = W == nvalid;
        // phase-A rank (a_bits): the `ca` smallest of the pool, same tie rule
        const uint32_t nA = a_bits ? min(ca, W) : W;
        allA = nA == nvalid;   // (all implies allA)
        if (!allA) {
            uint32_t bW = 0, belowW = 0, bA = 0, belowA = 0;
            if (!all) b1.find(W, bW, belowW);
            b1.find(nA, bA, belowA);
            uint32_t b2W = 0, below2W = 0, b2A = 0, below2A = 0;
            if (sh > 0) {   // at most 64 sub-bins each
                const uint32_t mask = (1u << sh) - 1u;
                sweep([&](const uint4 (&r)[4], uint32_t) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const uint32_t c[4] = {r[k].x, r[k].y, r[k].z, r[k].w};
#pragma unroll
                        for (int j = 0; j < 8; j++) {
                            const uint32_t kk = (c[j >> 1] >> (16 * (j & 1))) & 0xffffu;
                            if (kk == 0xffffu) continue;
                            const uint32_t bin = (kk - kmn) >> sh;
                            if (!all && bin == bW) atomicAdd(&h2[(kk - kmn) & mask], 1u);
                            if (bin == bA) atomicAdd(&h3[(kk - kmn) & mask], 1u);
                        }
                    }
                });
                g.sync();
                BinTab<1> b2, b3;
                if (!all) {
                    b2.load(h2);
                    b2.find(W - belowW, b2W, below2W);
                }
                b3.load(h3);
                b3.find(nA - belowA, b2A, below2A);
            }
            if (!all) {
                T1 = kmn + (bW << sh) + b2W;
                need = W - belowW - below2W;   // keys equal to T1 to take, in slot order
            }
            TA = kmn + (bA << sh) + b2A;
            needA = nA - belowA - below2A;
        }
    }
    if (t == 0) {
        pool_want[q] = W;
        // members outside A have keys >= TA (A takes every key below it)
        if (a_min) a_min[q] = (a_bits && min(ca, W) < W) ? TA : 0xffffu;
    }
    // collect: round r = words r*T .. r*T + T-1 in thread order (slot order)
    uint32_t tie_run = 0, tieA_run = 0, mark_run = 0;
    auto collect = [&](const uint4 (&r)[4], uint32_t w) {
        uint32_t live = 0, lt = 0, eq = 0, ltA = 0, eqA = 0;
        if (w < nwords)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                // TA <= T1, and T1 < 0xffff unless every live key is taken
                if (unit_min(r[k]) > T1) continue;
                uint32_t lv, l, e, la, ea;
                unit_masks2(r[k], 4 * w + k, n, T1, TA, lv, l, e, la, ea);
                live |= lv << (8 * k);
                lt |= l << (8 * k);
                eq |= e << (8 * k);
                ltA |= la << (8 * k);
                eqA |= ea << (8 * k);
            }
        // tie counts of this round (each at most 32 * T) packed in one scan
        uint32_t tt;
        const uint32_t pk = g.excl((all ? 0u : __popc(eq)) | ((allA ? 0u : __popc(eqA)) << 16), &tt);
        const uint32_t tr = tie_run + (pk & 0xffffu), trA = tieA_run + (pk >> 16);
        tie_run += tt & 0xffffu;
        tieA_run += tt >> 16;
        const uint32_t sel = all ? live : lt | (tr < need ? lowest_bits(eq, need - tr) : 0u);
        if (BITS) {   // one mask store per word
            if (w < nwords) {
                sel_bits[w] = sel;
                if (a_bits) a_bits[w] = allA ? live : ltA | (trA < needA ? lowest_bits(eqA, needA - trA) : 0u);
            }
            return;
        }
        uint32_t mt;
        uint32_t pre = mark_run + g.excl(__popc(sel), &mt);
        mark_run += mt;
        if (sel) {   // pool positions follow slot order
            uint32_t bits = sel;
            uint32_t pj = map ? tab.find(32 * w + (__ffs(bits) - 1)) : 0u;
            while (bits) {
                const uint32_t j = __ffs(bits) - 1;
                bits &= bits - 1;
                const size_t o = (size_t)q * want + pre++;
                pool_entry[o] = map ? tab.entry_from(pj, 32 * w + j) : 32 * w + j;
                if (pool_key) pool_key[o] = kp[32 * w + j];   // (from memory: no dynamic
                                                                 // index into the registers)
            }
        }
    };
#pragma unroll
    for (int i = 0; i < RW; i++)
        if (T * i < nwords) collect(reg[i], t + T * i);   // (uniform over the group)
    for (uint32_t r0 = T * RW; r0 < nwords; r0 += T) {
        const uint32_t w = r0 + t;
        uint4 v[4];
        if (w < nwords) {
            load_word_raw(kp, w, v);
            word_tail(v, w, n);
        }
        collect(v, w);
    }
}

constexpr uint32_t kWaveTab = 256;   // probe tables held in LDS up to this n_probe

// NW wavefronts per query (4 / NW queries per 256-thread workgroup), RW
// register-held 32-slot words per thread; four workgroups per CU.
template <bool BITS, int NW, int RW>
__global__ __launch_bounds__(256, 4) void k_pool_select(const PoolParams p, uint32_t nq) {
    constexpr int G = 4 / NW;
    __shared__ uint32_t hist[G][1152];
    __shared__ uint32_t sx[G][5 * NW];
    __shared__ uint32_t s_pbase[BITS ? 1 : G][BITS ? 1 : kWaveTab], s_ebase[BITS ? 1 : G][BITS ? 1 : kWaveTab];
    const uint32_t gi = threadIdx.x / (64 * NW);
    const uint32_t t = PoolGroup<NW>::tid();
    const uint32_t q = blockIdx.x * G + gi;
    if (q >= nq) return;   // whole group: its barriers involve only its own threads
    const uint64_t b = p.seg_off ? p.seg_off[q] : (uint64_t)q * p.seg_stride;   // multiple of 32
    const uint32_t n = p.seg_cnt ? p.seg_cnt[q] : p.seg_len;
    if (BITS && p.long_from && n > p.long_from) return;   // (the chunked select's)
    ProbeTab tab;
    tab.np = (!BITS && p.map) ? p.np_q[q] : 0u;
    tab.qp_row = p.qp_slot + (size_t)q * p.npr;
    tab.probe_row = p.probe_post + (size_t)q * p.npr;
    tab.post_off = p.post_off;
    tab.seg = b;
    tab.pbase = tab.ebase = nullptr;
    PoolGroup<NW> g;
    g.sx = sx[gi];
    if (!BITS && tab.np && tab.np <= kWaveTab) {
        for (uint32_t j = t; j < tab.np; j += 64 * NW) {
            s_pbase[gi][j] = (uint32_t)(tab.qp_row[j] - b);
            s_ebase[gi][j] = p.post_off[tab.probe_row[j]];
        }
        g.sync();
        tab.pbase = s_pbase[gi];
        tab.ebase = s_ebase[gi];
    }
    pool_group<BITS, NW, RW>(p.keys + b, q, n, p.want, p.pool_entry, p.pool_key, p.pool_want,
                             !BITS && p.map, tab, hist[gi], g,
                             BITS ? p.sel_bits + b / 32 : nullptr,
                             BITS && p.a_bits ? p.a_bits + b / 32 : nullptr, p.ca,
                             BITS ? p.a_min : nullptr);
}

// ---------------------------------------------------------------- long candidate lists
// Segments far above the batch's average work per pool-select workgroup (the
// Zipf-sized postings of a skewed index give some queries over a million
// candidates, which one workgroup would select long after the others are
// done) are selected by many workgroups, one per kSelChunk slots.  A chunk's `stride` smallest (key, slot) values
// contain every pool member in the chunk (the pool is the `want` <= stride
// smallest of the segment), and the chunk lists of a query, concatenated in
// chunk order, are in slot order, so the pool and phase A taken from them
// with the same rule (keys, ties in slot order) are the segment's.
__global__ void k_pool_chunk_count(const PoolLongParams p) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.nq) return;
    const uint32_t n = p.seg_cnt[q];
    p.chunk_cnt[q] = n > p.long_from ? (n + kSelChunk - 1) / kSelChunk : 0u;
}

__global__ __launch_bounds__(256, 4) void k_pool_chunk(const PoolLongParams p) {
    __shared__ uint32_t hist[1152];
    __shared__ uint32_t sx[20];
    const uint32_t c = blockIdx.x, tid = threadIdx.x;
    if (c >= p.chunk_off[p.nq]) return;   // (the grid is a bound)
    uint32_t lo = 0, hi = p.nq;            // the query whose chunk range holds c
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (p.chunk_off[mid] <= c) lo = mid; else hi = mid;
    }
    const uint32_t q = lo, j = c - p.chunk_off[q];
    const uint64_t b = p.seg_off[q] + (uint64_t)j * kSelChunk;   // a multiple of 32
    const uint32_t m = min(kSelChunk, p.seg_cnt[q] - j * kSelChunk);
    // the chunk's words of the segment's masks start empty (the merge sets bits)
    for (uint32_t w = tid; w < (m + 31) / 32; w += 256) {
        p.sel_bits[b / 32 + w] = 0u;
        p.a_bits[b / 32 + w] = 0u;
    }
    ProbeTab tab;
    tab.np = 0;
    tab.pbase = tab.ebase = nullptr;
    PoolGroup<4> g;
    g.sx = sx;
    pool_group<false, 4, 3>(p.keys + b, c, m, p.stride, p.lslot, p.lkey, p.lcnt, false, tab, hist, g,
                            nullptr, nullptr, 0, nullptr);
    __syncthreads();
    const uint32_t cnt = p.lcnt[c];
    for (uint32_t i = tid; i < p.stride; i += 256) {
        const size_t o = (size_t)c * p.stride + i;
        if (i < cnt) p.lslot[o] += j * kSelChunk;   // chunk slot -> segment slot
        else p.lkey[o] = 0xffffu;                   // "no key" to the stride
    }
}

__global__ __launch_bounds__(256, 4) void k_pool_merge(const PoolLongParams p) {
    __shared__ uint32_t hist[1152];
    __shared__ uint32_t sx[20];
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    if (q >= p.nq || p.chunk_cnt[q] == 0) return;   // whole workgroup
    const uint32_t c0 = p.chunk_off[q];
    const size_t base = (size_t)c0 * p.stride;   // a multiple of 32
    const uint32_t n = p.chunk_cnt[q] * p.stride;
    ProbeTab tab;
    tab.np = 0;
    tab.pbase = tab.ebase = nullptr;
    PoolGroup<4> g;
    g.sx = sx;
    pool_group<true, 4, 3>(p.lkey + base, q, n, p.want, nullptr, nullptr, p.pool_want, false, tab,
                           hist, g, p.lsel + base / 32, p.la + base / 32, p.ca, p.a_min);
    __syncthreads();
    // the members' bits at their segment slots
    uint32_t *sel = p.sel_bits + p.seg_off[q] / 32, *a = p.a_bits + p.seg_off[q] / 32;
    for (uint32_t w = tid; w < n / 32; w += 256) {
        uint32_t ms = p.lsel[base / 32 + w];
        const uint32_t ma = p.la[base / 32 + w];
        while (ms) {
            const uint32_t i = __ffs(ms) - 1;
            ms &= ms - 1;
            const uint32_t slot = p.lslot[base + 32 * w + i];
            atomicOr(&sel[slot >> 5], 1u << (slot & 31));
            if ((ma >> i) & 1u) atomicOr(&a[slot >> 5], 1u << (slot & 31));
        }
    }
}

// ---------------------------------------------------------------- admission
// Two-phase scans: the first phase (each query's nearest probes, up to `want`
// entries) stores dense keys; k_admit takes the `want`-th smallest live key
// among them as the query's bound, an upper bound of its pool's largest key
// (the pool is the `want` smallest keys of a superset); the second phase lists
// only keys at most the bound; k_pool_list selects the pool from both.

// bin (of the wave's 256-bin histogram h) holding the need-th smallest value
// (1 <= need <= total); *below = values in lower bins, *cnt = values in it
__device__ __forceinline__ uint32_t wave_bin(const uint32_t *h, uint32_t need, uint32_t *below,
                                             uint32_t *cnt) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t c[4], sum = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        c[i] = h[4 * lane + i];
        sum += c[i];
    }
    const uint32_t incl = wave_incl_scan(sum);
    const uint64_t m = __ballot(incl >= need);
    const uint32_t f = (uint32_t)(__ffsll((unsigned long long)m) - 1);
    uint32_t b = 0, run = incl - sum, cb = 0;
    if (lane == f) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (run + c[i] >= need) {
                b = 4 * lane + i;
                cb = c[i];
                break;
            }
            run += c[i];
        }
    }
    *below = __shfl(run, f, 64);
    *cnt = __shfl(cb, f, 64);
    return __shfl(b, f, 64);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Wave-level radix select over the values in [vmin, vmax] that visit(f)
// passes (f(v) once per counted value; lanes may make different numbers of
// calls): the need-th smallest, or, when a digit's bin holds exactly the
// values still needed, the largest value that bin could hold (clamped to
// vmax).  Either way at least `need` counted values (exactly `need` when they
// are distinct) lie at or below the result, and fewer below it.  Digits are
// 8 bits of v - vmin from the top of the range down, so the first digit
// already spreads the values over the bins (few same-bin LDS atomics).
template <class Visit>
__device__ __forceinline__ uint32_t wave_select(uint32_t *h, uint32_t need, uint32_t vmin,
                                                uint32_t vmax, Visit &&visit) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t span = vmax - vmin;
    const int bits = span ? 32 - __clz(span) : 0;
    int sh = bits > 8 ? bits - 8 : 0, fixed = 32;   // offset bits >= fixed are decided
    uint32_t prefix = 0;
    for (;;) {
#pragma unroll
        for (int i = 0; i < 4; i++) h[4 * lane + i] = 0u;
        wave_lds_sync();
        visit([&](uint32_t v) {
            if (v < vmin || v > vmax) return;
            const uint32_t o = v - vmin;
            if (fixed < 32 && (o >> fixed) != (prefix >> fixed)) return;
            atomicAdd(&h[(o >> sh) & 255u], 1u);
        });
        wave_lds_sync();
        uint32_t below, cb;
        const uint32_t b = wave_bin(h, need, &below, &cb);
        prefix |= b << sh;
        need -= below;
        wave_lds_sync();   // every lane has read the histogram
        if (cb == need) return vmin + min(prefix | ((1u << sh) - 1u), span);
        if (sh == 0) return vmin + prefix;
        fixed = sh;
        sh = sh > 8 ? sh - 8 : 0;
    }
}

__device__ __forceinline__ void wave_minmax(uint32_t &mn, uint32_t &mx) {
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, off, 64));
        mx = max(mx, (uint32_t)__shfl_xor(mx, off, 64));
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// the eight 16-bit keys of unit u (u < nu) of a segment, every lane alike
template <class F>
__device__ __forceinline__ void visit_units(const uint4 *kp, uint32_t nu, F &&f) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t u0 = 0; u0 < nu; u0 += 128) {   // two loads in flight per lane
        uint4 v[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint32_t u = u0 + 64 * c + lane;
            v[c] = u < nu ? kp[u] : make_uint4(~0u, ~0u, ~0u, ~0u);
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint32_t u = u0 + 64 * c + lane;
            const uint32_t w4[4] = {v[c].x, v[c].y, v[c].z, v[c].w};
#pragma unroll
            for (int t = 0; t < 8; t++) f((w4[t >> 1] >> (16 * (t & 1))) & 0xffffu, 8 * u + t);
        }
    }
}

// one wave per query, four per workgroup (no workgroup barriers)
__global__ __launch_bounds__(256) void k_admit(const AdmitParams p) {
    __shared__ uint32_t hist[4][256];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t q = blockIdx.x * 4 + wv;
    if (q >= p.nq) return;   // whole wave
    const uint4 *kp = reinterpret_cast<const uint4 *>(p.keys + p.seg_off[q]);
    const uint32_t nu = p.s1[q] / 8;   // first-phase slots: a multiple of the scan alignment (8)
    uint32_t n = 0, mn = ~0u, mx = 0u;
    visit_units(kp, nu, [&](uint32_t k, uint32_t) {
        if (k == 0xffffu) return;
        n++;
        mn = min(mn, k);
        mx = max(mx, k);
    });
    n = wave_sum(n);
    wave_minmax(mn, mx);
    uint32_t tau = 0xfffeu;   // fewer live keys than `want`: admit every key
    if (n >= p.want)
        tau = wave_select(hist[wv], p.want, mn, mx, [&](auto &&f) {
            visit_units(kp, nu, [&](uint32_t k, uint32_t) {
                if (k != 0xffffu) f(k);
            });
        });
    if (lane == 0) {
        p.tau[q] = tau;
        p.cnt[q] = 0u;   // the second phase's list of the query starts empty
    }
}

hipError_t launch_admit(const AdmitParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    if (p.want == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_admit, dim3((p.nq + 3) / 4), dim3(256), 0, s, p);
    return hipGetLastError();
}

constexpr uint32_t kPoolLds = 8192;     // a query's pool candidates held in LDS (else re-read)

struct PoolListLds {
    uint32_t v[kPoolLds];                       // candidates, key << 16 | slot
    uint32_t h[256];                            // histogram
    uint32_t w[8], r[8];
};

// bin of L.h holding the need-th smallest value (1 <= need <= total), over
// the workgroup: *below = values in lower bins, *cnt = values in it
__device__ __forceinline__ uint32_t block_bin(PoolListLds &L, uint32_t need, uint32_t *below,
                                              uint32_t *cnt) {
    const uint32_t t = threadIdx.x, c = L.h[t];
    uint32_t tot;
    const uint32_t ex = block_scan_fast(c, L.w, &tot);
    if (ex < need && ex + c >= need) {
        L.r[0] = t;
        L.r[1] = ex;
        L.r[2] = c;
    }
    __syncthreads();
    const uint32_t b = L.r[0];
    *below = L.r[1];
    *cnt = L.r[2];
    __syncthreads();   // (L.r and the histogram are rewritten next)
    return b;
}

// One workgroup per query.  Candidates: first-phase keys at most tau, then
// the query's second-phase list (contiguous).  A key histogram (bins of the
// keys' range, a finer pass when it exceeds 256 keys) gives the pool's
// boundary key K (its `want`-th smallest) and phase A's boundary key KA (its
// min(ca, pool)-th smallest); pool members: keys below K, and of the keys
// equal to K those with the smallest slots (a radix select over the slots,
// only when not all of them are taken); phase A: members with keys below KA
// when there are k of them, else up to KA.
__global__ __launch_bounds__(256) void k_pool_list(const PoolListParams p, uint32_t k) {
    __shared__ PoolListLds L;
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t q = blockIdx.x;
    const uint64_t seg = p.seg_off[q];
    const uint4 *kp = reinterpret_cast<const uint4 *>(p.keys + seg);
    const uint32_t nu = p.s1[q] / 8, tau = p.tau[q];
    const uint32_t *lst = p.adm_list + seg;
    const uint32_t na2 = p.adm_cnt[q];
    uint32_t kmin = 0xffffu, kmax = 0u;
    // 1. first-phase candidates: thread per 8-key unit, positions by a scan
    uint32_t n1 = 0;
    for (uint32_t u0 = 0; u0 < nu; u0 += 256) {
        const uint32_t u = u0 + tid;
        uint32_t w4[4] = {~0u, ~0u, ~0u, ~0u};
        if (u < nu) {
            const uint4 v = kp[u];
            w4[0] = v.x;
            w4[1] = v.y;
            w4[2] = v.z;
            w4[3] = v.w;
        }
        uint32_t m = 0;
#pragma unroll
        for (int t = 0; t < 8; t++)
            m |= (uint32_t)(((w4[t >> 1] >> (16 * (t & 1))) & 0xffffu) <= tau) << t;
        uint32_t tot;
        uint32_t pos = n1 + block_scan_fast((uint32_t)__popc(m), L.w, &tot);
#pragma unroll
        for (int t = 0; t < 8; t++)
            if ((m >> t) & 1u) {
                const uint32_t kk = (w4[t >> 1] >> (16 * (t & 1))) & 0xffffu;
                if (pos < kPoolLds) L.v[pos] = (kk << 16) | (8 * u + t);
                pos++;
                kmin = min(kmin, kk);
                kmax = max(kmax, kk);
            }
        n1 += tot;
    }
    // 2. the second-phase list after them (four loads in flight per thread)
    const uint32_t n = n1 + na2;
    const bool lds = n <= kPoolLds;
    auto visit_list = [&](auto &&f) {
        for (uint32_t i0 = 0; i0 < na2; i0 += 1024) {
            uint32_t v[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t i = i0 + 256 * c + tid;
                v[c] = i < na2 ? lst[i] : 0u;
            }
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (i0 + 256 * c + tid < na2) f(i0 + 256 * c + tid, v[c]);
        }
    };
    visit_list([&](uint32_t i, uint32_t v) {
        if (lds) L.v[n1 + i] = v;
        kmin = min(kmin, v >> 16);
        kmax = max(kmax, v >> 16);
    });
    // key range over the workgroup
    wave_minmax(kmin, kmax);
    if (lane == 0) {
        L.w[4 + (tid >> 6)] = kmin;
        L.h[tid >> 6] = kmax;
    }
    __syncthreads();
    kmin = min(min(L.w[4], L.w[5]), min(L.w[6], L.w[7]));
    kmax = max(max(L.h[0], L.h[1]), max(L.h[2], L.h[3]));
    __syncthreads();
    // every candidate once, f(v)
    auto visit = [&](auto &&f) {
        if (lds) {
            for (uint32_t i = tid; i < n; i += 256) f(L.v[i]);
        } else {
            for (uint32_t u = tid; u < nu; u += 256) {
                const uint4 v = kp[u];
                const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    const uint32_t kk = (w4[t >> 1] >> (16 * (t & 1))) & 0xffffu;
                    if (kk <= tau) f((kk << 16) | (8 * u + t));
                }
            }
            visit_list([&](uint32_t, uint32_t v) { f(v); });
        }
    };
    // histogram of sel(v) (values whose bin is below 256)
    auto hist = [&](auto &&sel) {
        L.h[tid] = 0u;
        __syncthreads();
        visit([&](uint32_t v) {
            const uint32_t b = sel(v);
            if (b < 256u) atomicAdd(&L.h[b], 1u);
        });
        __syncthreads();
    };
    const uint32_t c = min(n, p.want), need_a = min(c, p.ca);
    uint32_t kst = 0x10000u, vst = ~0u, ka = 0x10000u, nA = c;   // (n <= want: every candidate)
    bool a_lt = false;   // phase A: keys below ka (else up to ka)
    if (n > 0) {
        const uint32_t span = kmax - kmin;
        const int sh = span > 255 ? 32 - __clz(span) - 8 : 0;   // (span < 2^16: sh <= 8)
        const uint32_t lowm = (1u << sh) - 1u;
        hist([&](uint32_t v) { return ((v >> 16) - kmin) >> sh; });
        const bool bound = n > p.want;
        uint32_t bl_s = 0, cb_s = 0, bl_a, cb_a, bs = 0;
        if (bound) bs = block_bin(L, p.want, &bl_s, &cb_s);
        uint32_t ba = block_bin(L, need_a, &bl_a, &cb_a);
        // the finer digit of a coarse bin: the key holding the need-th value
        auto fine = [&](uint32_t coarse, uint32_t need, uint32_t &bl, uint32_t &cb) -> uint32_t {
            hist([&](uint32_t v) {
                const uint32_t o = (v >> 16) - kmin;
                return (o >> sh) == coarse ? (o & lowm) : 256u;
            });
            uint32_t l2, c2;
            const uint32_t b2 = block_bin(L, need - bl, &l2, &c2);
            bl += l2;
            cb = c2;
            return (coarse << sh) + b2;
        };
        if (sh > 0) {
            if (bound) bs = fine(bs, p.want, bl_s, cb_s);
            ba = fine(ba, need_a, bl_a, cb_a);
        }
        ka = kmin + ba;
        if (bound) {
            kst = kmin + bs;
            uint32_t m = p.want - bl_s;   // keys equal to kst the pool takes
            if (m < cb_s) {   // the m smallest slots among them: two 8-bit digits
                const uint32_t base = kst << 16;
                hist([&](uint32_t v) { return (v >> 16) == kst ? (v >> 8) & 255u : 256u; });
                uint32_t l1, c1;
                const uint32_t s_hi = block_bin(L, m, &l1, &c1);
                m -= l1;
                if (m < c1) {
                    const uint32_t hi8 = (base >> 8) | s_hi;
                    hist([&](uint32_t v) { return (v >> 8) == hi8 ? v & 255u : 256u; });
                    uint32_t l2, c2;
                    vst = base | (s_hi << 8) | block_bin(L, m, &l2, &c2);
                } else {
                    vst = base | (s_hi << 8) | 255u;
                }
            } else {
                vst = (kst << 16) | 0xffffu;
            }
        }
        // phase A: keys below ka when k of them exist (its size then at most
        // ca), else keys up to ka (all members when ka is the pool's boundary)
        a_lt = bl_a >= k;
        nA = a_lt ? bl_a : ka < kst ? bl_a + cb_a : c;
    }
    // members: phase A first, then the rest (each part in any order)
    if (tid == 0) {
        L.r[4] = 0u;
        L.r[5] = nA;
        L.r[6] = 0xffffu;
    }
    __syncthreads();
    uint32_t *out = p.pool + (size_t)q * p.want;
    uint32_t amin = 0xffffu;
    visit([&](uint32_t v) {
        const uint32_t kk = v >> 16;
        if (!(kk < kst || (kk == kst && v <= vst))) return;
        if (a_lt ? kk < ka : kk <= ka) {
            out[atomicAdd(&L.r[4], 1u)] = v;
        } else {
            out[atomicAdd(&L.r[5], 1u)] = v;
            amin = min(amin, kk);
        }
    });
    for (int off = 32; off > 0; off >>= 1) amin = min(amin, (uint32_t)__shfl_xor(amin, off, 64));
    if (lane == 0 && amin != 0xffffu) atomicMin(&L.r[6], amin);
    __syncthreads();
    if (tid == 0) {
        p.pool_want[q] = c;
        p.pool_na[q] = nA;
        p.a_min[q] = L.r[6];
    }
}

hipError_t launch_pool_list(const PoolListParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    if (p.ca == 0 || p.ca > p.want || p.k == 0 || p.k > p.ca)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pool_list, dim3(p.nq), dim3(256), 0, s, p, p.k);
    return hipGetLastError();
}

// ---------------------------------------------------------------- rerank
// Exact f32 distance, four lanes per pair (lane r of the quad runs chain r over
// units r, r+4, ...), H pairs per quad in flight; canonical 4-chain order.
__device__ __forceinline__ float quad_sum(float c) {
    const float c01 = c + __shfl_xor(c, 1, 64);
    return c01 + __shfl_xor(c01, 2, 64);
}

// H rows per quad; UC units of each row are loaded together per lane
template <int H, int UC>
__device__ __forceinline__ void quad_exact_h(uint32_t metric, const float4 *q4,
                                             const float4 *const x4[H], uint32_t units, uint32_t r,
                                             float qn, float out[H]) {
    float acc[H], b2[H];
#pragma unroll
    for (int h = 0; h < H; h++) acc[h] = b2[h] = 0.0f;
    const bool l2 = metric == M_L2 || metric == M_L2SQ;
    // UC units of each row are loaded together, then consumed in chain
    // order (units is a multiple of 4, so the bound is uniform over the quad)
    for (uint32_t u0 = r; u0 < units; u0 += 4 * UC) {
        float4 xc[UC][H];
#pragma unroll
        for (int c = 0; c < UC; c++)
#pragma unroll
            for (int h = 0; h < H; h++)
                xc[c][h] = u0 + 4 * c < units ? x4[h][u0 + 4 * c] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int c = 0; c < UC; c++) {
        const uint32_t u = u0 + 4 * c;
        if (u >= units) break;
        const float4 a = q4[u];
#pragma unroll
        for (int h = 0; h < H; h++) {
            const float4 b = xc[c][h];
            if (l2) {
                float t = a.x - b.x; acc[h] = __builtin_fmaf(t, t, acc[h]);
                t = a.y - b.y; acc[h] = __builtin_fmaf(t, t, acc[h]);
                t = a.z - b.z; acc[h] = __builtin_fmaf(t, t, acc[h]);
                t = a.w - b.w; acc[h] = __builtin_fmaf(t, t, acc[h]);
            } else {
                acc[h] = __builtin_fmaf(a.x, b.x, acc[h]);
                acc[h] = __builtin_fmaf(a.y, b.y, acc[h]);
                acc[h] = __builtin_fmaf(a.z, b.z, acc[h]);
                acc[h] = __builtin_fmaf(a.w, b.w, acc[h]);
                b2[h] = __builtin_fmaf(b.x, b.x, b2[h]);
                b2[h] = __builtin_fmaf(b.y, b.y, b2[h]);
                b2[h] = __builtin_fmaf(b.z, b.z, b2[h]);
                b2[h] = __builtin_fmaf(b.w, b.w, b2[h]);
            }
        }
        }
    }
#pragma unroll
    for (int h = 0; h < H; h++) {
        const float a = quad_sum(acc[h]);
        float d;
        if (metric == M_L2SQ) {
            d = a;
        } else if (metric == M_L2) {
            d = sqrtf(a);
        } else if (metric == M_DOT) {
            d = 1.0f - a;
        } else {
            const float bb = quad_sum(b2[h]);
            if (qn == 0.0f && bb == 0.0f) d = 0.0f;
            else if (a == 0.0f) d = 1.0f;
            else {
                const float rr = 1.0f - a / sqrtf(qn * bb);
                d = rr > 0.0f ? rr : 0.0f;
            }
        }
        out[h] = d;
    }
}

__device__ __forceinline__ void bitonic_lds(uint64_t *buf, uint32_t m) {
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < (m >> 1); i += 256) {
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

constexpr int kRerankH = 2, kRerankUC = 4;   // rows per quad, row units per lane in flight

__global__ __launch_bounds__(256) void k_rerank_pool(const RerankPoolParams p) {
    __shared__ uint64_t s_keys[kMaxK];
    extern __shared__ __align__(16) float s_qv[];   // dpad floats
    // XCD x takes a contiguous range of the queries
    const uint32_t x = blockIdx.x % kXcds, jx = blockIdx.x / kXcds;
    const uint32_t per = (p.nq + kXcds - 1) / kXcds;
    const uint32_t idx = x * per + jx;
    if (jx >= per || idx >= p.nq) return;
    const uint32_t q = idx;
    const uint32_t tid = threadIdx.x, r = tid & 3, quad = tid >> 2, units = p.dpad / 4;
    const uint32_t c = min(p.pool_want[q], p.want);
    for (uint32_t i = tid; i < p.dpad; i += 256) s_qv[i] = p.qv[(size_t)q * p.dpad + i];
    const float qn = p.qnorm ? p.qnorm[q] : 0.0f;
    __syncthreads();
    const uint32_t *pe = p.pool_entry + (size_t)q * p.want;
    for (uint32_t i0 = 0; i0 < c; i0 += 64 * kRerankH) {
        uint32_t ii[kRerankH], ee[kRerankH];
        const float4 *xr[kRerankH];
#pragma unroll
        for (int h = 0; h < kRerankH; h++) {
            ii[h] = i0 + 64 * h + quad;
            ee[h] = ii[h] < c ? pe[ii[h]] : 0u;
            xr[h] = reinterpret_cast<const float4 *>(p.raw_e + (size_t)ee[h] * p.dpad);
        }
        float d[kRerankH];
        quad_exact_h<kRerankH, kRerankUC>(p.metric, reinterpret_cast<const float4 *>(s_qv), xr, units, r,
                                          qn, d);
        if (r == 0) {
#pragma unroll
            for (int h = 0; h < kRerankH; h++)
                if (ii[h] < c) {
                    const uint32_t u = __float_as_uint(d[h]);
                    const uint32_t fk = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
                    s_keys[ii[h]] = ((uint64_t)fk << 32) | (p.rank ? p.rank[ee[h]] : ee[h]);
                }
        }
    }
    const uint32_t kk = min(p.k, c);
    if (tid == 0) p.out_cnt[q] = kk;
    if (c <= 512) {   // short pools: rank by counting (keys are distinct), scatter
        __syncthreads();
        for (uint32_t i = tid; i < c; i += 256) {
            const uint64_t v = s_keys[i];
            uint32_t rk = 0;
            for (uint32_t j = 0; j < c; j++) rk += s_keys[j] < v;
            if (rk < kk) p.out[(size_t)q * p.k + rk] = v;
        }
        return;
    }
    uint32_t m = 2;
    while (m < c) m <<= 1;
    for (uint32_t i = c + tid; i < m; i += 256) s_keys[i] = kNoKey;
    __syncthreads();
    bitonic_lds(s_keys, m);
    for (uint32_t i = tid; i < kk; i += 256) p.out[(size_t)q * p.k + i] = s_keys[i];
}

// ---------------------------------------------------------------- pruned rerank
// One workgroup per query (XCD-contiguous ranges of queries) over its
// pool of `c` (entry, 16-bit key) members, in pool order:
//   1. A = the `ca` smallest keys (ties in pool order), by two radix passes
//   2. A is reranked exactly; its k-th exact distance bounds the answer
//   3. members outside A whose key proves an exact distance above that bound
//      are dropped (prune), the rest reranked
//   4. the k smallest exact keys of A and the survivors are written.
// Dropped members cannot enter the top k, so the result is that of reranking
// the whole pool.  Without prune every member is reranked.
constexpr uint32_t kPrCount = 512;    // lists up to this size are ranked by counting
constexpr int kPrH = 2, kPrUC = 4;    // rows per quad, row units per lane in flight

__device__ __forceinline__ uint32_t blk_excl(uint32_t v, uint32_t *s_w) {
    uint32_t tot;
    return block_scan_fast(v, s_w, &tot);
}

// exact keys of `cnt` entries listed in LDS -> keys_out[0..cnt)
__device__ __forceinline__ void rerank_list(const RerankPrunedParams &p, const uint32_t *ents,
                                            uint32_t cnt, const float *s_qv, float qn,
                                            uint64_t *keys_out) {
    const uint32_t tid = threadIdx.x, r = tid & 3, quad = tid >> 2, units = p.dpad / 4;
    for (uint32_t i0 = 0; i0 < cnt; i0 += 64 * kPrH) {
        uint32_t ii[kPrH], ee[kPrH];
        const float4 *xr[kPrH];
#pragma unroll
        for (int h = 0; h < kPrH; h++) {
            ii[h] = i0 + 64 * h + quad;
            ee[h] = ii[h] < cnt ? ents[ii[h]] : ents[0];
            xr[h] = reinterpret_cast<const float4 *>(p.raw_e + (size_t)ee[h] * p.dpad);
        }
        float d[kPrH];
        quad_exact_h<kPrH, kPrUC>(p.metric, reinterpret_cast<const float4 *>(s_qv), xr, units, r, qn, d);
        if (r == 0) {
#pragma unroll
            for (int h = 0; h < kPrH; h++)
                if (ii[h] < cnt) keys_out[ii[h]] = ((uint64_t)fkey_m(d[h]) << 32) | p.rank[ee[h]];
        }
    }
}

// the k smallest of keys[0..cnt) (distinct) in ascending order -> out[0..min(k, cnt))
// emit(position, key) for the k smallest keys
template <class Emit>
__device__ __forceinline__ void smallest_k(uint64_t *keys, uint32_t cnt, uint32_t k, Emit &&emit) {
    const uint32_t tid = threadIdx.x;
    const uint32_t kk = min(k, cnt);
    if (cnt <= kPrCount) {   // rank by counting, scatter
        for (uint32_t i = tid; i < cnt; i += 256) {
            const uint64_t v = keys[i];
            uint32_t c = 0;
            for (uint32_t j = 0; j < cnt; j++) c += keys[j] < v;
            if (c < kk) emit(c, v);
        }
    } else {
        uint32_t m = 2;
        while (m < cnt) m <<= 1;
        for (uint32_t i = cnt + tid; i < m; i += 256) keys[i] = kNoKey;
        __syncthreads();
        bitonic_lds(keys, m);
        for (uint32_t i = tid; i < kk; i += 256) emit(i, keys[i]);
    }
}

template <uint32_t CAP>
struct RerankLds {
    uint64_t ex[CAP];            // exact keys: A, then survivors
    uint32_t ent[CAP];           // entries: A, then survivors
    uint32_t nzw[CAP], nzm[CAP], nzp[CAP];   // non-empty mask words: index, mask, first member
    uint32_t pbase[kWaveTab], ebase[kWaveTab];
    uint32_t w[8];
    uint32_t m[16];
    uint64_t kth;
};

// Members of the masks mask_of(w) over the query's words, in slot order, as
// entries out[0..): non-empty words are listed with the position of their
// first member, then each thread takes members i = tid, tid + 256, ... (word
// by binary search, bit by rank), so the work is even however the members
// cluster.  Returns the member count.  Thread t owns words [w0, w1), read B
// at a time in the counting pass (cheap mask_of: plain loads in flight
// together), then again up to its last non-empty word.
template <uint32_t CAP, int B = 1, class MaskOf>
__device__ __forceinline__ uint32_t rebuild_members(RerankLds<CAP> &L, uint32_t w0, uint32_t w1,
                                                    MaskOf &&mask_of, const ProbeTab &tab,
                                                    uint32_t *out, uint32_t cap) {
    const uint32_t tid = threadIdx.x;
    uint32_t nz = 0, cnt = 0;
    for (uint32_t w = w0; w < w1; w += B) {
        uint32_t m[B];
#pragma unroll
        for (int i = 0; i < B; i++) m[i] = w + i < w1 ? mask_of(w + i) : 0u;
#pragma unroll
        for (int i = 0; i < B; i++) {
            nz += m[i] != 0u;
            cnt += __popc(m[i]);
        }
    }
    // one scan of both counts (each total is at most the pool size, < 2^16)
    const uint32_t pk = blk_excl(nz | (cnt << 16), L.w);
    uint32_t zpos = pk & 0xffffu, ppos = pk >> 16;
    for (uint32_t w = w0, left = nz; w < w1 && left; w++) {
        const uint32_t m = mask_of(w);
        left -= m != 0u;
        if (m && zpos < CAP) {
            L.nzw[zpos] = w;
            L.nzm[zpos] = m;
            L.nzp[zpos] = ppos;
            zpos++;
            ppos += __popc(m);
        }
    }
    if (tid == 255) {   // the last thread's prefix + count = totals
        L.m[14] = zpos;
        L.m[15] = ppos;
    }
    __syncthreads();
    const uint32_t nzt = L.m[14], total = min(L.m[15], cap);
    for (uint32_t i = tid; i < total; i += 256) {
        uint32_t lo = 0, hi = nzt;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (L.nzp[mid] <= i) lo = mid; else hi = mid;
        }
        uint32_t m = L.nzm[lo];
        for (uint32_t r = i - L.nzp[lo]; r; r--) m &= m - 1;   // drop the lower members
        const uint32_t slot = 32 * L.nzw[lo] + (__ffs(m) - 1);
        uint32_t pj = tab.find(slot);
        out[i] = tab.entry_from(pj, slot);
    }
    __syncthreads();
    return total;
}

// (up to 1,024 members: at most 96 VGPRs, five workgroups a CU instead of
// four, no spills; the largest pools are bound by their LDS)
template <uint32_t CAP>
__global__ __launch_bounds__(256, CAP <= 1024 ? 5 : 1) void k_rerank_pruned(const RerankPrunedParams p) {
    __shared__ RerankLds<CAP> L;
    extern __shared__ __align__(16) float s_qv[];   // dpad floats
    enum { EXQ, XNQ, THR };
    const uint32_t x = blockIdx.x % kXcds, jx = blockIdx.x / kXcds;
    const uint32_t per = (p.nq + kXcds - 1) / kXcds;
    const uint32_t idx = x * per + jx;
    if (jx >= per || idx >= p.nq) return;   // whole workgroup
    const uint32_t q = idx;
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t c = min(p.pool_want[q], p.want);
    uint64_t *out = p.out + (size_t)q * p.k;
    // phase A: the `ca` smallest keys, or with list pools the members
    // listed first (at least min(ca, c) of them)
    const uint32_t nA = p.pool ? min(p.pool_na[q], c) : min(p.ca, c);
    const bool prune = p.prune && nA < c && nA >= p.k;
    const uint64_t b = p.seg_off[q];
    const uint32_t n = p.seg_cnt[q];
    const uint32_t nwords = (n + 31) / 32, Mw = (nwords + 255) / 256;
    const uint32_t w0 = tid * Mw, w1 = min(w0 + Mw, nwords);
    const uint32_t *sel = p.sel_bits ? p.sel_bits + b / 32 : nullptr;
    const uint32_t *abits = p.a_bits ? p.a_bits + b / 32 : nullptr;
    // the thread's mask words, loaded with the table below (up to 4 in registers)
    const bool cached = Mw <= 4;
    uint32_t sm[4] = {0u, 0u, 0u, 0u}, am[4] = {0u, 0u, 0u, 0u};
    if (cached && !p.pool)
#pragma unroll
        for (uint32_t i = 0; i < 4; i++)
            if (w0 + i < w1) {
                sm[i] = sel[w0 + i];
                am[i] = prune ? abits[w0 + i] : 0u;
            }
    auto pick = [&](const uint32_t (&v)[4], uint32_t w) {
        const uint32_t i = w - w0;
        return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
    };
    auto sel_w = [&](uint32_t w) { return cached ? pick(sm, w) : sel[w]; };
    auto a_w = [&](uint32_t w) { return cached ? pick(am, w) : abits[w]; };
    if (tid == 0) {
        L.m[EXQ] = 0u;
        L.m[XNQ] = 0u;
        L.m[THR] = 0xffffu;
        L.kth = kNoKey;   // no bound unless phase A yields its k-th key
        if (p.ids)
            p.counts[q] = min(p.k, c);
        else
            p.out_cnt[q] = min(p.k, c);
    }
    // the k smallest: keys, or (with p.ids) the caller's ids and distances
    // with the row padded by UINT64_MAX / +inf
    auto emit = [&](uint32_t pos, uint64_t v) {
        if (p.ids) {
            p.ids[(size_t)q * p.k + pos] = p.rank_id[(uint32_t)v];
            p.dists[(size_t)q * p.k + pos] = key_float_m((uint32_t)(v >> 32));
        } else {
            out[pos] = v;
        }
    };
    if (p.ids)
        for (uint32_t i = min(p.k, c) + tid; i < p.k; i += 256) {
            p.ids[(size_t)q * p.k + i] = ~0ull;
            p.dists[(size_t)q * p.k + i] = __builtin_inff();
        }
    for (uint32_t i = tid; i < p.dpad; i += 256) s_qv[i] = p.qv[(size_t)q * p.dpad + i];
    const float qn = p.qnorm ? p.qnorm[q] : 0.0f;
    // probe table (segment-relative first slot, first entry of each probe)
    ProbeTab tab;
    tab.np = p.np_q[q];
    tab.qp_row = p.qp_slot + (size_t)q * p.npr;
    tab.probe_row = p.probe_post + (size_t)q * p.npr;
    tab.post_off = p.post_off;
    tab.seg = b;
    tab.pbase = tab.ebase = nullptr;
    float exq = 0.0f, xnq = 0.0f;
    for (uint32_t j = tid; j < tab.np; j += 256) {
        const uint32_t post = tab.probe_row[j];
        if (tab.np <= kWaveTab) {
            L.pbase[j] = (uint32_t)(tab.qp_row[j] - b);
            L.ebase[j] = p.post_off[post];
        }
        if (prune) {
            exq = fmaxf(exq, p.post_ex[post]);
            xnq = fmaxf(xnq, p.post_xn[post]);
        }
    }
    if (tab.np <= kWaveTab) {
        tab.pbase = L.pbase;
        tab.ebase = L.ebase;
    }
    // list pools: members i0 .. i1 - 1 with key at most thr, as entries (any order)
    const uint32_t *pool = p.pool ? p.pool + (size_t)q * p.want : nullptr;
    auto list_members = [&](uint32_t i0, uint32_t i1, uint32_t thr, uint32_t *dst) -> uint32_t {
        if (tid == 0) L.m[13] = 0u;
        __syncthreads();   // (also publishes the probe table)
        for (uint32_t i = i0 + tid; i < i1; i += 256) {
            const uint32_t v = pool[i];
            if ((v >> 16) > thr) continue;
            const uint32_t slot = v & 0xffffu, j = tab.find(slot);
            dst[atomicAdd(&L.m[13], 1u)] = tab.first_entry(j) + (slot - tab.rel(j));
        }
        __syncthreads();
        const uint32_t cnt = L.m[13];
        __syncthreads();
        return cnt;
    };
    if (!prune) {   // rerank the whole pool (rebuild_members' barriers publish the table)
        const uint32_t cnt = pool ? list_members(0, c, 0xffffu, L.ent)
                                  : rebuild_members(L, w0, w1, [&](uint32_t w) { return sel_w(w); },
                                                    tab, L.ent, c);
        rerank_list(p, L.ent, cnt, s_qv, qn, L.ex);
        __syncthreads();
        smallest_k(L.ex, cnt, p.k, emit);
        if (p.stats && tid == 0) p.stats[q] = cnt;
        return;
    }
    for (int off = 32; off > 0; off >>= 1) {
        exq = fmaxf(exq, __shfl_xor(exq, off, 64));
        xnq = fmaxf(xnq, __shfl_xor(xnq, off, 64));
    }
    __syncthreads();   // L.m initialised
    if (lane == 0) {   // non-negative floats order as their bits
        atomicMax(&L.m[EXQ], __float_as_uint(exq));
        atomicMax(&L.m[XNQ], __float_as_uint(xnq));
    }
    // 1. phase A (the `ca` smallest keys, marked by pool selection): rerank,
    // then its k-th exact key (keys are distinct: the rank is in the low bits)
    const uint32_t na = pool ? list_members(0, nA, 0xffffu, L.ent)
                             : rebuild_members(L, w0, w1, [&](uint32_t w) { return a_w(w); }, tab,
                                               L.ent, nA);
    rerank_list(p, L.ent, na, s_qv, qn, L.ex);
    __syncthreads();
    if (na <= kPrCount) {   // ranks kept (in nzp, free until the survivor pass)
        for (uint32_t i = tid; i < na; i += 256) {
            const uint64_t v = L.ex[i];
            uint32_t cnt = 0;
            for (uint32_t j = 0; j < na; j++) cnt += L.ex[j] < v;
            if (cnt == p.k - 1) L.kth = v;
            L.nzp[i] = cnt;
        }
    } else {
        uint32_t mA = 2;
        while (mA < na) mA <<= 1;
        for (uint32_t i = na + tid; i < mA; i += 256) L.ex[i] = kNoKey;
        __syncthreads();
        bitonic_lds(L.ex, mA);
        if (tid == 0) L.kth = L.ex[p.k - 1];
    }
    __syncthreads();

    // 2. bound: a key above thr proves an exact distance above A's k-th
    if (tid == 0) {
        uint32_t thr = 0xffffu;
        const float kth = key_float_m((uint32_t)(L.kth >> 32));
        if (kth >= 0.0f && kth < __builtin_inff()) {
            const double u = 1.0 / 8388608.0;   // 2^-23
            double ksq = (double)kth;
            if (p.metric == M_L2) ksq = ksq * ksq * (1.0 + 1.0 / 1048576.0);
            // the f32 exact sum is at most (dim + 8) 2^-24 relative below the true value
            const double s = sqrt(ksq * (1.0 + (p.dim + 16) * u));
            const double eqd = (double)p.eq16[q] * (1.0 + 1e-5) + 1e-30;
            const double r = s + eqd + (double)__uint_as_float(L.m[EXQ]) * (1.0 + 1e-6);
            const double qn16 = p.qn16[q], xn16 = __uint_as_float(L.m[XNQ]);
            // f32 evaluation of |x16|^2 + |q16|^2 - 2 q16.x16 (exact products)
            const double e1 = (p.kdim + 16) * u * (qn16 + xn16 + 2.0 * sqrt(qn16 * xn16)) * 2.0;
            const double v = (r * r + e1) * (1.0 + 1e-9);
            if (v < 3.0e38) {
                float vf = (float)v;   // round up
                if ((double)vf < v) vf = __uint_as_float(__float_as_uint(vf) + 1u);
                thr = min(0x8000u | (__float_as_uint(vf) >> 16), 0xfffeu);
            }
        }
        L.m[THR] = thr;
    }
    __syncthreads();
    const uint32_t thr = L.m[THR];
    // 3. survivors: pool members outside A whose key is at most thr (none
    // when thr is below every such key)
    if (p.a_min && thr < p.a_min[q]) {   // A's k smallest, by the ranks found above
        for (uint32_t i = tid; i < na; i += 256) {
            const uint32_t rk = na <= kPrCount ? L.nzp[i] : i;   // (else A is sorted)
            if (rk < p.k) emit(rk, L.ex[i]);
        }
        if (p.stats && tid == 0) p.stats[q] = na;
        return;
    }
    const uint16_t *kp = p.keys + b;
    const uint32_t nb = pool ? list_members(nA, c, thr, L.ent + na) : rebuild_members(L, w0, w1, [&](uint32_t w) {
        const uint32_t rest = sel_w(w) & ~a_w(w);
        if (!rest) return 0u;
        uint4 r[4];
        load_word_raw(kp, w, r);
        uint32_t live, lt, eq;
        word_masks(r, w, n, thr, live, lt, eq);
        return rest & (lt | eq);
    }, tab, L.ent + na, c - na);
    rerank_list(p, L.ent + na, nb, s_qv, qn, L.ex + na);
    __syncthreads();
    smallest_k(L.ex, na + nb, p.k, emit);
    if (p.stats && tid == 0) p.stats[q] = na + nb;
}

// ---------------------------------------------------------------- f16 stage
// RaBitQ pools: every member's distance on its f16 row (four lanes per
// member, f32 accumulate) gives a proven lower bound of its exact f32
// distance:
//   L2     |q - x| >= |q - x16| - ex  (ex: the query's probed postings' max |x - f16(x)|)
//   dot    q.x <= q.x16 + |q| ex
//   cosine the same over |q| |x|
// widened for the f32 evaluation of both sums ((dim + 16) 2^-23 relative).
// Members are reranked exactly in bound order, a chunk at a time, until the
// next bound exceeds the k-th exact distance: the result is the exact rerank
// of the whole pool.
constexpr uint32_t kTop16 = 256;   // largest k of the f16 stage

template <uint32_t CAP>
struct Rerank16Lds {
    RerankLds<CAP> b;
    uint64_t top[2][kTop16];       // exact top k so far (sorted), merge target
    uint32_t list[64 * kPrH];      // entries of one exact chunk
    uint64_t chunk[64 * kPrH];
    uint32_t tc;
};

// LONG: rows of at least 256 dims (8+ units a lane), whose loads are unrolled
// eight deep (more registers, fewer waves: a loss for short rows)
template <uint32_t CAP, bool LONG>
__global__ __launch_bounds__(256) void k_rerank_f16(const RerankPrunedParams p) {
    __shared__ Rerank16Lds<CAP> S;
    RerankLds<CAP> &L = S.b;
    extern __shared__ __align__(16) float s_qv[];   // dpad floats
    const uint32_t x = blockIdx.x % kXcds, jx = blockIdx.x / kXcds;
    const uint32_t per = (p.nq + kXcds - 1) / kXcds;
    const uint32_t idx = x * per + jx;
    if (jx >= per || idx >= p.nq) return;   // whole workgroup
    const uint32_t q = idx;
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t c = min(p.pool_want[q], p.want);
    const uint64_t b = p.seg_off[q];
    const uint32_t n = p.seg_cnt[q];
    const uint32_t nwords = (n + 31) / 32, Mw = (nwords + 255) / 256;
    const uint32_t w0 = tid * Mw, w1 = min(w0 + Mw, nwords);
    const uint32_t *sel = p.sel_bits + b / 32;
    for (uint32_t i = tid; i < p.dpad; i += 256) s_qv[i] = p.qv[(size_t)q * p.dpad + i];
    const float qn = p.qnorm[q];
    ProbeTab tab;
    tab.np = p.np_q[q];
    tab.qp_row = p.qp_slot + (size_t)q * p.npr;
    tab.probe_row = p.probe_post + (size_t)q * p.npr;
    tab.post_off = p.post_off;
    tab.seg = b;
    tab.pbase = tab.ebase = nullptr;
    float exq = 0.0f;
    for (uint32_t j = tid; j < tab.np; j += 256) {
        const uint32_t post = tab.probe_row[j];
        if (tab.np <= kWaveTab) {
            L.pbase[j] = (uint32_t)(tab.qp_row[j] - b);
            L.ebase[j] = p.post_off[post];
        }
        exq = fmaxf(exq, p.post_ex[post]);
    }
    if (tab.np <= kWaveTab) {
        tab.pbase = L.pbase;
        tab.ebase = L.ebase;
    }
    for (int off = 32; off > 0; off >>= 1) exq = fmaxf(exq, __shfl_xor(exq, off, 64));
    if (tid == 0) {
        L.m[0] = 0u;
        S.tc = 0u;
    }
    __syncthreads();
    if (lane == 0) atomicMax(&L.m[0], __float_as_uint(exq));
    // 1. the pool's members (rebuild_members' barriers publish L.m and the table)
    const uint32_t cnt = rebuild_members<CAP, 8>(L, w0, w1, [&](uint32_t w) { return sel[w]; }, tab,
                                         L.ent, c);
    const float ex = __uint_as_float(L.m[0]);
    const float gam = (float)(p.dim + 16) * (1.0f / 8388608.0f);
    const float qa = sqrtf(qn);
    // 2. f16-row distance -> lower bound; key = (bound, member index)
    {
        const uint32_t r = tid & 3, quad = tid >> 2, units = p.dpad / 8;
        for (uint32_t i0 = 0; i0 < cnt; i0 += 64 * kPrH) {
#pragma unroll
            for (int h = 0; h < kPrH; h++) {
                const uint32_t i = i0 + 64 * h + quad;
                const uint32_t e = i < cnt ? L.ent[i] : L.ent[0];
                const uint4 *xr = reinterpret_cast<const uint4 *>(p.raw16 + (size_t)e * p.dpad);
                float acc = 0.0f;
                auto unit = [&](uint32_t u) {
                    const half8_t hv = as_half8(xr[u]);
                    const float4 qa4 = reinterpret_cast<const float4 *>(s_qv)[2 * u];
                    const float4 qb4 = reinterpret_cast<const float4 *>(s_qv)[2 * u + 1];
                    const float qq[8] = {qa4.x, qa4.y, qa4.z, qa4.w, qb4.x, qb4.y, qb4.z, qb4.w};
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const float xv = (float)hv[j];
                        if (p.metric == M_L2 || p.metric == M_L2SQ) {
                            const float t = qq[j] - xv;
                            acc = __builtin_fmaf(t, t, acc);
                        } else {
                            acc = __builtin_fmaf(qq[j], xv, acc);
                        }
                    }
                };
                if constexpr (LONG) {
#pragma unroll 8
                    for (uint32_t u = r; u < units; u += 4) unit(u);
                } else {
                    for (uint32_t u = r; u < units; u += 4) unit(u);
                }
                acc = quad_sum(acc);
                if (r == 0 && i < cnt) {
                    float lb;
                    if (p.metric == M_L2 || p.metric == M_L2SQ) {
                        const float sd = sqrtf(fmaxf(acc, 0.0f) / (1.0f + gam)) - ex;
                        lb = sd > 0.0f ? (p.metric == M_L2 ? sd * sqrtf(1.0f - gam) * (1.0f - 2.4e-7f)
                                                           : sd * sd * (1.0f - gam))
                                       : 0.0f;
                    } else {
                        const float xa = sqrtf(p.xn[e]);
                        const float slack = ex + gam * (2.0f * xa + ex);
                        if (p.metric == M_DOT) {
                            lb = (1.0f - acc) - qa * slack - 2.4e-7f * (1.0f + qa * (xa + ex));
                        } else if (xa > 0.0f && qa > 0.0f) {
                            lb = (1.0f - acc / (qa * xa)) - slack / xa - 4.0f * gam - 2.4e-7f;
                        } else {
                            lb = -__builtin_inff();
                        }
                    }
                    L.ex[i] = ((uint64_t)fkey_m(lb) << 32) | i;
                }
            }
        }
    }
    // 3. members in bound order
    uint32_t m2 = 2;
    while (m2 < cnt) m2 <<= 1;
    __syncthreads();
    for (uint32_t i = cnt + tid; i < m2; i += 256) L.ex[i] = kNoKey;
    __syncthreads();
    if (cnt > 1) bitonic_lds(L.ex, m2);
    __syncthreads();
    // 4. exact rerank in bound order until the next bound exceeds the k-th; the
    // first chunk holds k plus a margin (the bound is usually past the k-th
    // exact distance by then), later chunks a full pass of the workgroup
    uint32_t done = 0;
    const uint32_t chunk = 64 * kPrH;
    const uint32_t first = min(chunk, max(32u, (p.k + p.k / 2 + 31) / 32 * 32));
    while (done < cnt) {
        const uint32_t tc = S.tc;
        if (tc >= p.k) {
            const float nb = key_float_m((uint32_t)(L.ex[done] >> 32));
            const float kth = key_float_m((uint32_t)(S.top[0][p.k - 1] >> 32));
            if (nb > kth) break;   // (uniform: every thread reads the same words)
        }
        const uint32_t n1 = min(cnt - done, done ? chunk : first);
        for (uint32_t i = tid; i < n1; i += 256) S.list[i] = L.ent[(uint32_t)L.ex[done + i]];
        __syncthreads();
        rerank_list(p, S.list, n1, s_qv, qn, S.chunk);
        __syncthreads();
        const uint32_t tot = tc + n1;
        for (uint32_t i = tid; i < tot; i += 256) {   // keys are distinct (rank in the low bits)
            const uint64_t v = i < tc ? S.top[0][i] : S.chunk[i - tc];
            uint32_t rk = 0;
            for (uint32_t j = 0; j < tc; j++) rk += S.top[0][j] < v;
            for (uint32_t j = 0; j < n1; j++) rk += S.chunk[j] < v;
            if (rk < p.k) S.top[1][rk] = v;
        }
        __syncthreads();
        const uint32_t ntc = min(p.k, tot);
        for (uint32_t i = tid; i < ntc; i += 256) S.top[0][i] = S.top[1][i];
        if (tid == 0) S.tc = ntc;
        done += n1;
        __syncthreads();
    }
    const uint32_t tc = S.tc;
    for (uint32_t i = tid; i < p.k; i += 256) {   // the row, padded by UINT64_MAX / +inf
        const uint64_t v = i < tc ? S.top[0][i] : kNoKey;
        p.ids[(size_t)q * p.k + i] = i < tc ? p.rank_id[(uint32_t)v] : ~0ull;
        p.dists[(size_t)q * p.k + i] = i < tc ? key_float_m((uint32_t)(v >> 32)) : __builtin_inff();
    }
    if (tid == 0) {
        p.counts[q] = tc;
        if (p.stats) p.stats[q] = done;
    }
}

// ---------------------------------------------------------------- launchers
hipError_t launch_scan_mfma(const MixedParams &p, uint32_t grid, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    const bool f = p.qmask != nullptr;   // label filter
    const bool l2 = p.metric == M_L2 || p.metric == M_L2SQ;
    // K steps: 32 dims (f16: dim <= 32, 64, 96, 128, 256) or 64 dims (int8: dim <= 64 ... 512)
    switch (p.ks) {
#define SPF_SCAN_MFMA_L(KS_, F_, I_)                                                                \
        if (l2) hipLaunchKernelGGL((k_scan_mfma<KS_, F_, I_, true>), dim3(grid), dim3(256), 0, s, p); \
        else hipLaunchKernelGGL((k_scan_mfma<KS_, F_, I_, false>), dim3(grid), dim3(256), 0, s, p);
#define SPF_SCAN_MFMA(KS_)                                                                          \
        if (p.int8) {                                                                               \
            if (f) { SPF_SCAN_MFMA_L(KS_, true, true) } else { SPF_SCAN_MFMA_L(KS_, false, true) }  \
        } else {                                                                                    \
            if (f) { SPF_SCAN_MFMA_L(KS_, true, false) } else { SPF_SCAN_MFMA_L(KS_, false, false) } \
        }                                                                                           \
        break;
    case 1: SPF_SCAN_MFMA(1)
    case 2: SPF_SCAN_MFMA(2)
    case 3: SPF_SCAN_MFMA(3)
    case 4: SPF_SCAN_MFMA(4)
    case 5: case 6: case 7: case 8: SPF_SCAN_MFMA(8)
#undef SPF_SCAN_MFMA
#undef SPF_SCAN_MFMA_L
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The wave scan: NH = qg / 16 query halves per wave (qg 32 or 64), one
// 64-thread workgroup per wave.
hipError_t launch_scan_wave(const MixedParams &p, uint32_t grid, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    const bool f = p.qmask != nullptr;
    const bool l2 = p.metric == M_L2 || p.metric == M_L2SQ;
    const bool adm = p.adm_list != nullptr;
    if (p.qg != 32 && p.qg != 64) return hipErrorInvalidValue;
#define SPF_WAVE_A(KS_, NH_, F_, I_, L_)                                                            \
    if (adm) hipLaunchKernelGGL((k_scan_wave<KS_, NH_, F_, I_, L_, true>), dim3(grid), dim3(64), 0, s, p); \
    else hipLaunchKernelGGL((k_scan_wave<KS_, NH_, F_, I_, L_, false>), dim3(grid), dim3(64), 0, s, p);
#define SPF_WAVE_L(KS_, NH_, F_, I_)                                                                \
    if (l2) { SPF_WAVE_A(KS_, NH_, F_, I_, true) } else { SPF_WAVE_A(KS_, NH_, F_, I_, false) }
#define SPF_WAVE_F(KS_, NH_)                                                                        \
    if (p.int8) {                                                                                   \
        if (f) { SPF_WAVE_L(KS_, NH_, true, true) } else { SPF_WAVE_L(KS_, NH_, false, true) }     \
    } else {                                                                                        \
        if (f) { SPF_WAVE_L(KS_, NH_, true, false) } else { SPF_WAVE_L(KS_, NH_, false, false) }   \
    }
#define SPF_WAVE(KS_)                                                                               \
    if (p.qg == 32) { SPF_WAVE_F(KS_, 2) } else { SPF_WAVE_F(KS_, 4) }                              \
    break;
    switch (p.ks) {
    case 1: SPF_WAVE(1)
    case 2: SPF_WAVE(2)
    case 3: SPF_WAVE(3)
    case 4: SPF_WAVE(4)
    case 5: case 6: case 7: case 8: SPF_WAVE(8)
    default: return hipErrorInvalidValue;
    }
#undef SPF_WAVE
#undef SPF_WAVE_F
#undef SPF_WAVE_L
#undef SPF_WAVE_A
    return hipGetLastError();
}

hipError_t launch_pool_long(const PoolLongParams &p, uint64_t max_chunks, hipStream_t s) {
    if (p.nq == 0 || max_chunks == 0) return hipSuccess;
    if (p.stride % 32 || p.stride < p.want || p.stride > kSelChunk || p.ca > p.want ||
        p.long_from < kSelChunk)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pool_chunk_count, dim3((p.nq + 255) / 256), dim3(256), 0, s, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = launch_excl_scan_u32(p.chunk_cnt, p.chunk_off, p.nq, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool_chunk, dim3((uint32_t)max_chunks), dim3(256), 0, s, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pool_merge, dim3(p.nq), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_pool_select(const PoolParams &p, uint32_t nq, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    if (p.sel_bits)   // candidate pools: a workgroup per query, 24,576 keys in registers
        hipLaunchKernelGGL((k_pool_select<true, 4, 3>), dim3(nq), dim3(256), 0, s, p, nq);
    else if (p.seg_len <= 8192)   // probe pools: a wave per query, 4,096 keys in registers
        hipLaunchKernelGGL((k_pool_select<false, 1, 2>), dim3((nq + 3) / 4), dim3(256), 0, s, p, nq);
    else   // more centroids: a workgroup per query, 16,384 in registers (measured: 16,384
           // centroids 0.19 -> 0.175 ms a batch; at 8,192 the wave per query is faster)
        hipLaunchKernelGGL((k_pool_select<false, 4, 2>), dim3(nq), dim3(256), 0, s, p, nq);
    return hipGetLastError();
}

hipError_t launch_rerank_pool(const RerankPoolParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    const uint32_t grid = (p.nq + kXcds - 1) / kXcds * kXcds;
    hipLaunchKernelGGL(k_rerank_pool, dim3(grid), dim3(256), p.dpad * sizeof(float), s, p);
    return hipGetLastError();
}

hipError_t launch_rerank_pruned(const RerankPrunedParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    if (p.k > p.ca || p.ca > p.want || p.want > kMaxK) return hipErrorInvalidValue;
    const uint32_t grid = (p.nq + kXcds - 1) / kXcds * kXcds;
    if (p.want <= 512)   // smaller pools: half the LDS, twice the workgroups per CU
        hipLaunchKernelGGL(k_rerank_pruned<512>, dim3(grid), dim3(256), p.dpad * sizeof(float), s, p);
    else if (p.want <= 1024)
        hipLaunchKernelGGL(k_rerank_pruned<1024>, dim3(grid), dim3(256), p.dpad * sizeof(float), s, p);
    else
        hipLaunchKernelGGL(k_rerank_pruned<kMaxK>, dim3(grid), dim3(256), p.dpad * sizeof(float), s, p);
    return hipGetLastError();
}

hipError_t launch_rerank_f16(const RerankPrunedParams &p, hipStream_t s) {
    if (p.nq == 0) return hipSuccess;
    if (p.k > kTop16 || p.want > kMaxK) return hipErrorInvalidValue;
    const uint32_t grid = (p.nq + kXcds - 1) / kXcds * kXcds;
    const size_t lds = p.dpad * sizeof(float);
#define SPF_RR16(L_)                                                                        \
    if (p.want <= 512)   /* smaller pools: half the LDS, more workgroups per CU */            \
        hipLaunchKernelGGL((k_rerank_f16<512, L_>), dim3(grid), dim3(256), lds, s, p);        \
    else if (p.want <= 1024)                                                                  \
        hipLaunchKernelGGL((k_rerank_f16<1024, L_>), dim3(grid), dim3(256), lds, s, p);       \
    else                                                                                      \
        hipLaunchKernelGGL((k_rerank_f16<kMaxK, L_>), dim3(grid), dim3(256), lds, s, p);
    if (p.dpad >= 256) {
        SPF_RR16(true)
    } else {
        SPF_RR16(false)
    }
#undef SPF_RR16
    return hipGetLastError();
}

hipError_t launch_probe_mfma(const ProbeMfmaParams &p, hipStream_t s) {
    const uint32_t nch = (p.n_post + kEChunk - 1) / kEChunk;
    const uint32_t grid = ((p.nq + 15) / 16) * nch;
    if (grid == 0) return hipSuccess;
    switch (p.ks) {   // 32-dim K steps
    case 1: hipLaunchKernelGGL(k_probe_mfma<1>, dim3(grid), dim3(256), 0, s, p); break;
    case 2: hipLaunchKernelGGL(k_probe_mfma<2>, dim3(grid), dim3(256), 0, s, p); break;
    case 3: hipLaunchKernelGGL(k_probe_mfma<3>, dim3(grid), dim3(256), 0, s, p); break;
    case 4: hipLaunchKernelGGL(k_probe_mfma<4>, dim3(grid), dim3(256), 0, s, p); break;
    case 5: case 6: case 7: case 8:
        hipLaunchKernelGGL(k_probe_mfma<kMixKS>, dim3(grid), dim3(256), 0, s, p); break;
    default:   // long rows (dim <= 1024): K steps rounded up to a multiple of kMixKS
        if (p.ks <= 16) hipLaunchKernelGGL(k_probe_mfma<16>, dim3(grid), dim3(256), 0, s, p);
        else if (p.ks <= 24) hipLaunchKernelGGL(k_probe_mfma<24>, dim3(grid), dim3(256), 0, s, p);
        else if (p.ks <= 32) hipLaunchKernelGGL(k_probe_mfma<32>, dim3(grid), dim3(256), 0, s, p);
        else return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace spf
