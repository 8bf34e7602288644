// Internal interface between the host library (spf_host.cpp) and the gfx950
// kernels (spf_kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spf {

enum CodeKind : uint32_t { CK_F32 = 0, CK_F16 = 1, CK_U8 = 2 };
enum Metric : uint32_t { M_L2 = 0, M_L2SQ = 1, M_COS = 2, M_DOT = 3 };

constexpr uint32_t kWG = 256;        // threads per workgroup (4 wave64)
constexpr uint32_t kSelCap = 4096;   // LDS keys held by the selection kernel
constexpr uint32_t kMaxK = 2048;     // largest k / rerank pool a query may ask for
constexpr uint32_t kEChunk = 1024;   // entries of one posting per work item
constexpr uint64_t kNoKey = ~0ull;   // "no candidate" (deleted / padding)
constexpr uint32_t kXcds = 8;        // MI355X: workgroup i runs on XCD i % 8

// Work-item walk for a grid whose size is a multiple of kXcds: XCD x takes the
// contiguous item range [x*chunk, (x+1)*chunk), so all items of one posting
// (adjacent in item order) share one XCD's L2 instead of being spread round-robin.
struct XcdWalk {
    uint32_t item, end, step;
    __device__ __forceinline__ XcdWalk(uint32_t n_items) {
        const uint32_t x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
        step = gridDim.x / kXcds;
        const uint32_t chunk = (n_items + kXcds - 1) / kXcds;
        const uint32_t start = x * chunk;
        end = min(n_items, start + chunk);
        item = start + j;
    }
    __device__ __forceinline__ bool more() const { return item < end; }
    __device__ __forceinline__ void next() { item += step; }
};

// grid size for an item walk: a multiple of kXcds, at most `cap`
inline uint32_t xcd_grid(uint64_t n_items, uint32_t cap) {
    uint64_t g = n_items < cap ? n_items : cap;
    g = (g + kXcds - 1) / kXcds * kXcds;
    return (uint32_t)(g ? g : 0);
}

// One scan launch: distances between staged queries and code rows held in
// the tiled layout (16-byte unit u of entry e at ((e/64)*units + u)*64 + e%64).
struct ScanParams {
    const uint4 *data;
    uint32_t units;
    uint32_t n_entries;
    const uint32_t *low;        // low 32 key bits per entry (rank of its id); null = entry index
    const uint32_t *del;        // tombstone bit per entry; null = none
    const float *xnorm;         // cosine: |x|^2 per entry (canonical order)
    const uint32_t *xsq;        // int8 L2: sum c^2 per entry
    const float *qv;            // [nq][dpad] query values as the metric consumes them
    const uint32_t *q8;         // [nq][dpad/4] int8-encoded query, 4 codes per word
    const float *qnorm;         // [nq] |q|^2
    const uint32_t *qsq;        // [nq] sum q8^2
    uint32_t dpad;              // padded dim (multiple of 16)
    float scale, offset;        // int8 codec
    uint64_t *out;              // candidate keys (fkey(dist) << 32 | low)
    // dense mode (centroid probe): all queries x entries [0, n_entries)
    uint32_t nq;
    uint32_t n_echunks;
    uint64_t out_stride;
    // listed mode (posting scan): work items built by the planner
    const uint32_t *item_off;   // [n_post + 1]
    const uint32_t *item_post;  // [n_items] posting of each work item
    uint32_t n_post;
    uint32_t n_items;           // grid bound on the items (the count itself when n_items_dev is null)
    const uint32_t *n_items_dev;   // device count of the items (null = n_items)
    const uint32_t *post_off;   // [n_post + 1] entry offsets
    const uint32_t *pair_off;   // [n_post + 1] (query, slot) pairs per posting
    const uint32_t *pair_q;
    const uint64_t *pair_slot;  // candidate-buffer position of the posting's first entry
    uint32_t echunk;            // entries per work item (listed mode)
    // label filter (listed mode): entry label, per-query mask; null = unfiltered
    const uint8_t *label_e;
    const uint64_t *qmask;
};

// entry of label l passes mask m (labels >= 64 never do)
__device__ __forceinline__ bool label_ok(uint64_t m, uint32_t l) {
    return l < 64u && ((m >> l) & 1ull);
}

struct PrepParams {
    const float *q;             // [nq][dim] caller queries
    uint32_t nq, dim, dpad;
    uint32_t round_f16;         // F16 L2 scan: query rounded to f16 (src/sq.rs:51-55)
    uint32_t want_q8;           // Int8 L2 scan: query encoded (src/sq.rs:208-220)
    float scale, offset;
    float *qv_probe;            // [nq][dpad] f32 query (zero padded)
    float *qv_scan;             // [nq][dpad] (may alias qv_probe when not rounding)
    uint32_t *q8;               // [nq][dpad/4]
    float *qnorm;               // [nq]
    uint32_t *qsq;              // [nq]
    uint16_t *q16;              // [nq][dpad32] f16 query for the matrix-core scan (null = skip)
    uint32_t dpad32;            // dim rounded up to 32
    float *qn16;                // [nq] |f16(q)|^2
    float *eq16;                // [nq] |q - f16(q)| (null = skip)
    const float *mu;            // [dim] centring vector of the bf16 probe (null = skip)
    uint16_t *qbf;              // [nq][dpad32] bf16(q - mu)
    float *qnbf;                // [nq] |bf16(q - mu)|^2
    uint32_t *zero;             // zeroed alongside (the planner's counters; null = none)
    uint32_t zero_n;
    // int8 matrix-core scan: the query as symmetric int8 (scale max|q| / 127)
    // and the affine terms of q.decode(c) = a * sum q8 (c - 128) + b
    int8_t *q8m;                // [nq][dpad64] (null = skip)
    uint32_t dpad64;            // dim rounded up to 64
    float *q8a, *q8b;           // [nq]
};

// Routing-graph probe (one wave per query): beam search over the centroid
// graph with the reference's heap order; writes probe keys like the exact probe.
struct GraphParams {
    const float *vec;           // [n_nodes][dpad] node vectors, zero padded
    const uint32_t *adj_off;    // [n_nodes + 1]
    const uint32_t *adj;        // neighbor lists in stored order
    const uint8_t *live;        // [n_nodes] 0 = tombstoned
    const int32_t *node_post;   // [n_nodes] device posting index, -1 = none
    uint32_t n_seeds, seed0, seed1;
    uint32_t metric, dpad;
    uint32_t k;                 // results asked (the unclamped probe count)
    uint32_t wb;                // working beam
    uint32_t wcap, fcap, rcap;  // LDS entries of w, frontier, results (even)
    uint32_t bitmap_words;      // visited bitmap words (multiple of 4)
    uint32_t filtered;          // tombstones present: expanded beam, live results
    uint32_t max_it;
    float early;                // early-termination factor (inf = none)
    const float *qv;            // [nq][dpad] queries
    const float *qnorm;         // [nq] |q|^2
    uint32_t nq, npr;           // out row stride
    uint64_t *out;              // [nq][npr] fkey(dist) << 32 | posting
    uint32_t *out_cnt;          // [nq]
    uint32_t *overflow;         // queries whose frontier outgrew fcap
    // per-query escalation: a launch walks only the queries whose `only` flag
    // is set (null = every query) and sets `aborted` for those whose frontier
    // outgrew fcap (cleared for the others), which the next launch, with a
    // larger frontier, walks again
    const uint32_t *only;
    uint32_t *aborted;
    // global mode (state too large for LDS): bitmap and heaps of query
    // q0 + blockIdx.x at scratch + blockIdx.x * scratch_words
    uint32_t *scratch;          // null = LDS mode
    uint64_t scratch_words;
    uint32_t q0;
};

// bytes of the visited bitmap and the three heaps of one query
inline size_t graph_state_bytes(const GraphParams &p) {
    return (size_t)p.bitmap_words * 4 + (size_t)(p.wcap + p.fcap + (p.filtered ? p.rcap : 0)) * 8;
}
inline size_t graph_lds_bytes(const GraphParams &p) {
    return (p.scratch ? 0 : graph_state_bytes(p)) + (size_t)p.dpad * 4 + 64 * 8;
}

// Matrix-core centroid probe in bf16 with f32 accumulate, 16-bit keys of
// L2: |c-mu|^2 + |q-mu|^2 - 2(q-mu).(c-mu); dot: 1 - q.c; cosine: 1 - q.c / (|q||c|).
struct ProbeMfmaParams {
    const uint4 *qbf;           // [nq][q_units] bf16 rows
    uint32_t q_units, ks, nq;
    const float *qnbf;          // L2: |bf16(q - mu)|^2
    const float *qnorm;         // cosine: exact |q|^2
    const uint4 *cbf;           // tiled bf16 centroids (centred for L2)
    uint32_t c_units, n_post;
    const float *cnbf;          // [n_post] L2: |bf16(c - mu)|^2; cosine: 1 / |c|
    uint32_t metric;
    uint16_t *out;              // [nq][out_stride]
    uint32_t out_stride;
};

// Mixed-precision scan / rerank over the same posting-major work items.
struct MixedParams {
    const uint32_t *item_off;   // [n_post + 1]
    const uint32_t *item_post;  // [n_items]
    uint32_t n_post, n_items;   // n_items: bound, or the count when n_items_dev is null
    const uint32_t *n_items_dev;
    const uint32_t *post_off;   // [n_post + 1]
    const uint32_t *pair_off;   // [n_post + 1]
    const uint32_t *pair_q;
    const uint64_t *pair_slot;
    const uint4 *q16;           // [nq][q16_units] f16 rows
    uint32_t q16_units, ks;     // 16-byte units per query row; 32-dim K steps
    const float *qn16;
    const uint4 *codes16;       // tiled f16 codes
    uint32_t units;             // 16-byte units per f16 code
    const float *xn16;          // |x16|^2 per entry
    const uint32_t *del;
    uint32_t metric;            // key: L2 |q16-x16|^2, DOT 1-q16.x16, COSINE 1-cos(q16,x16)
    uint32_t echunk;            // entries per work item (64 x tiles per wave)
    uint32_t qg;                // queries per work item (a multiple of 32, at most kMixQ)
    // probe slot ranges start at multiples of 8 (plan align 8): keys leave
    // through an LDS tile as aligned 16-byte rows, the posting's last item
    // filling its range's tail with "no key"
    const uint8_t *label_e;     // label filter: entry label, per-query mask (null = off)
    const uint64_t *qmask;
    uint16_t *out16;            // 16-bit candidate keys
    // int8 postings (int8 MFMA, int32 sums): codes16 / units / xn16 / q16 /
    // q16_units / ks then hold the u8 codes, |decode(c)|^2, the int8 query
    // rows and their 64-dim K steps; per query the affine terms and |q|^2
    uint32_t int8;
    const float *q8a, *q8b, *qnorm;
    // two-phase plan: item_post holds virtual postings (v >= n_post is posting
    // v - n_post); the launch walks items [*item_lo_dev, *n_items_dev)
    const uint32_t *item_lo_dev;   // null = 0
    // admission (second phase, wave scan): a candidate whose key is at most
    // its query's bound tau[q] is appended to the query's list, which starts
    // at adm_list + seg_off[q] (room for every slot of the segment), as
    // key << 16 | slot within the segment, its length counted in adm_cnt[q];
    // nothing else is stored
    const uint32_t *tau;
    const uint64_t *seg_off;
    uint32_t *adm_list;         // [slots] (null = dense keys into out16)
    uint32_t *adm_cnt;          // [nq]
};

// Admission bound of each query: the `want`-th smallest live key among its
// first-phase slots [seg_off[q], seg_off[q] + s1[q]), 0xfffe when they hold
// fewer.
struct AdmitParams {
    const uint16_t *keys;
    const uint64_t *seg_off;
    const uint32_t *s1;
    uint32_t nq, want;
    uint32_t *tau;
    uint32_t *cnt;              // [nq] second-phase list lengths, zeroed here
};

// Pool of each query from the first phase's dense keys (at most tau) and the
// second phase's list: the `want` smallest (key, slot) values as
// key << 16 | slot, phase A (at least min(k, pool) of the members with the
// smallest keys, at most ca when the keys allow) first.
struct PoolListParams {
    const uint16_t *keys;       // first-phase dense keys
    const uint64_t *seg_off;
    const uint32_t *s1;
    const uint32_t *tau;
    const uint32_t *adm_list;   // the second phase's lists (from seg_off[q])
    const uint32_t *adm_cnt;    // [nq] their lengths
    uint32_t nq, want, ca, k;
    uint32_t *pool;             // [nq][want]
    uint32_t *pool_want;        // [nq]
    uint32_t *pool_na;          // [nq] phase-A members
    uint32_t *a_min;            // [nq] smallest key outside phase A (0xffff: none)
};

// Matrix-core scan work items: (posting, `mfma_chunk(ks)` entries, up to
// kMixQ queries).  ks = 32-dim K steps.
constexpr uint32_t kMixQ = 128;
inline uint32_t mfma_tiles(uint32_t ks) { return ks <= 4 ? 4u : 2u; }   // 16-entry tiles per wave
inline uint32_t mfma_chunk(uint32_t ks) { return 64u * mfma_tiles(ks); }
// Entries per work item of the wave scan (k_scan_wave: one wave, 32 or 64 queries).
constexpr uint32_t kWaveChunk = 256;

// Exact rerank of each query's pool (selection masks over the candidate
// keys from pool selection).  With prune, only the `ca` smallest keys and the members whose
// key does not prove an exact distance above the k-th of those are reranked
// (L2 metrics over f16 keys); the result equals reranking the whole pool.
struct RerankPrunedParams {
    const uint16_t *keys;       // candidate keys (segments at seg_off)
    const uint64_t *seg_off;    // [nq] first slot (multiple of 32)
    const uint32_t *seg_cnt;    // [nq]
    const uint32_t *sel_bits;   // pool membership, one mask per 32-slot word
    const uint32_t *a_bits;     // members of phase A (the `ca` smallest keys)
    const uint32_t *a_min;      // [nq] lower bound of the other members' keys (null = none)
    const uint32_t *pool_want;  // [nq] pool size
    const uint64_t *qp_slot;    // [nq][npr] absolute first slot of each probed posting
    const uint32_t *post_off;
    uint32_t nq, want, k, ca, metric, dpad, dim;
    const uint32_t *np_q;       // [nq]
    const uint32_t *probe_post; // [nq][npr]
    uint32_t npr;
    const float *qv, *qnorm;    // [nq][dpad], [nq]
    const float *raw_e;         // [n_entries][dpad]
    const uint32_t *rank;       // entry -> rank
    uint32_t prune;
    uint32_t kdim;              // terms in each f16 dot product (dim rounded to 32)
    const float *qn16, *eq16;   // per query |q16|^2, |q - q16|
    const float *post_ex;       // per posting max |x - x16|
    const float *post_xn;       // per posting max |x16|^2
    uint64_t *out;              // [nq][k] (dist key << 32 | rank)
    uint32_t *out_cnt;
    // final outputs instead (when ids is set): rank -> id, key -> distance
    const uint64_t *rank_id;
    uint64_t *ids;
    float *dists;
    uint32_t *counts;
    uint32_t *stats;            // optional [nq]: raw rows gathered per query
    // f16 stage (RaBitQ pools): every member scored on its f16 row, ordered
    // by a proven lower bound of its exact distance, reranked exactly in that
    // order until the bound passes the k-th exact distance
    const uint16_t *raw16;      // [n_entries][dpad] f16 rows (null = no f16 stage)
    const float *xn;            // [n_entries] |x|^2
    // pool as lists (admission): [nq][want] key << 16 | slot, phase A's
    // pool_na[q] members first; sel_bits / a_bits / keys unused
    const uint32_t *pool;
    const uint32_t *pool_na;
};

// RaBitQ: per-pair query prep, exact search, mixed estimator scan.
struct RqParams {
    const float *qv;            // [nq][dpad] f32 queries
    uint32_t dpad, dim, padded, words;   // words = padded / 64
    const float *cent;          // [n_post][dpad] centroid rows
    const float *signs;         // [3][padded]
    const uint32_t *probe_post; // [nq][npr]
    const uint32_t *np_q;       // [nq]
    uint32_t npr, nq;
    uint64_t *planes;           // [nq*npr][4*words]
    float *consts;              // [nq*npr][8]
    const uint32_t *post_off;   // [n_post + 1]
    const uint4 *codes;         // tiled RaBitQ codes
    uint32_t units;
    const uint32_t *del;
    const uint32_t *rank;
    const float *raw_e;         // [n_entries][dpad]
    uint32_t k, metric;
    uint64_t *out;              // [nq][k]
    uint32_t *out_cnt;
    // mixed scan over posting-major items
    uint32_t n_items;           // bound, or the count when n_items_dev is null
    const uint32_t *n_items_dev;
    const uint32_t *item_post, *item_off, *pair_off, *pair_src;
    const uint64_t *pair_slot;
    uint16_t *out16;
    const uint8_t *label_e;     // label filter: entry label, per-query mask (null = off)
    const uint64_t *qmask;
    // mixed precision: keys of the metric's estimate from the L2 estimate
    // (cosine on the unit query: 1 - (1 + |x|^2 - est) / (2 |x|); dot:
    // 1 - (|q|^2 + |x|^2 - est) / 2), the prep's norm as a wave sum
    uint32_t mixed;
    const float *xn;            // [n_entries] |x|^2 of the raw rows
    const float *qnorm;         // [nq] |q|^2
    // matrix-core estimator: the prep also writes each pair's 4-bit levels
    // as int8 rows ([nq*npr][padded]); the scan takes sum(bit * level) from
    // int8 MFMA (items of up to kRqMfmaQ pairs over 64 * rq_mfma_tiles entries)
    uint8_t *levels;            // null = bit planes only (popcount scan)
};

constexpr uint32_t kRqMfmaQ = 64;   // pairs per matrix-core estimator item
inline uint32_t rq_mfma_tiles(uint32_t words) { return words <= 16 ? 4u : 2u; }

// Per entry of a RaBitQ index: the f16 copy of its raw row, |x|^2, and the
// posting's max |x - f16(x)| (the f16 rerank stage's bound), rounded up.
struct RqRowsParams {
    const float *raw_e;         // [n][dpad]
    uint32_t n, dpad, dim, n_post;
    const uint32_t *post_off;   // [n_post + 1]
    uint16_t *raw16;            // [n][dpad] f16 bits
    float *xn;                  // [n]
    uint32_t *post_ex;          // [n_post] f32 bits, atomicMax (zeroed by the caller)
};

struct RqDecodeParams {
    const uint4 *codes;
    uint32_t units, padded, dpad, dim, out_units;
    const float *signs;
    const float *cent;          // [n_post][dpad]
    const uint32_t *entry_post; // posting of each entry
    uint4 *out;                 // tiled f32 codes (out_units per entry)
};

// Pool selection over 16-bit keys: the `want` smallest per query, written as
// entry indices (slot -> entry through the query's probe table).
struct PoolParams {
    const uint16_t *keys;
    const uint64_t *seg_off;    // [nq + 1] 32-aligned segments (null: q * seg_stride)
    const uint32_t *seg_cnt;    // [nq] (null: seg_len)
    uint64_t seg_stride;
    uint32_t seg_len;
    uint32_t map;               // 1: write entries via the probe table; 0: write slots
    uint32_t want, npr;
    const uint32_t *np_q;       // [nq]
    const uint64_t *qp_slot;    // [nq][npr] absolute slot of each probed posting
    const uint32_t *probe_post; // [nq][npr]
    const uint32_t *post_off;   // [n_post + 1]
    uint32_t *pool_entry;       // [nq][want]
    uint16_t *pool_key;         // [nq][want] keys of the members (null = not written)
    uint32_t *pool_want;        // [nq]
    uint32_t *sel_bits;         // non-null: write per-word selection masks instead
    uint32_t *a_bits;           // with sel_bits: masks of the `ca` smallest pool members
    uint32_t ca;
    uint32_t *a_min;            // with a_bits, optional [nq]: a lower bound of the keys of
                                // pool members outside phase A (0xffff: there are none)
    uint32_t long_from;         // with sel_bits: segments above this many slots are left to
                                // the chunked select (launch_pool_long); 0 = none
};

// Candidate pools of segments above kSelChunk slots (launch_pool_long): each
// chunk of kSelChunk slots contributes its `stride` smallest (key, slot)
// values in slot order to a list (keys padded with "no key" to the stride);
// the pool and phase A are selected from a query's lists, concatenated in
// chunk order (= slot order), and their bits set in the segment's masks.
struct PoolLongParams {
    const uint16_t *keys;
    const uint64_t *seg_off;    // [nq] 32-aligned segments
    const uint32_t *seg_cnt;    // [nq]
    uint32_t nq, stride;        // stride: want rounded up to 32
    uint32_t want, ca;
    uint32_t long_from;         // segments above this many slots (>= kSelChunk) are chunked
    uint32_t *chunk_cnt;        // [nq] chunks of each query (0: selected whole)
    uint32_t *chunk_off;        // [nq + 1]
    uint16_t *lkey;             // [chunks][stride]
    uint32_t *lslot;            // [chunks][stride] slot within the segment
    uint32_t *lcnt;             // [chunks]
    uint32_t *lsel, *la;        // [chunks * stride / 32] masks over the lists
    uint32_t *sel_bits, *a_bits, *pool_want, *a_min;
};

// Exact f32 rerank of each query's pool and the final top-k, one workgroup per
// query.
struct RerankPoolParams {
    const uint32_t *pool_entry; // [nq][want]
    const uint32_t *pool_want;  // [nq]
    uint32_t nq, want, k, metric, dpad;
    const float *qv;            // [nq][dpad]
    const float *qnorm;         // [nq]
    const float *raw_e;         // [n_entries][dpad] rows in entry order
    const uint32_t *rank;       // per entry (null: the entry index itself)
    uint64_t *out;              // [nq][k]
    uint32_t *out_cnt;          // [nq]
};

struct PlanParams {
    uint32_t nq, n_post, npr;   // npr = probes requested per query (after clamp)
    const uint64_t *probe_sel;  // [nq][npr] sorted probe keys
    const uint32_t *probe_cnt;  // [nq]
    float probe_ratio;
    const uint32_t *post_len;   // [n_post]
    uint32_t *np_q;             // [nq] probes kept after the probe_ratio cut
    uint32_t *probe_post;       // [nq][npr]
    uint64_t *cand_len;         // [nq + 1] segment length, rounded up to 32 slots
    uint32_t *cand_cnt;         // [nq] candidates actually written in the segment
    const uint64_t *cand_off;   // [nq + 1]
    uint32_t *post_cnt;         // [n_post + 1] pairs per posting
    uint32_t *qp_rel;           // [nq * npr] slot of each pair within its query's segment
    uint32_t *prank;            // [nq * npr] rank of each pair within its posting
    const uint32_t *pair_off;   // [n_post + 1]
    uint64_t *qp_slot;          // [nq * npr] slot of each (query, probe) pair
    uint32_t *pair_q;           // [nq * npr] posting-major pairs
    uint64_t *pair_slot;        // [nq * npr]
    uint32_t *pair_src;         // [nq * npr] (query * npr + probe) of each posting-major pair
    uint32_t *item_cnt;         // [n_post + 1]
    const uint32_t *item_off;   // [n_post + 1]
    uint32_t *item_post;        // [n_items]
    uint32_t n_items;           // item table size, or a bound on it when n_items_dev is set
    const uint32_t *n_items_dev;   // device count (item_off[n_post]); null = n_items
    uint32_t G;                 // queries per work item
    uint32_t echunk;            // entries per work item
    uint32_t align;             // each probe's candidate slots start at a multiple of this
    uint32_t *cand_real;        // [nq] candidates without alignment slots (null = skip)
    // pair ranks without global atomics: queries in blocks of qb, per-block
    // posting counts in LDS, then exclusive over blocks (0 = global atomics)
    uint32_t qb;
    uint32_t *blk_cnt;          // [ceil(nq / qb)][n_post]
    // two phases (split = real posting count, n_post = 2 split): a query's
    // probes whose preceding probes hold fewer than `want` entries form the
    // first phase, the rest are ranked under virtual posting post + split
    uint32_t split, want;
    uint32_t *m1;               // [nq] first-phase probes (probes j < m1)
    uint32_t *s1;               // [nq] their slots (a prefix of the segment)
};

constexpr uint32_t kRankQB = 256;              // queries per ranking block
constexpr uint32_t kRankMaxPost = 16384;       // LDS counters per ranking block

struct RerankParams {
    const uint64_t *sel;        // [nq][want] sorted pool keys
    const uint32_t *sel_cnt;    // [nq]
    uint32_t want, k, nq;
    const float *raw;           // [n][dpad] raw rows in entry order
    const uint32_t *rank_entry; // entry of each rank
    const float *qv;            // [nq][dpad] f32 queries
    const float *qnorm;
    uint32_t dim, dpad, metric;
    uint64_t *out;              // [nq][k]
    uint32_t *out_cnt;
};

// kernels launched from host code (spf_kernels.hip)
hipError_t launch_prep(const PrepParams &p, hipStream_t s);
hipError_t launch_scan(uint32_t code_kind, uint32_t metric, bool dense, uint32_t G,
                       const ScanParams &p, uint32_t grid, hipStream_t s);
// top K of each segment; segments longer than `chunk` keys in chunks (the top
// K of each chunk into tmp, then of the chunk results); max_chunks bounds the
// chunk count (0 = no segment that long).  Segments by seg_off, or seg_stride
// apart when seg_off is null.
hipError_t launch_select_long(const uint64_t *keys, const uint64_t *seg_off,
                              const uint32_t *seg_cnt, uint64_t seg_stride, uint32_t nseg,
                              uint32_t chunk, uint32_t K, uint64_t *out, uint32_t *out_cnt,
                              uint32_t *chunk_cnt, uint32_t *chunk_off, uint64_t *tmp,
                              uint64_t max_chunks, hipStream_t s);
constexpr uint32_t kSelChunk = 24576;      // candidate lists: keys per chunk (in registers)
constexpr uint32_t kProbeChunk = 24576;    // dense probe rows: keys per chunk (in registers)
// The matrix-core scan's slot ranges start at multiples of this many 16-bit
// keys (16 bytes: its store width) and are reserved up to the next one.
// (Whole 128-byte lines, 64 keys, measured no faster and add 13% slots.)
constexpr uint32_t kScanAlign = 8;
hipError_t launch_select(const uint64_t *keys, const uint64_t *seg_off, const uint32_t *seg_cnt,
                         uint64_t seg_stride, uint32_t nseg, uint32_t K, uint64_t *out,
                         uint32_t *out_cnt, hipStream_t s);
hipError_t launch_plan_q(const PlanParams &p, hipStream_t s);
hipError_t launch_scatter_pairs(const PlanParams &p, hipStream_t s);
hipError_t launch_items(const PlanParams &p, hipStream_t s);
hipError_t launch_item_table(const PlanParams &p, hipStream_t s);
hipError_t launch_excl_scan_u64(const uint64_t *in, uint64_t *out, uint32_t n, hipStream_t s);
hipError_t launch_pair_rank(const PlanParams &p, hipStream_t s);
hipError_t launch_excl_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t s);
hipError_t launch_rerank(const RerankParams &p, hipStream_t s);
hipError_t launch_graph_probe(const GraphParams &p, uint32_t n, hipStream_t s);
hipError_t launch_graph_walk(const GraphParams &p, uint32_t n, hipStream_t s);
hipError_t launch_nearest_out(const uint64_t *probe_sel, const uint32_t *probe_cnt, uint32_t n,
                              const uint32_t *block_sorted, uint32_t *out_block, float *out_dist,
                              hipStream_t s);
hipError_t launch_finalize(const uint64_t *keys, uint32_t stride, const uint32_t *cnt, uint32_t nq,
                           uint32_t k,
                           const uint64_t *rank_id, uint64_t *out_ids, float *out_dists,
                           uint32_t *out_counts, hipStream_t s);
hipError_t launch_entry_norms(uint32_t code_kind, const uint4 *data, uint32_t units,
                              uint32_t n, uint32_t dim, float scale, float offset, float *xnorm,
                              uint32_t *xsq, hipStream_t s);

hipError_t launch_selftest_mfma(float *d_out_f16, float *d_out_bf16, hipStream_t s);
hipError_t launch_scan_mfma(const MixedParams &p, uint32_t grid, hipStream_t s);
hipError_t launch_scan_wave(const MixedParams &p, uint32_t grid, hipStream_t s);
hipError_t launch_pool_select(const PoolParams &p, uint32_t nq, hipStream_t s);
hipError_t launch_pool_long(const PoolLongParams &p, uint64_t max_chunks, hipStream_t s);
hipError_t launch_admit(const AdmitParams &p, hipStream_t s);
hipError_t launch_pool_list(const PoolListParams &p, hipStream_t s);
hipError_t launch_rerank_pool(const RerankPoolParams &p, hipStream_t s);
hipError_t launch_rerank_pruned(const RerankPrunedParams &p, hipStream_t s);
hipError_t launch_rerank_f16(const RerankPrunedParams &p, hipStream_t s);
hipError_t launch_probe_mfma(const ProbeMfmaParams &p, hipStream_t s);
hipError_t launch_rq_prep(const RqParams &p, uint32_t n_pairs, hipStream_t s);
hipError_t launch_rq_search_exact(const RqParams &p, hipStream_t s);
hipError_t launch_rq_decode(const RqDecodeParams &p, uint32_t n, hipStream_t s);
hipError_t launch_rq_scan(const RqParams &p, uint32_t grid, hipStream_t s);
hipError_t launch_rq_rows(const RqRowsParams &p, hipStream_t s);
hipError_t launch_rq_scan_mfma(const RqParams &p, uint32_t grid, hipStream_t s);

// creation / incremental update (spf_update.hip); src maps: bit 31 set = the
// upload, else the old array, low bits = source entry
hipError_t launch_gather_tiled(const uint4 *old_t, const uint4 *up_t, const uint32_t *src,
                               uint32_t n, uint32_t units, uint4 *dst, hipStream_t s);
hipError_t launch_gather_rows(const float *old_r, const float *up_r, const uint32_t *src,
                              uint32_t n, uint32_t dpad, uint32_t up_stride, uint32_t dim,
                              float *dst, hipStream_t s);
hipError_t launch_tiled_f16(const uint4 *f32t, uint32_t units32, uint32_t n, uint4 *f16t,
                            uint32_t units16, hipStream_t s);
hipError_t launch_post_bounds(const float *raw_e, const uint4 *codes16, uint32_t units16,
                              uint32_t n, uint32_t dpad, uint32_t dim, const uint32_t *post_off,
                              uint32_t n_post, uint32_t *post_ex, uint32_t *post_xn,
                              hipStream_t s);
hipError_t launch_entry_post(const uint32_t *post_off, uint32_t n_post, uint32_t n, uint32_t *out,
                             hipStream_t s);
hipError_t launch_id_entry(const uint32_t *entry_id, uint32_t n, uint32_t *id_entry, uint64_t span,
                           hipStream_t s);
hipError_t launch_iota_u64(uint64_t *p, uint64_t n, hipStream_t s);

// Working memory of the no-allocation search: a file-scope device array of
// this size (spf_noalloc.hip), one per device.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
// Device address of the scratch array on `device` (which must be current),
// looked up once per device.
hipError_t noalloc_scratch(int device, void **ptr);

}  // namespace spf
