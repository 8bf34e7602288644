/*
 * CPU oracle for the SPFresh batch-search path.  TEST INFRASTRUCTURE ONLY:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it as the
 * checker; the product library never links or calls it.
 *
 * It restates, in plain C compiled with -ffp-contract=off, the reference's
 * search algorithm:
 *   Inner::search            src/spfresh.rs:671-727
 *   Inner::nearest (exact)   src/spfresh.rs:358-372
 *   Inner::nearest (graph)   src/spfresh.rs:374-378, IncrementalDiskANN::
 *                            search_with_dists src/incremental.rs:934-1003,
 *                            UnifiedView src/incremental.rs:357-480,
 *                            tombstone_config src/incremental.rs:489-499,
 *                            beam_search src/lib.rs:198-303
 *   Inner::code_dist         src/spfresh.rs:344-354
 *   Inner::search_rabitq     src/spfresh.rs:731-770
 *   quantized_distance_from_codes src/quantized.rs:13-54
 *   F16 / Int8 codecs        src/sq.rs:51-59, 128-172, 208-226
 *   RaBitQ                   src/rabitq.rs:31-256
 *   Distance impls           src/metric.rs:16-39
 *
 * Canonical choices where the reference leaves bits to a third-party SIMD
 * crate or to container internals (numkong 7.8.2 accumulation order, BinaryHeap
 * tie order, HashMap probe order) are fixed here and shared with the device
 * kernels:
 *   - every f32 reduction (sum of squares, dot product) runs as four fused
 *     multiply-add chains from +0, chain r taking dimensions i with
 *     (i / 4) % 4 == r in increasing order, combined as (c0 + c1) + (c2 + c3);
 *   - equal distances are ordered by id (posting candidates) or by reference
 *     block index (centroid probe).
 */
#ifndef SPF_ORACLE_H
#define SPF_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SPF_O_L2 = 0, SPF_O_L2SQ = 1, SPF_O_COSINE = 2, SPF_O_DOT = 3 };
enum { SPF_O_NONE = 0, SPF_O_F16 = 1, SPF_O_INT8 = 2, SPF_O_RABITQ = 3 };

/* Centroid routing graph (the reference's IncrementalDiskANN over posting
 * centroids: DANN base + delta layer + tombstones), in unified node ids: base
 * nodes 0..base_n-1, then delta nodes (UnifiedView, src/incremental.rs:357-487). */
typedef struct spf_o_graph {
    uint32_t n_nodes;
    uint32_t n_seeds;           /* entry points: base medoid, then base_n + delta entry */
    uint32_t seeds[2];
    uint32_t centroid_beam;     /* SPFreshConfig::centroid_beam */
    uint64_t n_tombstones;      /* size of the tombstone set (tombstone_config) */
    const float *vectors;       /* [n_nodes * dim] */
    const uint32_t *adj_off;    /* [n_nodes + 1] */
    const uint32_t *adj;        /* neighbor lists in stored order, PAD removed */
    const uint8_t *live;        /* [n_nodes] 0 = tombstoned */
    const int64_t *node_post;   /* [n_nodes] posting index of the node's cid, -1 = none */
} spf_o_graph;

typedef struct spf_o_index {
    uint32_t dim;
    uint32_t metric;
    uint32_t quant;
    uint32_t n_post;            /* live postings */
    uint32_t code_bytes;        /* bytes per stored code */
    uint32_t rerank_size;       /* SPFreshConfig::rerank_size */
    float probe_ratio;          /* SPFreshConfig::probe_ratio */
    float int8_scale, int8_offset;
    uint32_t rq_padded;
    const uint32_t *block_ids;  /* [n_post] reference block index of each posting */
    const float *centroids;     /* [n_post * dim] */
    const uint64_t *post_off;   /* [n_post + 1] entry offsets */
    const uint64_t *ids;        /* [n_entries] */
    const uint8_t *codes;       /* [n_entries * code_bytes] */
    const float *raw;           /* [n_raw * dim] raw vectors by id (may be NULL) */
    uint64_t n_raw;
    const uint8_t *deleted;     /* bitmap over ids < n_raw (may be NULL) */
    const float *rq_signs;      /* [3 * rq_padded] */
    const uint8_t *labels;      /* [n_raw] label of each id for filtered search (may be NULL) */
    const spf_o_graph *graph;   /* routing graph used above 4,096 live postings (may be NULL) */
} spf_o_index;

/* ---- scalar building blocks (exported so tests can pin each one) ---- */
uint16_t spf_o_f32_to_f16(float x);
float spf_o_f16_to_f32(uint16_t h);
float spf_o_l2sq(const float *a, const float *b, uint32_t d);
float spf_o_dot(const float *a, const float *b, uint32_t d);
float spf_o_eval(uint32_t metric, const float *a, const float *b, uint32_t d);
void spf_o_int8_encode(const float *v, uint32_t d, float scale, float offset, uint8_t *out);
float spf_o_int8_l2sq(const uint8_t *a, const uint8_t *b, uint32_t d, float scale);
void spf_o_fht(float *v, uint32_t n);
void spf_o_rabitq_encode(const float *v, const float *centroid, uint32_t d, uint32_t padded,
                         const float *signs, uint8_t *code);
void spf_o_rabitq_decode(const uint8_t *code, const float *centroid, uint32_t d, uint32_t padded,
                         const float *signs, float *out);
/* query prep: planes [4 * padded/64] u64, consts [5] = nq2, s, k1, k2, k3 */
void spf_o_rabitq_query(const float *q, const float *centroid, uint32_t d, uint32_t padded,
                        const float *signs, uint64_t *planes, float *consts);
void spf_o_rabitq_dist(const uint64_t *planes, const float *consts, const uint8_t *code,
                       uint32_t padded, float *est, float *err);
void spf_o_rotate(float *v, uint32_t padded, const float *signs);
void spf_o_unrotate(float *v, uint32_t padded, const float *signs);

/* ---- search ---- */
/* Probed postings for one query: writes up to n (posting index, dist) pairs
 * closest first after the probe_ratio cut; returns the count. */
uint32_t spf_o_probe(const spf_o_index *ix, const float *q, uint32_t n_probe,
                     uint32_t *out_post, float *out_dist);
/* IncrementalDiskANN::search_with_dists over the routing graph: up to k
 * (unified node, distance) pairs in the reference's order; returns the count. */
uint32_t spf_o_graph_search(const spf_o_graph *g, uint32_t metric, uint32_t dim, const float *q,
                            uint32_t k, uint32_t beam, uint32_t *out_node, float *out_dist);
/* One query (Inner::search).  Writes <= k results, returns the count. */
uint32_t spf_o_search(const spf_o_index *ix, const float *q, uint32_t k, uint32_t n_probe,
                      uint64_t *out_ids, float *out_dists);
/* nq queries on n_threads host threads (one task per query, like rayon
 * par_iter).  Rows are k wide; counts[nq]. */
void spf_o_search_batch(const spf_o_index *ix, const float *q, uint32_t nq, uint32_t k,
                        uint32_t n_probe, uint32_t n_threads, uint64_t *out_ids,
                        float *out_dists, uint32_t *out_counts);

/* Label-filtered search: Filter::LabelIn on one label field
 * (src/filtered.rs:120-135) applied where the scan skips tombstones
 * (src/spfresh.rs:707,745): an entry takes part only if bit labels[id] of the
 * query's mask is set (labels >= 64 never match).  Same outputs as above. */
uint32_t spf_o_search_filtered(const spf_o_index *ix, const float *q, uint64_t mask, uint32_t k,
                               uint32_t n_probe, uint64_t *out_ids, float *out_dists);
void spf_o_search_batch_filtered(const spf_o_index *ix, const float *q, const uint64_t *masks,
                                 uint32_t nq, uint32_t k, uint32_t n_probe, uint32_t n_threads,
                                 uint64_t *out_ids, float *out_dists, uint32_t *out_counts);

/* Planted-mistake switch for the tests that show the oracle checks have teeth.
 * 0 = faithful.  1 = drop the last dimension from f32 L2 sums.  2 = ignore
 * tombstones.  3 = skip the probe_ratio cut.  4 = round f16 by truncation.
 * 5 = ignore the label filter.  6 = graph search stops on best > worst instead
 * of best >= worst.  7 = graph search ignores the delta layer's entry point. */
void spf_o_set_fault(int fault);

#ifdef __cplusplus
}
#endif
#endif
