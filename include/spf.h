/*
 * spf.h -- C ABI of the MI355X (gfx950) SPFresh batch-search library
 * (libspf_hip.so).  Plain pointers and sizes only.
 *
 * The library replaces the search side of the reference crate:
 *   SPFresh::search_batch     src/spfresh.rs:897-908   -> spf_search_batch
 *   SPFresh::search_with_dists src/spfresh.rs:886-888  -> spf_search_batch (nq = 1)
 *   SPFresh::search           src/spfresh.rs:890-895   -> spf_search_batch (ids only)
 *   SPFresh::delete           src/spfresh.rs:867-874   -> spf_delete (device tombstones)
 * and is fed the state those calls read under the RwLock read guard
 * (Inner, src/spfresh.rs:165-195): live postings and centroids, (id, code)
 * entries, the raw vector store, tombstones, quantizer state and
 * SPFreshConfig::{probe_ratio, rerank_size}.
 *
 * Ownership: the library owns all device memory behind the opaque handle; the
 * caller owns every host buffer.  No pointer escapes a call.  Searches on one
 * handle may run concurrently from several host threads (each call takes its
 * own workspace and stream); spf_delete and spf_index_destroy must not overlap
 * a search on the same handle (the reference takes the write lock there).
 */
#ifndef SPF_H
#define SPF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define SPF_OK 0
#define SPF_EINVAL (-1)        /* bad argument (message in spf_last_error) */
#define SPF_EDIM (-2)          /* query dim != index dim: the reference panics with
                                  "Query dim {} != index dim {}" (src/spfresh.rs:672-678) */
#define SPF_EUNSUPPORTED (-3)  /* no device path (PQ postings): caller keeps its CPU search */
#define SPF_EDEVICE (-4)       /* HIP runtime error */
#define SPF_ENOMEM (-5)        /* device allocation failed, or a request above the
                                  server's device-memory budget */
#define SPF_EAGAIN (-6)        /* spf_server_wait: request not complete yet */

/* metrics: the Distance impls of src/metric.rs:16-39 */
#define SPF_METRIC_L2 0        /* DistL2    sqrt(sum (a-b)^2) */
#define SPF_METRIC_L2SQ 1      /* DistL2Sq  sum (a-b)^2   (Python "l2") */
#define SPF_METRIC_COSINE 2    /* DistCosine (angular)     (Python "cosine") */
#define SPF_METRIC_DOT 3       /* DistDot   1 - a.b        (Python "dot") */

/* posting codecs: QuantizerState (src/quantized.rs:150-155); PQ is not served */
#define SPF_QUANT_NONE 0       /* f32 codes = raw vector */
#define SPF_QUANT_F16 1        /* src/sq.rs:36-108 */
#define SPF_QUANT_INT8 2       /* src/sq.rs:121-234, global scale/offset, u8 codes */
#define SPF_QUANT_RABITQ 3     /* src/rabitq.rs, residual codes per posting */
#define SPF_QUANT_PQ 4         /* -> SPF_EUNSUPPORTED */

/* search precision */
#define SPF_PREC_EXACT 0       /* reference semantics, canonical f32 order: ids match
                                  the restated reference (ties by id) */
#define SPF_PREC_MIXED 1       /* reduced-precision probe and scan on matrix cores,
                                  exact f32 rerank of a candidate pool */

typedef struct spf_index spf_index;

/* Centroid routing graph: the reference's IncrementalDiskANN over the posting
 * centroids ({path}.centroids: DANN base + delta layer + tombstones,
 * src/incremental.rs:1213-1233, src/lib.rs:90-98), which Inner::nearest
 * searches instead of scanning every centroid above 4,096 live postings
 * (src/spfresh.rs:20,374-378).  Node ids are unified: base nodes
 * 0..base_n-1, then delta nodes (UnifiedView, src/incremental.rs:357-487). */
typedef struct spf_graph_desc {
    uint32_t n_nodes;
    uint32_t n_seeds;               /* entry points: base medoid, then base_n + delta entry */
    uint32_t seeds[2];
    uint32_t centroid_beam;         /* SPFreshConfig::centroid_beam */
    uint32_t reserved;
    uint64_t n_tombstones;          /* size of the graph's tombstone set */
    const float *vectors;           /* [n_nodes * dim] node vectors */
    const uint32_t *adj_off;        /* [n_nodes + 1] */
    const uint32_t *adj;            /* [adj_off[n_nodes]] neighbor lists, stored order, no PAD */
    const uint8_t *live;            /* [n_nodes] 0 = tombstoned */
    const int64_t *node_post;       /* [n_nodes] posting (index into this descriptor's
                                       postings) of the node's cid, -1 = none */
} spf_graph_desc;

typedef struct spf_index_desc {
    uint32_t dim;
    uint32_t metric;                /* SPF_METRIC_* */
    uint32_t quantizer;             /* SPF_QUANT_* */
    uint32_t n_postings;            /* live postings (Inner::cid_block.len()) */
    const uint32_t *block_ids;      /* [n_postings] reference block index per posting;
                                       distinct; orders equal centroid distances */
    const float *centroids;         /* [n_postings * dim]  block_centroid of each */
    const uint64_t *posting_offsets;/* [n_postings + 1] CSR offsets into the entries */
    const uint64_t *entry_ids;      /* [n_entries] ids, each id at most once */
    const uint8_t *codes;           /* [n_entries * code_bytes] posting codes */
    uint32_t code_bytes;            /* code_size (src/spfresh.rs:99-107) */
    const float *raw;               /* [n_raw * dim] raw store by id (Inner::raw);
                                       required for rerank / RaBitQ, else may be NULL */
    uint64_t n_raw;                 /* next_id */
    const uint64_t *deleted_ids;    /* tombstones (may be NULL) */
    uint64_t n_deleted;
    float int8_scale, int8_offset;  /* Int8Quantizer {scale, offset} */
    uint32_t rabitq_padded;         /* RaBitQ::padded */
    const float *rabitq_signs;      /* [3 * rabitq_padded] RaBitQ::signs */
    float probe_ratio;              /* SPFreshConfig::probe_ratio (inf = off) */
    uint32_t rerank_size;           /* SPFreshConfig::rerank_size */
    int32_t device;                 /* HIP device ordinal */
    const uint8_t *labels;          /* [n_raw] label of each id for filtered search (NULL = none) */
    const spf_graph_desc *graph;    /* routing graph (NULL = exact probe at any posting count) */
} spf_index_desc;

#define SPF_FLAG_PROFILE 1u     /* time each stage with HIP events on the call's stream;
                                   the call then waits for its work (spf_last_profile) */
#define SPF_FLAG_FULL_RERANK 2u /* MIXED: rerank every pool member exactly instead of
                                   skipping members whose proven lower bound exceeds the
                                   k-th exact distance (same results, more work) */
#define SPF_FLAG_SYNC_PLAN 4u   /* size the candidate buffers from the exact counts, read
                                   back with one host synchronisation mid-batch (the
                                   default when the index-derived bound exceeds
                                   plan_bound_mib) */
#define SPF_FLAG_NO_GRAPH 8u    /* do not capture / replay the call's launch chain */
#define SPF_FLAG_RQ_POPCOUNT 16u /* MIXED over RaBitQ codes: the estimator's integer sum by
                                   popcounts over bit planes on the vector units instead of
                                   int8 matrix-core products (the same integers, same result) */
#define SPF_FLAG_ADMIT 32u      /* MIXED over f32 / f16 / int8 codes: scan each query's
                                   nearest probes first (until they hold rerank_size entries),
                                   bound its pool by their keys, then list only the keys
                                   within that bound, per query, instead of storing every
                                   candidate's key (the same pools, same result; pays off
                                   when few queries share a posting and selecting pools from
                                   long candidate lists dominates the scan) */

/* Planning.  A batch's candidate count is known only after its probe.  By
 * default the library sizes the candidate buffers from a bound derived from
 * the index (the n_probe longest postings per query) whenever those buffers
 * stay within plan_bound_mib: the whole batch is then enqueued without a host
 * round trip, and a repeated call (same arguments and device buffers) replays
 * its captured launch chain as one hipGraph.  Above the limit (Zipf-sized
 * postings, very large batches), and for profiled or routing-graph calls, the
 * exact counts are read back once mid-batch.  Results are the same either way. */
typedef struct spf_search_opts {
    uint32_t precision;             /* SPF_PREC_* (default EXACT) */
    uint32_t rerank_size;           /* pool for MIXED (0 = index rerank_size, min k) */
    uint32_t flags;                 /* SPF_FLAG_* */
    uint32_t plan_bound_mib;        /* limit for planning on the device (0 = 2048 MiB) */
    uint32_t reserved[4];
} spf_search_opts;

/* stages of one batch search, in order */
#define SPF_STAGE_PREP 0
#define SPF_STAGE_PROBE 1
#define SPF_STAGE_PROBE_SELECT 2
#define SPF_STAGE_PLAN 3
#define SPF_STAGE_SCAN 4
#define SPF_STAGE_SELECT 5
#define SPF_STAGE_RERANK 6
#define SPF_STAGE_FINALIZE 7
#define SPF_N_STAGES 8

typedef struct spf_profile {
    float stage_ms[SPF_N_STAGES];   /* HIP-event time of each stage */
    uint64_t candidates;            /* distances computed by the posting scan */
    uint64_t scan_code_bytes;       /* candidates * code bytes: the scan's algorithmic reads */
    uint64_t probe_pairs;           /* (query, centroid) distances of the probe */
    uint32_t scan_items;            /* posting-scan work items */
    uint32_t rerank_rows;           /* MIXED: raw rows gathered by the exact rerank */
    uint32_t pool_members;          /* MIXED: candidates in the rerank pools */
    uint32_t admitted;              /* MIXED with admission: keys the second phase listed */
    /* the posting scan's algorithmic device-memory bytes under its work items:
     * every item reads its entries' codes (plus per-entry norms, tombstone
     * bits, labels) and its queries' rows once, and writes one key per
     * candidate slot (with admission: one per first-phase slot and one
     * list entry per admitted key) */
    uint64_t scan_bytes;
    uint64_t scan_flops;            /* 2 * (padded dim) * candidates of the posting scan */
    uint64_t probe_flops;           /* 2 * (padded dim) * queries * centroids of a dense probe */
    /* the posting scan's minimal HBM bytes: each probed posting's entries read
     * once (a posting's work items all run on one XCD, whose L2 serves the
     * rereads of its other items), each query's row read once, and the same
     * key (or list) bytes written as scan_bytes */
    uint64_t scan_hbm_bytes;
} spf_profile;

typedef struct spf_index_info {
    uint32_t dim, metric, quantizer, n_postings;
    uint64_t n_entries, n_raw, n_deleted;
    uint64_t device_bytes;          /* HBM held by the index */
    uint32_t max_posting;           /* longest posting (SPFreshStats::max_posting) */
    uint32_t reserved;
    uint64_t graph_replays;         /* search calls served by replaying a captured chain */
} spf_index_info;

/* Build the device mirror of a search-ready index (all host pointers are
 * copied; none is kept). */
int spf_index_create(const spf_index_desc *desc, spf_index **out);
void spf_index_destroy(spf_index *ix);
int spf_index_get_info(const spf_index *ix, spf_index_info *out);

/* Incremental sync of the mirror after mutations of the reference index
 * (Inner::insert / split / reassign / gc / compact, src/spfresh.rs:416-669),
 * which all write postings through mark_dirty (src/spfresh.rs:283-289):
 * the blocks written since the last sync (new content, or freed), the raw
 * rows and labels of the ids appended by insert (src/spfresh.rs:425-433), the
 * ids tombstoned since (SPFresh::delete, :867-874) and, after compact
 * renumbers the centroid graph (:645-669), the new routing graph.  Postings
 * not named keep their device entries (gathered into the new layout on the
 * device); the result equals spf_index_create on the mutated state.  Must not
 * overlap a search on the same handle.  SPF_EUNSUPPORTED when the mirror's
 * entry keys cannot follow (ids beyond 32 bits): create the mirror anew. */
typedef struct spf_block_update {
    uint32_t block;                 /* reference block index */
    uint32_t live;                  /* 0: the block was freed (free_posting, :406-414) */
    const float *centroid;          /* [dim] block_centroid[block] (live blocks) */
    uint64_t n_entries;             /* entries in posting order (Inner::entry, :267-270) */
    const uint64_t *entry_ids;      /* [n_entries] */
    const uint8_t *codes;           /* [n_entries * code_bytes] */
} spf_block_update;

typedef struct spf_index_update_desc {
    const spf_block_update *blocks; /* blocks written since the last sync */
    uint32_t n_blocks;
    uint32_t reserved0;
    uint64_t n_raw;                 /* next_id after the mutations (never below the mirror's) */
    const float *raw_rows;          /* [(n_raw - previous n_raw) * dim] rows of the appended ids */
    const uint8_t *labels;          /* [n_raw - previous n_raw] their labels (labelled mirrors) */
    const uint64_t *deleted_ids;    /* ids tombstoned since the last sync (may be NULL) */
    uint64_t n_deleted;
    const spf_graph_desc *graph;    /* routing graph after the mutations, node_post holding each
                                       node's reference block index (-1 = none); NULL = the
                                       nodes keep their blocks */
    uint32_t reserved[4];
} spf_index_update_desc;

int spf_index_update(spf_index *ix, const spf_index_update_desc *u);

/* Tombstone ids (additive, ids >= n_raw ignored), like SPFresh::delete. */
int spf_delete(spf_index *ix, const uint64_t *ids, uint64_t n);

/* Batch search from host buffers.  queries: nq x dim f32 row-major.
 * Outputs are nq rows of k: ids (UINT64_MAX pad), dists (+inf pad), counts[nq].
 * n_probe 0 is clamped to 1 (src/spfresh.rs:679).  k must be >= 1. */
int spf_search_batch(spf_index *ix, const float *queries, uint32_t nq, uint32_t dim,
                     uint32_t k, uint32_t n_probe, const spf_search_opts *opts,
                     uint64_t *out_ids, float *out_dists, uint32_t *out_counts);

/* Same, with queries and outputs already in device memory, enqueued on
 * `stream` (a hipStream_t).  The call returns once the work is enqueued (after
 * one mid-batch host synchronisation only when the call plans with
 * SPF_FLAG_SYNC_PLAN semantics, see spf_search_opts); outputs are final when
 * `stream` completes.  stream NULL = the library's own stream, and the call
 * returns with the outputs final.  Calls on different streams may overlap;
 * each call's working buffers stay reserved until its work completes. */
int spf_search_batch_device(spf_index *ix, const float *d_queries, uint32_t nq, uint32_t dim,
                            uint32_t k, uint32_t n_probe, const spf_search_opts *opts,
                            uint64_t *d_out_ids, float *d_out_dists, uint32_t *d_out_counts,
                            void *stream);

/* Label-filtered search (Filter::LabelIn on one label field,
 * src/filtered.rs:120-135, applied where the scan skips tombstones,
 * src/spfresh.rs:707,745): query i only sees entries whose id's label l
 * (spf_index_desc.labels) has bit l set in masks[i]; labels >= 64 never match.
 * Otherwise identical to spf_search_batch.  SPF_EINVAL if the index has no
 * labels.  masks: host array [nq]. */
int spf_search_batch_filtered(spf_index *ix, const float *queries, uint32_t nq, uint32_t dim,
                              uint32_t k, uint32_t n_probe, const spf_search_opts *opts,
                              const uint64_t *masks, uint64_t *out_ids, float *out_dists,
                              uint32_t *out_counts);

/* Same with device-resident queries, masks and outputs (see
 * spf_search_batch_device); d_masks NULL = unfiltered. */
int spf_search_batch_filtered_device(spf_index *ix, const float *d_queries, uint32_t nq,
                                     uint32_t dim, uint32_t k, uint32_t n_probe,
                                     const spf_search_opts *opts, const uint64_t *d_masks,
                                     uint64_t *d_out_ids, float *d_out_dists,
                                     uint32_t *d_out_counts, void *stream);

/* No-allocation form of spf_search_batch_filtered_device (d_masks NULL =
 * unfiltered).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  It calls no hipMalloc,
 * hipFree or other allocating routine: every working buffer is carved from a
 * fixed 64 MiB device array of the library (one per GPU, held until the
 * process ends), and the batch is planned on the device.  A batch whose
 * working-memory bound exceeds 64 MiB is refused with SPF_ENOMEM (split it);
 * profiling (SPF_EINVAL) and routing-graph probes (SPF_EUNSUPPORTED) are not
 * available.  All calls on one GPU share that array and each call overwrites
 * it: calls whose work overlaps on one GPU (different streams, or a call
 * enqueued before the previous call's work on another stream completed) race
 * and give wrong results; calls on one stream are ordered and safe.  Returns
 * once the work is enqueued on `stream` (NULL = the null stream). */
int spf_search_batch_device_noalloc(spf_index *ix, const float *d_queries, uint32_t nq,
                                    uint32_t dim, uint32_t k, uint32_t n_probe,
                                    const spf_search_opts *opts, const uint64_t *d_masks,
                                    uint64_t *d_out_ids, float *d_out_dists,
                                    uint32_t *d_out_counts, void *stream);

/* Insert routing: for each of n vectors, the reference block index of its
 * nearest live posting and the metric distance to that posting's centroid,
 * routed the way Inner::nearest(v, 1) routes it (src/spfresh.rs:358-378): up
 * to 4,096 live postings, or with no routing graph, exact over all live
 * centroids (equal distances -> lower block index); above 4,096 live postings
 * with a routing graph, the graph's beam search (its first result).
 * Replaces the parallel part of Inner::insert, `nearest(v, 1)` per vector of a
 * chunk (src/spfresh.rs:437-444).  out_block gets UINT32_MAX and out_dist +inf
 * when the index has no live posting. */
int spf_nearest_postings(spf_index *ix, const float *vectors, uint32_t n, uint32_t dim,
                         uint32_t *out_block, float *out_dist);

/* ---- Search server: continuous batching under a device-memory budget.
 * No counterpart in the reference (its SPFresh serves one call at a time
 * under a read lock, src/spfresh.rs:898); the serving loop of a filtered
 * SPFresh deployment.  Requests (query rows with one k / n_probe / options /
 * filter) queue in arrival order; each of n_slots workers (one HIP stream and
 * one working-memory set each) takes the oldest request plus every later one
 * with the same settings, up to max_batch rows, as one device batch, so
 * batches of several slots run on the device at once.  A slot starts a batch
 * only if the working memory all slots hold or have reserved plus the batch's
 * upper bound before its candidate buffers stays within hbm_budget_bytes;
 * once the probe has fixed the candidate count the batch reserves exactly what
 * those buffers add, waiting for running batches to finish if needed (a batch
 * that would wait on batches that are themselves waiting goes back to the
 * queue; one too large for the budget alone is split).  A single query that
 * cannot fit fails with SPF_ENOMEM (at submit when its bound already shows it).  Every row's result is
 * that of spf_search_batch(_filtered) on the same query.  The index must
 * outlive the server and must not be modified (spf_delete) while requests
 * are in flight. */
typedef struct spf_server spf_server;

typedef struct {
    uint32_t n_slots;           /* concurrent device batches; 0 = 3 */
    uint32_t max_batch;         /* rows per device batch; 0 = 1024 */
    uint64_t hbm_budget_bytes;  /* working memory of all slots; 0 = 8 GiB */
    uint32_t reserved[4];
} spf_server_opts;

typedef struct {
    uint64_t submit_ns, start_ns, done_ns; /* steady clock: submit, first batch start, done */
    uint32_t batches;                      /* device batches that carried its rows */
    uint32_t reserved[3];
} spf_request_info;

typedef struct {
    uint64_t requests_done, rows_done, batches, rejected;
    uint64_t peak_reserved_bytes; /* max working memory held + reserved at once */
    uint64_t budget_overruns;     /* batches whose memory exceeded their reservation (0) */
    uint64_t admission_waits;     /* times the oldest request did not fit the budget */
    uint64_t trims;               /* idle slots that gave their memory back */
    uint32_t max_concurrent;      /* most batches on the device at once */
    uint32_t requeues;            /* batches handed back while others held the memory */
    uint32_t splits;              /* batches handed back to run as smaller batches */
    uint32_t reserved[1];
} spf_server_stats;

int spf_server_create(spf_index *ix, const spf_server_opts *opts, spf_server **out);
/* Enqueue nq query rows (masks: NULL = unfiltered, else [nq] as in
 * spf_search_batch_filtered).  All host buffers must stay valid until
 * spf_server_wait returns for *ticket.  SPF_FLAG_PROFILE is ignored. */
int spf_server_submit(spf_server *sv, const float *queries, uint32_t nq, uint32_t dim,
                      uint32_t k, uint32_t n_probe, const spf_search_opts *opts,
                      const uint64_t *masks, uint64_t *out_ids, float *out_dists,
                      uint32_t *out_counts, uint64_t *ticket);
/* Wait for a request: timeout_us < 0 = until done; SPF_EAGAIN on timeout
 * (ticket stays valid); otherwise the request's status, and the ticket is
 * released.  info may be NULL. */
int spf_server_wait(spf_server *sv, uint64_t ticket, int64_t timeout_us, spf_request_info *info);
int spf_server_get_stats(spf_server *sv, spf_server_stats *out);
/* Completes every submitted request, then frees the server. */
void spf_server_destroy(spf_server *sv);

/* Launch-shape autotuner.  No counterpart in the reference (its search has no
 * device launch shapes).  Times the search of the given sample queries under
 * each candidate shape -- VALU scan (EXACT): 4 / 8 / 16 queries x 256 / 512 /
 * 1,024 / 2,048 entries per work item; matrix-core scan (MIXED): 32 / 64 / 128
 * queries per work item -- in interleaved rounds (median of five each), and
 * keeps the fastest on the index for later searches; another shape replaces
 * the default one only when at least 2% faster.  With cache_path (may be NULL), a line keyed by the device, the
 * index shape and the search settings is looked up first (nothing is timed
 * then) and appended after measuring.  Results never depend on the shape.
 * Must not run concurrently with searches on the same index. */
typedef struct {
    uint32_t scan_g, scan_chunk, mix_g;  /* shapes in effect after the call */
    uint32_t from_cache;                 /* 1: read from cache_path */
    float best_ms, default_ms;           /* sample search at the chosen / default shapes */
    uint32_t tried;                      /* shapes timed */
    uint32_t reserved[1];
} spf_tune_result;

int spf_autotune(spf_index *ix, const float *queries, uint32_t nq, uint32_t dim, uint32_t k,
                 uint32_t n_probe, const spf_search_opts *opts, const char *cache_path,
                 spf_tune_result *out);

/* Profile of the last search on this thread that set SPF_FLAG_PROFILE. */
int spf_last_profile(spf_profile *out);

/* Device self-check of the matrix-core fragment layout the mixed scan relies
 * on (a 16x16x32 f16 and bf16 product against the host).  SPF_OK or
 * SPF_EDEVICE with the first mismatch in spf_last_error. */
int spf_selftest(int device);

/* Thread-local message for the last error on this thread. */
const char *spf_last_error(void);
/* Library build string (arch, flags). */
const char *spf_version(void);

#ifdef __cplusplus
}
#endif
#endif
