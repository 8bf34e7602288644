/*
 * CPU oracle for the SPFresh batch-search path (test infrastructure; see
 * spf_oracle.h).  Compiled with -ffp-contract=off so that every operation
 * rounds exactly where the Rust source rounds (Rust never contracts a*b+c).
 */
#include "spf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

static int g_fault = 0;
void spf_o_set_fault(int fault) { g_fault = fault; }

/* ------------------------------------------------------------------------ */
/* total order on f32 bit patterns: identical to Rust's f32::total_cmp, which
 * the reference uses for every final sort (src/spfresh.rs:367,370,724,768). */
static inline uint32_t fkey(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* ---- f16 <-> f32, IEEE round-to-nearest-even.  The reference calls
 * numkong::cast for this (src/sq.rs:51-55, 96-99). ---- */
uint16_t spf_o_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t mant = x & 0x7fffffu;
    if (exp == 0xffu) {
        if (mant) return (uint16_t)(sign | 0x7e00u | (mant >> 13));
        return (uint16_t)(sign | 0x7c00u);
    }
    int e = (int)exp - 127 + 15;
    if (e >= 0x1f) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        uint32_t m = mant | 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t h = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (g_fault == 4) return (uint16_t)(sign | h);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (g_fault != 4 && (rem > 0x1000u || (rem == 0x1000u && (h & 1u)))) h++;
    return (uint16_t)(sign | h);
}

float spf_o_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t mant = (uint32_t)h & 0x3ffu;
    uint32_t x;
    float f;
    if (exp == 0) {
        f = (float)mant * 5.9604644775390625e-8f; /* mant * 2^-24, exact */
        return sign ? -f : f;
    }
    if (exp == 0x1fu) x = sign | 0x7f800000u | (mant << 13);
    else x = sign | ((exp + 112u) << 23) | (mant << 13);
    memcpy(&f, &x, 4);
    return f;
}

/* ---- Distances (src/metric.rs:16-39).  numkong's reduction order is not
 * public.  The canonical order: four FMA chains from +0, chain r taking the
 * dimensions i with (i / 4) % 4 == r in increasing i, combined as
 * (c0 + c1) + (c2 + c3). ---- */
static inline float combine4(const float c[4]) { return (c[0] + c[1]) + (c[2] + c[3]); }

float spf_o_l2sq(const float *a, const float *b, uint32_t d) {
    float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t n = (g_fault == 1 && d > 0) ? d - 1 : d;
    for (uint32_t i = 0; i < n; i++) {
        float t = a[i] - b[i];
        c[(i >> 2) & 3] = fmaf(t, t, c[(i >> 2) & 3]);
    }
    return combine4(c);
}

float spf_o_dot(const float *a, const float *b, uint32_t d) {
    float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < d; i++) c[(i >> 2) & 3] = fmaf(a[i], b[i], c[(i >> 2) & 3]);
    return combine4(c);
}

/* Angular distance: numkong's special cases for zero norms are restated as
 * 0 when both norms are zero and 1 when the dot product is zero. */
static float cosine_from(float ab, float a2, float b2) {
    if (a2 == 0.0f && b2 == 0.0f) return 0.0f;
    if (ab == 0.0f) return 1.0f;
    float r = 1.0f - ab / sqrtf(a2 * b2);
    return r > 0.0f ? r : 0.0f;
}

float spf_o_eval(uint32_t metric, const float *a, const float *b, uint32_t d) {
    switch (metric) {
    case SPF_O_L2: return sqrtf(spf_o_l2sq(a, b, d));
    case SPF_O_L2SQ: return spf_o_l2sq(a, b, d);
    case SPF_O_DOT: return 1.0f - spf_o_dot(a, b, d);
    case SPF_O_COSINE:
        return cosine_from(spf_o_dot(a, b, d), spf_o_dot(a, a, d), spf_o_dot(b, b, d));
    }
    return NAN;
}

/* ---- Int8 global-affine codec (src/sq.rs:128-172, 208-226). ---- */
static inline uint8_t sat_round_u8(float s) {
    if (!(s > 0.0f)) return 0;          /* negatives and NaN saturate to 0 */
    if (s >= 255.0f) return 255;
    return (uint8_t)rintf(s);           /* nearest, ties to even */
}

void spf_o_int8_encode(const float *v, uint32_t d, float scale, float offset, uint8_t *out) {
    float alpha = 1.0f / scale;
    float beta = -offset / scale;
    for (uint32_t i = 0; i < d; i++) {
        float s = v[i] * alpha + beta;
        out[i] = sat_round_u8(s);
    }
}

float spf_o_int8_l2sq(const uint8_t *a, const uint8_t *b, uint32_t d, float scale) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < d; i++) {
        int t = (int)a[i] - (int)b[i];
        acc += (uint64_t)(t * t);
    }
    return (float)acc * scale * scale;
}

static void int8_decode(const uint8_t *c, uint32_t d, float scale, float offset, float *out) {
    for (uint32_t i = 0; i < d; i++) out[i] = (float)c[i] * scale + offset;
}

/* ---- RaBitQ (src/rabitq.rs).  Pure scalar f32 arithmetic, restated in the
 * same operation order. ---- */
void spf_o_fht(float *v, uint32_t n) {
    for (uint32_t h = 1; h < n; h *= 2)
        for (uint32_t i = 0; i < n; i += 2 * h)
            for (uint32_t j = i; j < i + h; j++) {
                float x = v[j], y = v[j + h];
                v[j] = x + y;
                v[j + h] = x - y;
            }
}

void spf_o_rotate(float *v, uint32_t padded, const float *signs) {
    for (uint32_t r = 0; r < 3; r++) {
        for (uint32_t i = 0; i < padded; i++) v[i] *= signs[r * padded + i];
        spf_o_fht(v, padded);
    }
}

void spf_o_unrotate(float *v, uint32_t padded, const float *signs) {
    for (int r = 2; r >= 0; r--) {
        spf_o_fht(v, padded);
        for (uint32_t i = 0; i < padded; i++) v[i] *= signs[(uint32_t)r * padded + i];
    }
}

/* prep_with, src/rabitq.rs:136-149 */
static float rq_prep(const float *v, const float *c, uint32_t d, uint32_t padded,
                     const float *signs, float *r) {
    for (uint32_t i = 0; i < padded; i++) r[i] = i < d ? v[i] - c[i] : 0.0f;
    float acc = 0.0f;
    for (uint32_t i = 0; i < padded; i++) {
        float sq = r[i] * r[i];
        acc = acc + sq;
    }
    float norm = sqrtf(acc);
    if (norm > 0.0f)
        for (uint32_t i = 0; i < padded; i++) r[i] /= norm;
    spf_o_rotate(r, padded, signs);
    return norm;
}

/* encode_with + pack, src/rabitq.rs:218-256 */
void spf_o_rabitq_encode(const float *v, const float *centroid, uint32_t d, uint32_t padded,
                         const float *signs, uint8_t *code) {
    float *r = (float *)malloc(sizeof(float) * padded);
    float norm = rq_prep(v, centroid, d, padded, signs, r);
    float ip = 0.0f;
    for (uint32_t w = 0; w < padded / 64; w++) {
        uint64_t bits = 0;
        for (uint32_t i = 0; i < 64; i++) {
            float x = r[w * 64 + i];
            bits |= (uint64_t)(x > 0.0f) << i;
            ip += fabsf(x);
        }
        memcpy(code + 8 * w, &bits, 8);
    }
    ip /= sqrtf((float)padded);
    float fac = ip > 0.0f ? norm / ip : 0.0f;
    float n2 = norm * norm;
    memcpy(code + padded / 8, &n2, 4);
    memcpy(code + padded / 8 + 4, &fac, 4);
    free(r);
}

/* decode_with, src/rabitq.rs:224-238 */
void spf_o_rabitq_decode(const uint8_t *code, const float *centroid, uint32_t d, uint32_t padded,
                         const float *signs, float *out) {
    float *r = (float *)malloc(sizeof(float) * padded);
    float sd = sqrtf((float)padded);
    for (uint32_t w = 0; w < padded / 64; w++) {
        uint64_t bits;
        memcpy(&bits, code + 8 * w, 8);
        for (uint32_t i = 0; i < 64; i++)
            r[w * 64 + i] = ((bits >> i) & 1u) ? 1.0f / sd : -1.0f / sd;
    }
    spf_o_unrotate(r, padded, signs);
    float n2;
    memcpy(&n2, code + padded / 8, 4);
    float norm = sqrtf(n2);
    for (uint32_t i = 0; i < d; i++) {
        float t = norm * r[i];
        out[i] = centroid[i] + t;
    }
    free(r);
}

/* query_with, src/rabitq.rs:157-187.  consts = {nq2, s, k1, k2, k3}. */
void spf_o_rabitq_query(const float *q, const float *centroid, uint32_t d, uint32_t padded,
                        const float *signs, uint64_t *planes, float *consts) {
    float *r = (float *)malloc(sizeof(float) * padded);
    float nq = rq_prep(q, centroid, d, padded, signs, r);
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (uint32_t i = 0; i < padded; i++) {
        lo = fminf(lo, r[i]);
        hi = fmaxf(hi, r[i]);
    }
    const float levels = 15.0f;
    float delta = (hi - lo) / levels;
    float inv = delta > 0.0f ? 1.0f / delta : 0.0f;
    uint32_t words = padded / 64;
    memset(planes, 0, sizeof(uint64_t) * 4 * words);
    uint32_t sum_u = 0;
    for (uint32_t i = 0; i < padded; i++) {
        float t = roundf((r[i] - lo) * inv);
        uint32_t u = (t > 0.0f) ? (t >= 4294967295.0f ? 0xffffffffu : (uint32_t)t) : 0u;
        if (u > 15u) u = 15u;
        sum_u += u;
        for (uint32_t j = 0; j < 4; j++)
            if ((u >> j) & 1u) planes[j * words + i / 64] |= 1ull << (i % 64);
    }
    float sd = sqrtf((float)padded);
    float a = delta * (float)sum_u;
    float b = (float)padded * lo;
    consts[0] = nq * nq;
    consts[1] = 2.0f * nq;
    consts[2] = 2.0f * delta / sd;
    consts[3] = 2.0f * lo / sd;
    consts[4] = -(a + b) / sd;
    free(r);
}

/* distance_with_bound, src/rabitq.rs:191-215 */
void spf_o_rabitq_dist(const uint64_t *planes, const float *consts, const uint8_t *code,
                       uint32_t padded, float *est, float *err) {
    uint32_t words = padded / 64;
    uint32_t s = 0, c = 0;
    for (uint32_t i = 0; i < words; i++) {
        uint64_t b;
        memcpy(&b, code + 8 * i, 8);
        c += (uint32_t)__builtin_popcountll(b);
        s += (uint32_t)__builtin_popcountll(b & planes[i]) +
             ((uint32_t)__builtin_popcountll(b & planes[words + i]) << 1) +
             ((uint32_t)__builtin_popcountll(b & planes[2 * words + i]) << 2) +
             ((uint32_t)__builtin_popcountll(b & planes[3 * words + i]) << 3);
    }
    float n2, fac;
    memcpy(&n2, code + padded / 8, 4);
    memcpy(&fac, code + padded / 8 + 4, 4);
    float nq2 = consts[0], sq = consts[1], k1 = consts[2], k2 = consts[3], k3 = consts[4];
    float inner = k1 * (float)s + k2 * (float)c;
    inner = inner + k3;
    float sf = sq * fac;
    float e = (n2 + nq2) - sf * inner;
    *est = fmaxf(e, 0.0f);
    float ip = fac > 0.0f ? sqrtf(n2) / fac : 0.0f;
    if (ip > 0.0f) {
        float one = fmaxf(1.0f - ip * ip, 0.0f);
        float rad = sqrtf(one / ((float)padded - 1.0f));
        *err = sf * 1.9f * rad;
    } else {
        *err = INFINITY;
    }
}

/* ------------------------------------------------------------------------ */
typedef struct { uint32_t key; uint32_t aux; uint64_t id; float dist; } cand_t;

static int cand_cmp(const void *a, const void *b) {
    const cand_t *x = (const cand_t *)a, *y = (const cand_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return 0;
}

static inline int is_deleted(const spf_o_index *ix, uint64_t id) {
    if (g_fault == 2 || !ix->deleted || id >= ix->n_raw) return 0;
    return (ix->deleted[id >> 3] >> (id & 7)) & 1;
}

/* entry skipped by the scan: tombstoned, or (filtered) label not in the mask */
static inline int skipped(const spf_o_index *ix, uint64_t id, int filtered, uint64_t mask) {
    if (is_deleted(ix, id)) return 1;
    if (!filtered || !ix->labels || g_fault == 5) return 0;
    uint8_t l = id < ix->n_raw ? ix->labels[id] : 255;
    return l >= 64 || !((mask >> l) & 1u);
}

/* ---- routing graph (Inner::nearest above BRUTE_FORCE_CENTROIDS = 4096 live
 * postings, src/spfresh.rs:20,358-378) ---- */

/* Candidate (src/lib.rs:100-122) compares by dist alone; `a <= b` is
 * PartialOrd::le, false when either dist is NaN. */
typedef struct { float d; uint32_t id; } gcand_t;

/* std::collections::BinaryHeap restated: push = append + sift_up; pop = take
 * the last element, put it at the root, sift_down_to_bottom (always towards the
 * child that is not <= its sibling, i.e. the right one on ties), then sift_up.
 * `rev` = the heap of Reverse<Candidate> (a min-heap): a <= b becomes b <= a. */
typedef struct { gcand_t *v; uint32_t n; int rev; } gheap_t;

static inline int gle(const gheap_t *h, gcand_t a, gcand_t b) {
    return h->rev ? (b.d <= a.d) : (a.d <= b.d);
}

static uint32_t gheap_sift_up(gheap_t *h, uint32_t start, uint32_t pos) {
    gcand_t e = h->v[pos];
    while (pos > start) {
        uint32_t parent = (pos - 1) / 2;
        if (gle(h, e, h->v[parent])) break;
        h->v[pos] = h->v[parent];
        pos = parent;
    }
    h->v[pos] = e;
    return pos;
}

static void gheap_push(gheap_t *h, gcand_t c) {
    h->v[h->n] = c;
    h->n++;
    gheap_sift_up(h, 0, h->n - 1);
}

static gcand_t gheap_pop(gheap_t *h) {
    gcand_t item = h->v[h->n - 1];
    h->n--;
    if (h->n == 0) return item;
    gcand_t top = h->v[0];
    h->v[0] = item;
    /* sift_down_to_bottom(0) */
    uint32_t end = h->n, pos = 0, child = 1;
    gcand_t e = h->v[0];
    while (end >= 2 && child <= end - 2) {
        child += gle(h, h->v[child], h->v[child + 1]) ? 1u : 0u;
        h->v[pos] = h->v[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (child == end - 1) {
        h->v[pos] = h->v[child];
        pos = child;
    }
    h->v[pos] = e;
    gheap_sift_up(h, 0, pos);
    return top;
}

/* f32::total_cmp as an unsigned key */
static inline uint32_t total_key(float f) { return fkey(f); }

typedef struct { uint32_t key; uint32_t pos; gcand_t c; } gsort_t;
static int gsort_cmp(const void *a, const void *b) {   /* stable: ties by position */
    const gsort_t *x = (const gsort_t *)a, *y = (const gsort_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

static float graph_dist(const spf_o_graph *g, uint32_t metric, uint32_t dim, const float *q,
                        uint32_t node) {
    return spf_o_eval(metric, q, g->vectors + (size_t)node * dim, dim);
}

/* beam_search (src/lib.rs:198-303) with UnifiedView's seeds, neighbors and
 * liveness and tombstone_config (src/incremental.rs:383-393,441-487,489-499). */
uint32_t spf_o_graph_search(const spf_o_graph *g, uint32_t metric, uint32_t dim, const float *q,
                            uint32_t k, uint32_t beam, uint32_t *out_node, float *out_dist) {
    if (!g || g->n_seeds == 0 || g->n_nodes == 0) return 0;
    const int filtered = g->n_tombstones > 0;
    uint64_t wb = beam;
    double max_it = INFINITY, early = INFINITY;
    if (filtered) {
        uint64_t t = g->n_tombstones < beam ? g->n_tombstones : beam;
        uint64_t e1 = (uint64_t)beam * 2, e2 = (uint64_t)k + t;
        wb = e1 > e2 ? e1 : e2;
        max_it = (double)(wb * 2);
        early = 1.5;
    }
    uint8_t *visited = (uint8_t *)calloc(g->n_nodes, 1);
    gcand_t *wv = (gcand_t *)malloc(sizeof(gcand_t) * (size_t)(g->n_nodes + 1));
    gcand_t *fv = (gcand_t *)malloc(sizeof(gcand_t) * (size_t)(g->n_nodes + 1));
    gcand_t *res = (gcand_t *)malloc(sizeof(gcand_t) * (size_t)(k + 2));
    gheap_t w = {wv, 0, 0}, fr = {fv, 0, 1};
    uint32_t nres = 0;
    uint32_t n_seeds = (g_fault == 7 && g->n_seeds > 1) ? 1 : g->n_seeds;
    for (uint32_t i = 0; i < n_seeds; i++) {
        uint32_t sid = g->seeds[i];
        if (visited[sid]) continue;
        visited[sid] = 1;
        gcand_t c = {graph_dist(g, metric, dim, q, sid), sid};
        gheap_push(&fr, c);
        gheap_push(&w, c);
        if (filtered && g->live[sid]) res[nres++] = c;
    }
    if (filtered) {   /* sort_by(total_cmp) (stable), truncate(k) */
        for (uint32_t i = 1; i < nres; i++) {
            gcand_t c = res[i];
            uint32_t j = i;
            while (j > 0 && total_key(res[j - 1].d) > total_key(c.d)) {
                res[j] = res[j - 1];
                j--;
            }
            res[j] = c;
        }
        if (nres > k) nres = k;
    }
    double iterations = 0;
    while (fr.n > 0) {
        gcand_t best = fr.v[0];
        iterations += 1;
        if (iterations > max_it) break;
        if (filtered && nres >= k && nres > 0) {
            float worst = res[nres - 1].d;
            if (isfinite(early) && worst >= 0.0f && best.d > worst * (float)fmax(early, 1.0)) break;
        }
        if (w.n >= wb) {
            gcand_t worst = w.v[0];
            if (g_fault == 6 ? best.d > worst.d : best.d >= worst.d) break;
        }
        gcand_t cur = gheap_pop(&fr);
        for (uint32_t a = g->adj_off[cur.id]; a < g->adj_off[cur.id + 1]; a++) {
            uint32_t nb = g->adj[a];
            if (visited[nb]) continue;
            visited[nb] = 1;
            gcand_t c = {graph_dist(g, metric, dim, q, nb), nb};
            if (w.n < wb) {
                gheap_push(&w, c);
                gheap_push(&fr, c);
            } else if (c.d < w.v[0].d) {
                gheap_pop(&w);
                gheap_push(&w, c);
                gheap_push(&fr, c);
            }
            if (filtered && g->live[nb]) {
                uint32_t pos = nres;
                for (uint32_t i = 0; i < nres; i++)
                    if (c.d < res[i].d) {
                        pos = i;
                        break;
                    }
                if (pos < k) {
                    for (uint32_t i = nres; i > pos; i--) res[i] = res[i - 1];
                    res[pos] = c;
                    nres++;
                    if (nres > k) nres--;
                }
            }
        }
    }
    uint32_t cnt = 0;
    if (filtered) {
        for (uint32_t i = 0; i < nres; i++) {
            out_node[cnt] = res[i].id;
            out_dist[cnt] = res[i].d;
            cnt++;
        }
    } else {   /* w.into_vec(), stable sort by total_cmp, truncate(k) */
        gsort_t *s = (gsort_t *)malloc(sizeof(gsort_t) * (size_t)(w.n + 1));
        for (uint32_t i = 0; i < w.n; i++) {
            s[i].key = total_key(w.v[i].d);
            s[i].pos = i;
            s[i].c = w.v[i];
        }
        qsort(s, w.n, sizeof(gsort_t), gsort_cmp);
        for (uint32_t i = 0; i < w.n && cnt < k; i++) {
            out_node[cnt] = s[i].c.id;
            out_dist[cnt] = s[i].c.d;
            cnt++;
        }
        free(s);
    }
    free(visited);
    free(wv);
    free(fv);
    free(res);
    return cnt;
}

/* Inner::nearest over the graph (src/spfresh.rs:374-378): n results with
 * beam max(centroid_beam, n), mapped to postings (nodes without one dropped). */
static uint32_t graph_probe(const spf_o_index *ix, const float *q, uint32_t n, uint32_t *out_post,
                            float *out_dist) {
    const spf_o_graph *g = ix->graph;
    uint32_t beam = g->centroid_beam > n ? g->centroid_beam : n;
    uint32_t *nodes = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    float *dd = (float *)malloc(sizeof(float) * (size_t)(n + 1));
    uint32_t c = spf_o_graph_search(g, ix->metric, ix->dim, q, n, beam, nodes, dd);
    uint32_t m = 0;
    for (uint32_t i = 0; i < c; i++) {
        int64_t p = g->node_post[nodes[i]];
        if (p < 0) continue;
        out_post[m] = (uint32_t)p;
        out_dist[m] = dd[i];
        m++;
    }
    free(nodes);
    free(dd);
    return m;
}

/* Inner::nearest, exact branch (src/spfresh.rs:358-372), then the probe_ratio
 * take_while of Inner::search (src/spfresh.rs:683-687).  Ties in centroid
 * distance are ordered by reference block index. */
uint32_t spf_o_probe(const spf_o_index *ix, const float *q, uint32_t n_probe,
                     uint32_t *out_post, float *out_dist) {
    uint32_t n = n_probe < 1 ? 1 : n_probe;
    if (ix->n_post == 0) return 0;
    if (ix->graph && ix->n_post > 4096) {
        uint32_t *gp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
        float *gd = (float *)malloc(sizeof(float) * (size_t)(n + 1));
        uint32_t m = graph_probe(ix, q, n, gp, gd);
        uint32_t cnt = 0;
        float ratio = ix->probe_ratio;
        for (uint32_t i = 0; i < m; i++) {
            if (g_fault != 3 && isfinite(ratio) && gd[0] >= 0.0f && gd[i] > gd[0] * fmaxf(ratio, 1.0f))
                break;
            out_post[cnt] = gp[i];
            out_dist[cnt] = gd[i];
            cnt++;
        }
        free(gp);
        free(gd);
        return cnt;
    }
    cand_t *all = (cand_t *)malloc(sizeof(cand_t) * ix->n_post);
    for (uint32_t p = 0; p < ix->n_post; p++) {
        float dd = spf_o_eval(ix->metric, q, ix->centroids + (size_t)p * ix->dim, ix->dim);
        all[p].key = fkey(dd);
        all[p].id = ix->block_ids[p];
        all[p].aux = p;
        all[p].dist = dd;
    }
    qsort(all, ix->n_post, sizeof(cand_t), cand_cmp);
    if (n > ix->n_post) n = ix->n_post;
    float d0 = all[0].dist, ratio = ix->probe_ratio;
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n; i++) {
        float dd = all[i].dist;
        if (g_fault != 3 && isfinite(ratio) && d0 >= 0.0f && dd > d0 * fmaxf(ratio, 1.0f)) break;
        out_post[cnt] = all[i].aux;
        out_dist[cnt] = dd;
        cnt++;
    }
    free(all);
    return cnt;
}

/* Inner::code_dist (src/spfresh.rs:344-354) with quantized_distance_from_codes
 * (src/quantized.rs:13-54) for the non-RaBitQ-L2 cases. */
typedef struct {
    uint16_t *q16;   /* F16 prepared query */
    uint8_t *q8;     /* Int8 prepared query */
    float *tmp;      /* decode buffer (>= padded floats) */
} prep_t;

static float code_dist(const spf_o_index *ix, const float *q, const uint8_t *code, uint32_t post,
                       prep_t *pp) {
    uint32_t d = ix->dim;
    int l2 = ix->metric == SPF_O_L2, l2sq = ix->metric == SPF_O_L2SQ;
    switch (ix->quant) {
    case SPF_O_NONE:
        return spf_o_eval(ix->metric, q, (const float *)code, d);
    case SPF_O_F16: {
        if (l2 || l2sq) {
            float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            uint32_t n = (g_fault == 1 && d > 0) ? d - 1 : d;
            for (uint32_t i = 0; i < n; i++) {
                uint16_t h;
                memcpy(&h, code + 2 * i, 2);
                float t = spf_o_f16_to_f32(pp->q16[i]) - spf_o_f16_to_f32(h);
                c[(i >> 2) & 3] = fmaf(t, t, c[(i >> 2) & 3]);
            }
            float acc = combine4(c);
            return l2 ? sqrtf(acc) : acc;
        }
        for (uint32_t i = 0; i < d; i++) {
            uint16_t h;
            memcpy(&h, code + 2 * i, 2);
            pp->tmp[i] = spf_o_f16_to_f32(h);
        }
        return spf_o_eval(ix->metric, q, pp->tmp, d);
    }
    case SPF_O_INT8:
        if (l2 || l2sq) {
            float v = spf_o_int8_l2sq(pp->q8, code, d, ix->int8_scale);
            return l2 ? sqrtf(v) : v;
        }
        int8_decode(code, d, ix->int8_scale, ix->int8_offset, pp->tmp);
        return spf_o_eval(ix->metric, q, pp->tmp, d);
    case SPF_O_RABITQ:
        /* reached only for cosine / dot (L2 metrics take search_rabitq) */
        spf_o_rabitq_decode(code, ix->centroids + (size_t)post * d, d, ix->rq_padded, ix->rq_signs,
                            pp->tmp);
        return spf_o_eval(ix->metric, q, pp->tmp, d);
    }
    return NAN;
}

/* search_rabitq (src/spfresh.rs:731-770).  The size-k max-heap keeps its
 * entries sorted by (dist, id); an entry replaces the current k-th only when
 * its exact distance is strictly smaller, as in the reference. */
static uint32_t search_rabitq(const spf_o_index *ix, const float *q, uint32_t k,
                              const uint32_t *probes, uint32_t np, uint64_t *out_ids,
                              float *out_dists, int filtered, uint64_t mask) {
    uint32_t d = ix->dim, P = ix->rq_padded;
    uint64_t *planes = (uint64_t *)malloc(sizeof(uint64_t) * 4 * (P / 64));
    float consts[5];
    cand_t *heap = (cand_t *)malloc(sizeof(cand_t) * (k + 1));
    uint32_t cnt = 0;
    for (uint32_t pi = 0; pi < np; pi++) {
        uint32_t p = probes[pi];
        spf_o_rabitq_query(q, ix->centroids + (size_t)p * d, d, P, ix->rq_signs, planes, consts);
        for (uint64_t e = ix->post_off[p]; e < ix->post_off[p + 1]; e++) {
            uint64_t id = ix->ids[e];
            if (skipped(ix, id, filtered, mask)) continue;
            float est, err;
            spf_o_rabitq_dist(planes, consts, ix->codes + e * ix->code_bytes, P, &est, &err);
            float kth = cnt < k ? INFINITY : heap[cnt - 1].dist;
            if (est - err < kth) {
                float exact = spf_o_l2sq(q, ix->raw + id * d, d);
                cand_t c = {fkey(exact), 0, id, exact};
                if (cnt < k) {
                    heap[cnt++] = c;
                } else if (exact < kth) {
                    heap[cnt - 1] = c;
                } else {
                    continue;
                }
                /* restore sorted order (insertion) */
                for (uint32_t i = cnt - 1; i > 0 && cand_cmp(&heap[i], &heap[i - 1]) < 0; i--) {
                    cand_t t = heap[i];
                    heap[i] = heap[i - 1];
                    heap[i - 1] = t;
                }
            }
        }
    }
    for (uint32_t i = 0; i < cnt; i++) {
        float v = ix->metric == SPF_O_L2 ? sqrtf(heap[i].dist) : heap[i].dist;
        heap[i].dist = v;
        heap[i].key = fkey(v);
    }
    qsort(heap, cnt, sizeof(cand_t), cand_cmp);
    for (uint32_t i = 0; i < cnt; i++) {
        out_ids[i] = heap[i].id;
        out_dists[i] = heap[i].dist;
    }
    free(planes);
    free(heap);
    return cnt;
}

/* Inner::search (src/spfresh.rs:671-727).  The top-`want` heap is restated as
 * "the `want` smallest by (dist, id)". */
static uint32_t search_impl(const spf_o_index *ix, const float *q, uint32_t k, uint32_t n_probe,
                            uint64_t *out_ids, float *out_dists, int filtered, uint64_t mask) {
    uint32_t d = ix->dim;
    uint32_t n = n_probe < 1 ? 1 : n_probe;
    /* the graph's beam depends on the unclamped probe count */
    if (n > ix->n_post && !(ix->graph && ix->n_post > 4096)) n = ix->n_post;
    if (ix->n_post == 0 || k == 0) return 0;
    uint32_t *probes = (uint32_t *)malloc(sizeof(uint32_t) * n);
    float *pd = (float *)malloc(sizeof(float) * n);
    uint32_t np = spf_o_probe(ix, q, n, probes, pd);
    free(pd);
    if (ix->quant == SPF_O_RABITQ && (ix->metric == SPF_O_L2 || ix->metric == SPF_O_L2SQ)) {
        uint32_t r = search_rabitq(ix, q, k, probes, np, out_ids, out_dists, filtered, mask);
        free(probes);
        return r;
    }
    int rerank = ix->quant != SPF_O_NONE && ix->rerank_size > 0;
    uint32_t want = rerank ? (k > ix->rerank_size ? k : ix->rerank_size) : k;
    uint64_t total = 0;
    for (uint32_t i = 0; i < np; i++) total += ix->post_off[probes[i] + 1] - ix->post_off[probes[i]];
    cand_t *c = (cand_t *)malloc(sizeof(cand_t) * (total ? total : 1));
    prep_t pp = {0};
    uint32_t tmpn = ix->rq_padded > d ? ix->rq_padded : d;
    pp.tmp = (float *)malloc(sizeof(float) * tmpn);
    if (ix->quant == SPF_O_F16) {
        pp.q16 = (uint16_t *)malloc(2 * d);
        for (uint32_t i = 0; i < d; i++) pp.q16[i] = spf_o_f32_to_f16(q[i]);
    }
    if (ix->quant == SPF_O_INT8) {
        pp.q8 = (uint8_t *)malloc(d);
        spf_o_int8_encode(q, d, ix->int8_scale, ix->int8_offset, pp.q8);
    }
    uint64_t m = 0;
    for (uint32_t i = 0; i < np; i++) {
        uint32_t p = probes[i];
        for (uint64_t e = ix->post_off[p]; e < ix->post_off[p + 1]; e++) {
            uint64_t id = ix->ids[e];
            if (skipped(ix, id, filtered, mask)) continue;
            float dd = code_dist(ix, q, ix->codes + e * ix->code_bytes, p, &pp);
            c[m].key = fkey(dd);
            c[m].id = id;
            c[m].dist = dd;
            c[m].aux = 0;
            m++;
        }
    }
    qsort(c, m, sizeof(cand_t), cand_cmp);
    uint64_t keep = m < want ? m : want;
    if (rerank) {
        for (uint64_t i = 0; i < keep; i++) {
            float dd = spf_o_eval(ix->metric, q, ix->raw + c[i].id * d, d);
            c[i].dist = dd;
            c[i].key = fkey(dd);
        }
        qsort(c, keep, sizeof(cand_t), cand_cmp);
    }
    uint32_t out = (uint32_t)(keep < k ? keep : k);
    for (uint32_t i = 0; i < out; i++) {
        out_ids[i] = c[i].id;
        out_dists[i] = c[i].dist;
    }
    free(c);
    free(pp.tmp);
    free(pp.q16);
    free(pp.q8);
    free(probes);
    return out;
}

uint32_t spf_o_search(const spf_o_index *ix, const float *q, uint32_t k, uint32_t n_probe,
                      uint64_t *out_ids, float *out_dists) {
    return search_impl(ix, q, k, n_probe, out_ids, out_dists, 0, 0);
}

uint32_t spf_o_search_filtered(const spf_o_index *ix, const float *q, uint64_t mask, uint32_t k,
                               uint32_t n_probe, uint64_t *out_ids, float *out_dists) {
    return search_impl(ix, q, k, n_probe, out_ids, out_dists, 1, mask);
}

/* ---- batch: one task per query on a thread pool (src/spfresh.rs:897-908) ---- */
typedef struct {
    const spf_o_index *ix;
    const float *q;
    const uint64_t *masks;      /* NULL = unfiltered */
    uint32_t nq, k, n_probe;
    uint64_t *ids;
    float *dists;
    uint32_t *counts;
    uint32_t next;
    pthread_mutex_t mu;
} batch_t;

static void *batch_worker(void *arg) {
    batch_t *b = (batch_t *)arg;
    for (;;) {
        pthread_mutex_lock(&b->mu);
        uint32_t i = b->next;
        if (i < b->nq) b->next = i + 1;
        pthread_mutex_unlock(&b->mu);
        if (i >= b->nq) break;
        b->counts[i] = search_impl(b->ix, b->q + (size_t)i * b->ix->dim, b->k, b->n_probe,
                                   b->ids + (size_t)i * b->k, b->dists + (size_t)i * b->k,
                                   b->masks != NULL, b->masks ? b->masks[i] : 0);
    }
    return NULL;
}

static void batch_run(const spf_o_index *ix, const float *q, const uint64_t *masks, uint32_t nq,
                      uint32_t k, uint32_t n_probe, uint32_t n_threads, uint64_t *out_ids,
                      float *out_dists, uint32_t *out_counts) {
    batch_t b = {ix, q, masks, nq, k, n_probe, out_ids, out_dists, out_counts, 0,
                 PTHREAD_MUTEX_INITIALIZER};
    if (n_threads < 1) n_threads = 1;
    if (n_threads == 1) {
        batch_worker(&b);
        return;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
    for (uint32_t t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, batch_worker, &b);
    for (uint32_t t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    free(th);
}

void spf_o_search_batch(const spf_o_index *ix, const float *q, uint32_t nq, uint32_t k,
                        uint32_t n_probe, uint32_t n_threads, uint64_t *out_ids,
                        float *out_dists, uint32_t *out_counts) {
    batch_run(ix, q, NULL, nq, k, n_probe, n_threads, out_ids, out_dists, out_counts);
}

void spf_o_search_batch_filtered(const spf_o_index *ix, const float *q, const uint64_t *masks,
                                 uint32_t nq, uint32_t k, uint32_t n_probe, uint32_t n_threads,
                                 uint64_t *out_ids, float *out_dists, uint32_t *out_counts) {
    batch_run(ix, q, masks, nq, k, n_probe, n_threads, out_ids, out_dists, out_counts);
}
