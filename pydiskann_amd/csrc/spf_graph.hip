// Routing-graph probe: Inner::nearest above 4,096 live postings
// (src/spfresh.rs:20,374-378) searches the centroid graph, an
// IncrementalDiskANN (src/incremental.rs:934-1003) whose beam_search
// (src/lib.rs:198-303) is a sequential best-first walk.  One wave per query:
// the 64 lanes take the distances of a node's neighbors (canonical f32 order),
// lane 0 applies the heap updates in neighbor order, so the visit order, the
// heap layouts and therefore the tie order are the reference's.
//
// Heaps follow std::collections::BinaryHeap: push = append + sift_up; pop =
// last element to the root, sift_down_to_bottom (towards the child that is
// not <= its sibling), sift_up.  Candidate compares by distance alone
// (src/lib.rs:100-122).  `w` is a max-heap, the frontier a min-heap
// (Reverse<Candidate>).
#include "spf_internal.h"

namespace spf {

namespace {

struct GC {
    float d;
    uint32_t id;
};

__device__ __forceinline__ uint32_t total_key(float f) {   // f32::total_cmp order
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// a <= b in the heap's order (rev: Reverse<Candidate>)
__device__ __forceinline__ bool gle(bool rev, GC a, GC b) { return rev ? (b.d <= a.d) : (a.d <= b.d); }

__device__ void heap_sift_up(GC *h, bool rev, uint32_t start, uint32_t pos) {
    const GC e = h[pos];
    while (pos > start) {
        const uint32_t parent = (pos - 1) / 2;
        if (gle(rev, e, h[parent])) break;
        h[pos] = h[parent];
        pos = parent;
    }
    h[pos] = e;
}

__device__ void heap_push(GC *h, uint32_t &n, bool rev, GC c) {
    h[n] = c;
    n++;
    heap_sift_up(h, rev, 0, n - 1);
}

__device__ GC heap_pop(GC *h, uint32_t &n, bool rev) {
    const GC item = h[n - 1];
    n--;
    if (n == 0) return item;
    const GC top = h[0];
    h[0] = item;
    const uint32_t end = n;
    uint32_t pos = 0, child = 1;
    while (end >= 2 && child <= end - 2) {
        child += gle(rev, h[child], h[child + 1]) ? 1u : 0u;
        h[pos] = h[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (child == end - 1) {
        h[pos] = h[child];
        pos = child;
    }
    h[pos] = item;
    heap_sift_up(h, rev, 0, pos);
    return top;
}

// one node's metric distance by one lane, in the canonical order: chain r over
// 4-dim units u with u % 4 == r, combined as (c0 + c1) + (c2 + c3).  The row
// is read kNodeBatch units at a time, all loads of a batch in flight together
// (a walk step waits on these reads; one unit at a time, a 96-dim row would be
// 24 round trips to memory).
constexpr uint32_t kNodeBatch = 12;   // a multiple of 4: chain r = j & 3 in a batch
__device__ float node_dist(uint32_t metric, const float4 *q4, const float4 *x4, uint32_t units,
                           float qn) {
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, b2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const bool l2 = metric == M_L2 || metric == M_L2SQ;
    for (uint32_t u0 = 0; u0 < units; u0 += kNodeBatch) {
        float4 xb[kNodeBatch];
#pragma unroll
        for (uint32_t j = 0; j < kNodeBatch; j++)
            xb[j] = u0 + j < units ? x4[u0 + j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (uint32_t j = 0; j < kNodeBatch; j++) {
            if (u0 + j >= units) break;
            const uint32_t r = j & 3;
            const float4 a = q4[u0 + j], b = xb[j];
            if (l2) {
                float t = a.x - b.x; acc[r] = __builtin_fmaf(t, t, acc[r]);
                t = a.y - b.y; acc[r] = __builtin_fmaf(t, t, acc[r]);
                t = a.z - b.z; acc[r] = __builtin_fmaf(t, t, acc[r]);
                t = a.w - b.w; acc[r] = __builtin_fmaf(t, t, acc[r]);
            } else {
                acc[r] = __builtin_fmaf(a.x, b.x, acc[r]);
                acc[r] = __builtin_fmaf(a.y, b.y, acc[r]);
                acc[r] = __builtin_fmaf(a.z, b.z, acc[r]);
                acc[r] = __builtin_fmaf(a.w, b.w, acc[r]);
                b2[r] = __builtin_fmaf(b.x, b.x, b2[r]);
                b2[r] = __builtin_fmaf(b.y, b.y, b2[r]);
                b2[r] = __builtin_fmaf(b.z, b.z, b2[r]);
                b2[r] = __builtin_fmaf(b.w, b.w, b2[r]);
            }
        }
    }
    const float s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    switch (metric) {
    case M_L2SQ: return s;
    case M_L2: return sqrtf(s);
    case M_DOT: return 1.0f - s;
    default: {
        const float xn = (b2[0] + b2[1]) + (b2[2] + b2[3]);
        if (qn == 0.0f && xn == 0.0f) return 0.0f;
        if (s == 0.0f) return 1.0f;
        const float r = 1.0f - s / sqrtf(qn * xn);
        return r > 0.0f ? r : 0.0f;
    }
    }
}

}  // namespace

// State: visited bitmap | w heap (wcap) | frontier heap (fcap) | results
// (rcap, filtered mode), in LDS (GL = false) or in this query's global scratch
// (GL = true, when it does not fit); then in LDS the query (dpad floats) and
// the neighbor scratch (64).
template <bool GL>
__global__ __launch_bounds__(64) void k_graph_probe(const GraphParams p) {
    extern __shared__ __align__(16) uint32_t s_mem[];
    __shared__ uint32_t s_go, s_cur, s_wn, s_fn, s_rn, s_abort;
    const uint32_t q = p.q0 + blockIdx.x, lane = threadIdx.x;
    if (q >= p.nq) return;
    if (p.only && !p.only[q]) {   // walked by an earlier launch
        if (lane == 0 && p.aborted) p.aborted[q] = 0u;
        return;
    }
    uint32_t *vis = GL ? p.scratch + (size_t)blockIdx.x * p.scratch_words : s_mem;
    GC *w = reinterpret_cast<GC *>(vis + p.bitmap_words);
    GC *fr = w + p.wcap;
    GC *res = fr + p.fcap;
    float *qv = GL ? reinterpret_cast<float *>(s_mem)
                   : reinterpret_cast<float *>(res + (p.filtered ? p.rcap : 0));
    GC *nbs = reinterpret_cast<GC *>(qv + p.dpad);
    for (uint32_t i = lane; i < p.bitmap_words; i += 64) vis[i] = 0u;
    for (uint32_t i = lane; i < p.dpad; i += 64) qv[i] = p.qv[(size_t)q * p.dpad + i];
    const float qn = p.qnorm[q];
    const uint32_t units = p.dpad / 4;
    if (lane == 0) {
        s_wn = 0;
        s_fn = 0;
        s_rn = 0;
        s_abort = 0;
    }
    __syncthreads();
    const float4 *q4 = reinterpret_cast<const float4 *>(qv);

    // seeds: every new one enters both heaps (and the results when live)
    {
        const uint32_t sid = lane == 0 ? p.seed0 : p.seed1;
        const bool ok = lane < p.n_seeds && !(lane == 1 && p.seed1 == p.seed0);
        float d = 0.0f;
        if (ok) d = node_dist(p.metric, q4, reinterpret_cast<const float4 *>(p.vec + (size_t)sid * p.dpad),
                              units, qn);
        if (lane < 2) nbs[lane] = GC{d, ok ? sid : 0xffffffffu};
        __syncthreads();
        if (lane == 0) {
            uint32_t wn = 0, fn = 0, rn = 0;
            for (uint32_t i = 0; i < 2; i++) {
                const GC c = nbs[i];
                if (c.id == 0xffffffffu) continue;
                vis[c.id >> 5] |= 1u << (c.id & 31);
                heap_push(fr, fn, true, c);
                heap_push(w, wn, false, c);
                if (p.filtered && p.live[c.id]) res[rn++] = c;
            }
            if (p.filtered) {   // stable sort by total_cmp, truncate to k
                if (rn == 2 && total_key(res[1].d) < total_key(res[0].d)) {
                    const GC t = res[0];
                    res[0] = res[1];
                    res[1] = t;
                }
                if (rn > p.k) rn = p.k;
            }
            s_wn = wn;
            s_fn = fn;
            s_rn = rn;
        }
        __syncthreads();
    }

    uint32_t iterations = 0;
    for (;;) {
        if (lane == 0) {
            uint32_t go = s_fn > 0 && !s_abort;
            if (go) {
                const GC best = fr[0];
                iterations++;
                if (iterations > p.max_it) go = 0;
                if (go && p.filtered && s_rn >= p.k && s_rn > 0) {
                    const float worst = res[s_rn - 1].d;
                    if (isfinite(p.early) && worst >= 0.0f && best.d > worst * fmaxf(p.early, 1.0f))
                        go = 0;
                }
                if (go && s_wn >= p.wb && best.d >= w[0].d) go = 0;
                if (go) {
                    uint32_t fn = s_fn;
                    s_cur = heap_pop(fr, fn, true).id;
                    s_fn = fn;
                }
            }
            s_go = go;
        }
        __syncthreads();
        if (!s_go) break;
        const uint32_t a0 = p.adj_off[s_cur], a1 = p.adj_off[s_cur + 1];
        for (uint32_t base = a0; base < a1; base += 64) {
            const uint32_t n = min(64u, a1 - base);
            const bool valid = lane < n;
            const uint32_t nb = valid ? p.adj[base + lane] : 0xffffffffu;
            bool fresh = valid && !((vis[nb >> 5] >> (nb & 31)) & 1u);
            for (uint32_t j = 0; j < n; j++) {   // a repeat of an earlier slot is not new
                const uint32_t o = __shfl(nb, j, 64);
                if (j < lane && o == nb) fresh = false;
            }
            float d = 0.0f;
            bool keep = fresh;
            if (fresh) {
                d = node_dist(p.metric, q4, reinterpret_cast<const float4 *>(p.vec + (size_t)nb * p.dpad),
                              units, qn);
                atomicOr(&vis[nb >> 5], 1u << (nb & 31));
                // The full w heap's top only falls while this batch is applied
                // (a pop of the top, then a push of something below it), and so
                // does the k-th result: a neighbor at or above both now is
                // passed over in its turn, and is dropped here already.
                if (s_wn >= p.wb && !(d < w[0].d))
                    keep = p.filtered && p.live[nb] && (s_rn < p.k || (p.k && d < res[p.k - 1].d));
                nbs[lane] = GC{d, nb};
            }
            const uint64_t kept = __ballot(keep);
            __syncthreads();
            if (lane == 0 && !s_abort && kept) {
                uint32_t wn = s_wn, fn = s_fn, rn = s_rn;
                for (uint64_t km = kept; km; km &= km - 1) {   // in neighbor order
                    const GC c = nbs[__builtin_ctzll(km)];
                    bool push_f = false;
                    if (wn < p.wb) {
                        heap_push(w, wn, false, c);
                        push_f = true;
                    } else if (c.d < w[0].d) {
                        heap_pop(w, wn, false);
                        heap_push(w, wn, false, c);
                        push_f = true;
                    }
                    if (push_f) {
                        if (fn >= p.fcap) {
                            s_abort = 1;
                            break;
                        }
                        heap_push(fr, fn, true, c);
                    }
                    if (p.filtered && p.live[c.id]) {
                        uint32_t pos = rn;
                        for (uint32_t t = 0; t < rn; t++)
                            if (c.d < res[t].d) {
                                pos = t;
                                break;
                            }
                        if (pos < p.k) {
                            for (uint32_t t = rn; t > pos; t--) res[t] = res[t - 1];
                            res[pos] = c;
                            rn++;
                            if (rn > p.k) rn--;
                        }
                    }
                }
                s_wn = wn;
                s_fn = fn;
                s_rn = rn;
            }
            __syncthreads();
        }
    }

    // output: first k of the result order, nodes without a posting dropped
    uint64_t *out = p.out + (size_t)q * p.npr;
    if (lane == 0 && p.aborted) p.aborted[q] = s_abort;
    if (s_abort) {
        if (lane == 0) {
            p.out_cnt[q] = 0;
            atomicAdd(p.overflow, 1u);
        }
        return;
    }
    if (p.filtered) {
        if (lane == 0) {
            uint32_t m = 0;
            for (uint32_t i = 0; i < s_rn && m < p.npr; i++) {
                const int32_t post = p.node_post[res[i].id];
                if (post < 0) continue;
                out[m++] = ((uint64_t)total_key(res[i].d) << 32) | (uint32_t)post;
            }
            p.out_cnt[q] = m;
        }
        return;
    }
    // w.into_vec() stably sorted by total_cmp: rank of slot i = slots before
    // it in (key, slot) order; the frontier region holds slot-by-rank
    const uint32_t wn = s_wn;
    uint32_t *byrank = reinterpret_cast<uint32_t *>(fr);
    for (uint32_t i = lane; i < wn; i += 64) {
        const uint32_t ki = total_key(w[i].d);
        uint32_t r = 0;
        for (uint32_t j = 0; j < wn; j++) {
            const uint32_t kj = total_key(w[j].d);
            r += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
        }
        byrank[r] = i;
    }
    __syncthreads();
    if (lane == 0) {
        uint32_t m = 0;
        const uint32_t top = min(wn, p.k);
        for (uint32_t r = 0; r < top && m < p.npr; r++) {
            const GC c = w[byrank[r]];
            const int32_t post = p.node_post[c.id];
            if (post < 0) continue;
            out[m++] = ((uint64_t)total_key(c.d) << 32) | (uint32_t)post;
        }
        p.out_cnt[q] = m;
    }
}

// ---------------------------------------------------------------------------
// Wave-uniform walk: the same beam search with every lane of the wave running
// the heap updates in lockstep instead of lane 0 alone behind barriers.  The
// walk state (heap sizes, current node) sits in scalar registers, the w heap
// (at most 64 entries) in one VGPR pair, entry i in lane i (a sift moves
// every entry on its path at once, by shuffles), the frontier heap in LDS.
// The heap algorithms and the visit order are those of k_graph_probe, so
// results are identical.
// Used when the graph has no tombstones (no result list) and the beam is at
// most 64, with the visited bitmap and frontier in LDS.
//
// Distances take two lanes per neighbor (32 neighbors a pass): lane half h
// runs chains r = 2h and 2h + 1 of the canonical order over its units, one
// batch of loads in flight, and the halves combine as (c0 + c1) + (c2 + c3).

__device__ __forceinline__ float rl_f(float v, uint32_t l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}
__device__ __forceinline__ uint32_t rl_u(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
// the w heap (max-heap by distance) in lanes: wd / wi hold entry `lane`
struct LaneHeap {
    float wd;
    uint32_t wi;
    uint32_t lane;
    __device__ __forceinline__ GC get(uint32_t i) const { return GC{rl_f(wd, i), rl_u(wi, i)}; }
    // BinaryHeap sift_up of e from slot pos, every lane at once: lane j
    // (1 <= j <= depth) reads the j-th ancestor's distance, a ballot finds the
    // first at or above e (the walk up stops below it), the ancestors below
    // that move down one level (each takes its parent's entry), e takes the
    // freed slot
    __device__ __forceinline__ void sift_up(uint32_t pos, GC e) {
        const uint32_t I = pos + 1, dp = 31u - (uint32_t)__builtin_clz(I);   // 1-based slot
        const bool mine = lane >= 1 && lane <= dp;
        const float ad = __shfl(wd, mine ? (int)((I >> lane) - 1) : 0, 64);
        const int par = lane > 0 ? (int)((lane - 1) / 2) : 0;
        const float pd = __shfl(wd, par, 64);
        const uint32_t pi = __shfl(wi, par, 64);
        const uint64_t stop = __ballot(mine && e.d <= ad);
        const uint32_t J = stop ? (uint32_t)__builtin_ctzll(stop) : dp + 1;
        // this lane's slot: k levels above pos on its path (k = 0: pos)
        const int k = (int)dp - (int)(31u - (uint32_t)__builtin_clz(lane + 1));
        const bool on = k >= 0 && (I >> k) == lane + 1;
        if (on && (uint32_t)k + 1 < J) {
            wd = pd;
            wi = pi;
        } else if (on && (uint32_t)k + 1 == J) {
            wd = e.d;
            wi = e.id;
        }
    }
    __device__ __forceinline__ void push(uint32_t &n, GC c) {
        n++;
        sift_up(n - 1, c);
    }
    // BinaryHeap::pop (result unused): the descent's path is found by
    // readlane compares, then every entry on it moves up one level at once
    // (each takes its child toward the path's end), then the last entry is
    // sifted up from there
    __device__ __forceinline__ void pop(uint32_t &n) {
        const GC item = get(n - 1);
        n--;
        if (n == 0) return;
        const uint32_t end = n;
        uint32_t pos = 0, child = 1;
        while (end >= 2 && child <= end - 2) {
            const bool right = rl_f(wd, child) <= rl_f(wd, child + 1);
            pos = child + (right ? 1u : 0u);
            child = 2 * pos + 1;
        }
        if (child == end - 1) pos = child;
        const uint32_t I = pos + 1, dp = 31u - (uint32_t)__builtin_clz(I);
        const int k = (int)dp - (int)(31u - (uint32_t)__builtin_clz(lane + 1));
        const bool on = k >= 1 && (I >> k) == lane + 1;
        const int src = on ? (int)((I >> (k - 1)) - 1) : (int)lane;
        const float sd = __shfl(wd, src, 64);
        const uint32_t si = __shfl(wi, src, 64);
        if (on) {
            wd = sd;
            wi = si;
        }
        sift_up(pos, item);
    }
};

// the frontier (min-heap by distance) in LDS, run by the whole wave: loads are
// made uniform, lane 0 stores
__device__ __forceinline__ GC lds_get(const GC *h, uint32_t i) {
    const GC c = h[i];
    return GC{__int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(c.d))),
              (uint32_t)__builtin_amdgcn_readfirstlane((int)c.id)};
}
__device__ __forceinline__ void lds_set(GC *h, uint32_t i, GC c, uint32_t lane) {
    if (lane == 0) h[i] = c;
}
// BinaryHeap sift_up of e from slot pos, all ancestors read at once: lane j
// (1 <= j <= depth) holds the j-th ancestor; the walk up stops below the
// first one at or before e in the frontier's order (Reverse: a.d <= e.d), the
// ancestors below it move down one level, e takes the freed slot.
__device__ __forceinline__ void fr_sift_up(GC *h, uint32_t pos, GC e, uint32_t lane) {
    const uint32_t I = pos + 1, depth = 31u - (uint32_t)__builtin_clz(I);   // 1-based slot
    const bool mine = lane >= 1 && lane <= depth;
    GC v{0.0f, 0u};
    if (mine) v = h[(I >> lane) - 1];
    const uint64_t stop = __ballot(mine && v.d <= e.d);
    const uint32_t J = stop ? (uint32_t)__builtin_ctzll(stop) : depth + 1;
    if (mine && lane < J) h[(I >> (lane - 1)) - 1] = v;
    if (lane == 0) h[(I >> (J - 1)) - 1] = e;
}

// BinaryHeap pop (sift_down_to_bottom, then sift_up).  The descent reads the
// five levels below its position at once (lane t < 62: the node at depth
// k = log2(t + 2) below, path t + 2 - 2^k) and takes five steps from them.
constexpr uint32_t kPopLevels = 5;
__device__ __forceinline__ GC fr_pop(GC *h, uint32_t &n, uint32_t lane) {
    const GC item = lds_get(h, n - 1);
    n--;
    if (n == 0) return item;
    const GC top = lds_get(h, 0);
    const uint32_t end = n;
    const uint32_t k = 31u - (uint32_t)__builtin_clz(lane + 2), b = lane + 2 - (1u << k);
    uint32_t pos = 0, child = 1;
    while (end >= 2 && child <= end - 2) {
        const uint32_t idx = ((pos + 1) << k) + b - 1;
        GC v{0.0f, 0u};
        if (lane < 62 && idx < end) v = h[idx];
        uint32_t rel = 1;   // the position in the loaded subtree, 1-based (its root: pos)
        for (uint32_t l = 0; l < kPopLevels && end >= 2 && child <= end - 2; l++) {
            const GC c0{rl_f(v.d, 2 * rel - 2), rl_u(v.id, 2 * rel - 2)};
            const GC c1{rl_f(v.d, 2 * rel - 1), rl_u(v.id, 2 * rel - 1)};
            const bool right = c1.d <= c0.d;   // Reverse order: c0 <= c1 iff c1.d <= c0.d
            lds_set(h, pos, right ? c1 : c0, lane);
            pos = child + (right ? 1u : 0u);
            rel = 2 * rel + (right ? 1u : 0u);
            child = 2 * pos + 1;
        }
    }
    if (child == end - 1) {
        lds_set(h, pos, lds_get(h, child), lane);
        pos = child;
    }
    fr_sift_up(h, pos, item, lane);
    return top;
}

constexpr uint32_t kPairSteps = 6;   // 4-unit steps a lane pair loads at once (12 units a lane)

// the metric distance of node x by a lane pair; lane half h adds chains 2h, 2h + 1
__device__ float pair_dist(uint32_t metric, const float4 *q4, const float4 *x4, uint32_t units,
                           float qn, uint32_t h) {
    float acc[2] = {0.0f, 0.0f}, b2[2] = {0.0f, 0.0f};
    const bool l2 = metric == M_L2 || metric == M_L2SQ;
    for (uint32_t t0 = 0; 4 * t0 < units; t0 += kPairSteps) {
        float4 xb[kPairSteps][2];
#pragma unroll
        for (uint32_t t = 0; t < kPairSteps; t++)
#pragma unroll
            for (uint32_t e = 0; e < 2; e++) {
                const uint32_t u = 4 * (t0 + t) + 2 * h + e;
                xb[t][e] = u < units ? x4[u] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
        for (uint32_t t = 0; t < kPairSteps; t++)
#pragma unroll
            for (uint32_t e = 0; e < 2; e++) {
                const uint32_t u = 4 * (t0 + t) + 2 * h + e;
                if (u >= units) continue;
                const float4 a = q4[u], b = xb[t][e];
                if (l2) {
                    float d = a.x - b.x; acc[e] = __builtin_fmaf(d, d, acc[e]);
                    d = a.y - b.y; acc[e] = __builtin_fmaf(d, d, acc[e]);
                    d = a.z - b.z; acc[e] = __builtin_fmaf(d, d, acc[e]);
                    d = a.w - b.w; acc[e] = __builtin_fmaf(d, d, acc[e]);
                } else {
                    acc[e] = __builtin_fmaf(a.x, b.x, acc[e]);
                    acc[e] = __builtin_fmaf(a.y, b.y, acc[e]);
                    acc[e] = __builtin_fmaf(a.z, b.z, acc[e]);
                    acc[e] = __builtin_fmaf(a.w, b.w, acc[e]);
                    b2[e] = __builtin_fmaf(b.x, b.x, b2[e]);
                    b2[e] = __builtin_fmaf(b.y, b.y, b2[e]);
                    b2[e] = __builtin_fmaf(b.z, b.z, b2[e]);
                    b2[e] = __builtin_fmaf(b.w, b.w, b2[e]);
                }
            }
    }
    const float sh = acc[0] + acc[1], so = __shfl_xor(sh, 1, 64);
    const float s = h ? so + sh : sh + so;   // (c0 + c1) + (c2 + c3)
    switch (metric) {
    case M_L2SQ: return s;
    case M_L2: return sqrtf(s);
    case M_DOT: return 1.0f - s;
    default: {
        const float bh = b2[0] + b2[1], bo = __shfl_xor(bh, 1, 64);
        const float xn = h ? bo + bh : bh + bo;
        if (qn == 0.0f && xn == 0.0f) return 0.0f;
        if (s == 0.0f) return 1.0f;
        const float r = 1.0f - s / sqrtf(qn * xn);
        return r > 0.0f ? r : 0.0f;
    }
    }
}

// LDS: visited bitmap | frontier heap (fcap) | query (dpad floats)
__global__ __launch_bounds__(64) void k_graph_walk(const GraphParams p) {
    extern __shared__ __align__(16) uint32_t s_mem[];
    const uint32_t q = p.q0 + blockIdx.x, lane = threadIdx.x;
    if (q >= p.nq) return;
    if (p.only && !p.only[q]) {
        if (lane == 0 && p.aborted) p.aborted[q] = 0u;
        return;
    }
    uint32_t *vis = s_mem;
    GC *fr = reinterpret_cast<GC *>(vis + p.bitmap_words);
    float *qv = reinterpret_cast<float *>(fr + p.fcap);
    for (uint32_t i = lane; i < p.bitmap_words; i += 64) vis[i] = 0u;
    for (uint32_t i = lane; i < p.dpad; i += 64) qv[i] = p.qv[(size_t)q * p.dpad + i];
    const float qn = p.qnorm[q];
    const uint32_t units = p.dpad / 4, half = lane & 1, pr = lane >> 1;
    const float4 *q4 = reinterpret_cast<const float4 *>(qv);
    __syncthreads();

    LaneHeap w{0.0f, 0u, lane};
    uint32_t wn = 0, fn = 0;
    bool abort = false;
    {   // seeds, in order: lanes 0-1 seed 0, lanes 2-3 seed 1
        const uint32_t sid = pr == 0 ? p.seed0 : p.seed1;
        const bool ok = pr < p.n_seeds && !(pr == 1 && p.seed1 == p.seed0);
        const float d = pair_dist(p.metric, q4,
                                  reinterpret_cast<const float4 *>(p.vec + (size_t)(ok ? sid : p.seed0) * p.dpad),
                                  units, qn, half);
        for (uint32_t i = 0; i < 2; i++) {
            if (!rl_u(ok ? 1u : 0u, 2 * i)) continue;
            const GC c{rl_f(d, 2 * i), rl_u(sid, 2 * i)};
            if (lane == 0) vis[c.id >> 5] |= 1u << (c.id & 31);
            fn++;
            fr_sift_up(fr, fn - 1, c, lane);
            w.push(wn, c);
        }
    }

    uint32_t iterations = 0;
    while (fn > 0) {
        const GC best = lds_get(fr, 0);
        iterations++;
        if (iterations > p.max_it) break;
        if (wn >= p.wb && best.d >= w.get(0).d) break;
        const uint32_t cur = fr_pop(fr, fn, lane).id;
        const uint32_t a0 = p.adj_off[cur], a1 = p.adj_off[cur + 1];
        for (uint32_t base = a0; base < a1; base += 32) {
            const uint32_t n = min(32u, a1 - base);
            const bool valid = pr < n;
            const uint32_t nb = valid ? p.adj[base + pr] : 0u;
            const uint32_t bit = 1u << (nb & 31);
            bool fresh = valid && !(vis[nb >> 5] & bit);
            // mark visited; a neighbor listed twice shows up as a set bit in
            // a second lane's atomic result, then the later copies are not new
            const uint32_t old = (fresh && !half) ? atomicOr(&vis[nb >> 5], bit) : 0u;
            if (__ballot(fresh && !half && (old & bit))) {
                for (uint32_t j = 0; j < n; j++) {
                    const uint32_t o = __shfl(nb, 2 * j, 64);
                    if (j < pr && o == nb) fresh = false;
                }
            }
            float d = 0.0f;
            if (__ballot(fresh))
                d = pair_dist(p.metric, q4, reinterpret_cast<const float4 *>(p.vec + (size_t)nb * p.dpad),
                              units, qn, half);
            // the full w heap's top only falls while this pass is applied: a
            // neighbor at or above it now is passed over in its turn
            const bool keep = fresh && (wn < p.wb || d < w.get(0).d);
            for (uint64_t km = __ballot(keep && !half); km; km &= km - 1) {   // neighbor order
                const uint32_t l = __builtin_ctzll(km);
                const GC c{rl_f(d, l), rl_u(nb, l)};
                if (wn < p.wb) {
                    w.push(wn, c);
                } else if (c.d < w.get(0).d) {
                    w.pop(wn);
                    w.push(wn, c);
                } else {
                    continue;
                }
                if (fn >= p.fcap) {
                    abort = true;
                    break;
                }
                fn++;
                fr_sift_up(fr, fn - 1, c, lane);
            }
            if (abort) break;
        }
        if (abort) break;
    }

    uint64_t *out = p.out + (size_t)q * p.npr;
    if (lane == 0 && p.aborted) p.aborted[q] = abort ? 1u : 0u;
    if (abort) {
        if (lane == 0) {
            p.out_cnt[q] = 0;
            atomicAdd(p.overflow, 1u);
        }
        return;
    }
    // w.into_vec() stably sorted by total_cmp: lane i's rank among the slots,
    // entries by rank into the frontier region (fcap >= 64)
    uint32_t r = 0;
    const uint32_t ki = total_key(w.wd);
    for (uint32_t j = 0; j < wn; j++) {
        const uint32_t kj = total_key(rl_f(w.wd, j));
        r += (kj < ki || (kj == ki && j < lane)) ? 1u : 0u;
    }
    __syncthreads();
    if (lane < wn) fr[r] = GC{w.wd, w.wi};
    __syncthreads();
    if (lane == 0) {
        uint32_t m = 0;
        const uint32_t top = min(wn, p.k);
        for (uint32_t t = 0; t < top && m < p.npr; t++) {
            const GC c = fr[t];
            const int32_t post = p.node_post[c.id];
            if (post < 0) continue;
            out[m++] = ((uint64_t)total_key(c.d) << 32) | (uint32_t)post;
        }
        p.out_cnt[q] = m;
    }
}

// the wave-uniform walk over queries [p.q0, p.q0 + n) (LDS mode, no result list, wcap <= 64)
hipError_t launch_graph_walk(const GraphParams &p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (p.scratch || p.filtered || p.wcap > 64 || p.fcap < 64) return hipErrorInvalidValue;
    const size_t lds = (size_t)p.bitmap_words * 4 + (size_t)p.fcap * 8 + (size_t)p.dpad * 4;
    if (lds > 65536) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_graph_walk, dim3(n), dim3(64), lds, s, p);
    return hipGetLastError();
}

// one launch over queries [p.q0, p.q0 + n)
hipError_t launch_graph_probe(const GraphParams &p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (p.scratch)
        hipLaunchKernelGGL(k_graph_probe<true>, dim3(n), dim3(64), graph_lds_bytes(p), s, p);
    else
        hipLaunchKernelGGL(k_graph_probe<false>, dim3(n), dim3(64), graph_lds_bytes(p), s, p);
    return hipGetLastError();
}

}  // namespace spf
