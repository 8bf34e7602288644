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
// 4-dim units u with u % 4 == r, combined as (c0 + c1) + (c2 + c3)
__device__ float node_dist(uint32_t metric, const float4 *q4, const float4 *x4, uint32_t units,
                           float qn) {
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, b2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const bool l2 = metric == M_L2 || metric == M_L2SQ;
    for (uint32_t u = 0; u < units; u++) {
        const uint32_t r = u & 3;
        const float4 a = q4[u], b = x4[u];
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
            if (fresh)
                d = node_dist(p.metric, q4, reinterpret_cast<const float4 *>(p.vec + (size_t)nb * p.dpad),
                              units, qn);
            if (valid) nbs[lane] = GC{d, fresh ? nb : 0xffffffffu};
            __syncthreads();
            if (lane == 0 && !s_abort) {
                uint32_t wn = s_wn, fn = s_fn, rn = s_rn;
                for (uint32_t i = 0; i < n; i++) {
                    const GC c = nbs[i];
                    if (c.id == 0xffffffffu) continue;
                    vis[c.id >> 5] |= 1u << (c.id & 31);
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
