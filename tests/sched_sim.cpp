// CPU check of the search server's scheduler (pydiskann_amd/csrc/spf_sched.h)
// against a simulated executor: every row's result must be the function of
// that row alone, the simulated device memory of all slots must stay within
// the budget, batches must merge requests and run concurrently, and errors,
// timeouts, rejection and shutdown must behave as documented.
// Built and run by tests/test_sched.py (also under the thread sanitizer).
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "spf_sched.h"

using spf_sched::Batch;
using spf_sched::Key;
using spf_sched::Piece;
using spf_sched::Scheduler;

static int g_fail = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "CHECK failed line %d: %s\n", __LINE__, #c); \
            g_fail = 1;                                                 \
        }                                                               \
    } while (0)

struct SimExec {
    uint64_t budget;
    std::mutex mu;
    std::vector<uint64_t> held_b, cand_b;   // all buffers; the candidate part of them
    uint64_t over_budget = 0;
    uint32_t fail_probe = 999;   // n_probe that makes a batch fail
    int sleep_us = 200;

    explicit SimExec(int slots, uint64_t b) : budget(b), held_b(slots, 0), cand_b(slots, 0) {}
    int validate(const Key &key, uint32_t dim, std::string *err) {
        if (dim != 4) {
            *err = "dim";
            return SPF_EDIM;
        }
        if (key.k == 0) {
            *err = "k";
            return SPF_EINVAL;
        }
        return SPF_OK;
    }
    uint64_t bound(const Key &key, uint32_t nq) { return 4096 + (uint64_t)nq * (64 + 12 * key.k); }
    uint64_t held(int s) {
        std::lock_guard<std::mutex> g(mu);
        return held_b[s];
    }
    void trim(int s) {
        std::lock_guard<std::mutex> g(mu);
        held_b[s] = 0;
        cand_b[s] = 0;
    }
    // candidate bytes of a row, known only once the batch runs (row value
    // 1e9 marks a row too large for any budget)
    static uint64_t cand_bytes(const float *q) { return q[2] >= 1e9f ? (1ull << 40) : (uint64_t)q[2]; }
    void grow_to(int s, uint64_t v) {
        std::lock_guard<std::mutex> g(mu);
        held_b[s] = std::max(held_b[s], v);
        uint64_t t = 0;
        for (uint64_t x : held_b) t += x;
        if (t > budget) over_budget++;
    }
    template <class Grow>
    int run(int s, const Batch &b, std::string *err, Grow &grow) {
        // phase-1 buffers grow to this batch's need, as a workspace does
        grow_to(s, bound(b.key, b.nq) - 1024);
        uint64_t extra = 0;
        for (const Piece &p : b.pieces)
            for (uint32_t i = 0; i < p.n; i++) extra += cand_bytes(p.r->q + (size_t)(p.row0 + i) * 4);
        uint64_t more;   // candidate buffers grow to this batch's need
        {
            std::lock_guard<std::mutex> g(mu);
            more = extra > cand_b[s] ? extra - cand_b[s] : 0;
        }
        if (more)
            if (int g = grow(more)) return g;
        {
            std::lock_guard<std::mutex> g(mu);
            held_b[s] += more;
            cand_b[s] += more;
            uint64_t t = 0;
            for (uint64_t x : held_b) t += x;
            if (t > budget) over_budget++;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(sleep_us));
        if (b.key.n_probe == fail_probe) {
            *err = "simulated device error";
            return SPF_EDEVICE;
        }
        for (const Piece &p : b.pieces)
            for (uint32_t i = 0; i < p.n; i++) {
                const uint32_t row = p.row0 + i;
                const float *q = p.r->q + (size_t)row * 4;
                const uint64_t m = b.key.filtered ? p.r->masks[row] : 0;
                for (uint32_t j = 0; j < b.key.k; j++) {
                    p.r->ids[(size_t)row * b.key.k + j] =
                        (uint64_t)q[0] * 7 + j + b.key.n_probe * 100000ull + (m & 0xffff);
                    p.r->dists[(size_t)row * b.key.k + j] = q[1] + (float)j;
                }
                p.r->counts[row] = b.key.k;
            }
        return SPF_OK;
    }
};

struct Req {
    Key key;
    uint32_t nq;
    std::vector<float> q;
    std::vector<uint64_t> masks, ids;
    std::vector<float> d;
    std::vector<uint32_t> c;
    uint64_t ticket = 0;
    int rc_submit = 0;
};

static Req make_req(std::mt19937_64 &rng, uint32_t nq, uint32_t k, uint32_t n_probe, bool filt) {
    Req r;
    r.key.k = k;
    r.key.n_probe = n_probe;
    r.key.filtered = filt;
    r.nq = nq;
    r.q.resize((size_t)nq * 4);
    for (auto &v : r.q) v = (float)(rng() % 100000);
    for (uint32_t i = 0; i < nq; i++) r.q[(size_t)i * 4 + 2] = (float)(rng() % 2000);
    if (filt) {
        r.masks.resize(nq);
        for (auto &m : r.masks) m = rng();
    }
    r.ids.assign((size_t)nq * k, 0);
    r.d.assign((size_t)nq * k, 0);
    r.c.assign(nq, 0);
    return r;
}

static bool rows_ok(const Req &r) {
    for (uint32_t row = 0; row < r.nq; row++) {
        if (r.c[row] != r.key.k) return false;
        const uint64_t m = r.key.filtered ? r.masks[row] : 0;
        for (uint32_t j = 0; j < r.key.k; j++) {
            const uint64_t want = (uint64_t)r.q[(size_t)row * 4] * 7 + j +
                                  r.key.n_probe * 100000ull + (m & 0xffff);
            if (r.ids[(size_t)row * r.key.k + j] != want) return false;
            if (r.d[(size_t)row * r.key.k + j] != r.q[(size_t)row * 4 + 1] + (float)j) return false;
        }
    }
    return true;
}

static void mixed_load(int slots, uint64_t budget, int n_req, int n_threads, uint32_t seed,
                       bool expect_merge = true) {
    SimExec ex(slots, budget);
    std::vector<Req> reqs(n_req);
    {
        std::mt19937_64 rng(seed);
        for (auto &r : reqs) {
            const uint32_t nq = (uint32_t)(rng() % 257);   // 0..256 rows
            const uint32_t k = rng() % 2 ? 10 : 100;
            const uint32_t np = rng() % 2 ? 16 : 64;
            r = make_req(rng, nq, k, np, rng() % 2);
        }
    }
    Scheduler<SimExec> sch(&ex, slots, 1024, budget);
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; t++)
        th.emplace_back([&, t] {
            for (int i = t; i < n_req; i += n_threads) {
                Req &r = reqs[i];
                std::string err;
                r.rc_submit = sch.submit(r.key, r.q.data(), r.nq, 4,
                                         r.key.filtered ? r.masks.data() : nullptr, r.ids.data(),
                                         r.d.data(), r.c.data(), &r.ticket, &err);
            }
            for (int i = t; i < n_req; i += n_threads) {
                Req &r = reqs[i];
                CHECK(r.rc_submit == SPF_OK);
                spf_sched::Request info;
                std::string err;
                CHECK(sch.wait(r.ticket, -1, &info, &err) == SPF_OK);
                CHECK(info.t_done >= info.t_start && info.t_start >= info.t_submit);
                CHECK(rows_ok(r));
            }
        });
    for (auto &t : th) t.join();
    const spf_sched::Stats st = sch.stats();
    uint64_t rows = 0;
    for (auto &r : reqs) rows += r.nq;
    CHECK(st.requests_done == (uint64_t)n_req);
    CHECK(st.rows_done == rows);
    CHECK(st.overruns == 0);
    CHECK(st.peak_reserved <= budget);
    CHECK(ex.over_budget == 0);
    if (expect_merge) CHECK(st.batches < (uint64_t)n_req);   // requests share batches
    if (slots > 1) CHECK(st.max_concurrent > 1);
    std::printf("mixed slots=%d budget=%llu: %llu batches for %d requests, max concurrent %u, "
                "peak reserved %llu, waits %llu, trims %llu\n",
                slots, (unsigned long long)budget, (unsigned long long)st.batches, n_req,
                st.max_concurrent, (unsigned long long)st.peak_reserved,
                (unsigned long long)st.admission_waits, (unsigned long long)st.trims);
}

static void edge_cases() {
    SimExec ex(2, 1 << 20);
    Scheduler<SimExec> sch(&ex, 2, 64, 1 << 20);
    std::mt19937_64 rng(7);
    std::string err;
    // a single query above the budget is refused at submit
    Req big = make_req(rng, 1, 100000, 16, false);
    CHECK(sch.submit(big.key, big.q.data(), 1, 4, nullptr, big.ids.data(), big.d.data(),
                     big.c.data(), &big.ticket, &err) == SPF_ENOMEM);
    // argument errors come from the executor's check
    Req r0 = make_req(rng, 3, 10, 16, false);
    CHECK(sch.submit(r0.key, r0.q.data(), 3, 5, nullptr, r0.ids.data(), r0.d.data(),
                     r0.c.data(), &r0.ticket, &err) == SPF_EDIM);
    // a failing batch fails its requests only
    Req bad = make_req(rng, 100, 10, 999, false);
    Req good = make_req(rng, 100, 10, 16, false);
    CHECK(sch.submit(bad.key, bad.q.data(), bad.nq, 4, nullptr, bad.ids.data(), bad.d.data(),
                     bad.c.data(), &bad.ticket, &err) == SPF_OK);
    CHECK(sch.submit(good.key, good.q.data(), good.nq, 4, nullptr, good.ids.data(),
                     good.d.data(), good.c.data(), &good.ticket, &err) == SPF_OK);
    CHECK(sch.wait(bad.ticket, -1, nullptr, &err) == SPF_EDEVICE);
    CHECK(err == "simulated device error");
    CHECK(sch.wait(good.ticket, -1, nullptr, &err) == SPF_OK);
    CHECK(rows_ok(good));
    // consumed and unknown tickets
    CHECK(sch.wait(good.ticket, 0, nullptr, &err) == SPF_EINVAL);
    // empty request completes at once
    Req empty = make_req(rng, 0, 10, 16, false);
    CHECK(sch.submit(empty.key, empty.q.data(), 0, 4, nullptr, empty.ids.data(), empty.d.data(),
                     empty.c.data(), &empty.ticket, &err) == SPF_OK);
    CHECK(sch.wait(empty.ticket, 0, nullptr, &err) == SPF_OK);
    // timeout leaves the ticket valid
    ex.sleep_us = 20000;
    Req slow = make_req(rng, 200, 10, 16, false);
    CHECK(sch.submit(slow.key, slow.q.data(), slow.nq, 4, nullptr, slow.ids.data(),
                     slow.d.data(), slow.c.data(), &slow.ticket, &err) == SPF_OK);
    CHECK(sch.wait(slow.ticket, 0, nullptr, &err) == SPF_EAGAIN);
    spf_sched::Request info;
    CHECK(sch.wait(slow.ticket, -1, &info, &err) == SPF_OK);
    CHECK(info.batches >= 2);   // 200 rows over batches of at most 64
    CHECK(rows_ok(slow));
    ex.sleep_us = 200;
}

static void drain_on_destroy() {
    SimExec ex(2, 1 << 24);
    std::mt19937_64 rng(9);
    std::vector<Req> reqs;
    for (int i = 0; i < 20; i++) reqs.push_back(make_req(rng, 50, 10, 16, i % 2));
    {
        Scheduler<SimExec> sch(&ex, 2, 128, 1 << 24);
        std::string err;
        for (auto &r : reqs)
            CHECK(sch.submit(r.key, r.q.data(), r.nq, 4, r.key.filtered ? r.masks.data() : nullptr,
                             r.ids.data(), r.d.data(), r.c.data(), &r.ticket, &err) == SPF_OK);
    }   // destroyed without waiting: every request still runs
    for (auto &r : reqs) CHECK(rows_ok(r));
}

static void tight_budget_trims() {
    // budget for about one 100-row k=100 batch: slots must give memory back
    SimExec probe(1, 0);
    Key k100;
    k100.k = 100;
    const uint64_t budget = probe.bound(k100, 100) + probe.bound(k100, 20);
    mixed_load(3, budget, 60, 3, 11, false);
}

static void candidate_heavy() {
    // rows whose candidate bytes dominate: batches must wait, yield or split
    SimExec probe(1, 0);
    Key k10;
    k10.k = 10;
    const uint64_t budget = probe.bound(k10, 256) * 3 + 300000;
    SimExec ex(3, budget);
    std::mutex mu;
    std::vector<Req> reqs;
    std::mt19937_64 rng(21);
    for (int i = 0; i < 80; i++) {
        Req r = make_req(rng, 1 + (uint32_t)(rng() % 256), 10, 16, i % 2);
        for (uint32_t j = 0; j < r.nq; j++) r.q[(size_t)j * 4 + 2] = (float)(rng() % 4000);
        reqs.push_back(std::move(r));
    }
    Req huge = make_req(rng, 3, 10, 16, false);
    huge.q[4 + 2] = 1e9f;   // row 1 can never fit
    {
        Scheduler<SimExec> sch(&ex, 3, 1024, budget);
        std::vector<std::thread> th;
        for (int t = 0; t < 4; t++)
            th.emplace_back([&, t] {
                for (int i = t; i < (int)reqs.size(); i += 4) {
                    Req &r = reqs[i];
                    std::string err;
                    CHECK(sch.submit(r.key, r.q.data(), r.nq, 4,
                                     r.key.filtered ? r.masks.data() : nullptr, r.ids.data(),
                                     r.d.data(), r.c.data(), &r.ticket, &err) == SPF_OK);
                }
                for (int i = t; i < (int)reqs.size(); i += 4) {
                    std::string err;
                    CHECK(sch.wait(reqs[i].ticket, -1, nullptr, &err) == SPF_OK);
                    CHECK(rows_ok(reqs[i]));
                }
            });
        std::string err;
        CHECK(sch.submit(huge.key, huge.q.data(), huge.nq, 4, nullptr, huge.ids.data(),
                         huge.d.data(), huge.c.data(), &huge.ticket, &err) == SPF_OK);
        CHECK(sch.wait(huge.ticket, -1, nullptr, &err) == SPF_ENOMEM);
        for (auto &t : th) t.join();
        const spf_sched::Stats st = sch.stats();
        CHECK(st.overruns == 0);
        CHECK(st.peak_reserved <= budget);
        CHECK(st.splits > 0);
        std::printf("candidate-heavy: %llu batches, requeues %u, splits %u, waits %llu, "
                    "peak %llu of %llu\n",
                    (unsigned long long)st.batches, st.requeues, st.splits,
                    (unsigned long long)st.admission_waits,
                    (unsigned long long)st.peak_reserved, (unsigned long long)budget);
    }
    CHECK(ex.over_budget == 0);
}

int main() {
    mixed_load(3, 64ull << 20, 400, 6, 1);
    mixed_load(1, 64ull << 20, 100, 2, 2);
    tight_budget_trims();
    edge_cases();
    drain_on_destroy();
    candidate_heavy();
    if (g_fail) return 1;
    std::printf("sched ok\n");
    return 0;
}
