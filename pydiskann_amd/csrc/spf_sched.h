// Continuous-batching scheduler of the search server.
//
// Requests (a block of query rows with one k / n_probe / precision / filter
// setting) enter a FIFO.  Each slot -- one worker thread owning one device
// stream and one working-memory set -- forms its next batch from the head
// request plus every later request with the same setting, up to max_batch
// rows, and a request's rows may be split over several batches and slots.
// Admission is by a device-memory budget: a slot may start a batch only if the
// bytes every slot holds or has reserved, plus the batch's upper bound, stay
// within the budget; idle slots give their memory back when that lets the
// head request start, or at least doubles its batch.  A request is complete
// when all its rows are back.
//
// Free of HIP so the same code builds with a host compiler and is tested on
// the CPU (thread sanitizer included) against a simulated executor.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "spf.h"

namespace spf_sched {

// run() results that hand a batch's rows back to the queue instead of
// completing them: kRequeue (memory needed mid-batch is held by batches that
// are themselves waiting), kSplit (the batch alone exceeds the budget; its
// requests continue in batches half as large)
constexpr int kRequeue = 1, kSplit = 2;

struct Key {
    uint32_t k = 0, n_probe = 0, precision = 0, rerank = 0, flags = 0;
    bool filtered = false;
    bool operator==(const Key &o) const {
        return k == o.k && n_probe == o.n_probe && precision == o.precision &&
               rerank == o.rerank && flags == o.flags && filtered == o.filtered;
    }
};

struct Request {
    uint64_t ticket = 0;
    Key key;
    uint32_t nq = 0, dim = 0;
    const float *q = nullptr;
    const uint64_t *masks = nullptr;
    uint64_t *ids = nullptr;
    float *dists = nullptr;
    uint32_t *counts = nullptr;
    std::deque<std::pair<uint32_t, uint32_t>> todo;   // row ranges not handed to a batch
    uint32_t left = 0;       // rows in todo
    uint32_t inflight = 0;   // batches holding rows of this request
    uint32_t batches = 0;
    uint32_t row_cap = UINT32_MAX;   // rows per batch after a split
    int status = SPF_OK;
    std::string err;
    uint64_t t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
};

struct Piece {
    Request *r;
    uint32_t row0, n;
};

struct Batch {
    Key key;
    uint32_t nq = 0;
    std::vector<Piece> pieces;
};

struct Stats {
    uint64_t requests_done = 0, rows_done = 0, batches = 0, rejected = 0;
    uint64_t peak_reserved = 0, overruns = 0, admission_waits = 0, trims = 0;
    uint32_t max_concurrent = 0, requeues = 0, splits = 0;
};

inline uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Exec provides:
//   int validate(const Key &, uint32_t dim, std::string *err)   argument check
//   uint64_t bound(const Key &, uint32_t nq)    bytes a batch of nq rows may add to a
//                                               slot before its candidate buffers
//   uint64_t held(int slot)                     bytes the slot holds now
//   void trim(int slot)                         free the slot's buffers (slot idle)
//   int run(int slot, const Batch &, std::string *err, Grow &grow)
//       run and write the rows back; before allocating its candidate buffers
//       it calls grow(bytes), and returns what grow returned if that is not 0
template <class Exec>
class Scheduler {
public:
    // the extra reservation of a running batch (its candidate buffers)
    struct Grow {
        Scheduler *sch;
        int slot;
        int operator()(uint64_t extra) { return sch->grow(slot, extra); }
    };

    Scheduler(Exec *ex, int n_slots, uint32_t max_batch, uint64_t budget)
        : ex_(ex), max_batch_(max_batch), budget_(budget), reserved_(n_slots, 0),
          running_(n_slots, false), waiting_(n_slots, false) {
        for (int s = 0; s < n_slots; s++) threads_.emplace_back([this, s] { worker(s); });
    }

    // Runs every submitted request to completion, then joins the workers.
    ~Scheduler() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        work_cv_.notify_all();
        for (auto &t : threads_) t.join();
        for (auto &kv : reqs_) delete kv.second;
    }

    int submit(const Key &key, const float *q, uint32_t nq, uint32_t dim, const uint64_t *masks,
               uint64_t *ids, float *dists, uint32_t *counts, uint64_t *ticket,
               std::string *err) {
        std::unique_lock<std::mutex> lk(mu_);
        int rc = ex_->validate(key, dim, err);
        if (rc) {
            stats_.rejected++;
            return rc;
        }
        if (ex_->bound(key, 1) > budget_) {
            stats_.rejected++;
            *err = "a single query of this request needs more device memory than the budget";
            return SPF_ENOMEM;
        }
        Request *r = new Request();
        r->ticket = next_ticket_++;
        r->key = key;
        r->nq = nq;
        r->dim = dim;
        r->q = q;
        r->masks = masks;
        r->ids = ids;
        r->dists = dists;
        r->counts = counts;
        r->t_submit = now_ns();
        reqs_[r->ticket] = r;
        *ticket = r->ticket;
        if (nq == 0) {
            r->t_start = r->t_done = r->t_submit;
            r->done = true;
            stats_.requests_done++;
            return SPF_OK;
        }
        r->todo.push_back({0u, nq});
        r->left = nq;
        pending_.push_back(r);
        lk.unlock();
        work_cv_.notify_all();
        return SPF_OK;
    }

    // timeout_us < 0: wait until done.  SPF_EAGAIN if not done in time (the
    // ticket stays valid); otherwise the request's own status, and the
    // ticket is consumed.
    int wait(uint64_t ticket, int64_t timeout_us, Request *info, std::string *err) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) {
            *err = "unknown ticket";
            return SPF_EINVAL;
        }
        Request *r = it->second;
        auto ready = [r] { return r->done; };
        if (timeout_us < 0)
            done_cv_.wait(lk, ready);
        else if (!done_cv_.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
            *err = "request not complete";
            return SPF_EAGAIN;
        }
        if (info) {
            info->ticket = r->ticket;
            info->t_submit = r->t_submit;
            info->t_start = r->t_start;
            info->t_done = r->t_done;
            info->batches = r->batches;
        }
        const int st = r->status;
        if (st) *err = r->err;
        reqs_.erase(it);
        delete r;
        return st;
    }

    Stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        return stats_;
    }

private:
    uint64_t reserved_except(int s) const {
        uint64_t t = 0;
        for (size_t i = 0; i < reserved_.size(); i++)
            if ((int)i != s) t += reserved_[i];
        return t;
    }

    void note_peak() {
        uint64_t tot = 0;
        for (uint64_t v : reserved_) tot += v;
        stats_.peak_reserved = std::max(stats_.peak_reserved, tot);
    }

    // give back the memory of idle slots (and of `self` when it is idle)
    void trim_idle(int self) {
        for (size_t t = 0; t < reserved_.size(); t++)
            if (((int)t == self || !running_[t]) && reserved_[t]) {
                ex_->trim((int)t);
                reserved_[t] = ex_->held((int)t);
                stats_.trims++;
            }
    }

    // put a request back into the queue in ticket (arrival) order
    void enqueue(Request *r) {
        if (std::find(pending_.begin(), pending_.end(), r) != pending_.end()) return;
        auto at = std::find_if(pending_.begin(), pending_.end(),
                               [r](Request *o) { return o->ticket > r->ticket; });
        pending_.insert(at, r);
    }

    // Called with mu_ held.  Fills b and reserves its memory, or returns false
    // (nothing to do, or the head request does not fit the budget yet).
    bool form(int s, Batch *b) {
        if (pending_.empty()) return false;
        // a running batch waiting for its candidate memory goes first
        for (size_t t = 0; t < waiting_.size(); t++)
            if (waiting_[t]) {
                stats_.admission_waits++;
                return false;
            }
        const Key key = pending_.front()->key;
        uint64_t avail = 0;
        uint32_t cap = max_batch_;
        for (Request *r : pending_)
            if (r->key == key) {
                avail += r->left;
                cap = std::min(cap, r->row_cap);
            }
        const uint32_t want = (uint32_t)std::min<uint64_t>(avail, std::max(cap, 1u));
        // largest row count in [0, want] whose bound fits beside `held` bytes
        auto rows_fitting = [&](uint64_t held) -> uint32_t {
            if (held + ex_->bound(key, 1) > budget_) return 0;
            uint32_t lo = 1, hi = want;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo + 1) / 2;
                if (held + ex_->bound(key, mid) <= budget_)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            return lo;
        };
        uint64_t base = reserved_[s];   // idle: the bytes this slot holds
        uint32_t lo = rows_fitting(reserved_except(s) + base);
        if (lo < want) {
            // idle slots (this one included) give their memory back when that
            // at least doubles the batch, or lets it start at all
            uint64_t busy = 0;
            for (size_t t = 0; t < reserved_.size(); t++)
                if ((int)t != s && running_[t]) busy += reserved_[t];
            const uint32_t freed = rows_fitting(busy);
            if (freed > 0 && (lo == 0 || freed >= 2 * lo)) {
                trim_idle(s);
                base = reserved_[s];
                lo = rows_fitting(reserved_except(s) + base);
            }
        }
        if (lo == 0) {
            stats_.admission_waits++;
            return false;
        }
        b->key = key;
        b->nq = lo;
        b->pieces.clear();
        uint32_t rows = lo;
        const uint64_t t = now_ns();
        for (Request *r : pending_) {
            if (!rows) break;
            if (!(r->key == key)) continue;
            bool took = false;
            while (rows && !r->todo.empty()) {
                auto &rg = r->todo.front();
                const uint32_t n = std::min(rows, rg.second - rg.first);
                b->pieces.push_back({r, rg.first, n});
                rg.first += n;
                if (rg.first == rg.second) r->todo.pop_front();
                r->left -= n;
                rows -= n;
                took = true;
            }
            if (took) {
                if (!r->t_start) r->t_start = t;
                r->inflight++;
                r->batches++;
            }
        }
        pending_.erase(std::remove_if(pending_.begin(), pending_.end(),
                                      [](Request *r) { return r->left == 0; }),
                       pending_.end());
        reserved_[s] = base + ex_->bound(key, lo);
        note_peak();
        return true;
    }

    // A running batch asks for `extra` bytes (its candidate buffers).  Granted
    // when they fit; otherwise the slot waits while some other batch is still
    // running and not itself waiting; if every other running batch waits too,
    // this one yields (kRequeue); alone and still too large, it is split.
    int grow(int s, uint64_t extra) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            if (reserved_except(s) + reserved_[s] + extra <= budget_) {
                reserved_[s] += extra;
                waiting_[s] = false;
                note_peak();
                return 0;
            }
            bool progress = false;   // another batch that will finish on its own
            for (size_t t = 0; t < running_.size(); t++)
                if ((int)t != s && running_[t] && !waiting_[t]) progress = true;
            if (!progress) {
                trim_idle(-1);
                if (reserved_except(s) + reserved_[s] + extra <= budget_) continue;
                waiting_[s] = false;
                bool others = false;
                for (size_t t = 0; t < running_.size(); t++)
                    if ((int)t != s && running_[t]) others = true;
                return others ? kRequeue : kSplit;
            }
            waiting_[s] = true;
            stats_.admission_waits++;
            work_cv_.wait(lk);
        }
    }

    void worker(int s) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            Batch b;
            while (!form(s, &b)) {
                if (stop_ && pending_.empty()) return;
                work_cv_.wait(lk);
            }
            running_[s] = true;
            n_running_++;
            stats_.max_concurrent = std::max(stats_.max_concurrent, n_running_);
            lk.unlock();
            std::string err;
            Grow g{this, s};
            int rc = ex_->run(s, b, &err, g);
            lk.lock();
            running_[s] = false;
            waiting_[s] = false;
            n_running_--;
            const uint64_t held = ex_->held(s);
            if (held > reserved_[s]) stats_.overruns++;
            reserved_[s] = held;
            const uint64_t t = now_ns();
            if (rc == kSplit && b.nq == 1) {
                rc = SPF_ENOMEM;
                err = "a single query needs more device memory than the budget";
            }
            if (rc == kRequeue || rc == kSplit) {   // rows back to the queue, in order
                if (rc == kRequeue)
                    stats_.requeues++;
                else
                    stats_.splits++;
                for (auto it = b.pieces.rbegin(); it != b.pieces.rend(); ++it) {
                    Request *r = it->r;
                    if (r->status != SPF_OK) continue;   // failed meanwhile: rows dropped
                    r->todo.push_front({it->row0, it->row0 + it->n});
                    r->left += it->n;
                    if (rc == kSplit) r->row_cap = std::max(1u, b.nq / 2);
                }
                Request *last = nullptr;
                for (const Piece &p : b.pieces) {   // a request's pieces are adjacent
                    Request *r = p.r;
                    if (r == last) continue;
                    last = r;
                    r->inflight--;
                    r->batches--;
                    if (r->left)
                        enqueue(r);
                    else if (r->inflight == 0 && !r->done) {
                        r->done = true;
                        r->t_done = t;
                        stats_.requests_done++;
                    }
                }
            } else {
                stats_.batches++;
                stats_.rows_done += b.nq;
                Request *last = nullptr;
                for (const Piece &p : b.pieces) {
                    Request *r = p.r;
                    if (r != last) r->inflight--;
                    last = r;
                    if (rc && r->status == SPF_OK) {
                        r->status = rc;
                        r->err = err;
                        if (r->left) {   // hand out no more of its rows
                            r->todo.clear();
                            r->left = 0;
                            pending_.erase(std::remove(pending_.begin(), pending_.end(), r),
                                           pending_.end());
                        }
                    }
                    if (r->left == 0 && r->inflight == 0 && !r->done) {
                        r->done = true;
                        r->t_done = t;
                        stats_.requests_done++;
                    }
                }
            }
            done_cv_.notify_all();
            work_cv_.notify_all();
        }
    }

    Exec *ex_;
    const uint32_t max_batch_;
    const uint64_t budget_;
    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<Request *> pending_;
    std::unordered_map<uint64_t, Request *> reqs_;
    std::vector<uint64_t> reserved_;
    std::vector<bool> running_, waiting_;
    std::vector<std::thread> threads_;
    uint32_t n_running_ = 0;
    uint64_t next_ticket_ = 1;
    bool stop_ = false;
    Stats stats_;
};

}  // namespace spf_sched
