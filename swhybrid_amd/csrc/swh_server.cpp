// swh_server.cpp -- continuous-batching search service on a resident database.
//
// A loop thread repeats rounds.  Each round it admits waiting requests (in
// arrival order) while their device buffers fit the memory budget and a slot
// is free, then enqueues ONE scan launch (plus the rescore) covering the next
// database slice of admitted requests, waits for it, and hands the requests
// that finished their last slice to a completion thread (top-k on the host
// from the pinned copy of the score vector).  Scores of a request depend only
// on its query and the database: each task scans its own block ranges and
// writes its own buffers, whatever else shares the launch.
//
// Live queue (the default on a database of one layout).  Instead of rounds,
// one persistent scan launch stays on the device while there is work, and
// the loop feeds it (swh_live.h): a request admitted while others run gets its
// task at once and its items are published between theirs, shortest
// remaining work first (with aging), keeping a short lead of queued items
// ahead of the workgroups.  The kernel writes each request's scores straight
// into its page-locked score copy, so a request whose items are all done is
// complete while the launch scans the others, and the loop makes no call
// while the launch runs that could wait for it (a copy that needs the CUs
// the launch holds would).  The launch ends when nothing is left to publish
// (or a request overflowed int16: the int32 rescore runs when it has ended),
// and the next one starts when work arrives.
//
// Rounds.  With opts.queue = SWH_QUEUE_ROUNDS (and always on a small database
// of two layouts) the loop runs rounds.  By default a round holds about round_ms of device work: each
// request keeps a cursor over the database's pair and lane blocks, requests
// are served shortest remaining work first (with aging, so none waits
// forever), and each one's slice covers the same share of its remaining pair
// and lane work.  A short request that arrives while a long one is on the
// device therefore runs in the next round instead of after the long one; a
// long one advances a slice per round.  With opts.slices = n every request is
// cut into n fixed slices instead (n = 1: the whole database per round).
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "swh_host.h"
#include "swh_live.h"
#include "swh_plan.h"

using namespace swhh;

namespace {

struct Request {
    uint64_t id = 0;
    std::vector<uint8_t> query;
    int k = 0;
    bool want_scores = false;
    int slot = -1;
    int next_slice = 0;                // fixed slices: the next one to run
    // budgeted rounds: the next block of each kind (lane blocks start after
    // the pair blocks; swh_server_submit sets lane_pos to the first of them)
    int pair_pos = 0, lane_pos = 0;
    size_t bytes = 0;                 // device bytes charged against the budget
    swh_request_info info{};
    std::vector<int32_t> top_score;
    std::vector<int64_t> top_index;
    std::vector<int32_t> scores;      // copy of the score vector if asked for
    bool done = false;
    // live queue: the request's items in launch order and how far they went
    std::vector<uint2> items;
    std::vector<double> rest;         // rest[i]: padded cells of items [i, end)
    int group = 0;                    // lanes per group of its group items
    size_t next_item = 0, items_done = 0;
    int task = -1;                    // its task in the running launch
    bool ready = false;               // its task's prep item is done
    bool overflowed = false;          // an item of the running launch overflowed
    bool scanned = false;             // every item done
};

}  // namespace

struct swh_server {
    swh_ctx* ctx = nullptr;
    swh_db* db = nullptr;
    swh_server_opts opts{};
    double t0 = 0;
    // slices of the block ranges
    std::vector<int> pair_cut, lane_cut;
    size_t max_slice_seqs = 0;
    bool budgeted = false;       // rounds of about round_ms of work (else fixed slices)
    double rate = 9.0e9;         // padded cells per ms the rounds have run at (estimate)
    // per-slot device buffers and pinned score copies
    DevBuf scores, ovf, prof, counters;
    PinnedBuf h_prof, h_scores;
    size_t prof_slot_elems = 0;
    std::vector<int> free_slots;
    size_t bytes_in_use = 0;
    // request state
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_complete;
    std::deque<std::shared_ptr<Request>> waiting;
    std::vector<std::shared_ptr<Request>> active;
    std::deque<std::shared_ptr<Request>> to_complete;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_id = 1;
    bool stop = false;        // no new admissions; the loop ends once nothing is active
    bool loop_done = false;   // the loop thread has returned
    int waiters = 0;          // threads inside swh_server_wait (destroy waits for them)
    std::string loop_error;
    swh_server_stats stats{};
    std::thread loop, completer;
    // live queue
    bool live = false;
    std::unique_ptr<LiveScan> lq;
    int32_t* d_scores = nullptr;        // device address of h_scores
};

namespace {

// Device work per budgeted round: a request that arrives while others are on
// the device waits at most about this long to join them.
// 160 ms: on the 5M-sequence database with Poisson requests of log-normal
// length (mean 400), 40 / 80 / 160 ms rounds against whole-request rounds gave
// p50 / p99 of 73 / 377, 78 / 324, 77 / 316 against 76 / 266 ms at 5 QPS,
// 1,316 / 1,933, 1,229 / 1,780, 1,201 / 1,761 against 1,558 / 2,354 ms at 20
// QPS, and 8,545, 9,031, 9,218 against 9,233 GCUPS at 50 QPS: a round costs a
// few ms of launch tail and host work, and a lane item of the longest lane
// blocks of a long query runs for longer than a short round anyway.
constexpr double kDefaultRoundMs = 160.0;
constexpr double kDefaultWaitWeight = 0.5;

double since(const swh_server* s) { return now_ms() - s->t0; }

size_t request_bytes(const swh_server* s, int qlen) {
    return s->db->nseq * sizeof(int32_t) + s->max_slice_seqs * sizeof(uint32_t) +
           (size_t)kProfRows * profile_stride(qlen) * sizeof(int16_t);
}

// Split [b0, b1) into n parts of about equal padded work.
std::vector<int> cut_blocks(const swh_db* db, int b0, int b1, int n) {
    std::vector<int> cut(n + 1, b0);
    cut[n] = b1;
    const uint64_t base = db->blk_padded[b0];
    const uint64_t total = db->blk_padded[b1] - base;
    int b = b0;
    for (int i = 1; i < n; ++i) {
        const uint64_t target = base + total * (uint64_t)i / n;
        while (b < b1 && db->blk_padded[b + 1] <= target) ++b;
        cut[i] = std::max(b, cut[i - 1]);
    }
    return cut;
}

// Budgeted round: the slices of swh_plan.h's plan_budgeted, priced at the
// rate the rounds have run at, with two lane items per workgroup slot of a
// full grid (2 workgroups per CU, 4 lane blocks per item) per round.
template <typename AddTask>
void plan_round(swh_server* s, const std::vector<std::shared_ptr<Request>>& batch,
                AddTask&& add_task) {
    const swh_db* db = s->db;
    const double now = since(s);
    std::vector<PlanRequest> reqs;
    for (const auto& rp : batch)
        reqs.push_back(PlanRequest{(int)rp->query.size(), rp->pair_pos, rp->lane_pos,
                                   now - rp->info.submit_ms});
    const std::vector<PlanSlice> plan =
        plan_budgeted(db->blk_padded.data(), db->n_pairs, db->n_blocks, reqs,
                      s->opts.round_ms * s->rate, s->rate, 16 * std::max(s->ctx->num_cus, 1),
                      s->opts.wait_weight);
    for (const PlanSlice& p : plan)
        add_task(batch[p.request], p.pair_begin, p.pair_end, p.lane_begin, p.lane_end, p.last);
}

// What a round ended with: work done, a clean stop (nothing active, the
// server is stopping), or a device failure (loop_error says which).
enum class Round { kDone, kStop, kFail };

// Admit waiting requests in arrival order while a slot is free and their
// device bytes fit the budget (the first always fits when nothing is active).
// Caller holds s->mu.
void admit(swh_server* s, std::vector<std::shared_ptr<Request>>* admitted) {
    while (!s->stop && !s->waiting.empty() && !s->free_slots.empty()) {
        auto& r = s->waiting.front();
        if (s->bytes_in_use + r->bytes > s->opts.memory_budget && !s->active.empty()) break;
        r->slot = s->free_slots.back();
        s->free_slots.pop_back();
        s->bytes_in_use += r->bytes;
        r->info.admit_ms = since(s);
        s->active.push_back(r);
        if (admitted) admitted->push_back(r);
        s->waiting.pop_front();
    }
    s->stats.peak_inflight = std::max<int>(s->stats.peak_inflight, (int)s->active.size());
    s->stats.waiting = (int)s->waiting.size();
}

// A request whose scores are in its pinned copy goes to the completion
// thread.  Caller holds s->mu.
void hand_over(swh_server* s, const std::shared_ptr<Request>& r) {
    s->active.erase(std::find(s->active.begin(), s->active.end(), r));
    s->to_complete.push_back(r);
    s->cv_complete.notify_one();
}

// One round: admit, launch the next slice of every active request, wait.
Round run_round(swh_server* s) {
    swh_ctx* c = s->ctx;
    swh_db* db = s->db;
    std::vector<std::shared_ptr<Request>> batch;
    std::vector<std::shared_ptr<Request>> admitted;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        // (with every slot held by requests awaiting completion there is
        // nothing to do until the completion thread frees one)
        s->cv_work.wait(lk, [&] {
            return s->stop || !s->active.empty() ||
                   (!s->waiting.empty() && !s->free_slots.empty());
        });
        admit(s, &admitted);
        if (s->active.empty()) return s->stop ? Round::kStop : Round::kDone;
        batch = s->active;
    }
    const double t_round = now_ms();
    (void)hipSetDevice(c->device);
    // profiles of newly admitted requests
    for (auto& r : admitted) {
        int16_t* hp = (int16_t*)s->h_prof.p + (size_t)r->slot * s->prof_slot_elems;
        build_profile(c->matrix, r->query.data(), (int)r->query.size(), hp);
        const size_t n = (size_t)kProfRows * profile_stride((int)r->query.size());
        if (hipMemcpyAsync((int16_t*)s->prof.p + (size_t)r->slot * s->prof_slot_elems, hp,
                           n * sizeof(int16_t), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            s->loop_error = "profile upload failed";
            return Round::kFail;
        }
    }
    const size_t nseq = db->nseq;
    // whole-database rounds of a small database run on the layout that suits
    // the round's request count (slices are block ranges of the first layout)
    if (s->opts.slices == 1 && !s->budgeted) db = pick_layout(c, db, (int)batch.size());
    std::vector<Task> tasks;
    std::vector<int> qlens;
    std::vector<std::shared_ptr<Request>> ran;   // the requests of this round, task order
    std::vector<bool> last;                      // ... and whether it is their last slice
    auto add_task = [&](const std::shared_ptr<Request>& rp, int pb, int pe, int lb, int le,
                        bool is_last) {
        Request& r = *rp;
        Task T{};
        T.prof = (const int16_t*)s->prof.p + (size_t)r.slot * s->prof_slot_elems;
        T.qlen = (int)r.query.size();
        T.scores = (int32_t*)s->scores.p + (size_t)r.slot * nseq;
        T.ovf = (uint32_t*)s->ovf.p + (size_t)r.slot * s->max_slice_seqs;
        T.pair_begin = pb;
        T.pair_end = pe;
        T.lane_begin = lb;
        T.lane_end = le;
        tasks.push_back(T);
        qlens.push_back(T.qlen);
        ran.push_back(rp);
        last.push_back(is_last);
    };
    if (!s->budgeted) {
        for (auto& rp : batch) {
            const Request& r = *rp;
            if (s->opts.slices == 1)
                add_task(rp, 0, db->n_pairs, db->n_pairs, db->n_blocks, true);
            else
                add_task(rp, s->pair_cut[r.next_slice], s->pair_cut[r.next_slice + 1],
                         s->lane_cut[r.next_slice], s->lane_cut[r.next_slice + 1],
                         r.next_slice + 1 == s->opts.slices);
        }
    } else {
        plan_round(s, batch, add_task);
    }
    const Shape sh = choose_shape(
        c, qlens.data(), (int)qlens.size(), s->opts.search, db,
        range_residues(db, tasks[0].pair_begin, tasks[0].pair_end) +
            range_residues(db, tasks[0].lane_begin, tasks[0].lane_end));
    swh_stats st{};
    int rc = enqueue_tasks(c, db, tasks.data(), (int)tasks.size(), sh, s->opts.search,
                           (uint32_t*)s->counters.p, ctx_res(c), &st);
    if (rc != SWH_OK) {
        s->loop_error = g_err;
        return Round::kFail;
    }
    // requests on their last slice: copy the scores out
    std::vector<std::shared_ptr<Request>> finished;
    for (size_t i = 0; i < ran.size(); ++i)
        if (last[i]) {
            const Request& r = *ran[i];
            finished.push_back(ran[i]);
            if (hipMemcpyAsync((int32_t*)s->h_scores.p + (size_t)r.slot * nseq,
                               (int32_t*)s->scores.p + (size_t)r.slot * nseq,
                               nseq * sizeof(int32_t), hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess) {
                s->loop_error = "score copy failed";
                return Round::kFail;
            }
        }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        s->loop_error = "round failed on the device";
        return Round::kFail;
    }
    const double round_ms = now_ms() - t_round;
    if (round_ms > 0 && st.padded_cells > 0) {   // the rate the next budget is priced at
        const double r = (double)st.padded_cells / round_ms;
        s->rate = std::min(2.0e10, std::max(1.0e8, 0.7 * s->rate + 0.3 * r));
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.rounds += 1;
    s->stats.busy_ms += round_ms;
    s->stats.cells += st.cells;
    for (size_t i = 0; i < ran.size(); ++i) {
        Request& r = *ran[i];
        const Task& T = tasks[i];
        r.next_slice += 1;
        r.pair_pos = T.pair_end;
        r.lane_pos = T.lane_end;
        r.info.rounds += 1;
        r.info.cells += (uint64_t)r.query.size() * (range_residues(db, T.pair_begin, T.pair_end) +
                                                    range_residues(db, T.lane_begin, T.lane_end));
    }
    for (auto& r : finished) hand_over(s, r);
    return Round::kDone;
}

// ---------------------------------------------------------------------------
// Live queue.

// Items queued ahead of the workgroups: enough that none waits for the loop
// between two polls, few enough that a request arriving now starts soon.
uint32_t live_lead(const LiveScan& L) { return (uint32_t)std::max(64, L.grid() / 2); }
constexpr int kLivePublishChunk = 32;          // items of one request per pick
constexpr auto kLivePoll = std::chrono::microseconds(100);
constexpr double kLiveCloseWaitMs = 120000.0;  // a launch ends within its longest item

// Profile and items of a newly admitted request.
void plan_live_request(swh_server* s, Request& r) {
    const int qlen = (int)r.query.size();
    int16_t* hp = (int16_t*)s->h_prof.p + (size_t)r.slot * s->prof_slot_elems;
    build_profile(s->ctx->matrix, r.query.data(), qlen, hp);
    r.group = s->lq->group_lanes_of(qlen);
    std::vector<double> cells;
    s->lq->plan_items(qlen, r.group, r.items, cells);
    r.rest.assign(cells.size() + 1, 0.0);
    for (size_t i = cells.size(); i-- > 0;) r.rest[i] = r.rest[i + 1] + cells[i];
    r.next_item = r.items_done = 0;
}

// A request whose scores are final goes to the completion thread.
void finish_live(swh_server* s, const std::shared_ptr<Request>& r) {
    r->info.cells = (uint64_t)r->query.size() * s->db->residues;
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.cells += r->info.cells;
    hand_over(s, r);
}

// One live launch: from the first admitted request until nothing is left to
// publish (or a rescore is due), feeding the kernel while it runs.
Round run_epoch(swh_server* s) {
    swh_ctx* c = s->ctx;
    LiveScan& L = *s->lq;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_work.wait(lk, [&] {
            return s->stop || !s->active.empty() ||
                   (!s->waiting.empty() && !s->free_slots.empty());
        });
        admit(s, nullptr);
        if (s->active.empty()) return s->stop ? Round::kStop : Round::kDone;
    }
    (void)hipSetDevice(c->device);
    const double t_epoch = now_ms();
    if (L.begin() != SWH_OK) {
        s->loop_error = g_err;
        return Round::kFail;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rounds += 1;   // launches begun (a request may finish before its launch ends)
    }
    std::vector<std::shared_ptr<Request>> mine;   // requests with a task in this launch
    bool drain = false;                           // publish nothing more: end the launch
    double padded_done = 0;
    auto fail_epoch = [&](const std::string& why) {
        s->loop_error = why;
        (void)L.close_and_wait(kLiveCloseWaitMs);
        return Round::kFail;
    };
    const uint32_t target = (uint32_t)L.grid() + live_lead(L);
    for (;;) {
        // admission; a task in this launch for every request without one
        std::vector<std::shared_ptr<Request>> need;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            admit(s, nullptr);
            for (auto& r : s->active)
                if (r->task < 0 && !r->scanned) need.push_back(r);
        }
        for (auto& r : need) {
            if (drain || L.room() == 0) break;
            if (r->items.empty()) plan_live_request(s, *r);
            const int16_t* hp = (const int16_t*)s->h_prof.p + (size_t)r->slot * s->prof_slot_elems;
            LiveTask t{hp, (int)r->query.size(), r->group, s->d_scores + (size_t)r->slot * s->db->nseq,
                       (uint32_t*)s->ovf.p + (size_t)r->slot * s->max_slice_seqs};
            const int q = L.add_task(t);
            if (q < 0) {   // this launch's arena is full: the next launch takes it
                drain = true;
                break;
            }
            r->task = q;
            r->ready = false;
            r->info.rounds += 1;
            mine.push_back(r);
            L.publish(q, make_uint2(kPrepItem, 0), r.get());   // stages it in HBM
        }
        L.flush();
        // completed items
        L.poll([&](void* tag, uint32_t kind, bool ovf) {
            Request& r = *(Request*)tag;
            if (kind & kPrepItem) {
                r.ready = true;
                return;
            }
            ++r.items_done;
            if (ovf) r.overflowed = true;
        });
        // (the kernel writes scores straight into the request's page-locked
        // copy: a request whose items are done is complete, with no copy
        // that would have to wait for the launch)
        for (auto& r : mine) {
            if (r->scanned || r->items_done < r->items.size()) continue;
            r->scanned = true;
            if (r->overflowed) drain = true;   // rescored once the launch has ended
            else finish_live(s, r);
        }
        // publish: shortest remaining work first, with aging
        if (!drain) {
            const double now = since(s);
            while (L.outstanding() < target && L.room() > 0) {
                Request* best = nullptr;
                double best_key = 0;
                for (auto& r : mine) {
                    if (!r->ready || r->next_item >= r->items.size()) continue;
                    const double key = r->rest[r->next_item] -
                                       s->opts.wait_weight * (now - r->info.submit_ms) * s->rate;
                    if (!best || key < best_key) {
                        best = r.get();
                        best_key = key;
                    }
                }
                if (!best) break;
                const uint32_t n = std::min<uint32_t>(
                    {(uint32_t)kLivePublishChunk, target - L.outstanding(), L.room(),
                     (uint32_t)(best->items.size() - best->next_item)});
                for (uint32_t k = 0; k < n; ++k) {
                    padded_done += best->rest[best->next_item] - best->rest[best->next_item + 1];
                    L.publish(best->task, best->items[best->next_item++], best);
                }
            }
            L.flush();
        }
        if (L.failed()) return fail_epoch(g_err);
        // the launch ends once nothing it holds is left and nothing can join
        if (L.outstanding() == 0) {
            bool more = false;
            for (auto& r : mine)
                if (r->next_item < r->items.size()) more = true;
            if (!drain && !more) {
                // (a waiting request that the budget keeps out waits for
                // bytes that completions free: not a reason to stay)
                std::lock_guard<std::mutex> lk(s->mu);
                for (auto& r : s->active)
                    if (r->task < 0 && !r->scanned) more = true;
                if (!s->stop && !s->waiting.empty() && !s->free_slots.empty() &&
                    (s->active.empty() ||
                     s->bytes_in_use + s->waiting.front()->bytes <= s->opts.memory_budget))
                    more = true;
            }
            if (drain || !more) break;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_work.wait_for(lk, kLivePoll);
    }
    if (L.close_and_wait(kLiveCloseWaitMs) != SWH_OK) return fail_epoch(g_err);
    // int32 rescore of what overflowed, then those requests' scores
    std::vector<int> tasks;
    for (auto& r : mine)
        if (r->overflowed) tasks.push_back(r->task);
    if (!tasks.empty()) {
        if (L.rescore(tasks.data(), (int)tasks.size()) != SWH_OK ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            s->loop_error = "int32 rescore failed on the device";
            return Round::kFail;
        }
        for (auto& r : mine) {
            if (!r->overflowed) continue;
            r->overflowed = false;
            if (r->scanned) finish_live(s, r);   // else the overflowed ones are exact now
        }
    }
    for (auto& r : mine) {
        r->task = -1;
        r->ready = false;
    }
    const double ms = now_ms() - t_epoch;
    if (ms > 0 && padded_done > 0)
        s->rate = std::min(2.0e10, std::max(1.0e8, 0.7 * s->rate + 0.3 * padded_done / ms));
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.busy_ms += ms;
    return Round::kDone;
}

void loop_main(swh_server* s) {
    for (;;) {
        const Round r = s->live ? run_epoch(s) : run_round(s);
        const bool ok = r != Round::kFail;
        std::lock_guard<std::mutex> lk(s->mu);
        if (!ok || (s->stop && s->active.empty())) {
            if (!ok)
                for (auto& r : s->active) {   // fail what is in flight
                    r->info.status = SWH_ERR_HIP;
                    s->to_complete.push_back(r);
                }
            s->active.clear();
            for (auto& r : s->waiting) {
                r->info.status = ok ? SWH_ERR_STATE : SWH_ERR_HIP;
                s->to_complete.push_back(r);
            }
            s->waiting.clear();
            s->stop = true;
            s->loop_done = true;
            s->cv_complete.notify_one();
            return;
        }
    }
}

// Top-k of finished requests from their pinned score copies, then the slot
// and its bytes go back to the pool.
void complete_main(swh_server* s) {
    const size_t nseq = s->db->nseq;
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_complete.wait(lk, [&] { return !s->to_complete.empty() || s->loop_done; });
            if (s->to_complete.empty()) return;   // the loop has ended and all is delivered
            r = s->to_complete.front();
            s->to_complete.pop_front();
        }
        if (r->info.status == SWH_OK && r->slot >= 0) {
            const int32_t* sc = (const int32_t*)s->h_scores.p + (size_t)r->slot * nseq;
            r->top_score.assign(r->k, 0);
            r->top_index.assign(r->k, 0);
            swh_topk(sc, nseq, r->k, r->top_score.data(), r->top_index.data());
            if (r->want_scores) r->scores.assign(sc, sc + nseq);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->slot >= 0) {
            s->free_slots.push_back(r->slot);
            s->bytes_in_use -= r->bytes;
            r->slot = -1;
        }
        r->info.done_ms = since(s);
        r->info.latency_ms = r->info.done_ms - r->info.submit_ms;
        r->done = true;
        s->stats.completed += 1;
        s->cv_done.notify_all();
        s->cv_work.notify_one();   // a slot came free
    }
}

}  // namespace

extern "C" {

int swh_server_create(swh_ctx* c, swh_db* db, const swh_server_opts* opts, swh_server** out) {
    if (!c || !db || !out) return fail(SWH_ERR_ARG, "NULL argument");
    if (db->ctx != c) return fail(SWH_ERR_ARG, "database belongs to another context");
    *out = nullptr;
    swh_server_opts o{};
    if (opts) o = *opts;
    if (!o.max_inflight) o.max_inflight = 64;
    if (!o.memory_budget) o.memory_budget = uint64_t(4) << 30;
    if (!o.max_qlen) o.max_qlen = 8192;
    if (!o.max_k) o.max_k = 100;
    if (o.round_ms == 0) o.round_ms = kDefaultRoundMs;
    if (o.wait_weight == 0) o.wait_weight = kDefaultWaitWeight;
    if (o.max_inflight < 1 || o.slices < 0 || o.max_qlen < 1 || o.max_k < 1 ||
        !(o.round_ms > 0) || !(o.wait_weight >= 0) || o.queue < SWH_QUEUE_AUTO ||
        o.queue > SWH_QUEUE_ROUNDS)
        return fail(SWH_ERR_ARG, "bad server options");
    // the live queue on a database of one layout, or when asked for (then on
    // the first layout); otherwise rounds: budgeted on a database of one
    // layout, whole rounds on a small database (two layouts, a whole round of
    // a few milliseconds)
    const bool live = o.queue == SWH_QUEUE_LIVE ||
                      (o.queue == SWH_QUEUE_AUTO && o.slices == 0 && !db->alt);
    const bool budgeted = !live && o.slices == 0 && !db->alt;
    if (o.slices == 0) o.slices = 1;
    int rc = check_opts(o.search);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    auto s = std::make_unique<swh_server>();
    // one server per context: its loop thread owns the context's stream and
    // launch buffers until swh_server_destroy
    swh_server* none = nullptr;
    if (!c->server.compare_exchange_strong(none, s.get()))
        return fail(SWH_ERR_STATE, "a server is already attached to this context");
    auto detach = [&] {
        c->server.store(nullptr);
        db->served.store(false);
    };
    if (db->served.exchange(true)) {
        c->server.store(nullptr);
        return fail(SWH_ERR_STATE, "another server searches this database");
    }
    s->ctx = c;
    s->db = db;
    s->opts = o;
    s->budgeted = budgeted;
    s->live = live;
    s->t0 = now_ms();
    s->pair_cut = cut_blocks(db, 0, db->n_pairs, o.slices);
    s->lane_cut = cut_blocks(db, db->n_pairs, db->n_blocks, o.slices);
    for (int i = 0; i < o.slices; ++i) {
        const size_t n = (size_t)(s->pair_cut[i + 1] - s->pair_cut[i]) * 2 +
                         (size_t)(s->lane_cut[i + 1] - s->lane_cut[i]) * kLaneSeqs;
        s->max_slice_seqs = std::max(s->max_slice_seqs, n);
    }
    if (budgeted || live)   // a slice (a task) may cover the whole first layout
        s->max_slice_seqs = (size_t)db->n_pairs * 2 +
                            (size_t)(db->n_blocks - db->n_pairs) * kLaneSeqs;
    if (o.slices == 1 && db->alt)   // a round may run on the second layout
        s->max_slice_seqs = std::max<size_t>(
            s->max_slice_seqs,
            (size_t)db->alt->n_pairs * 2 +
                (size_t)(db->alt->n_blocks - db->alt->n_pairs) * kLaneSeqs);
    s->max_slice_seqs = std::max<size_t>(s->max_slice_seqs, 1);
    // slots: as many as the budget holds for the longest admissible query
    const size_t per_slot = request_bytes(s.get(), o.max_qlen);
    const int slots = (int)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)o.max_inflight, o.memory_budget / per_slot));
    s->prof_slot_elems = (size_t)kProfRows * profile_stride(o.max_qlen);
    const size_t nseq = std::max<size_t>(db->nseq, 1);
    // (the live queue's kernel writes scores into the page-locked copies)
    hipError_t e = live ? hipSuccess : s->scores.reserve((size_t)slots * nseq * sizeof(int32_t));
    if (e == hipSuccess) e = s->ovf.reserve((size_t)slots * s->max_slice_seqs * sizeof(uint32_t));
    if (e == hipSuccess) e = s->prof.reserve((size_t)slots * s->prof_slot_elems * sizeof(int16_t));
    if (e == hipSuccess) e = s->counters.reserve((16 + (size_t)slots) * sizeof(uint32_t));
    if (e == hipSuccess) e = s->h_prof.reserve((size_t)slots * s->prof_slot_elems * sizeof(int16_t));
    if (e == hipSuccess) e = s->h_scores.reserve((size_t)slots * nseq * sizeof(int32_t));
    if (e != hipSuccess) {
        s->scores.release();
        s->ovf.release();
        s->prof.release();
        s->counters.release();
        s->h_prof.release();
        s->h_scores.release();
        detach();
        return fail(SWH_ERR_NOMEM, "server buffers: %s", hipGetErrorString(e));
    }
    if (live && (size_t)kProfRows * profile_stride(o.max_qlen) * sizeof(int16_t) >
                    LiveScan::kArenaBytes) {
        s->scores.release();
        s->ovf.release();
        s->prof.release();
        s->counters.release();
        s->h_prof.release();
        s->h_scores.release();
        detach();
        return fail(SWH_ERR_ARG, "max_qlen %d: its profile exceeds the live queue's %zu MB arena",
                    o.max_qlen, LiveScan::kArenaBytes >> 20);
    }
    if (live) {
        // one launch shape for every request: the one a 512-residue query
        // (within the caller's strip and group options) would run with
        const int rep = std::min(512, o.max_qlen);
        const Shape sh = choose_shape(c, &rep, 1, o.search, db, db->residues);
        s->lq = std::make_unique<LiveScan>();
        rc = s->lq->init(c, db, sh, o.search);
        if (rc == SWH_OK && hipHostGetDevicePointer((void**)&s->d_scores, s->h_scores.p, 0) !=
                                hipSuccess)
            rc = fail(SWH_ERR_HIP, "no device address for the score copies");
        if (rc != SWH_OK) {
            const std::string why = g_err;
            s->lq.reset();
            s->scores.release();
            s->ovf.release();
            s->prof.release();
            s->counters.release();
            s->h_prof.release();
            s->h_scores.release();
            detach();
            return fail(rc, "%s", why.c_str());
        }
    }
    for (int i = slots - 1; i >= 0; --i) s->free_slots.push_back(i);
    swh_server* raw = s.release();
    raw->loop = std::thread(loop_main, raw);
    raw->completer = std::thread(complete_main, raw);
    *out = raw;
    return SWH_OK;
}

int swh_server_submit(swh_server* s, const uint8_t* query, int qlen, int k, int want_scores,
                      uint64_t* id) {
    if (!s || !id) return fail(SWH_ERR_ARG, "NULL argument");
    int rc = check_query(query, qlen, 0);
    if (rc) return rc;
    if (qlen > s->opts.max_qlen)
        return fail(SWH_ERR_CAPACITY, "query length %d above the server's max_qlen %d", qlen,
                    s->opts.max_qlen);
    if (k < 0 || k > s->opts.max_k) return fail(SWH_ERR_ARG, "k = %d outside 0..%d", k, s->opts.max_k);
    auto r = std::make_shared<Request>();
    r->query.assign(query, query + qlen);
    r->k = k;
    r->want_scores = want_scores != 0;
    r->bytes = request_bytes(s, qlen);
    r->pair_pos = 0;
    r->lane_pos = s->db->n_pairs;
    if (r->bytes > s->opts.memory_budget)
        return fail(SWH_ERR_CAPACITY, "request needs %zu device bytes, above the server's "
                    "memory_budget %llu", r->bytes, (unsigned long long)s->opts.memory_budget);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(SWH_ERR_STATE, "server is stopping");
    r->id = s->next_id++;
    r->info.submit_ms = since(s);
    s->waiting.push_back(r);
    s->all[r->id] = r;
    *id = r->id;
    s->cv_work.notify_one();
    return SWH_OK;
}

int swh_server_wait(swh_server* s, uint64_t id, int timeout_ms, int32_t* top_score,
                    int64_t* top_index, int32_t* scores, swh_request_info* info) {
    if (!s) return fail(SWH_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(SWH_ERR_ARG, "unknown request %llu", (unsigned long long)id);
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->done; };
    s->waiters += 1;
    bool got = true;
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else got = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    s->waiters -= 1;
    if (s->waiters == 0) s->cv_done.notify_all();   // a destroy may wait for this
    if (!got)
        return fail(SWH_ERR_STATE, "request %llu not done after %d ms", (unsigned long long)id,
                    timeout_ms);
    s->all.erase(id);
    if (info) *info = r->info;
    if (r->info.status != SWH_OK)
        return fail(r->info.status, "request %llu failed: %s", (unsigned long long)id,
                    s->loop_error.empty() ? "server stopped" : s->loop_error.c_str());
    if (r->k > 0 && (!top_score || !top_index)) return fail(SWH_ERR_ARG, "top-k outputs are NULL");
    for (int t = 0; t < r->k; ++t) {
        top_score[t] = r->top_score[t];
        top_index[t] = r->top_index[t];
    }
    if (scores && r->want_scores) std::memcpy(scores, r->scores.data(), r->scores.size() * 4);
    return SWH_OK;
}

int swh_server_get_stats(swh_server* s, swh_server_stats* out) {
    if (!s || !out) return fail(SWH_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->waiting = (int)s->waiting.size();
    return SWH_OK;
}

int swh_server_destroy(swh_server* s) {
    if (!s) return SWH_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->loop.joinable()) s->loop.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->cv_complete.notify_all();
    }
    if (s->completer.joinable()) s->completer.join();
    {
        // every request is done now (admitted ones finished, waiting ones
        // failed with SWH_ERR_STATE): let threads blocked in swh_server_wait
        // return before the server goes away
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.notify_all();
        s->cv_done.wait(lk, [&] { return s->waiters == 0; });
    }
    (void)hipSetDevice(s->ctx->device);
    s->lq.reset();
    s->scores.release();
    s->ovf.release();
    s->prof.release();
    s->counters.release();
    s->h_prof.release();
    s->h_scores.release();
    s->db->served.store(false);
    s->ctx->server.store(nullptr);
    delete s;
    return SWH_OK;
}

}  // extern "C"
