// swh_host.h -- private host-side definitions shared by the sources of
// libswh.so (contexts, the resident database, launch planning).  Not part of
// the public ABI (include/swh.h).
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "swh.h"
#include "swh_internal.h"

namespace swhh {

using namespace swh;

extern thread_local std::string g_err;

inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return ::swhh::fail(SWH_ERR_HIP, "%s failed: %s (%s:%d)", #expr,               \
                                hipGetErrorString(_e), __FILE__, __LINE__);                \
    } while (0)

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

template <typename F>
inline void parallel_for(int64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> ts;
    int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        ts.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : ts) t.join();
}

// Device buffer that only grows.  A fixed buffer is a view of memory the
// buffer does not own (a part of the static scratch array): it never
// allocates, a larger request fails, and release() only forgets it.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool fixed = false;
    // Grows to want, or to 1.5x the old size when that is more (hipFree waits
    // for the device to be idle).
    hipError_t reserve_grow(size_t want) {
        if (want <= bytes) return hipSuccess;
        return reserve(std::max(want, bytes + bytes / 2));
    }
    static DevBuf view(void* at, size_t n) {
        DevBuf b;
        b.p = at;
        b.bytes = n;
        b.fixed = true;
        return b;
    }
    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (fixed) return hipErrorOutOfMemory;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p && !fixed) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// Device buffers of destroyed databases, kept by their context for the next
// database it builds.  hipFree waits for the whole device (an implicit device
// synchronisation), so freeing them would hold a database build on one
// thread behind a search running on another: the worker pipeline builds chunk
// k + 1 while chunk k is searched, and destroys chunk k - 1's database first.
struct BufCache {
    static constexpr size_t kMaxBytes = size_t(8) << 30;
    static constexpr size_t kMaxBuffers = 32;
    std::mutex mu;
    std::vector<DevBuf> spare;   // oldest first
    size_t total = 0;
    // d gets at least want bytes: its own buffer, the smallest spare one that
    // is not much larger, or a new allocation.
    hipError_t take(DevBuf& d, size_t want) {
        if (want <= d.bytes) return hipSuccess;
        {
            std::lock_guard<std::mutex> g(mu);
            int best = -1;
            for (size_t i = 0; i < spare.size(); ++i) {
                const size_t b = spare[i].bytes;
                if (b >= want && b <= 2 * want + (size_t(64) << 20) &&
                    (best < 0 || b < spare[(size_t)best].bytes))
                    best = (int)i;
            }
            if (best >= 0) {
                DevBuf got = spare[(size_t)best];
                spare.erase(spare.begin() + best);
                total -= got.bytes;
                if (d.p) {   // d's own, too small: back to the spares
                    spare.push_back(d);
                    total += d.bytes;
                }
                d = got;
                trim();
                return hipSuccess;
            }
        }
        return d.reserve(want);
    }
    void give(DevBuf& d) {
        if (!d.p) return;
        std::lock_guard<std::mutex> g(mu);
        spare.push_back(d);
        total += d.bytes;
        d.p = nullptr;
        d.bytes = 0;
        trim();
    }
    void release() {
        std::lock_guard<std::mutex> g(mu);
        for (auto& b : spare) b.release();
        spare.clear();
        total = 0;
    }

  private:
    void trim() {   // caller holds mu
        while (!spare.empty() && (spare.size() > kMaxBuffers || total > kMaxBytes)) {
            total -= spare.front().bytes;
            spare.front().release();
            spare.erase(spare.begin());
        }
    }
};

struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool fixed = false;   // allocated once at set-up: a larger request fails
    // Grows to want, or to 1.5x the old size when that is more: freeing
    // page-locked memory waits for the device to be idle, so buffers refilled
    // beside a running search (the next database's tables) should rarely grow.
    hipError_t reserve_grow(size_t want) {
        if (want <= bytes) return hipSuccess;
        return reserve(std::max(want, bytes + bytes / 2));
    }
    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (fixed) return hipErrorOutOfMemory;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

inline int check_scoring(const int32_t* matrix, int gap_open, int gap_extend) {
    if (!matrix) return fail(SWH_ERR_ARG, "matrix is NULL");
    for (int i = 0; i < 1024; ++i)
        if (matrix[i] < -128 || matrix[i] > 127)
            return fail(SWH_ERR_ARG, "matrix entry %d = %d outside [-128, 127]", i, matrix[i]);
    if (gap_open < 0 || gap_extend < 0 || gap_open + gap_extend > 32767)
        return fail(SWH_ERR_ARG, "invalid gap penalties open=%d extend=%d", gap_open,
                    gap_extend);
    return SWH_OK;
}


}  // namespace swhh


using swhh::BufCache;
using swhh::DevBuf;
using swhh::PinnedBuf;
using swh::BlockDesc;

// A captured launch sequence and what it was captured for.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    std::vector<long long> key;
    swh_stats st{};
    void* desc = nullptr;
    void* border = nullptr;
    void* h_desc = nullptr;
    std::vector<char> desc_bytes;   // the pinned descriptor tables the graph uploads
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
};

struct swh_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;     // every launch and copy of a search
    hipStream_t upload = nullptr;     // database packing (may overlap a search)
    int32_t matrix[1024];
    int gap_open = 10, gap_extend = 2, max_score = 0;
    DevBuf border, counters, prof, scores, ovf, desc;
    PinnedBuf h_prof, h_scores, h_desc, h_ctr;
    size_t desc_used = 0;             // bytes of h_desc the last enqueue uploaded
    hipEvent_t ev[8] = {};
    size_t border_budget = size_t(8) << 30;   // bytes of HBM for strip borders
    GraphCache graph;
    BufCache cache;                     // buffers of destroyed databases
    std::mutex stage_mu;                // one database upload at a time
    DevBuf stage_raw, stage_soff;       // its staging buffers, reused
    PinnedBuf h_stage;                  // ... and the page-locked staging of its tables
    // Buffers that a build or a search outgrew (grow_beside).  Freeing one
    // waits for the whole device (hipFree and hipHostFree synchronise it),
    // which would hold a build behind a search running on another thread, or
    // a search behind a build: the worker pipeline overlaps the two.  They
    // are kept until the context is destroyed, up to kRetiredMax bytes.
    std::mutex retired_mu;
    std::vector<DevBuf> retired_dev;
    std::vector<PinnedBuf> retired_host;
    size_t retired_bytes = 0;
    static constexpr size_t kRetiredMax = size_t(4) << 30;
    std::vector<swh_tune_entry> tune;   // empty: built-in shape choice
    uint64_t tune_gen = 0;              // bumped whenever the table changes
    // A server's loop thread launches on this context's stream with its
    // descriptor, border and counter buffers; while one is attached every
    // other call that would use them is refused (SWH_ERR_STATE).
    std::atomic<swh_server*> server{nullptr};
    std::atomic<int> live_dbs{0};       // databases (layouts) made on it and not destroyed
    // The allocation-free search (swh_search_noalloc): this device's static
    // scratch array and a page-locked descriptor staging area of fixed size,
    // both set up by swh_ctx_create.
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    PinnedBuf na_h_desc;
};

struct swh_db {
    swh_ctx* ctx = nullptr;
    uint64_t uid = 0;           // unique per database ever created in the process
    std::vector<uint64_t> blk_real;     // prefix sums over blocks of real residues
    std::vector<uint64_t> blk_padded;   // ... and of padded residues (rows x sequences)
    uint64_t nseq = 0, residues = 0, padded = 0;
    uint64_t dev_bytes = 0;     // device bytes the layout uses (its buffers may be larger)
    int n_pairs = 0;            // group-mode pair blocks, blocks [0, n_pairs)
    int n_blocks = 0;           // all blocks; lane blocks are [n_pairs, n_blocks)
    uint32_t max_rows_lane = 0, max_rows_pair = 0;
    uint32_t group_min_len = 0;   // sequences at least this long are in pair blocks
    swh_db* alt = nullptr;        // small databases: the same sequences with more pairs
    uint64_t padded_pair = 0, padded_lane = 0;   // padded residues per layout
    uint64_t residues_pair = 0;                  // real residues in pair blocks
    std::vector<BlockDesc> blocks;
    DevBuf words, d_blocks, perm, sorted_len, sorted_blk;
    std::atomic<bool> served{false};   // a server searches it: not destroyable
};


namespace swhh {

// One query of a launch, with device buffers that outlive the launch.
struct Task {
    const int16_t* prof;        // device profile [24][profile_stride(qlen)]
    int qlen;
    int32_t* scores;            // device, caller order (nseq entries)
    uint32_t* ovf;              // device overflow list, >= sequences in the ranges
    int pair_begin, pair_end;   // pair blocks scanned by this task
    int lane_begin, lane_end;   // lane blocks scanned by this task
};

struct Shape {
    int Wl, Wg;
    bool pair;   // lane scan reads the residue-pair table
};

// Lane-scan configurations of a tuning table's lane_cost columns.
struct LaneCfg {
    int width;
    bool pair;
};
constexpr LaneCfg kLaneCfgs[SWH_TUNE_LANE_CONFIGS] = {
    {16, true}, {24, true}, {16, false}, {24, false}, {32, false}};


int profile_stride(int qlen);
void build_profile(const int32_t* matrix, const uint8_t* q, int qlen, int16_t* out);
uint64_t range_residues(const swh_db* db, int b0, int b1);
uint64_t range_padded(const swh_db* db, int b0, int b1);
int check_opts(const swh_search_opts& o);
int check_query(const uint8_t* q, int qlen, int i);
// db_residues: the residues each query of the launch searches in db.
Shape choose_shape(const swh_ctx* c, const int* qlens, int nq, const swh_search_opts& o,
                   const swh_db* db, uint64_t db_residues);
// Lanes per group in group mode within max_passes passes over the query
// (1: the fewest passes), at least min_lanes.
int group_lanes_for(int qlen, int W, int max_passes, int min_lanes);
// Cost-weighted lane columns of a query at strip width W.
double lane_cols(int qlen, int W);
// The layout of db (itself or its small-database second layout) that a launch
// of nq queries over the whole database runs on.
swh_db* pick_layout(const swh_ctx* c, swh_db* db, int nq);

// What one launch sequence owns: its stream, descriptor staging and strip
// borders.  Two of them let two launch sequences be in flight at once.
struct LaunchRes {
    hipStream_t stream;
    DevBuf* desc;
    PinnedBuf* h_desc;
    DevBuf* border;
    size_t* desc_used;   // bytes of h_desc the last enqueue uploaded
    uint64_t border_budget;   // bytes the strip borders may take (the grid shrinks to fit)
};
inline LaunchRes ctx_res(swh_ctx* c) {
    return LaunchRes{c->stream, &c->desc, &c->h_desc, &c->border, &c->desc_used,
                     c->border_budget};
}
// SWH_ERR_STATE (with a message naming what) while a server is attached to c.
int refuse_if_served(const swh_ctx* c, const char* what);
// Grows b to at least want (1.5 times its old size when that is more, unless
// exact; a fixed buffer cannot grow) and keeps its old allocation in c's
// retired list instead of freeing it now.
hipError_t grow_beside(swh_ctx* c, DevBuf& b, size_t want, bool exact = false);
hipError_t grow_beside(swh_ctx* c, PinnedBuf& b, size_t want, bool exact = false);
// Frees c's retired buffers (waits for the device).
void free_retired(swh_ctx* c);

int enqueue_tasks(swh_ctx* c, swh_db* db, const Task* t, int nt, const Shape& sh,
                  const swh_search_opts& o, uint32_t* counters, const LaunchRes& lr,
                  swh_stats* st);
// Work items of nt tasks (lane descriptors lane_d[i], group descriptors
// group_d[i], n_items set) in launch order, longest first: x = task (|
// kLaneItem for a lane item), y = item of the task.
void order_items(const swh_db* db, const Shape& sh, const QueryDesc* lane_d,
                 const QueryDesc* group_d, int nt, uint2* out);
// Group-mode shape helpers of the launch planning (swh_capi.cpp).
int group_min_lanes(const swh_ctx* c, uint32_t rows_max, uint64_t border_budget);
int group_pass_budget(const swh_ctx* c, uint64_t launch_cells, uint32_t rows_max, int W);
int group_passes(const swh_ctx* c, const swh_db* db, uint64_t launch_cells, int qlen, int W,
                 int max_passes, int min_lanes);
int group_lanes(const swh_ctx* c, int qlen, int Wg, const swh_search_opts& o, int max_passes,
                int min_lanes);
// An int32 rescore launch and the strip-border bytes it needs.
struct RescoreLaunch {
    RescoreArgs args;
    int grid;
    size_t lds;
    uint64_t per_wg;   // border bytes per workgroup
};
RescoreLaunch plan_rescore(const swh_ctx* c, const swh_db* db, int max_pass_r, int max_gpw_r,
                           uint64_t border_budget, const QueryDesc* d_queries, int nq,
                           uint32_t* work_counter);
// Search every query over the whole database; scores_host gets nq rows of nseq.
int run_queries(swh_ctx* c, swh_db* db, const uint8_t* const* qs, const int* qlens, int nq,
                const swh_search_opts& o, int32_t* scores_host, swh_stats* stats);

}  // namespace swhh
