// swh_live.cpp -- host side of the live-queue scan (see swh_live.h).
#include "swh_live.h"

#include <chrono>
#include <thread>

namespace swhh {

namespace {

constexpr size_t kDescBlock = kLiveDescStride * sizeof(QueryDesc);   // 256 bytes per task
constexpr size_t kProfAlign = 4096;
// A workgroup with no item leaves after this long (the host closes an idle
// launch long before; this only bounds a launch whose host side is gone).
constexpr double kIdleSeconds = 20.0;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

static_assert(sizeof(QueryDesc) * kLiveDescStride == 256, "task descriptor blocks are 256 bytes");

int LiveScan::init(swh_ctx* c, swh_db* db, const Shape& sh, const swh_search_opts& o) {
    release();
    c_ = c;
    db_ = db;
    sh_ = sh;
    o_ = o;
    const int exact_thr = std::min(65535 - c->max_score, kScoreLimitF16);
    thr_ = o.overflow_threshold > 0 ? std::min(o.overflow_threshold, exact_thr) : exact_thr;
    // strip borders: any task may sweep its query in several strips (lane
    // items) and in several passes (group items, at least min_lanes_ lanes a
    // group then, so at most 64 / min_lanes_ groups per wave keep a border)
    min_lanes_ = group_min_lanes(c, db->max_rows_pair, c->border_budget);
    const uint64_t stride_l = (uint64_t)db->max_rows_lane * kWave;
    const uint64_t stride_g =
        db->n_pairs > 0 ? (uint64_t)(kWave / min_lanes_) * db->max_rows_pair : 0;
    stride_ = std::max<uint64_t>({stride_l, stride_g, 1});
    // LDS: a whole profile segment (single-profile lanes) and a whole pass of
    // 64 lanes (group items)
    ps_l_ = kSegCols + kProfPad;
    ps_g_ = kWave * sh.Wg + kProfPad;
    lds_ = std::max(lane_lds_bytes(sh.Wl, sh.pair, ps_l_),
                    (size_t)kProfRows * ps_g_ * sizeof(int16_t));
    const uint64_t per_wg = stride_ * sizeof(uint2) * kWavesPerWG;
    grid_ = c->num_cus * scan_live_max_wg_per_cu(sh.Wl, sh.Wg, sh.pair, lds_);
    grid_ = (int)std::max<uint64_t>(1, std::min<uint64_t>(grid_, c->border_budget / per_wg));
    HIP_TRY(c->border.reserve(per_wg * (uint64_t)grid_));
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    idle_ticks_ = (uint64_t)(kIdleSeconds * 1000.0 * std::max(khz, 1));
    void* p = nullptr;
    HIP_TRY(hipHostMalloc(&p, 4096, hipHostMallocCoherent | hipHostMallocMapped));
    ctl_ = (uint64_t*)p;
    HIP_TRY(hipHostMalloc(&p, kRingCap * sizeof(uint2), hipHostMallocCoherent | hipHostMallocMapped));
    ring_ = (uint2*)p;
    HIP_TRY(hipHostMalloc(&p, kRingCap * sizeof(uint32_t),
                          hipHostMallocCoherent | hipHostMallocMapped));
    done_ = (uint32_t*)p;
    HIP_TRY(hipHostMalloc(&p, kMaxTasks * kDescBlock + kArenaBytes,
                          hipHostMallocNonCoherent | hipHostMallocMapped));
    arena_ = (char*)p;
    HIP_TRY(hipHostGetDevicePointer((void**)&d_ctl_, ctl_, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&d_ring_, ring_, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&d_done_, done_, 0));
    HIP_TRY(hipHostGetDevicePointer((void**)&d_arena_, arena_, 0));
    std::memset(ctl_, 0, 4096);
    std::memset(ring_, 0, kRingCap * sizeof(uint2));
    std::memset(done_, 0, kRingCap * sizeof(uint32_t));
    HIP_TRY(dev_arena_.reserve(kMaxTasks * kDescBlock + kArenaBytes));
    HIP_TRY(ctr_.reserve((2 + (size_t)kMaxTasks) * sizeof(uint32_t)));
    HIP_TRY(hipEventCreateWithFlags(&ev_end_, hipEventDisableTiming));
    tag_.assign(kRingCap, nullptr);
    kind_.assign(kRingCap, 0);
    consumed_.assign(kRingCap, 1);
    return SWH_OK;
}

void LiveScan::release() {
    if (running_) (void)close_and_wait(60000.0);
    if (ctl_) (void)hipHostFree(ctl_);
    if (ring_) (void)hipHostFree(ring_);
    if (done_) (void)hipHostFree(done_);
    if (arena_) (void)hipHostFree(arena_);
    ctl_ = d_ctl_ = nullptr;
    ring_ = d_ring_ = nullptr;
    done_ = d_done_ = nullptr;
    arena_ = d_arena_ = nullptr;
    ctr_.release();
    dev_arena_.release();
    if (ev_end_) (void)hipEventDestroy(ev_end_);
    ev_end_ = nullptr;
}

int LiveScan::group_lanes_of(int qlen) const {
    const uint64_t cells = (uint64_t)qlen * db_->residues;
    const int passes = group_passes(c_, db_, cells, qlen, sh_.Wg,
                                    group_pass_budget(c_, cells, db_->max_rows_pair, sh_.Wg),
                                    min_lanes_);
    int G = group_lanes(c_, qlen, sh_.Wg, o_, passes, min_lanes_);
    const int strips = (qlen + sh_.Wg - 1) / sh_.Wg;
    // a query swept in several passes keeps a border per group: no more
    // groups per wave than the launch's border stride holds
    if (G < strips && G < min_lanes_) G = std::min(min_lanes_, kWave);
    return G;
}

void LiveScan::plan_items(int qlen, int G, std::vector<uint2>& items,
                          std::vector<double>& cells) const {
    QueryDesc L{}, Gd{};
    L.qpad = ((qlen + 7) / 8) * 8;
    L.group = 1;
    L.blk_begin = db_->n_pairs;
    L.blk_end = db_->n_blocks;
    L.n_items = (db_->n_blocks - db_->n_pairs + kWavesPerWG - 1) / kWavesPerWG;
    Gd.qpad = ((qlen + sh_.Wg - 1) / sh_.Wg) * sh_.Wg;
    Gd.group = G;
    Gd.blk_begin = 0;
    Gd.blk_end = db_->n_pairs;
    const int per_item = kWavesPerWG * (kWave / G);
    Gd.n_items = (db_->n_pairs + per_item - 1) / per_item;
    items.resize((size_t)L.n_items + Gd.n_items);
    order_items(db_, sh_, &L, &Gd, 1, items.data());
    const int passes = (Gd.qpad / sh_.Wg + G - 1) / G;
    cells.resize(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        const uint2 it = items[i];
        if (it.x & kLaneItem) {
            const int b0 = L.blk_begin + (int)it.y * kWavesPerWG;
            const int b1 = std::min(b0 + kWavesPerWG, L.blk_end);
            cells[i] = (double)range_padded(db_, b0, b1) * L.qpad;
        } else {
            const int p0 = (int)it.y * per_item;
            const int p1 = std::min(p0 + per_item, Gd.blk_end);
            cells[i] = (double)range_padded(db_, p0, p1) * passes * G * sh_.Wg;
        }
    }
}

int LiveScan::begin() {
    if (running_) return fail(SWH_ERR_STATE, "live launch already running");
    ctl_[0] = 0;
    status()[0] = 0;
    published_ = completed_ = oldest_ = 0;
    std::fill(consumed_.begin(), consumed_.end(), 1);
    ntasks_ = 0;
    arena_used_ = kMaxTasks * kDescBlock;
    __atomic_thread_fence(__ATOMIC_SEQ_CST);
    HIP_TRY(launch_reset((uint32_t*)ctr_.p, 2 + kMaxTasks, c_->stream));
    ScanArgs a{};
    a.words = (const uint32_t*)db_->words.p;
    a.blocks = (const BlockDesc*)db_->d_blocks.p;
    a.perm = (const uint32_t*)db_->perm.p;
    a.lane_q = (const QueryDesc*)dev_arena_.p;
    a.group_q = a.lane_q + 1;
    a.desc_stride = kLiveDescStride;
    a.items = nullptr;
    a.work_counter = (uint32_t*)ctr_.p;
    a.border = (uint2*)c_->border.p;
    a.border_stride = stride_;
    a.border_rows = db_->max_rows_pair;
    a.n_queries = kMaxTasks;
    a.n_items = 0;
    a.ps_lane = ps_l_;
    a.ps_group = ps_g_;
    a.gap_oe = c_->gap_open + c_->gap_extend;
    a.gap_e = c_->gap_extend;
    a.live.ctl = d_ctl_;
    a.live.ring = d_ring_;
    a.live.done = d_done_;
    a.live.status = (uint32_t*)(d_ctl_ + 8);
    a.live.cap = kRingCap;
    a.live.idle_ticks = idle_ticks_;
    a.live.host_arena = d_arena_;
    a.live.dev_arena = (char*)dev_arena_.p;
    HIP_TRY(launch_scan_live(sh_.Wl, sh_.Wg, sh_.pair, a, grid_, lds_, c_->stream));
    HIP_TRY(hipEventRecord(ev_end_, c_->stream));
    t_begin_ = now_ms();
    running_ = true;
    return SWH_OK;
}

int LiveScan::add_task(const LiveTask& t) {
    const int ps = profile_stride(t.qlen);
    const size_t pbytes = (size_t)kProfRows * ps * sizeof(int16_t);
    if (ntasks_ >= kMaxTasks || arena_used_ + pbytes > kMaxTasks * kDescBlock + kArenaBytes)
        return -1;
    const int q = ntasks_++;
    const size_t off = arena_used_;
    arena_used_ = align_up(arena_used_ + pbytes, kProfAlign);
    std::memcpy(arena_ + off, t.prof, pbytes);
    QueryDesc base{};
    base.prof = (const int16_t*)((char*)dev_arena_.p + off);
    base.scores = t.scores;
    base.ovf_list = t.ovf;
    base.ovf_count = (uint32_t*)ctr_.p + 2 + q;
    base.qlen = t.qlen;
    base.pstride = ps;
    base.ovf_threshold = thr_;
    QueryDesc* d = (QueryDesc*)(arena_ + (size_t)q * kDescBlock);
    d[0] = base;   // lane items
    d[0].qpad = ((t.qlen + 7) / 8) * 8;
    d[0].group = 1;
    d[0].blk_begin = db_->n_pairs;
    d[0].blk_end = db_->n_blocks;
    d[1] = base;   // group items
    d[1].qpad = ((t.qlen + sh_.Wg - 1) / sh_.Wg) * sh_.Wg;
    d[1].group = t.group;
    d[1].blk_begin = 0;
    d[1].blk_end = db_->n_pairs;
    d[2] = base;   // the int32 rescore
    d[2].qpad = ((t.qlen + kRescoreWidth - 1) / kRescoreWidth) * kRescoreWidth;
    d[2].group = std::min(kWave, d[2].qpad / kRescoreWidth);
    d[3] = QueryDesc{};
    return q;
}

void LiveScan::publish(int task, uint2 item, void* tag) {
    const uint32_t s = published_ & (kRingCap - 1);
    tag_[s] = tag;
    kind_[s] = item.x;
    consumed_[s] = 0;
    done_[s] = 0;
    ring_[s] = make_uint2(item.x | (uint32_t)task, item.y);
    ++published_;
}

void LiveScan::flush() {
    // the ring entries and cleared done flags before the count that shows them
    __atomic_store_n(ctl_, (uint64_t)published_, __ATOMIC_RELEASE);
}

bool LiveScan::failed() {
    if (!running_) return false;
    const uint32_t st = __atomic_load_n(status(), __ATOMIC_ACQUIRE);
    const hipError_t e = hipEventQuery(ev_end_);
    if (st == 0 && e == hipErrorNotReady) return false;
    fail(SWH_ERR_HIP, "live launch ended early after %.1f ms (%s; a workgroup %s its idle limit; "
         "%u of %u items done, %d tasks)", now_ms() - t_begin_,
         e == hipSuccess ? "the kernel ended" : hipGetErrorString(e),
         st ? "reached" : "did not reach", completed_, published_, ntasks_);
    return true;
}

int LiveScan::close_and_wait(double timeout_ms) {
    if (!running_) return SWH_OK;
    __atomic_store_n(ctl_, (uint64_t)published_ | (uint64_t(1) << 32), __ATOMIC_RELEASE);
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t e;
    while ((e = hipEventQuery(ev_end_)) == hipErrorNotReady) {
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0).count();
        if (ms > timeout_ms)
            return fail(SWH_ERR_HIP, "live launch did not end %.0f ms after close", ms);
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    running_ = false;
    if (e != hipSuccess)
        return fail(SWH_ERR_HIP, "live launch failed: %s", hipGetErrorString(e));
    if (status()[0] != 0)
        return fail(SWH_ERR_HIP, "live launch: a workgroup waited %.0f s for an item", kIdleSeconds);
    return SWH_OK;
}

int LiveScan::rescore(const int* tasks, int n) {
    if (n <= 0) return SWH_OK;
    if (running_) return fail(SWH_ERR_STATE, "rescore before the live launch ended");
    std::vector<QueryDesc> rd((size_t)n);
    int max_pass_r = 0, max_gpw_r = 0;
    for (int i = 0; i < n; ++i) {
        rd[(size_t)i] = ((const QueryDesc*)(arena_ + (size_t)tasks[i] * kDescBlock))[2];
        const QueryDesc& R = rd[(size_t)i];
        max_pass_r = std::max(max_pass_r, R.group * kRescoreWidth);
        if (R.qpad > R.group * kRescoreWidth) max_gpw_r = std::max(max_gpw_r, kWave / R.group);
    }
    const size_t bytes = (size_t)n * sizeof(QueryDesc);
    HIP_TRY(c_->h_desc.reserve(bytes));
    HIP_TRY(c_->desc.reserve(bytes));
    std::memcpy(c_->h_desc.p, rd.data(), bytes);
    HIP_TRY(hipMemcpyAsync(c_->desc.p, c_->h_desc.p, bytes, hipMemcpyHostToDevice, c_->stream));
    RescoreLaunch rl = plan_rescore(c_, db_, max_pass_r, max_gpw_r, c_->border_budget,
                                    (const QueryDesc*)c_->desc.p, n, (uint32_t*)ctr_.p + 1);
    HIP_TRY(c_->border.reserve(std::max<uint64_t>(16, rl.per_wg * (uint64_t)rl.grid)));
    rl.args.border = (uint2*)c_->border.p;
    HIP_TRY(launch_reset((uint32_t*)ctr_.p + 1, 1, c_->stream));
    HIP_TRY(launch_rescore(rl.args, rl.grid, rl.lds, c_->stream));
    return SWH_OK;
}

}  // namespace swhh
