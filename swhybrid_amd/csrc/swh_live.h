// swh_live.h -- host side of the live-queue scan (k_scan<..., LIVE = true>):
// one persistent launch that the serving loop keeps feeding with work items
// while it runs (LiveQueue in swh_internal.h describes the device side).
//
// A launch (an epoch) runs until close(): tasks (one query over every block of
// the database's first layout) are added to it at any time, their items are
// published one ticket at a time, and completed tickets are polled from
// page-locked memory, so a task whose items are done is complete while the
// kernel scans the others.  Tasks' descriptors and profiles are written by the
// host into a page-locked arena and staged into a device arena by each task's
// prep item, so scans read them from HBM, not across PCIe (every strip's pair
// table reads the profile again); each launch fills both arenas from the
// start, so no cache line the kernel has read can be stale.  Nothing the
// kernel waits for depends on a copy engine or another launch: it only polls
// host memory.
#pragma once
#include <vector>

#include "swh_host.h"

namespace swhh {

struct LiveTask {
    const int16_t* prof;   // host profile [24][profile_stride(qlen)]
    int qlen;
    int group;             // lanes per group of its group items
    int32_t* scores;       // device, caller order
    uint32_t* ovf;         // device overflow list, >= sequences of the layout
};

class LiveScan {
  public:
    static constexpr uint32_t kRingCap = 1u << 16;   // tickets in flight at most
    static constexpr int kMaxTasks = 4096;            // tasks per launch
    static constexpr size_t kArenaBytes = size_t(64) << 20;   // profiles per launch

    // Buffers and the launch shape for db; c's stream and strip borders are
    // the caller's for the life of this object.
    int init(swh_ctx* c, swh_db* db, const Shape& sh, const swh_search_opts& o);
    void release();
    ~LiveScan() { release(); }

    const Shape& shape() const { return sh_; }
    int grid() const { return grid_; }
    // Lanes per group of a query's group items in this launch shape.
    int group_lanes_of(int qlen) const;
    // A query's items in launch order, longest first (x: kLaneItem for a
    // lane item, else 0; y: item of the task), and the padded cells of each.
    void plan_items(int qlen, int group, std::vector<uint2>& items,
                    std::vector<double>& cells) const;

    // One launch: zero the counters, reset the queue, start the kernel.
    int begin();
    // A task's descriptors and profile, written into this launch's arena; the
    // task index, or -1 when the arena or the task table is full.
    int add_task(const LiveTask& t);
    // Tickets that may be published now (ring slots whose last ticket was polled).
    uint32_t room() const { return kRingCap - (published_ - oldest_); }
    // Queue one item of task (not visible to the kernel before flush()).  A
    // task's first item must be its prep item (kPrepItem); its other items
    // may follow once that one is done.
    void publish(int task, uint2 item, void* tag);
    void flush();
    // f(tag, item.x, overflowed) for every ticket completed since the last call.
    template <typename F>
    void poll(F&& f);
    uint32_t outstanding() const { return published_ - completed_; }
    uint32_t published() const { return published_; }
    // The launch ended before close(), or a workgroup gave up waiting
    // (g_err says which).
    bool failed();
    // Close the queue and wait (at most timeout_ms) for the launch to end.
    int close_and_wait(double timeout_ms);
    // After close_and_wait: the int32 rescore of tasks' overflow lists,
    // enqueued on the stream.
    int rescore(const int* tasks, int n);
    hipStream_t stream() const { return c_->stream; }

  private:
    swh_ctx* c_ = nullptr;
    swh_db* db_ = nullptr;
    Shape sh_{};
    swh_search_opts o_{};
    int grid_ = 0;
    size_t lds_ = 0;
    uint64_t stride_ = 0;
    int ps_l_ = 0, ps_g_ = 0;
    int min_lanes_ = 1;
    int thr_ = 0;
    uint64_t idle_ticks_ = 0;
    // host-coherent control block, ring and done flags
    uint64_t* ctl_ = nullptr;    // [0] published | closed << 32, [8] status
    uint2* ring_ = nullptr;
    uint32_t* done_ = nullptr;
    uint64_t* d_ctl_ = nullptr;   // ... and their device addresses
    uint2* d_ring_ = nullptr;
    uint32_t* d_done_ = nullptr;
    // page-locked arena: task descriptor blocks, then profiles; prep items
    // copy each task's part to the same offsets of the device arena
    char* arena_ = nullptr;
    char* d_arena_ = nullptr;   // its device address
    DevBuf dev_arena_;
    size_t arena_used_ = 0;
    DevBuf ctr_;                 // [0] tickets, [1] rescore items, [2 + q] task q's overflow count
    hipEvent_t ev_end_ = nullptr;
    double t_begin_ = 0;
    bool running_ = false;
    int ntasks_ = 0;
    uint32_t published_ = 0, completed_ = 0, oldest_ = 0;
    uint32_t* status() const { return (uint32_t*)(ctl_ + 8); }
    std::vector<void*> tag_;
    std::vector<uint32_t> kind_;   // item.x of each ring slot's ticket
    std::vector<uint8_t> consumed_;
};

template <typename F>
void LiveScan::poll(F&& f) {
    for (uint32_t t = oldest_; t != published_; ++t) {
        const uint32_t s = t & (kRingCap - 1);
        if (consumed_[s]) continue;
        const uint32_t v = __atomic_load_n(done_ + s, __ATOMIC_ACQUIRE);
        if ((v >> 1) != t + 1) continue;
        consumed_[s] = 1;
        ++completed_;
        f(tag_[s], kind_[s], (v & 1u) != 0);
    }
    while (oldest_ != published_ && consumed_[oldest_ & (kRingCap - 1)]) ++oldest_;
}

}  // namespace swhh
