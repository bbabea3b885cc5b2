// swh_internal.h -- shared between the gfx950 kernels and the host side of
// libswh.so.  Not part of the public ABI.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace swh {

constexpr int kWave = 64;            // wavefront width on CDNA4
constexpr int kWavesPerWG = 4;       // 256-thread workgroups
constexpr int kLaneSeqs = 2 * kWave; // sequences per lane-mode block (lo/hi halves)
constexpr int kProfRows = 24;        // residue codes 0..23 (23 = pad)
// Database rows per group-mode step of k_scan<WL, WG>: two where the
// kernel's register budget holds a second row's profile (256 VGPRs at
// WL >= 24; at WL = 16 the 168 VGPRs of three workgroups per CU hold it beside
// 8-column groups only).  With two, the rows' dependent chains interleave.
constexpr int group_step_rows(int wl, int wg) { return (wl == 16 && wg == 16) ? 1 : 2; }
constexpr int kSegCols = 1344;       // query columns per LDS profile segment (multiple of 16, 24 and 32;
                                     // 24 rows x (1344+8) int16 = 64,896 B of LDS)
constexpr int kProfPad = 8;          // extra int16 per LDS row: rows start 16 B apart mod 128 B
constexpr int kRescoreWidth = 8;    // int32 rescore: query columns per lane and pass
constexpr uint32_t kNoSeq = 0xffffffffu;
// The int16 scan is exact for scores below this (its three-way max reads the
// biased patterns as fp16 values, which order only below 0x7C00); sequences
// that reach it are rescored in int32.
constexpr int kScoreLimitF16 = 31744;

// Per-block descriptor of the packed database.  Lane-mode blocks hold 128
// sequences: lane l carries sorted sequences first+l (low half) and
// first+64+l (high half); word layout ((row/4)*64 + lane)*2 + half, 4 residues
// per 32-bit word.  Group-mode blocks hold 2 sequences (low/high half) that a
// whole wavefront works on; word layout (row/4)*2 + half.
struct BlockDesc {
    uint64_t word_off;   // first 32-bit word of the block
    uint32_t rows;       // padded length, multiple of 4
    uint32_t first;      // sorted position of the block's first sequence
};

// One query of a launch, as one kernel sees it.  Kernels take a table of
// these, so one persistent launch serves many queries through one device work
// queue; query q owns n_items items of its table's kind.
struct QueryDesc {
    const int16_t* prof;     // profile [24][pstride] int16, zero beyond qlen
    int32_t* scores;         // caller order, one per database sequence
    uint32_t* ovf_list;      // sorted positions sent to the int32 rescore
    uint32_t* ovf_count;
    int qlen;
    int qpad;                // qlen rounded up to this kernel's strip width
    int pstride;             // profile row stride (int16), >= qpad
    int group;               // group kernel: lanes per group (1..64)
    int n_items;
    int ovf_threshold;       // score >= threshold -> int32 rescore
    int blk_begin, blk_end;  // blocks of this query (pair blocks in the group
                             // table, lane blocks in the lane table)
};

// Live queue of the serving scan (k_scan with LIVE = true): the host appends
// work items while the kernel runs.  A workgroup draws ticket t from the work
// counter; its item is ring[t & (cap - 1)] once the low half of ctl (items
// published so far) exceeds t.  Until then it polls ctl, one 64-bit read with
// sleeps growing to about 100 us between reads (512 workgroups polling page-
// locked memory every few microseconds took the PCIe bandwidth the working
// ones needed), and it leaves without an item once the high half (closed) is
// set and t is still unpublished, or after idle_ticks of the 100 MHz
// real-time counter, when it also sets status[0].
// After an item the workgroup writes its results back to memory (a
// system-scope release) and stores ((t + 1) << 1 | overflowed) into
// done[t & (cap - 1)], so the host sees a task's scores while the kernel runs
// on.  ctl, ring, done and status are host-coherent page-locked memory,
// accessed with system-scope atomics.  A task's descriptor block and profile
// are written by the host into a page-locked arena; the task's first item (a
// prep item) copies them to the same offsets of a device arena, and the host
// publishes the task's scan items once that item is done, so scans read them
// from HBM.  Each launch fills both arenas from the start: nothing the kernel
// reads there was read earlier in the launch (no stale cache line holds it).
struct LiveQueue {
    const uint64_t* ctl;       // published | closed << 32
    const uint2* ring;
    uint32_t* done;
    uint32_t* status;
    uint32_t cap;              // a power of two
    uint64_t idle_ticks;
    const char* host_arena;    // device address of the page-locked arena
    char* dev_arena;           // the device arena (same layout)
};
// QueryDescs per task in a live launch: lane, group and rescore descriptor
// and one unused, 256 bytes, so no two tasks share a cache line.
constexpr int kLiveDescStride = 4;

struct ScanArgs {
    const uint32_t* words;     // packed residues
    const BlockDesc* blocks;   // block table: pair blocks, then lane blocks
    const uint32_t* perm;      // sorted position -> caller index (kNoSeq for fillers)
    const QueryDesc* group_q;  // per query: group-item range, G, qpad for the group width
    const QueryDesc* lane_q;   // per query: lane-item range, qpad for the lane width
    int desc_stride;           // QueryDescs from one query's descriptor to the next's
    const uint2* items;        // work queue order: x = query | kLaneItem, y = item of the query
    uint32_t* work_counter;
    uint2* border;             // per wave slot, border_stride uint2
    uint64_t border_stride;
    uint32_t border_rows;      // group items: rows reserved per group (longest pair)
    int n_queries;
    int n_items;
    int ps_lane, ps_group;     // LDS profile row strides (int16)
    int gap_oe, gap_e;
    LiveQueue live;            // k_scan<..., LIVE = true> only
};

// Work items of the search kernel: group items (G lanes cooperate on one long
// sequence pair; lane g of a group holds query columns [pass*G*W + g*W, +W)
// and works on row t-g at step t, its last column's H/F reaching lane g+1 one
// step later through a wave-wide DPP shift; a wave holds floor(64/G) groups,
// an item is 4 waves of pairs) and lane items (4 consecutive lane blocks of
// one query, one per wave), handed out longest first across every query of
// the launch.  Rescore items are 256 overflow-list entries.
constexpr uint32_t kLaneItem = 0x80000000u;
constexpr uint32_t kPrepItem = 0x40000000u;   // live queue: stage a task's descriptors and profile

struct RescoreArgs {
    const uint32_t* words;
    const BlockDesc* blocks;
    const uint32_t* perm;
    const uint32_t* sorted_len;   // real length per sorted position
    const uint32_t* sorted_blk;   // block of each sorted position
    const QueryDesc* queries;     // qpad = qlen rounded to 8, group = lanes per sequence;
                                  // items from the overflow counts
    uint32_t* work_counter;
    uint2* border;                // per wave slot, border_stride uint2
    uint64_t border_stride;
    uint32_t border_rows;         // rows reserved per lane group (longest sequence)
    int n_queries;
    int n_lane_blocks_first;      // index of the first lane-mode block
    int ps;
    int gap_oe, gap_e;
};

// Device scratch of the allocation-free search entry point: one static
// array per device (kScratchBytes), whose address is looked up once per
// device with that device current.  The descriptor tables of such a search
// are staged through kNoAllocDescBytes of page-locked memory that the
// context allocates when it is created.
constexpr size_t kScratchBytes = size_t(256) << 20;
constexpr size_t kNoAllocDescBytes = size_t(1) << 20;
constexpr int kNoAllocMaxQlen = 8192;
hipError_t scratch_address(int device, void** out, size_t* bytes);

// Launchers (swh_kernels.hip).
hipError_t launch_reset(uint32_t* p, int n, hipStream_t s);
// [24][pstride] int16 profile s(q_j, r) (zero beyond qlen) of a query held
// in device memory, from a 32x32 int32 matrix in device memory.
hipError_t launch_build_profile(const int32_t* matrix, const uint8_t* query, int qlen,
                                int pstride, int16_t* out, hipStream_t s);
// Interleaves raw residues (sequence id's codes at raw + soff[id]) into the
// block layout; one workgroup per block.
hipError_t launch_pack(const BlockDesc* blocks, int n_blocks, int n_pairs, const uint32_t* perm,
                       const uint32_t* sorted_len, const uint8_t* raw, const uint64_t* soff,
                       uint32_t* words, hipStream_t s);
// pair: the lane scan reads a per-strip residue-pair table (widths 16, 24).
hipError_t launch_scan(int lane_width, int group_width, bool pair, const ScanArgs& a, int grid,
                       size_t lds_bytes, hipStream_t s);
int scan_max_wg_per_cu(int lane_width, int group_width, bool pair, size_t lds_bytes);
// The same scan fed through a.live (swh_kernels_live.hip).
hipError_t launch_scan_live(int lane_width, int group_width, bool pair, const ScanArgs& a,
                            int grid, size_t lds_bytes, hipStream_t s);
int scan_live_max_wg_per_cu(int lane_width, int group_width, bool pair, size_t lds_bytes);
hipError_t launch_rescore(const RescoreArgs& a, int grid, size_t lds_bytes, hipStream_t s);
int rescore_max_wg_per_cu(size_t lds_bytes);
// LDS the lane path of the search kernel needs for a strip width.
size_t lane_lds_bytes(int lane_width, bool pair, int ps_lane);

}  // namespace swh
