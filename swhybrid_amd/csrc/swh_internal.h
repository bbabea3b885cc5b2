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
constexpr int kSegCols = 1344;       // query columns per LDS profile segment (multiple of 16, 24 and 32;
                                     // 24 rows x (1344+8) int16 = 64,896 B of LDS)
constexpr int kProfPad = 8;          // extra int16 per LDS row: rows start 16 B apart mod 128 B
constexpr uint32_t kNoSeq = 0xffffffffu;

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

struct ScanArgs {
    const uint32_t* words;     // packed residues
    const BlockDesc* blocks;   // block table
    const uint32_t* perm;      // sorted position -> caller index (kNoSeq for fillers)
    const int16_t* prof;       // query profile [24][qpad], zero beyond qlen
    int32_t* scores;           // caller order
    uint32_t* ovf_list;        // sorted positions to rescore
    uint32_t* ovf_count;
    uint32_t* work_counter;
    uint2* border;             // per wave slot: rows * 64 uint2 (H pair, F pair)
    uint64_t border_stride;    // uint2 elements per wave slot
    int blk_begin, blk_end;    // block range of this launch
    int n_items;               // work items (4 blocks each)
    int qlen, qpad, ps;        // query length, padded length, LDS row stride (int16)
    int pstride;               // row stride of prof (int16), >= qpad
    int gap_oe, gap_e;
    int ovf_threshold;         // score >= threshold -> int32 rescore
};

// Group mode: G lanes cooperate on one sequence pair.  Lane g of a group holds
// query columns [pass*G*W + g*W, +W) and works on row t-g at step t; the H/F
// of its last column reach lane g+1 one step later through a wave-wide DPP
// shift.  A wave holds floor(64/G) groups; a workgroup item is 4 waves.
struct GroupArgs {
    const uint32_t* words;
    const BlockDesc* blocks;   // pair blocks (2 sequences each)
    const uint32_t* perm;
    const int16_t* prof;
    int32_t* scores;
    uint32_t* ovf_list;
    uint32_t* ovf_count;
    uint32_t* work_counter;
    uint2* border;             // per wave slot: groups_per_wave * border_rows uint2
    uint64_t border_stride;    // uint2 per wave slot
    uint32_t border_rows;      // rows reserved per group (>= longest pair block)
    int blk_begin, blk_end;    // pair-block range
    int group;                 // G, lanes per group (1..64)
    int n_items;
    int qlen, qpad, ps, pstride;
    int gap_oe, gap_e;
    int ovf_threshold;
};

struct RescoreArgs {
    const uint32_t* words;
    const BlockDesc* blocks;
    const uint32_t* perm;
    const uint32_t* sorted_len;   // real length per sorted position
    const uint32_t* sorted_blk;   // block of each sorted position
    const int16_t* prof;
    int32_t* scores;
    const uint32_t* ovf_list;
    const uint32_t* ovf_count;
    uint32_t* work_counter;
    uint2* border;
    uint64_t border_stride;
    int n_lane_blocks_first;      // index of the first lane-mode block
    int qpad, ps, pstride;
    int gap_oe, gap_e;
};

// Launchers (swh_kernels.hip).
hipError_t launch_reset(uint32_t* p, int n, hipStream_t s);
hipError_t launch_lane_scan(int width, const ScanArgs& a, int grid, size_t lds_bytes,
                            hipStream_t s);
hipError_t launch_group_scan(int width, const GroupArgs& a, int grid, size_t lds_bytes,
                             hipStream_t s);
int group_scan_max_wg_per_cu(int width, size_t lds_bytes);
hipError_t launch_rescore(const RescoreArgs& a, int grid, size_t lds_bytes, hipStream_t s);
int lane_scan_max_wg_per_cu(int width, size_t lds_bytes);
int rescore_max_wg_per_cu(size_t lds_bytes);

}  // namespace swh
