// CpuKernel.h -- inter-sequence Smith-Waterman on the host's AVX2 units, the
// scan the CPU worker runs beside the GPU.
//
// The same arithmetic as the reference's AVX2 path (src/AVXKernel.cpp:196-218
// recurrence, 9 operations per cell, zero floor by saturation): 32 database
// sequences per 256-bit vector in biased int8, overflowed lanes again 16 per
// vector in biased int16; what overflows int16 is returned as -1 for the
// caller to rescore in int32.  Written for this host library; the profile is
// a per-4-row table of every query residue's scores against the 32 lanes'
// database residues, built with byte shuffles.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace swhy {
namespace cpu {

bool avx2_available();

// Query columns per pass of the scan when the query is longer (0: the
// default, measured on the host, see kDefaultQuerySegment; -1: never split).
struct Scoring {
    int32_t matrix[32 * 32];   // row = query code, column = database code
    int gap_open = 10, gap_extend = 2;
    int query_segment = 0;
};

// Default query segment: 4096 columns, the reference's own bound (mlen,
// src/AVXKernel.cpp:437), which keeps a thread's column state at 256 KB.  On
// a 2.7 GHz Xeon with 2 MB of L2 per core, 60k database sequences, query
// lengths 5478 and 9000, 1 and 8 threads, segments of 2048-4096 columns run
// level with the unsplit scan (within 1%), 1024 at -1%, 512 at -3%, 256 at
// -5 to -9%.
constexpr int kDefaultQuerySegment = 4096;

// The segment length a scan of a qlen-residue query uses (qlen: unsplit).
int query_segment(const Scoring& sc, int qlen);

// Per-thread working memory, sized on first use.
struct Scratch {
    std::vector<uint8_t> inter;     // interleaved residues of one lane group
    std::vector<uint8_t> colstate;  // H and E per query column
    std::vector<uint8_t> prof;      // profile of 4 rows
    std::vector<uint8_t> edge;      // (H, F) per row at a query segment's last column
    std::vector<const void*> qp;    // profile rows of each query column
    std::vector<uint32_t> ovf;      // overflowed sequences of a call
    std::vector<int32_t> ovf_scores;
    std::vector<int32_t> s8;
    std::vector<uint32_t> all;
};

// Scores n sequences: sequence k has lens[k] codes at seqs[k] (above 23
// treated as 23).  Lane groups are formed in the given order, each run to its
// longest member, so similar lengths should be adjacent.  out[k] = exact
// score, or -1 where the int16 pass may have saturated.  Query codes <= 23.
void score_seqs(const Scoring& sc, const uint8_t* query, int qlen, const uint8_t* const* seqs,
                const uint32_t* lens, uint32_t n, int32_t* out, Scratch& scratch);

// score_seqs over n sequences with a team of `threads` threads (the caller's
// included): longest first, in pieces of whole 32-lane groups of about 256 KB
// of residues handed out from an atomic counter (the reference's thread team
// over batches with a shared counter, src/AVXKernel.cpp:408-490).  out[k] for
// sequence k as in score_seqs.
void score_team(const Scoring& sc, const uint8_t* query, int qlen, const uint8_t* const* seqs,
                const uint32_t* lens, size_t n, int32_t* out, int threads);

// A database prepared for the scan: sequences longest first, 32 per group,
// each group interleaved row by row (byte 32 * row + lane is the code of the
// group's lane-th sequence at that row, 23 beyond its end; rows a multiple of
// 4) -- the layout of the reference's AVXWorker::packDB (src/AVXWorker.cpp:
// 67-302), without its transpose defect.
struct PackedDb {
    std::vector<uint8_t> data;
    std::vector<uint64_t> group_off;    // byte offset of each group
    std::vector<uint32_t> group_rows;   // rows of each group
    std::vector<uint32_t> lane_seq;     // sequence of each (group, lane), or ~0u
    const uint8_t* const* seqs = nullptr;   // the caller's sequences (int16 pass)
    const uint32_t* lens = nullptr;
    size_t n = 0;
};
void pack_db(const uint8_t* const* seqs, const uint32_t* lens, size_t n, PackedDb& db,
             int threads);

// The scan over a packed database with a team of `threads` threads taking
// groups from an atomic counter (the reference's OpenMP team over batches,
// src/AVXKernel.cpp:408-490): int8 pass, then the lanes that may have
// saturated again in int16 (the RecalcWorker's int16 rescore,
// src/RecalcWorker.cpp:184-257).  out[k] for sequence k as in score_seqs.
void scan_db(const Scoring& sc, const uint8_t* query, int qlen, const PackedDb& db, int32_t* out,
             int threads);

// The same for n sequences of length len stored back to back at seqs.
void score_batch(const Scoring& sc, const uint8_t* query, int qlen, const uint8_t* seqs,
                 uint32_t n, uint32_t len, int32_t* out, Scratch& scratch);

}  // namespace cpu
}  // namespace swhy
