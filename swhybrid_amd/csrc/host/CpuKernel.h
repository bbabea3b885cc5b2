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
#include <cstdint>
#include <vector>

namespace swhy {
namespace cpu {

bool avx2_available();

struct Scoring {
    int32_t matrix[32 * 32];   // row = query code, column = database code
    int gap_open = 10, gap_extend = 2;
};

// Per-thread working memory, sized on first use.
struct Scratch {
    std::vector<uint8_t> inter;     // interleaved residues of one lane group
    std::vector<uint8_t> colstate;  // H and E per query column
    std::vector<uint8_t> prof;      // profile of 4 rows
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

// The same for n sequences of length len stored back to back at seqs.
void score_batch(const Scoring& sc, const uint8_t* query, int qlen, const uint8_t* seqs,
                 uint32_t n, uint32_t len, int32_t* out, Scratch& scratch);

}  // namespace cpu
}  // namespace swhy
