/*
 * sw_oracle.h -- CPU restatement of SWhybrid's Smith-Waterman search arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (swhybrid_amd/, the HIP
 * library, bench.py's timed GPU leg) links, loads or calls this code.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker.
 *
 * Coordinates follow the reference kernels: the database sequence runs along
 * the rows (i), the query along the columns (j).  Residue codes are 0..23
 * (DataPool.cpp:33-47); the substitution matrix is a flat 32x32 int table
 * indexed [query code * 32 + db code] (Param.h:110-115).  Gap penalties are
 * the reference's pair (gapOpen, gapExtend), the first gap residue costs
 * gapOpen + gapExtend (Param.h:109).
 */
#ifndef SW_ORACLE_H_
#define SW_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Exact int32 Gotoh local alignment score. */
int32_t orc_gotoh_i32(const uint8_t* query, int qlen, const uint8_t* db, int dlen,
                      const int32_t* matrix, int gap_open, int gap_extend);

/* Scalar emulation of one lane of the reference's biased saturating scan.
 * bits = 8 restates ONE_CELL_UPDATE (AVXKernel.cpp:196-218) with 0x80 = 0;
 * bits = 16 restates ONE_CELL_UPDATE_16 (AVXKernel.cpp:556-577) with 0x8000 = 0.
 * Returns the lane's final unbiased S (S ^ 0x80 / S ^ 0x8000). */
int32_t orc_sat_scan(const uint8_t* query, int qlen, const uint8_t* db, int dlen,
                     const int32_t* matrix, int gap_open, int gap_extend, int bits);

/* Overflow limit of the 8-bit pass: 128 - max(matrix) (Worker.cpp:62-70). */
int32_t orc_limit8(const int32_t* matrix);

/* The reference's whole cascade for a database: int8 pass, lanes with
 * score >= orc_limit8 are rescored by the int16 pass (AVXKernel.cpp:475-485,
 * RecalcWorker.cpp:184-257).  Sequences are given as a concatenation with
 * per-sequence offsets/lengths.  out_scores[i] is the final score of
 * sequence i, out_overflow8[i] (may be NULL) is 1 if it went to the rescore.
 * Parallel over sequences with OpenMP (nthreads <= 0: all). */
void orc_search(const uint8_t* query, int qlen, const uint8_t* residues,
                const uint64_t* offsets, const uint32_t* lengths, int64_t nseq,
                const int32_t* matrix, int gap_open, int gap_extend,
                int32_t* out_scores, uint8_t* out_overflow8, int nthreads);

/* Exact int32 scores only (Gotoh), OpenMP over sequences. */
void orc_search_i32(const uint8_t* query, int qlen, const uint8_t* residues,
                    const uint64_t* offsets, const uint32_t* lengths, int64_t nseq,
                    const int32_t* matrix, int gap_open, int gap_extend,
                    int32_t* out_scores, int nthreads);

/* Top-k with ResultList semantics (ResultList.h:37-79): the list starts as k
 * (score 0, index 0) entries, scores pushed in index order, a push is
 * rejected when score <= bottom, and an accepted push is inserted before the
 * first strictly smaller score (ties keep the earlier index first).
 * Writes exactly k entries. */
void orc_topk(const int32_t* scores, int64_t n, int k, int32_t* out_score,
              int64_t* out_index);

#ifdef __cplusplus
}
#endif
#endif
