/*
 * sw_oracle.c -- CPU restatement of SWhybrid's Smith-Waterman search path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the thing
 * measured or shipped.  See sw_oracle.h for the conventions.
 *
 * What each function restates (reference paths relative to the repository
 * root of turbo0628/SWhybrid):
 *   orc_gotoh_i32   the recurrence every reference kernel computes
 *                   (ONE_CELL_UPDATE, src/AVXKernel.cpp:196-218, applied by
 *                   avxSearchTile, src/AVXKernel.cpp:220-289), in plain int32
 *                   without saturation.
 *   orc_sat_scan    one SIMD lane of avxSearchTile / donormal16 with the
 *                   biased saturating encoding (0x80 / 0x8000 = score 0,
 *                   src/AVXKernel.cpp:443,460-464 and :903-908), in the
 *                   reference's own operation order.
 *   orc_limit8      score_limit_7 = 128 - max(matrix) (src/Worker.cpp:62-70).
 *   orc_search      avxCompute's overflow test (src/AVXKernel.cpp:475-485)
 *                   followed by the RecalcWorker int16 rescore
 *                   (src/RecalcWorker.cpp:184-257 -> avxCompute16).
 *   orc_topk        ResultList<int>::push (src/ResultList.h:37-79).
 *
 * Deliberate difference: avxCompute16 hardcodes gap 12/2
 * (src/AVXKernel.cpp:914-915); orc_search passes the caller's gaps to the
 * int16 rescore.  For the default 10/2 both agree.
 */
#include "sw_oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEG_INF (INT32_MIN / 4)

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

int32_t orc_gotoh_i32(const uint8_t* query, int qlen, const uint8_t* db, int dlen,
                      const int32_t* matrix, int gap_open, int gap_extend)
{
    if (qlen <= 0 || dlen <= 0) return 0;
    const int32_t goe = gap_open + gap_extend;
    const int32_t ge = gap_extend;
    /* H[j] holds H(i-1, j) before the update of row i and H(i, j) after;
       E[j] the vertical gap state E(i-1, j) -> E(i, j). */
    int32_t* H = (int32_t*)malloc(sizeof(int32_t) * (size_t)qlen);
    int32_t* E = (int32_t*)malloc(sizeof(int32_t) * (size_t)qlen);
    for (int j = 0; j < qlen; ++j) { H[j] = 0; E[j] = NEG_INF; }
    int32_t best = 0;
    for (int i = 0; i < dlen; ++i) {
        const int32_t* srow = matrix + db[i];          /* column d_i, stride 32 */
        int32_t diag = 0;                                /* H(i-1, -1) = 0 */
        int32_t hleft = 0;                               /* H(i, -1) = 0 */
        int32_t f = NEG_INF;                             /* F(i, -1) */
        for (int j = 0; j < qlen; ++j) {
            int32_t up = H[j];
            int32_t e = imax(E[j] - ge, up - goe);
            f = imax(f - ge, hleft - goe);
            int32_t h = diag + srow[(size_t)query[j] * 32];
            h = imax(h, e);
            h = imax(h, f);
            h = imax(h, 0);
            E[j] = e;
            H[j] = h;
            diag = up;
            hleft = h;
            if (h > best) best = h;
        }
    }
    free(H);
    free(E);
    return best;
}

/* Saturating helpers on a signed lane of `bits` width. */
static inline int32_t sat(int32_t v, int32_t lo, int32_t hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

int32_t orc_sat_scan(const uint8_t* query, int qlen, const uint8_t* db, int dlen,
                     const int32_t* matrix, int gap_open, int gap_extend, int bits)
{
    const int32_t lo = bits == 8 ? -128 : -32768;
    const int32_t hi = bits == 8 ? 127 : 32767;
    const int32_t bias = -lo;
    const int32_t Q = gap_open + gap_extend;     /* gapoe, src/Param.h:109 */
    const int32_t R = gap_extend;
    if (qlen <= 0) return 0;
    int32_t* Hs = (int32_t*)malloc(sizeof(int32_t) * (size_t)qlen);
    int32_t* Es = (int32_t*)malloc(sizeof(int32_t) * (size_t)qlen);
    for (int j = 0; j < qlen; ++j) { Hs[j] = lo; Es[j] = lo; }  /* memset 0x80 */
    int32_t S = lo;
    for (int i = 0; i < dlen; ++i) {
        int32_t diag = lo;       /* vH[i] = vZero for the first column */
        int32_t F = lo;          /* vF[i] = vZero */
        for (int j = 0; j < qlen; ++j) {
            int32_t P = matrix[(size_t)query[j] * 32 + db[i]];
            int32_t E = Es[j];
            /* ONE_CELL_UPDATE, in the reference's order */
            int32_t h = sat(diag + P, lo, hi);   /* adds  H = H + p[q]   */
            h = imax(h, F);                      /* max   H = max(H, F)  */
            h = imax(h, E);                      /* max   H = max(H, E)  */
            S = imax(S, h);                      /* max   S = max(S, H)  */
            F = sat(F - R, lo, hi);              /* subs  F = F - R      */
            E = sat(E - R, lo, hi);              /* subs  E = E - R      */
            int32_t n = h;                       /* N = H                */
            h = sat(h - Q, lo, hi);              /* subs  H = H - Q      */
            E = imax(h, E);                      /* max   E = max(H, E)  */
            F = imax(h, F);                      /* max   F = max(H, F)  */
            diag = Hs[j];
            Hs[j] = n;
            Es[j] = E;
        }
    }
    free(Hs);
    free(Es);
    return S + bias;                             /* S ^ 0x80 / S ^ 0x8000 */
}

int32_t orc_limit8(const int32_t* matrix)
{
    int32_t hi = 0;
    for (int i = 0; i < 32 * 32; ++i)
        if (matrix[i] > hi) hi = matrix[i];
    return 128 - hi;
}

void orc_search(const uint8_t* query, int qlen, const uint8_t* residues,
                const uint64_t* offsets, const uint32_t* lengths, int64_t nseq,
                const int32_t* matrix, int gap_open, int gap_extend,
                int32_t* out_scores, uint8_t* out_overflow8, int nthreads)
{
    const int32_t limit = orc_limit8(matrix);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t s = 0; s < nseq; ++s) {
        const uint8_t* d = residues + offsets[s];
        int dlen = (int)lengths[s];
        int32_t sc = orc_sat_scan(query, qlen, d, dlen, matrix, gap_open, gap_extend, 8);
        uint8_t ovf = 0;
        if (sc >= limit) {
            ovf = 1;
            sc = orc_sat_scan(query, qlen, d, dlen, matrix, gap_open, gap_extend, 16);
        }
        out_scores[s] = sc;
        if (out_overflow8) out_overflow8[s] = ovf;
    }
}

void orc_search_i32(const uint8_t* query, int qlen, const uint8_t* residues,
                    const uint64_t* offsets, const uint32_t* lengths, int64_t nseq,
                    const int32_t* matrix, int gap_open, int gap_extend,
                    int32_t* out_scores, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t s = 0; s < nseq; ++s)
        out_scores[s] = orc_gotoh_i32(query, qlen, residues + offsets[s], (int)lengths[s],
                                      matrix, gap_open, gap_extend);
}

void orc_topk(const int32_t* scores, int64_t n, int k, int32_t* out_score,
              int64_t* out_index)
{
    if (k <= 0) return;
    /* The list is a sorted array of k entries, pre-filled with (0, 0). */
    for (int t = 0; t < k; ++t) { out_score[t] = 0; out_index[t] = 0; }
    int32_t bottom = 0;
    for (int64_t i = 0; i < n; ++i) {
        int32_t sc = scores[i];
        if (sc <= bottom) continue;
        int pos = 0;
        while (pos < k && out_score[pos] >= sc) ++pos;    /* first strictly smaller */
        for (int t = k - 1; t > pos; --t) {
            out_score[t] = out_score[t - 1];
            out_index[t] = out_index[t - 1];
        }
        out_score[pos] = sc;
        out_index[pos] = i;
        bottom = out_score[k - 1];
    }
}
