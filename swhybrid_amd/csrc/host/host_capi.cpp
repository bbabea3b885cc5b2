// host_capi.cpp -- C entry point that runs the whole worker pipeline
// (DataPool -> HIPWorker(s) -> RecalcWorker -> merged ResultList), the way the
// reference's main() drives WorkerPool (src/main.cpp:7-16), for callers that
// hold the database in memory.
#include <string>
#include <vector>

#include "DataPool.h"
#include "Param.h"
#include "WorkerPool.h"
#include "swhybrid_host.h"

namespace {
thread_local std::string g_err;
}

extern "C" {

const char* swhy_last_error(void) { return g_err.c_str(); }

int swhy_pool_search(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                     uint64_t nseq, const uint8_t* query, int qlen, const int32_t* matrix,
                     int gap_open, int gap_extend, int topk, uint64_t chunk_bytes,
                     int n_workers, int32_t* out_top_score, int64_t* out_top_index,
                     int64_t* out_top_pool_index, int32_t* out_scores, double* out_seconds,
                     double* out_gcups) {
    if (!query || qlen <= 0 || !matrix || topk < 0 || n_workers < 1 ||
        (nseq > 0 && (!residues || !offsets || !lengths)) ||
        (topk > 0 && (!out_top_score || !out_top_index))) {
        g_err = "swhy_pool_search: bad arguments";
        return -1;
    }
    DataPool pool(residues, offsets, lengths, nseq, query, (size_t)qlen);
    std::vector<int> m(matrix, matrix + 1024);
    Param params(m.data(), gap_open, gap_extend, topk, n_workers);
    std::vector<int> sink;
    if (out_scores) sink.assign(nseq, -1);
    int rc;
    double secs = 0, gcups = 0;
    std::vector<Result<int>> top;
    {
        WorkerPool wp(&pool, &params, chunk_bytes ? chunk_bytes : (size_t(256) << 20));
        rc = wp.run(out_scores ? sink.data() : nullptr);
        if (rc != 0) {
            g_err = "swhy_pool_search: " + wp.error();
            return -2;
        }
        secs = wp.seconds();
        gcups = wp.gcups();
        top = wp.result().getResult();
    }
    for (int t = 0; t < topk; ++t) {
        const bool real = t < (int)top.size() && top[t].score > 0;
        out_top_score[t] = t < (int)top.size() ? top[t].score : 0;
        out_top_index[t] = real ? (int64_t)pool.originalIndex(top[t].idx) : 0;
        if (out_top_pool_index) out_top_pool_index[t] = real ? (int64_t)top[t].idx : 0;
    }
    if (out_scores)
        for (uint64_t p = 0; p < nseq; ++p) out_scores[pool.originalIndex(p)] = sink[p];
    if (out_seconds) *out_seconds = secs;
    if (out_gcups) *out_gcups = gcups;
    return 0;
}

}  // extern "C"
