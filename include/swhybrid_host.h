/*
 * swhybrid_host.h -- C entry point of the host worker pipeline
 * (libswhybrid_host.so): an in-memory DataPool feeds HIPWorker threads (the
 * Worker plugin that reaches libswh.so), a RecalcWorker collects anything a
 * worker could not score, and the per-worker top-k lists are merged.
 *
 * Replaces the reference's process entry (src/main.cpp:7-16: Param, DataPool,
 * WorkerPool constructed from argv, results printed by ~WorkerPool,
 * src/WorkerPool.cpp:84-115) with a call that returns the results.
 */
#ifndef SWHYBRID_HOST_H_
#define SWHYBRID_HOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* swhy_last_error(void);

/* Search one query through the worker pipeline.
 *   chunk_bytes    DataPool chunk budget per fill (0 = 256 MB, the reference's
 *                  CUDAWorker bufferSize, src/CUDAWorker.cu:44)
 *   n_workers      HIPWorkers on devices 0..n_workers-1
 *   out_top_*      topk entries: score, caller index, pool index; ties are
 *                  ordered by pool index (the order sequences are handed out),
 *                  padded with (0, 0)
 *   out_scores     optional: every sequence's exact score, caller order
 * Returns 0 on success, negative with swhy_last_error() set otherwise. */
int swhy_pool_search(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                     uint64_t nseq, const uint8_t* query, int qlen, const int32_t* matrix,
                     int gap_open, int gap_extend, int topk, uint64_t chunk_bytes,
                     int n_workers, int32_t* out_top_score, int64_t* out_top_index,
                     int64_t* out_top_pool_index, int32_t* out_scores, double* out_seconds,
                     double* out_gcups);

#ifdef __cplusplus
}
#endif
#endif
