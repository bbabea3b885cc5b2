// WorkerPool.h -- builds the device workers, runs them over one DataPool and
// merges their results.
//
// Mirrors the reference's WorkerPool (src/WorkerPool.h:38-58,
// src/WorkerPool.cpp:10-155): workers are created per device and started on
// their own threads; after they join, the RecalcWorker runs, the per-worker
// top-k lists are merged (workers in creation order, then the recalc list) and
// GCUPS = padded residues x query length / (last stop - first start).  The
// reference did all of this in its constructor/destructor; here run() does it
// explicitly and errors come back as a status instead of exit().
#pragma once
#include <string>
#include <vector>

#include "CPUWorker.h"
#include "DataPool.h"
#include "PoolHIPWorker.h"
#include "Param.h"
#include "RecalcWorker.h"
#include "ResultList.h"

class WorkerPool {
  public:
    // params->hipNum PoolHIPWorkers (chunks of chunkBytes) and, when
    // params->cpuNum > 0, one CPUWorker with that many threads (chunks of
    // cpuChunkBytes), in that order (the reference creates the CPU worker
    // first, src/WorkerPool.cpp:47-52; the order only decides ties).
    // chunkBytes: each HIPWorker's chunk; firstChunkBytes (0 = chunkBytes):
    // its first chunk, each later one doubling up to chunkBytes
    WorkerPool(DataPool* loader, Param* params, size_t chunkBytes, size_t cpuChunkBytes = 0,
               size_t firstChunkBytes = 0);
    ~WorkerPool();
    // Start every worker, wait for them, rescore, merge.  0 on success.
    int run(int* scoreSink);
    const ResultList<int>& result() const { return resultList_; }
    double seconds() const { return seconds_; }
    double gcups() const { return gcups_; }
    double kernelSeconds() const;
    // per kind: padded residues searched and seconds spent computing
    uint64_t residues(bool cpu) const;
    double busySeconds(bool cpu) const;
    // the device workers' phases summed, and their swh_db_create time
    Worker::Phases devicePhases() const;
    double devicePackSeconds() const;
    uint64_t recalcCount() const { return recalcWorker_->getNum(); }
    const std::string& error() const { return error_; }

  private:
    DataPool* loader_;
    Param* params_;
    RecalcWorker* recalcWorker_;
    std::vector<Worker*> workerList_;
    std::vector<bool> isCpu_;
    ResultList<int> resultList_;
    double seconds_ = 0, gcups_ = 0;
    std::string error_;
};
