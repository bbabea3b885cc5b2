// CPUWorker.h -- the host's AVX2 cores as a worker beside HIPWorker.
//
// The reference's hybrid design (src/WorkerPool.cpp:25-78: an AVXWorker or
// SSEWorker always sits next to the device workers, all pulling chunks from
// one DataPool) with this library's own AVX2 scan (CpuKernel.h):
//   alloc      checks the host for AVX2, fixes the scoring
//   packDB     copies the chunk and its batch table into the aux entry (the
//              chunk buffer is refilled while this entry is searched)
//   search     the chunk's sequences, longest first, in pieces of whole
//              32-lane groups (~256 KB each; groups span batches, whose long
//              tail holds few sequences of each length), taken by a team of
//              threads from an atomic counter: int8 pass, int16 pass for the
//              lanes that overflowed, and sequences beyond int16 go to the
//              RecalcWorker
//              (scored 0 here, as the reference's workers do,
//              src/AVXWorker.cpp:429-459)
#pragma once
#include <vector>

#include "CpuKernel.h"
#include "Worker.h"

class CPUWorker : public Worker {
  public:
    CPUWorker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int threads,
              size_t chunkBytes);
    ~CPUWorker() override;
    void setQuery(const char* query, size_t qlen) override;
    double busySeconds() const override { return busy_; }
    uint64_t pushedToRecalc() const { return pushed_; }

  private:
    void alloc() override;
    void free() override;
    void search() override;
    void packDB() override;
    void showPerformance() override;
    void who() override;

    int threads_;
    swhy::cpu::Scoring scoring_;
    std::vector<char> buf_[2];
    std::vector<batchInfo> batches_[2];
    std::vector<int> scores_[2];
    double busy_ = 0;
    uint64_t pushed_ = 0;
};
