// HIPWorker.h -- the MI355X device worker.
//
// Takes the place of the reference's CUDAWorker (src/CUDAWorker.h:12-75,
// src/CUDAWorker.cu:29-246) behind the same Worker interface, and reaches the
// gfx950 kernels only through the C ABI of libswh.so (include/swh.h):
//   alloc      swh_ctx_create on the worker's device
//   packDB     the DataPool chunk (filled into page-locked memory) becomes a
//              packed, HBM-resident database (swh_db_create: one DMA upload,
//              interleaved on the device), on the aux thread while the
//              previous chunk is searched
//   asyncCopy  nothing left to copy: packDB uploaded the chunk
//   search     swh_search -> final exact scores in compEntry->getScores();
//              overflowed lanes were rescored on the device, so nothing goes
//              to the RecalcWorker
//   free       releases the databases and the context
#pragma once
#include <vector>

#include "Worker.h"
#include "swh.h"

class HIPWorker : public Worker {
  public:
    HIPWorker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int deviceId,
              size_t chunkBytes);
    ~HIPWorker() override;
    void setQuery(const char* query, size_t qlen) override;

    double busySeconds() const override { return kernelMs_ / 1e3; }
    uint64_t cells() const { return cells_; }

  private:
    void alloc() override;
    void free() override;
    void asyncCopy() override {}
    void search() override;
    void packDB() override;
    char* allocChunk(size_t bytes) override;
    void freeChunk(char* p) override;
    void showPerformance() override;
    void who() override;

    int deviceId_;
    swh_ctx* ctx_ = nullptr;
    swh_db* dbs_[2] = {nullptr, nullptr};
    std::vector<int> scores_[2];
    std::vector<uint64_t> offsets_;
    std::vector<uint32_t> lengths_;
    double kernelMs_ = 0;
    uint64_t cells_ = 0;
};
