// HIPWorker.h -- the MI355X device worker.
//
// Takes the place of the reference's CUDAWorker (src/CUDAWorker.h:12-75,
// src/CUDAWorker.cu:29-246) behind the reference's own Worker interface
// (src/Worker.h:82-249), and reaches the gfx950 kernels only through the C ABI
// of libswh.so (include/swh.h).  It uses nothing but what that Worker
// declares -- query/queryLen, params, dbSeqBuf, info, infoSize, filledSize,
// fillBaseIndex, bufferSize, TotalAminoAcidResidue, entry/bufFlag/compEntry --
// so the same two files compile inside the reference's src/ against its
// Worker.h, and here against this pipeline's Worker.h (a superset):
//   setQuery   keeps its own copy of the query codes
//   alloc      swh_ctx_create on the worker's device
//   packDB     the DataPool chunk in dbSeqBuf becomes a packed, HBM-resident
//              database (swh_db_create uploads it -- straight from dbSeqBuf
//              when that is page-locked, else through the library's own
//              page-locked staging -- and interleaves it on the device), on
//              the aux thread while the previous chunk is searched
//   asyncCopy  nothing left to copy: packDB uploaded the chunk
//   search     swh_search -> final exact scores in compEntry->getScores();
//              overflowed lanes were rescored on the device, so nothing goes
//              to the RecalcWorker
//   free       releases the databases and the context
// dbSeqBuf belongs to the base class (the reference's Worker _mm_mallocs it);
// HIPWorker never frees it.
#ifndef SWHYBRID_HIPWORKER_H_
#define SWHYBRID_HIPWORKER_H_

#include <stdint.h>

#include <vector>

#include "Worker.h"
#include "swh.h"

class HIPWorker : public Worker {
  public:
    // The reference's CUDAWorker signature plus the chunk size (the reference
    // fixes 1 << 28, src/CUDAWorker.cu:44); it is capped at the base class's
    // chunk buffer, maxReaderBufferSize.
    HIPWorker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int deviceId,
              size_t chunkBytes = size_t(1) << 28);
    ~HIPWorker() override;
    void setQuery(const char* query, size_t qlen) override;

    // device time (scan + rescore kernels) and cells of this worker's searches
    double kernelSeconds() const { return kernelMs_ / 1e3; }
    uint64_t cells() const { return cells_; }
    // wall time of the swh_db_create calls (chunk upload and device packing)
    double packSeconds() const { return packSeconds_; }

  protected:
    void alloc() override;
    void free() override;
    void asyncCopy() override {}
    void search() override;
    void packDB() override;
    void showPerformance() override;
    void who() override;

  private:
    int deviceId_;
    swh_ctx* ctx_;
    swh_db* dbs_[2];
    std::vector<char> queryCodes_;
    std::vector<int> scores_[2];
    std::vector<uint64_t> offsets_;
    std::vector<uint32_t> lengths_;
    double kernelMs_;
    uint64_t cells_;
    double packSeconds_;
};

#endif  // SWHYBRID_HIPWORKER_H_
