// PoolHIPWorker.h -- HIPWorker with the hooks of this pipeline's Worker.
//
// HIPWorker itself is written against the reference's Worker (so it drops
// into the reference's src/ unchanged); this subclass adds what only this
// repository's Worker offers:
//   allocChunk / freeChunk  the DataPool chunk buffer in page-locked memory,
//                           so swh_db_create uploads it by DMA without the
//                           library's staging copy
//   search                  also writes every score into the pool's sink
//                           (pool order), for parity checks
//   busySeconds             device time of the worker's searches
//   fillBytes               a chunk schedule: the first chunk firstChunkBytes,
//                           each later one twice the one before, up to
//                           chunkBytes (0: every chunk chunkBytes)
#pragma once
#include <algorithm>

#include "HIPWorker.h"

class PoolHIPWorker : public HIPWorker {
  public:
    PoolHIPWorker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int deviceId,
                  size_t chunkBytes, size_t firstChunkBytes = 0)
        : HIPWorker(loader, params, recalcWorker, deviceId, chunkBytes),
          firstChunkBytes_(firstChunkBytes) {}
    ~PoolHIPWorker() override {
        freeChunk(dbSeqBuf);
        dbSeqBuf = nullptr;
    }
    double busySeconds() const override { return kernelSeconds(); }

  protected:
    char* allocChunk(size_t bytes) override {
        void* p = nullptr;
        if (swh_pinned_alloc(bytes, &p) != SWH_OK) throw "PoolHIPWorker: swh_pinned_alloc failed\n";
        return static_cast<char*>(p);
    }
    void freeChunk(char* p) override {
        if (p) (void)swh_pinned_free(p);
    }
    void search() override {
        HIPWorker::search();
        if (sink_) {
            const int* s = compEntry->getScores();
            for (size_t i = 0; i < compEntry->getResultSize(); ++i)
                sink_[compEntry->getBaseIdx() + i] = s[i];
        }
    }

    size_t fillBytes(int k) const override {
        if (firstChunkBytes_ == 0 || firstChunkBytes_ >= bufferSize) return bufferSize;
        size_t b = firstChunkBytes_;
        for (int i = 0; i < k && b < bufferSize; ++i) b *= 2;
        return std::min(b, bufferSize);
    }

  private:
    size_t firstChunkBytes_;
};
