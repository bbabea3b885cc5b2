// DataPool.h -- hands out database chunks to workers.
//
// Interface and chunk format of the reference's DataPool (src/DataPool.h:12-75,
// fillChunk src/DataPool.cpp:280-286): a chunk is a run of consecutive batches,
// each batch numSeqs sequences of one padded length stored back to back, and
// baseIndex is the pool index of the chunk's first sequence.  Two sources:
// the caller's sequences in memory (sorted longest first and grouped into
// batches of up to 128 sequences of equal length rounded up to a multiple of
// 4, padded with code 23), or a database in the reference's .seq/.map/.title
// files, whose batches are handed out in file order as the reference does.
#pragma once
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "DbFile.h"
#include "Defs.h"

class DataPool {
  public:
    DataPool(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
             uint64_t nseq, const uint8_t* query, size_t qlen);
    // Pool index = file index (fillers included).
    DataPool(const swhy::DbFile& db, const uint8_t* query, size_t qlen);

    // Copies the next batches into dst while they fit in bufferSize bytes
    // (codes as given: consumers clamp codes above 23 when they pack).
    // Returns false (infoSize = 0) when the pool is exhausted, or, for a CPU
    // worker (cpuFlag), once no more than the CPU reserve is left or once its
    // chunk would outlast the device workers' share of the rest.
    bool fillChunk(char* dst, batchInfo* seqBatchInfo, size_t& infoSize, size_t& aaSize,
                   size_t bufferSize, bool cpuFlag, size_t& baseIndex);

    const char* getQuery() const { return reinterpret_cast<const char*>(query_.data()); }
    size_t getQueryLen() const { return query_.size(); }
    size_t getTotalResidue() const { return totalPadded_; }
    uint64_t numSeqs() const { return order_.size(); }
    size_t maxBatchBytes() const { return maxBatchBytes_; }
    size_t maxInfoPerChunk(size_t bufferSize) const;

    // Caller's index of pool sequence poolIdx (the reference looks titles up
    // by pool index, src/DataPool.cpp:289-303): the file index for a file
    // database.
    uint64_t originalIndex(size_t poolIdx) const { return order_[poolIdx]; }
    void rewind();
    // Residue bytes at the end of the pool that CPU workers leave to the
    // device workers, so that a slow CPU chunk does not finish last.
    void setCpuReserve(size_t bytes) { cpuReserve_ = bytes; }
    // Threads that copy a large chunk (default: the host's cores, at most 16);
    // fewer leave cores to a CPU worker.
    void setFillThreads(unsigned n) { fillThreads_ = n; }

  private:
    struct Batch {
        uint64_t first;    // pool index of the first sequence
        unsigned num;
        unsigned len;      // padded length
    };
    const uint8_t* residues_;
    const uint64_t* offsets_;
    const uint32_t* lengths_;
    std::vector<uint32_t> order_;    // pool index -> caller index
    std::vector<Batch> batches_;
    std::vector<uint8_t> query_;
    size_t next_ = 0;
    size_t handedOut_ = 0;
    size_t handedOutCpu_ = 0, handedOutDev_ = 0;   // per kind of worker
    double t0_ = 0;                                // first fill, seconds
    size_t cpuReserve_ = 0;
    unsigned fillThreads_ = 0;   // 0: the default
    size_t totalPadded_ = 0;
    size_t maxBatchBytes_ = 0;
    std::mutex mu_;
};
