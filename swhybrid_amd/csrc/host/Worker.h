// Worker.h -- the device-worker plugin interface.
//
// Same virtual surface and call order as the reference's Worker
// (src/Worker.h:82-249, src/Worker.cpp:176-220), written anew:
//   threadControl: setQuery -> launch
//   launch: alloc; fill (DataPool chunk -> packDB); swapEntries; asyncCopy;
//           loop { aux thread: push previous scores + fill/pack next  ||  search }
//           then push the last scores, stop the timer, showPerformance, free.
// Two Entry slots double-buffer (chunk, map, scores).  After search() the
// scores of compEntry are final: getScores()[i] belongs to pool sequence
// getBaseIdx() + i.  Errors are thrown as const char* and reported by
// threadControl, as in the reference.
#pragma once
#include <pthread.h>

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "DataPool.h"
#include "Defs.h"
#include "Param.h"
#include "RecalcWorker.h"
#include "ResultList.h"

class Entry {
  public:
    void setBatchNum(int n) { batchNum_ = n; }
    int getBatchNum() const { return batchNum_; }
    void setBaseIdx(size_t b) { baseIdx_ = b; }
    size_t getBaseIdx() const { return baseIdx_; }
    void setEntry(char* buf, int* map, int* scores) {
        buf_ = buf;
        map_ = map;
        scores_ = scores;
    }
    char* getBuf() { return buf_; }
    int* getMap() { return map_; }
    int* getScores() { return scores_; }
    void setResultSize(size_t n) { resultSize_ = n; }
    size_t getResultSize() const { return resultSize_; }

  private:
    size_t baseIdx_ = 0;
    size_t resultSize_ = 0;
    int batchNum_ = 0;
    char* buf_ = nullptr;
    int* map_ = nullptr;
    int* scores_ = nullptr;
};

class Worker {
  public:
    Worker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int packSize,
           int batchSize, int typeSize);
    virtual ~Worker();

    void launch();
    virtual void alloc() {}
    virtual void free() {}
    virtual void asyncCopy() {}
    virtual void search() {}
    virtual void showPerformance() {}
    virtual void setQuery(const char* query, size_t qlen);

    ResultList<int>& getResult() { return resultList; }
    double getStartTime() const { return startTime_; }
    double getStopTime() const { return stopTime_; }
    size_t getResidueCount() const { return TotalAminoAcidResidue; }
    const std::string& lastError() const { return error_; }
    // host sink for every per-sequence score (pool order), for parity checks
    void setScoreSink(int* sink) { sink_ = sink; }
    // seconds spent computing (device kernels or host scan)
    virtual double busySeconds() const { return 0; }
    // Where the worker's time went: DataPool fills (packDB included) on the
    // worker thread before the loop and on the aux thread, search() calls,
    // waits for the aux thread after a search, and the chunks searched.
    struct Phases {
        double first_fill = 0, aux_fill = 0, search = 0, aux_wait = 0;
        int chunks = 0;
    };
    const Phases& phases() const { return phases_; }

    int startThread();
    void waitForThread();

  protected:
    virtual void who() {}
    virtual void packDB();
    // the chunk buffer the DataPool fills, allocated at the first fill with
    // bufferSize bytes (PoolHIPWorker: page-locked memory); a subclass that
    // overrides these frees dbSeqBuf in its own destructor
    virtual char* allocChunk(size_t bytes);
    virtual void freeChunk(char* p);
    // bytes the DataPool may put into fill k of this worker (k = 0 first;
    // at most bufferSize, at least one batch)
    virtual size_t fillBytes(int k) const {
        (void)k;
        return bufferSize;
    }
    bool fill();
    void auxWork();
    void swapEntries() {
        bufFlag = !bufFlag;
        auxEntry = entry + !bufFlag;
        compEntry = entry + bufFlag;
    }
    void startTimer() { startTime_ = now(); }
    void stopTimer() { stopTime_ = now(); }
    static double now() {
        return std::chrono::duration<double>(
                   std::chrono::steady_clock::now().time_since_epoch())
            .count();
    }

    // query and scoring: a subclass's setQuery points query at its own copy
    // (as the reference's workers do); Worker::setQuery keeps one in the base
    char* query = nullptr;
    int queryLen = 0;
    int matrix[32][32];
    int topNum;
    Param* params;
    RecalcWorker* recalcWorker;
    ResultList<int> resultList;

    // chunk staging filled by the DataPool
    char* dbSeqBuf = nullptr;
    batchInfo* info = nullptr;
    size_t infoCapacity = 0;
    size_t infoSize = 0;
    size_t filledSize = 0;
    size_t fillBaseIndex = 0;
    size_t bufferSize = 0;
    int fills_ = 0;   // fills so far
    size_t TotalAminoAcidResidue = 0;
    bool cpuFlag = false;
    int score_limit_7 = 0;   // 128 - max(matrix) (src/Worker.cpp:62-70)

    // double buffering
    bool loaderFlag = false;
    bool bufFlag = true;
    Entry entry[2];
    Entry* auxEntry = entry;
    Entry* compEntry = entry + 1;

    const int packSize;
    const int batchSize;
    const int typeSize;
    int* sink_ = nullptr;

  private:
    std::vector<char> queryStore_;   // Worker::setQuery's copy of the query
    static void* threadEntry(void* self);
    static void* auxEntryFunc(void* self);
    void threadControl();
    DataPool* loader_;
    pthread_t thread_{};
    pthread_t auxThread_{};
    double startTime_ = 0, stopTime_ = 0;
    Phases phases_;
    std::string error_;
};
