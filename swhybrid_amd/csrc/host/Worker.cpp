#include "Worker.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

Worker::Worker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int packSize,
               int batchSize, int typeSize)
    : topNum(params->topNum), params(params), recalcWorker(recalcWorker),
      resultList(params->topNum), packSize(packSize), batchSize(batchSize),
      typeSize(typeSize), loader_(loader) {
    int hi = 0;
    for (int i = 0; i < 32; ++i)
        for (int j = 0; j < 32; ++j) {
            matrix[i][j] = params->matrix[i][j];
            if (matrix[i][j] > hi) hi = matrix[i][j];
        }
    score_limit_7 = 128 - hi;
}

Worker::~Worker() {
    std::free(dbSeqBuf);   // subclasses with their own allocChunk cleared it
    delete[] info;
}

char* Worker::allocChunk(size_t bytes) { return (char*)std::malloc(bytes); }

void Worker::freeChunk(char* p) { std::free(p); }

void Worker::setQuery(const char* q, size_t qlen) {
    queryStore_.assign(q, q + qlen);
    query = queryStore_.data();
    queryLen = (int)qlen;
}

bool Worker::fill() {
    if (!dbSeqBuf) {
        bufferSize = std::max(bufferSize, loader_->maxBatchBytes());
        dbSeqBuf = allocChunk(bufferSize + 64);
        infoCapacity = loader_->maxInfoPerChunk(bufferSize) + 1;
        info = new batchInfo[infoCapacity];
        if (!dbSeqBuf) throw "worker: chunk buffer allocation failed\n";
    }
    const size_t cap = std::max(std::min(fillBytes(fills_++), bufferSize),
                                loader_->maxBatchBytes());
    bool ret = loader_->fillChunk(dbSeqBuf, info, infoSize, filledSize, cap, cpuFlag,
                                  fillBaseIndex);
    if (infoSize != 0) packDB();
    return ret;
}

// Base packing: interleave batchSize sequences per lane group (the layout the
// reference's SIMD workers use, dst[pos * batchSize + lane]).
void Worker::packDB() {
    Entry* e = entry + !bufFlag;
    TotalAminoAcidResidue += filledSize;
    e->setBaseIdx(fillBaseIndex);
    e->setBatchNum((int)infoSize);
    size_t n = 0;
    for (size_t b = 0; b < infoSize; ++b) n += info[b].numSeqs;
    e->setResultSize(n);
}

void Worker::launch() {
    alloc();
    double t = now();
    loaderFlag = fill();
    phases_.first_fill += now() - t;
    if (!loaderFlag) {
        free();
        return;
    }
    swapEntries();
    asyncCopy();
    startTimer();
    while (loaderFlag) {
        pthread_create(&auxThread_, nullptr, auxEntryFunc, this);
        t = now();
        try {
            search();
        } catch (...) {
            pthread_join(auxThread_, nullptr);   // never leave the aux thread behind
            free();
            throw;
        }
        const double t_searched = now();
        pthread_join(auxThread_, nullptr);
        phases_.search += t_searched - t;
        phases_.aux_wait += now() - t_searched;
        phases_.chunks += 1;
        if (!error_.empty()) {
            free();
            throw error_.c_str();
        }
        swapEntries();
        asyncCopy();
    }
    auxWork();
    stopTimer();
    showPerformance();
    free();
}

void Worker::auxWork() {
    Entry* e = auxEntry;
    if (e->getResultSize() != 0) {
        resultList.push(e->getScores(), e->getResultSize(), e->getBaseIdx());
        e->setResultSize(0);
    }
    const double t = now();
    loaderFlag = fill();
    phases_.aux_fill += now() - t;
}

void* Worker::auxEntryFunc(void* self) {
    Worker* w = static_cast<Worker*>(self);
    try {
        w->auxWork();
    } catch (const char* msg) {
        w->error_ = msg;
        w->loaderFlag = false;
    }
    return nullptr;
}

void Worker::threadControl() {
    try {
        setQuery(loader_->getQuery(), loader_->getQueryLen());
        launch();
    } catch (const char* msg) {
        error_ = msg;
        std::fprintf(stderr, "%s", msg);
    }
}

void* Worker::threadEntry(void* self) {
    static_cast<Worker*>(self)->threadControl();
    return nullptr;
}

int Worker::startThread() { return pthread_create(&thread_, nullptr, threadEntry, this); }

void Worker::waitForThread() { pthread_join(thread_, nullptr); }
