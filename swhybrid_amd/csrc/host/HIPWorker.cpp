#include "HIPWorker.h"

#include <chrono>
#include <cstdio>
#include <string>

namespace {
thread_local std::string g_msg;
const char* swh_fail(const char* what) {
    g_msg = std::string("HIPWorker: ") + what + ": " + swh_last_error() + "\n";
    return g_msg.c_str();
}
}  // namespace

HIPWorker::HIPWorker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int deviceId,
                     size_t chunkBytes)
    : Worker(loader, params, recalcWorker, 4, 128, sizeof(int)),
      deviceId_(deviceId),
      ctx_(NULL),
      kernelMs_(0),
      cells_(0),
      packSeconds_(0) {
    dbs_[0] = dbs_[1] = NULL;
    bufferSize = chunkBytes < (size_t)maxReaderBufferSize ? chunkBytes
                                                          : (size_t)maxReaderBufferSize;
}

HIPWorker::~HIPWorker() { free(); }

void HIPWorker::setQuery(const char* q, size_t qlen) {
    for (size_t j = 0; j < qlen; ++j)
        if ((unsigned char)q[j] > 23) throw "HIPWorker: query code above 23\n";
    queryCodes_.assign(q, q + qlen);
    query = queryCodes_.data();
    queryLen = (int)qlen;
}

void HIPWorker::alloc() {
    if (swh_ctx_create(deviceId_, params->plain_matrix, params->gapOpen, params->gapExtend,
                       &ctx_) != SWH_OK)
        throw swh_fail("swh_ctx_create");
    for (int i = 0; i < 2; ++i) entry[i].setEntry(NULL, NULL, NULL);
}

void HIPWorker::free() {
    for (int i = 0; i < 2; ++i) {
        if (dbs_[i]) swh_db_destroy(dbs_[i]);
        dbs_[i] = NULL;
    }
    if (ctx_) swh_ctx_destroy(ctx_);
    ctx_ = NULL;
}

// Chunk -> resident packed database in the aux entry's slot.
void HIPWorker::packDB() {
    const int slot = (int)!bufFlag;
    Entry* e = entry + slot;
    TotalAminoAcidResidue += filledSize;
    offsets_.clear();
    lengths_.clear();
    size_t n = 0;
    for (size_t b = 0; b < infoSize; ++b) n += info[b].numSeqs;
    offsets_.reserve(n);
    lengths_.reserve(n);
    uint64_t off = 0;
    for (size_t b = 0; b < infoSize; ++b)
        for (unsigned k = 0; k < info[b].numSeqs; ++k) {
            offsets_.push_back(off);
            lengths_.push_back(info[b].seqLen);
            off += info[b].seqLen;
        }
    if (dbs_[slot]) swh_db_destroy(dbs_[slot]);
    dbs_[slot] = NULL;
    swh_db_opts dopt = {};
    dopt.queries_per_launch = 1;   // one query per search: one layout per chunk
    const std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    if (swh_db_create(ctx_, reinterpret_cast<const uint8_t*>(dbSeqBuf), offsets_.data(),
                      lengths_.data(), lengths_.size(), &dopt, &dbs_[slot]) != SWH_OK)
        throw swh_fail("swh_db_create");
    packSeconds_ +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    scores_[slot].resize(lengths_.size());
    e->setEntry(dbSeqBuf, NULL, scores_[slot].data());
    e->setBaseIdx(fillBaseIndex);
    e->setBatchNum((int)infoSize);
    e->setResultSize((int)lengths_.size());
}

void HIPWorker::search() {
    const int slot = (int)bufFlag;
    swh_stats st = {};
    if (swh_search(ctx_, dbs_[slot], reinterpret_cast<const uint8_t*>(query), queryLen, NULL,
                   compEntry->getScores(), &st) != SWH_OK)
        throw swh_fail("swh_search");
    kernelMs_ += st.kernel_ms + st.rescore_ms;
    cells_ += st.cells;
}

void HIPWorker::showPerformance() {
    const double t = getStopTime() - getStartTime();
    std::fprintf(stderr, "HIP[%d] computing time %lfs, AARes %zu, GCUPs %lf\n", deviceId_, t,
                 (size_t)TotalAminoAcidResidue,
                 t > 0 ? TotalAminoAcidResidue * (double)queryLen / t / 1e9 : 0.0);
}

void HIPWorker::who() { std::fprintf(stderr, "HIP\n"); }
