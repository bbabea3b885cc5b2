#include "CPUWorker.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <thread>

CPUWorker::CPUWorker(DataPool* loader, Param* params, RecalcWorker* recalcWorker, int threads,
                     size_t chunkBytes)
    : Worker(loader, params, recalcWorker, 4, 32, sizeof(int)),
      threads_(threads > 0 ? threads : 1) {
    bufferSize = chunkBytes;
    cpuFlag = true;
}

CPUWorker::~CPUWorker() { free(); }

void CPUWorker::setQuery(const char* q, size_t qlen) {
    Worker::setQuery(q, qlen);
    for (size_t j = 0; j < qlen; ++j)
        if ((unsigned char)q[j] > 23) throw "CPUWorker: query code above 23\n";
}

void CPUWorker::alloc() {
    if (!swhy::cpu::avx2_available()) throw "CPUWorker: the host has no AVX2\n";
    std::memcpy(scoring_.matrix, params->plain_matrix, sizeof(scoring_.matrix));
    scoring_.gap_open = params->gapOpen;
    scoring_.gap_extend = params->gapExtend;
    for (int i = 0; i < 2; ++i) entry[i].setEntry(nullptr, nullptr, nullptr);
}

void CPUWorker::free() {
    for (int i = 0; i < 2; ++i) {
        std::vector<char>().swap(buf_[i]);
        std::vector<int>().swap(scores_[i]);
    }
}

void CPUWorker::packDB() {
    const int slot = (int)!bufFlag;
    Entry* e = entry + slot;
    TotalAminoAcidResidue += filledSize;
    buf_[slot].assign(dbSeqBuf, dbSeqBuf + filledSize);
    batches_[slot].assign(info, info + infoSize);
    size_t n = 0;
    for (size_t b = 0; b < infoSize; ++b) n += info[b].numSeqs;
    scores_[slot].assign(n, 0);
    e->setEntry(buf_[slot].data(), nullptr, scores_[slot].data());
    e->setBaseIdx(fillBaseIndex);
    e->setBatchNum((int)infoSize);
    e->setResultSize(n);
}

void CPUWorker::search() {
    const auto t0 = std::chrono::steady_clock::now();
    const int slot = (int)bufFlag;
    const std::vector<batchInfo>& bs = batches_[slot];
    const uint8_t* data = reinterpret_cast<const uint8_t*>(buf_[slot].data());
    int* out = scores_[slot].data();
    // every sequence of the chunk; the team takes them longest first in
    // pieces of whole 32-lane groups (batches of one length can be small, a
    // long tail of distinct lengths, so groups span batches)
    std::vector<const uint8_t*> ptr;
    std::vector<uint32_t> len;
    size_t boff = 0;
    for (const batchInfo& b : bs) {
        for (uint32_t k = 0; k < b.numSeqs; ++k) {
            const uint8_t* p = data + boff + (size_t)k * b.seqLen;
            uint32_t n = b.seqLen;
            while (n > 0 && p[n - 1] == 23) --n;   // trailing pads cost rows only
            ptr.push_back(p);
            len.push_back(n);
        }
        boff += (size_t)b.numSeqs * b.seqLen;
    }
    const size_t nseq = ptr.size();
    swhy::cpu::score_team(scoring_, reinterpret_cast<const uint8_t*>(query), queryLen, ptr.data(),
                          len.data(), nseq, out, threads_);
    const size_t base = compEntry->getBaseIdx();
    uint64_t pushed = 0;
    for (size_t i = 0; i < nseq; ++i)
        if (out[i] < 0) {   // beyond int16: exact int32 after the workers finish
            recalcWorker->pushOverflowSeq(reinterpret_cast<const char*>(ptr[i]), len[i], base + i);
            ++pushed;
            out[i] = 0;
        }
    pushed_ += pushed;
    if (sink_)
        for (size_t i = 0; i < compEntry->getResultSize(); ++i) sink_[base + i] = out[i];
    busy_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void CPUWorker::showPerformance() {
    const double t = getStopTime() - getStartTime();
    std::fprintf(stderr, "CPU[%d threads] computing time %lfs, AARes %zu, GCUPs %lf\n", threads_, t,
                TotalAminoAcidResidue,
                t > 0 ? TotalAminoAcidResidue * (double)queryLen / t / 1e9 : 0.0);
}

void CPUWorker::who() { std::fprintf(stderr, "CPU\n"); }
