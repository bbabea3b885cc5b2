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
    // every sequence of the chunk, longest first, cut into pieces of whole
    // 32-lane groups of about 256 KB of residues: batches of one length can
    // be small (a long tail of distinct lengths), so groups span batches
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
    std::vector<uint32_t> order(nseq);
    for (size_t i = 0; i < nseq; ++i) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return len[a] > len[b]; });
    struct Piece {
        size_t first, n;
    };
    std::vector<Piece> pieces;
    for (size_t i = 0; i < nseq;) {
        size_t n = 0, bytes = 0;
        while (i + n < nseq && (n == 0 || n % 32 != 0 || bytes < (size_t(256) << 10))) {
            bytes += len[order[i + n]];
            ++n;
        }
        pieces.push_back(Piece{i, n});
        i += n;
    }
    const uint8_t* q = reinterpret_cast<const uint8_t*>(query);
    const size_t base = compEntry->getBaseIdx();
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> pushed{0};
    auto work = [&]() {
        swhy::cpu::Scratch scratch;
        std::vector<int32_t> sc;
        std::vector<const uint8_t*> pp;
        std::vector<uint32_t> pl;
        for (size_t p = next++; p < pieces.size(); p = next++) {
            const Piece& pc = pieces[p];
            pp.resize(pc.n);
            pl.resize(pc.n);
            for (size_t k = 0; k < pc.n; ++k) {
                pp[k] = ptr[order[pc.first + k]];
                pl[k] = len[order[pc.first + k]];
            }
            sc.resize(pc.n);
            swhy::cpu::score_seqs(scoring_, q, queryLen, pp.data(), pl.data(), (uint32_t)pc.n,
                                  sc.data(), scratch);
            for (size_t k = 0; k < pc.n; ++k) {
                const uint32_t i = order[pc.first + k];
                int v = sc[k];
                if (v < 0) {   // beyond int16: exact int32 after the workers finish
                    recalcWorker->pushOverflowSeq(reinterpret_cast<const char*>(ptr[i]), len[i],
                                                  base + i);
                    ++pushed;
                    v = 0;
                }
                out[i] = v;
            }
        }
    };
    std::vector<std::thread> team;
    for (int t = 1; t < threads_; ++t) team.emplace_back(work);
    work();
    for (auto& t : team) t.join();
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
