#include "DataPool.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <thread>

static const unsigned kBatchWidth = 128;   // src/Defs.h:15

DataPool::DataPool(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                   uint64_t nseq, const uint8_t* query, size_t qlen)
    : residues_(residues), offsets_(offsets), lengths_(lengths), query_(query, query + qlen) {
    order_.resize(nseq);
    std::iota(order_.begin(), order_.end(), 0u);
    std::stable_sort(order_.begin(), order_.end(),
                     [&](uint32_t a, uint32_t b) { return lengths[a] > lengths[b]; });
    uint64_t i = 0;
    while (i < nseq) {
        unsigned len = std::max<unsigned>(4, (lengths[order_[i]] + 3) & ~3u);
        unsigned n = 0;
        while (i + n < nseq && n < kBatchWidth &&
               std::max<unsigned>(4, (lengths[order_[i + n]] + 3) & ~3u) == len)
            ++n;
        batches_.push_back(Batch{i, n, len});
        totalPadded_ += (size_t)n * len;
        maxBatchBytes_ = std::max(maxBatchBytes_, (size_t)n * len);
        i += n;
    }
}

DataPool::DataPool(const swhy::DbFile& db, const uint8_t* query, size_t qlen)
    : residues_(db.residue_data()), offsets_(db.offsets()), lengths_(db.lengths()),
      query_(query, query + qlen) {
    if (db.slots() >= UINT32_MAX) throw "DataPool: too many sequences\n";
    order_.resize(db.slots());
    std::iota(order_.begin(), order_.end(), 0u);
    for (const auto& b : db.batches()) {
        if (b.num == 0) continue;
        batches_.push_back(Batch{b.first, b.num, b.len});
        totalPadded_ += (size_t)b.num * b.len;
        maxBatchBytes_ = std::max(maxBatchBytes_, (size_t)b.num * b.len);
    }
}

size_t DataPool::maxInfoPerChunk(size_t bufferSize) const {
    // every batch holds at least 4 bytes per sequence and one sequence
    return std::min(batches_.size(), bufferSize / 4 + 1);
}

void DataPool::rewind() {
    std::lock_guard<std::mutex> g(mu_);
    next_ = 0;
    handedOut_ = handedOutCpu_ = handedOutDev_ = 0;
    t0_ = 0;
}

bool DataPool::fillChunk(char* dst, batchInfo* info, size_t& infoSize, size_t& aaSize,
                         size_t bufferSize, bool cpuFlag, size_t& baseIndex) {
    infoSize = 0;
    aaSize = 0;
    size_t b0, b1;
    {
        // claim a run of batches under the lock, copy it outside
        std::lock_guard<std::mutex> g(mu_);
        if (next_ >= batches_.size()) return false;
        const double t = std::chrono::duration<double>(
                             std::chrono::steady_clock::now().time_since_epoch()).count();
        if (t0_ == 0) t0_ = t;
        if (cpuFlag) {
            const size_t remaining = totalPadded_ - handedOut_;
            if (remaining <= cpuReserve_) return false;
            // Take a CPU chunk only if it should finish before the device
            // workers, at the rates both kinds have drawn chunks so far, finish
            // what would be left: a slow last CPU chunk lengthens the search.
            const double el = t - t0_;
            if (el > 0 && handedOutDev_ > 0 && handedOutCpu_ > 0) {
                const double dev_rate = handedOutDev_ / el, cpu_rate = handedOutCpu_ / el;
                const double chunk = (double)std::min(bufferSize, remaining);
                if (chunk / cpu_rate > (remaining - chunk) / dev_rate) return false;
            }
        }
        b0 = next_;
        size_t used = 0;
        while (next_ < batches_.size()) {
            const Batch& b = batches_[next_];
            const size_t bytes = (size_t)b.num * b.len;
            if (used + bytes > bufferSize && next_ > b0) break;
            if (bytes > bufferSize) throw "error: readBufferSize too small\n";
            used += bytes;
            ++next_;
        }
        b1 = next_;
        handedOut_ += used;
        (cpuFlag ? handedOutCpu_ : handedOutDev_) += used;
    }
    baseIndex = batches_[b0].first;
    // copy the claimed batches; a large chunk is split over a few threads
    std::vector<size_t> at(b1 - b0 + 1, 0);
    for (size_t bi = b0; bi < b1; ++bi)
        at[bi - b0 + 1] = at[bi - b0] + (size_t)batches_[bi].num * batches_[bi].len;
    auto copy = [&](size_t lo, size_t hi) {
        for (size_t bi = lo; bi < hi; ++bi) {
            const Batch& b = batches_[bi];
            char* out = dst + at[bi - b0];
            for (unsigned k = 0; k < b.num; ++k) {
                const uint32_t id = order_[b.first + k];
                const uint32_t n = std::min<uint32_t>(lengths_[id], b.len);
                std::memcpy(out + (size_t)k * b.len, residues_ + offsets_[id], n);
                std::memset(out + (size_t)k * b.len + n, kPadCode, b.len - n);
            }
        }
    };
    const size_t used = at.back();
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = used >= (size_t(32) << 20) ? (fillThreads_ ? fillThreads_ : hw) : 1u;
    if (nt == 1 || b1 - b0 < 2 * nt) {
        copy(b0, b1);
    } else {
        // split by bytes: thread t takes the batches starting in its share
        std::vector<size_t> cut(nt + 1, b0);
        cut[nt] = b1;
        size_t bi = b0;
        for (unsigned t = 1; t < nt; ++t) {
            const size_t target = used * t / nt;
            while (bi < b1 && at[bi - b0] < target) ++bi;
            cut[t] = bi;
        }
        std::vector<std::thread> team;
        for (unsigned t = 1; t < nt; ++t) team.emplace_back(copy, cut[t], cut[t + 1]);
        copy(cut[0], cut[1]);
        for (auto& th : team) th.join();
    }
    for (size_t bi = b0; bi < b1; ++bi) info[infoSize++] = batchInfo(batches_[bi].num, batches_[bi].len);
    aaSize = used;
    return true;
}
