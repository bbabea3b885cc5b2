#include "RecalcWorker.h"

#include <chrono>
#include <string>

#include "swh.h"

namespace {
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
thread_local std::string g_msg;
}  // namespace

RecalcWorker::RecalcWorker(Param* params, int device)
    : params_(params), device_(device), result_(params->topNum) {}

RecalcWorker::~RecalcWorker() {}

void RecalcWorker::setQuery(const char* q, size_t qlen) { query_.assign(q, q + qlen); }

// Same contract as src/RecalcWorker.cpp:270-343: the sequence is copied with
// codes above 23 clamped to 23 and remembered under its pool index.
void RecalcWorker::pushOverflowSeq(const char* seq, size_t len, size_t idx) {
    std::lock_guard<std::mutex> g(mu_);
    offsets_.push_back(residues_.size());
    lengths_.push_back((uint32_t)len);
    for (size_t j = 0; j < len; ++j) {
        unsigned char c = (unsigned char)seq[j];
        residues_.push_back(c > 23 ? 23 : c);
    }
    idx_.push_back(idx);
}

void RecalcWorker::launch() {
    t0_ = now_s();
    if (!idx_.empty()) {
        swh_ctx* ctx = nullptr;
        swh_db* db = nullptr;
        std::vector<int32_t> sc(idx_.size());
        if (swh_ctx_create(device_, params_->plain_matrix, params_->gapOpen, params_->gapExtend,
                           &ctx) != SWH_OK ||
            swh_db_create(ctx, residues_.data(), offsets_.data(), lengths_.data(),
                          lengths_.size(), nullptr, &db) != SWH_OK ||
            swh_search(ctx, db, reinterpret_cast<const uint8_t*>(query_.data()),
                       (int)query_.size(), nullptr, sc.data(), nullptr) != SWH_OK) {
            g_msg = std::string("RecalcWorker: ") + swh_last_error() + "\n";
            if (db) swh_db_destroy(db);
            if (ctx) swh_ctx_destroy(ctx);
            throw g_msg.c_str();
        }
        for (size_t i = 0; i < idx_.size(); ++i) result_.push(sc[i], idx_[i]);
        scores_ = sc;
        swh_db_destroy(db);
        swh_ctx_destroy(ctx);
    }
    t1_ = now_s();
}
