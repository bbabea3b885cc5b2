// host_capi.cpp -- C entry points of the host side: the whole worker
// pipeline (DataPool -> HIPWorker(s) -> RecalcWorker -> merged ResultList),
// the way the reference's main() drives WorkerPool (src/main.cpp:7-16), over
// a database in memory or in the reference's .seq/.map/.title files; the
// DBmaker; and FASTA query input.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "CpuKernel.h"
#include "DataPool.h"
#include "DbFile.h"
#include "Param.h"
#include "WorkerPool.h"
#include "swhybrid_host.h"

namespace {
thread_local std::string g_err;
// A HIPWorker's chunk schedule: its first chunk, then each one twice the one
// before, up to chunk.
struct ChunkPlan {
    size_t chunk, first;
};
// The default schedule, from the whole call (first chunk's fill and build
// included, which nothing overlaps) over the 5M-sequence database in host
// memory on MI355X.  A chunk is one launch: larger ones lose less to launch
// tails, and the reference's measure (first worker start to last stop) is
// best with 1 GB chunks at every query length.  But the first chunk's fill and
// build take 0.1 ms per MB before any search, and below query length 600 the
// host's fills (about 0.09 ms per MB with the build) outlast the searches.
// Whole call, GCUPS (256 MB / 512 MB / 1 GB fixed, 256 MB then 1 GB):
//   query 144:  1,218 / 1,273 / 1,039        464:  3,871 / 3,950 / 3,124
//   query 1000: 6,102-6,654 / 5,923-6,356 / 5,263-5,420 / 4,929
//   query 3000: 7,718-7,858 / 7,530 / 7,482-7,534 / 8,122
//   query 5478: 8,032-8,041 / 8,015-8,038 / 8,175-8,192 / 8,253
// A small first chunk holds only the longest sequences (the pool hands them
// out first), whose launch balances poorly, so it pays only where later
// chunks run long enough (from query length 2,500).
ChunkPlan default_chunks(int qlen) {
    constexpr size_t MB = size_t(1) << 20;
    if (qlen < 600) return ChunkPlan{512 * MB, 512 * MB};
    if (qlen < 2500) return ChunkPlan{256 * MB, 256 * MB};
    return ChunkPlan{1024 * MB, 256 * MB};
}
ChunkPlan chunk_plan(const swhy_pool_opts& o, int qlen) {
    ChunkPlan p = o.chunk_bytes ? ChunkPlan{o.chunk_bytes, o.chunk_bytes} : default_chunks(qlen);
    if (o.first_chunk_bytes) p.first = std::min<size_t>(o.first_chunk_bytes, p.chunk);
    return p;
}
}

namespace {

// Runs the worker pipeline over pool; top-k and scores by pool index
// translated through pool.originalIndex.
int run_pool(DataPool& pool, uint64_t nseq, const int32_t* matrix, int gap_open, int gap_extend,
             const swhy_pool_opts& o, int32_t* out_top_score, int64_t* out_top_index,
             int64_t* out_top_pool_index, int32_t* out_scores, swhy_pool_report* rep) {
    std::vector<int> m(matrix, matrix + 1024);
    Param params(m.data(), gap_open, gap_extend, o.topk, o.hip_workers, o.cpu_threads);
    std::vector<int> sink;
    if (out_scores) sink.assign(nseq, -1);
    std::vector<Result<int>> top;
    swhy_pool_report r{};
    const auto t0 = std::chrono::steady_clock::now();
    {
        const ChunkPlan plan = chunk_plan(o, (int)pool.getQueryLen());
        WorkerPool wp(&pool, &params, plan.chunk, o.cpu_chunk_bytes, plan.first);
        r.chunk_bytes = plan.chunk;
        r.first_chunk_bytes = plan.first;
        const int rc = wp.run(out_scores ? sink.data() : nullptr);
        if (rc != 0) {
            g_err = "worker pipeline: " + wp.error();
            return -2;
        }
        r.seconds = wp.seconds();
        r.gcups = wp.gcups();
        r.hip_residues = wp.residues(false);
        r.cpu_residues = wp.residues(true);
        r.hip_busy_s = wp.busySeconds(false);
        r.cpu_busy_s = wp.busySeconds(true);
        r.recalc = wp.recalcCount();
        const Worker::Phases ph = wp.devicePhases();
        r.dev_first_fill_s = ph.first_fill;
        r.dev_aux_fill_s = ph.aux_fill;
        r.dev_pack_s = wp.devicePackSeconds();
        r.dev_search_s = ph.search;
        r.dev_aux_wait_s = ph.aux_wait;
        r.dev_kernel_s = wp.kernelSeconds();
        r.dev_chunks = ph.chunks;
        top = wp.result().getResult();
    }
    r.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int t = 0; t < o.topk; ++t) {
        const bool real = t < (int)top.size() && top[t].score > 0;
        out_top_score[t] = t < (int)top.size() ? top[t].score : 0;
        out_top_index[t] = real ? (int64_t)pool.originalIndex(top[t].idx) : 0;
        if (out_top_pool_index) out_top_pool_index[t] = real ? (int64_t)top[t].idx : 0;
    }
    if (out_scores)
        for (uint64_t p = 0; p < nseq; ++p) out_scores[pool.originalIndex(p)] = sink[p];
    if (rep) *rep = r;
    return 0;
}

bool bad_opts(const swhy_pool_opts* o, int32_t* ts, int64_t* ti) {
    return !o || o->topk < 0 || o->hip_workers < 0 || o->cpu_threads < 0 ||
           o->hip_workers + o->cpu_threads == 0 || (o->topk > 0 && (!ts || !ti));
}

}  // namespace

struct swhy_dbfile {
    std::unique_ptr<swhy::DbFile> f;
};

struct swhy_fasta {
    std::vector<std::string> titles;
    std::vector<std::vector<uint8_t>> codes;
};

extern "C" {

const char* swhy_last_error(void) { return g_err.c_str(); }

int swhy_cpu_search(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                    uint64_t nseq, const uint8_t* query, int qlen, const int32_t* matrix,
                    int gap_open, int gap_extend, int nthreads, int query_segment,
                    int32_t* out_scores, double* out_seconds, double* out_pack_seconds) {
    if ((nseq > 0 && (!residues || !offsets || !lengths || !out_scores)) || !query || qlen <= 0 ||
        !matrix || nthreads < 0 || query_segment < -1 || gap_open < 0 || gap_extend < 0) {
        g_err = "swhy_cpu_search: bad arguments";
        return -1;
    }
    for (int j = 0; j < qlen; ++j)
        if (query[j] > 23) {
            g_err = "swhy_cpu_search: query code above 23";
            return -1;
        }
    for (int i = 0; i < 1024; ++i)
        if (matrix[i] < -128 || matrix[i] > 127) {
            g_err = "swhy_cpu_search: matrix entry outside [-128, 127]";
            return -1;
        }
    if (!swhy::cpu::avx2_available()) {
        g_err = "swhy_cpu_search: the host has no AVX2";
        return -3;
    }
    swhy::cpu::Scoring sc;
    std::memcpy(sc.matrix, matrix, sizeof(sc.matrix));
    sc.gap_open = gap_open;
    sc.gap_extend = gap_extend;
    sc.query_segment = query_segment;
    if (nthreads == 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<const uint8_t*> ptr(nseq);
    for (uint64_t i = 0; i < nseq; ++i) ptr[i] = residues + offsets[i];
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    swhy::cpu::PackedDb db;
    swhy::cpu::pack_db(ptr.data(), lengths, nseq, db, nthreads);
    const auto t1 = clock::now();
    swhy::cpu::scan_db(sc, query, qlen, db, out_scores, nthreads);
    const auto t2 = clock::now();
    if (out_pack_seconds) *out_pack_seconds = std::chrono::duration<double>(t1 - t0).count();
    if (out_seconds) *out_seconds = std::chrono::duration<double>(t2 - t1).count();
    return 0;
}

int swhy_pool_search(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                     uint64_t nseq, const uint8_t* query, int qlen, const int32_t* matrix,
                     int gap_open, int gap_extend, int topk, uint64_t chunk_bytes,
                     int n_workers, int32_t* out_top_score, int64_t* out_top_index,
                     int64_t* out_top_pool_index, int32_t* out_scores, double* out_seconds,
                     double* out_gcups) {
    if (!query || qlen <= 0 || !matrix || topk < 0 || n_workers < 1 ||
        (nseq > 0 && (!residues || !offsets || !lengths)) ||
        (topk > 0 && (!out_top_score || !out_top_index))) {
        g_err = "swhy_pool_search: bad arguments";
        return -1;
    }
    try {
        DataPool pool(residues, offsets, lengths, nseq, query, (size_t)qlen);
        const swhy_pool_opts o{topk, n_workers, 0, chunk_bytes, 0, 0};
        swhy_pool_report r{};
        const int rc = run_pool(pool, nseq, matrix, gap_open, gap_extend, o, out_top_score,
                                out_top_index, out_top_pool_index, out_scores, &r);
        if (out_seconds) *out_seconds = r.seconds;
        if (out_gcups) *out_gcups = r.gcups;
        return rc;
    } catch (const char* e) {
        g_err = std::string("swhy_pool_search: ") + e;
    } catch (const std::exception& e) {
        g_err = std::string("swhy_pool_search: ") + e.what();
    }
    return -2;
}

int swhy_dbfile_make(const char* fasta_path, const char* out_base, uint64_t flush_bytes,
                     uint64_t* out_records, uint64_t* out_slots) {
    if (!fasta_path || !out_base) {
        g_err = "swhy_dbfile_make: NULL path";
        return -1;
    }
    try {
        const swhy::DbMakeStats st =
            swhy::make_db_files(fasta_path, out_base, flush_bytes ? flush_bytes : (uint64_t(1) << 28));
        if (out_records) *out_records = st.records;
        if (out_slots) *out_slots = st.slots;
        return 0;
    } catch (const std::exception& e) {
        g_err = std::string("swhy_dbfile_make: ") + e.what();
        return -2;
    }
}

int swhy_dbfile_open(const char* base, swhy_dbfile** out) {
    if (!base || !out) {
        g_err = "swhy_dbfile_open: NULL argument";
        return -1;
    }
    *out = nullptr;
    try {
        auto* h = new swhy_dbfile();
        h->f.reset(new swhy::DbFile(base));
        *out = h;
        return 0;
    } catch (const std::exception& e) {
        g_err = std::string("swhy_dbfile_open: ") + e.what();
        return -2;
    }
}

int swhy_dbfile_info(const swhy_dbfile* h, uint64_t* records, uint64_t* slots,
                     uint64_t* residues, uint64_t* batches) {
    if (!h) {
        g_err = "swhy_dbfile_info: NULL handle";
        return -1;
    }
    if (records) *records = h->f->records();
    if (slots) *slots = h->f->slots();
    if (residues) *residues = h->f->residues();
    if (batches) *batches = h->f->batches().size();
    return 0;
}

int swhy_dbfile_arrays(const swhy_dbfile* h, const uint8_t** residues, const uint64_t** offsets,
                       const uint32_t** lengths) {
    if (!h || !residues || !offsets || !lengths) {
        g_err = "swhy_dbfile_arrays: NULL argument";
        return -1;
    }
    *residues = h->f->residue_data();
    *offsets = h->f->offsets();
    *lengths = h->f->lengths();
    return 0;
}

int swhy_dbfile_title(swhy_dbfile* h, uint64_t index, char* buf, uint64_t cap, uint64_t* len) {
    if (!h || (cap > 0 && !buf)) {
        g_err = "swhy_dbfile_title: NULL argument";
        return -1;
    }
    const std::string t = h->f->title(index);
    if (len) *len = t.size();
    if (cap > 0) {
        const uint64_t n = std::min<uint64_t>(cap - 1, t.size());
        std::memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
    return 0;
}

int swhy_dbfile_close(swhy_dbfile* h) {
    delete h;
    return 0;
}

int swhy_dbfile_pool_search(swhy_dbfile* h, const uint8_t* query, int qlen,
                            const int32_t* matrix, int gap_open, int gap_extend, int topk,
                            uint64_t chunk_bytes, int n_workers, int32_t* out_top_score,
                            int64_t* out_top_index, int32_t* out_scores, double* out_seconds,
                            double* out_gcups) {
    if (!h || !query || qlen <= 0 || !matrix || topk < 0 || n_workers < 1 ||
        (topk > 0 && (!out_top_score || !out_top_index))) {
        g_err = "swhy_dbfile_pool_search: bad arguments";
        return -1;
    }
    try {
        DataPool pool(*h->f, query, (size_t)qlen);
        const swhy_pool_opts o{topk, n_workers, 0, chunk_bytes, 0, 0};
        swhy_pool_report r{};
        const int rc = run_pool(pool, h->f->slots(), matrix, gap_open, gap_extend, o,
                                out_top_score, out_top_index, nullptr, out_scores, &r);
        if (out_seconds) *out_seconds = r.seconds;
        if (out_gcups) *out_gcups = r.gcups;
        return rc;
    } catch (const char* e) {
        g_err = std::string("swhy_dbfile_pool_search: ") + e;
    } catch (const std::exception& e) {
        g_err = std::string("swhy_dbfile_pool_search: ") + e.what();
    }
    return -2;
}

int swhy_pool_search_ex(const uint8_t* residues, const uint64_t* offsets,
                        const uint32_t* lengths, uint64_t nseq, const uint8_t* query, int qlen,
                        const int32_t* matrix, const int gap_open, int gap_extend,
                        const swhy_pool_opts* opts, int32_t* out_top_score,
                        int64_t* out_top_index, int32_t* out_scores, swhy_pool_report* report) {
    if (!query || qlen <= 0 || !matrix || (nseq > 0 && (!residues || !offsets || !lengths)) ||
        bad_opts(opts, out_top_score, out_top_index)) {
        g_err = "swhy_pool_search_ex: bad arguments";
        return -1;
    }
    try {
        DataPool pool(residues, offsets, lengths, nseq, query, (size_t)qlen);
        return run_pool(pool, nseq, matrix, gap_open, gap_extend, *opts, out_top_score,
                        out_top_index, nullptr, out_scores, report);
    } catch (const char* e) {
        g_err = std::string("swhy_pool_search_ex: ") + e;
    } catch (const std::exception& e) {
        g_err = std::string("swhy_pool_search_ex: ") + e.what();
    }
    return -2;
}

int swhy_dbfile_pool_search_ex(swhy_dbfile* h, const uint8_t* query, int qlen,
                               const int32_t* matrix, int gap_open, int gap_extend,
                               const swhy_pool_opts* opts, int32_t* out_top_score,
                               int64_t* out_top_index, int32_t* out_scores,
                               swhy_pool_report* report) {
    if (!h || !query || qlen <= 0 || !matrix || bad_opts(opts, out_top_score, out_top_index)) {
        g_err = "swhy_dbfile_pool_search_ex: bad arguments";
        return -1;
    }
    try {
        DataPool pool(*h->f, query, (size_t)qlen);
        return run_pool(pool, h->f->slots(), matrix, gap_open, gap_extend, *opts, out_top_score,
                        out_top_index, nullptr, out_scores, report);
    } catch (const char* e) {
        g_err = std::string("swhy_dbfile_pool_search_ex: ") + e;
    } catch (const std::exception& e) {
        g_err = std::string("swhy_dbfile_pool_search_ex: ") + e.what();
    }
    return -2;
}

int swhy_fasta_open(const char* path, swhy_fasta** out) {
    if (!path || !out) {
        g_err = "swhy_fasta_open: NULL argument";
        return -1;
    }
    *out = nullptr;
    try {
        swhy::FastaReader r(path);
        auto* h = new swhy_fasta();
        std::string title, letters;
        while (r.next(title, letters)) {
            std::vector<uint8_t> c(letters.size());
            for (size_t i = 0; i < letters.size(); ++i)
                c[i] = swhy::residue_code((unsigned char)letters[i]);
            h->titles.push_back(title);
            h->codes.push_back(std::move(c));
        }
        *out = h;
        return 0;
    } catch (const std::exception& e) {
        g_err = std::string("swhy_fasta_open: ") + e.what();
        return -2;
    }
}

int swhy_fasta_count(const swhy_fasta* h, uint64_t* n) {
    if (!h || !n) {
        g_err = "swhy_fasta_count: NULL argument";
        return -1;
    }
    *n = h->titles.size();
    return 0;
}

int swhy_fasta_record(const swhy_fasta* h, uint64_t i, const uint8_t** codes, uint64_t* len,
                      const char** title) {
    if (!h || i >= h->titles.size()) {
        g_err = "swhy_fasta_record: no such record";
        return -1;
    }
    if (codes) *codes = h->codes[i].data();
    if (len) *len = h->codes[i].size();
    if (title) *title = h->titles[i].c_str();
    return 0;
}

int swhy_fasta_close(swhy_fasta* h) {
    delete h;
    return 0;
}

}  // extern "C"
