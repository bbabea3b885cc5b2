// swhybrid -- protein database search from the command line, with the
// reference's flags and printout (src/Param.h:59-117, src/main.cpp:7-16,
// src/WorkerPool.cpp:107-155):
//   swhybrid -db <base> -q <queries.fasta> [-t top] [-o open] [-e extend]
//            [-g hip_workers] [-c cpu_threads] [-m blosum62|blosum50] [-batch]
// One query record runs through the worker pipeline (DataPool over the
// database files -> HIPWorkers, and a CPU worker when -c > 0); several
// records, or -batch, are searched in one persistent launch over the
// database held resident in HBM.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Matrices.h"
#include "swh.h"
#include "swhybrid_host.h"

namespace {

struct Args {
    const char* db = nullptr;
    const char* query = nullptr;
    int top = 10, open = 10, extend = 2, gpus = 1, cpus = 0;
    const char* matrix = "blosum62";
    bool batch = false;
};

[[noreturn]] void usage() {
    std::fprintf(stderr,
                 "usage: swhybrid -db <protein database base> -q <query FASTA> [-t top] "
                 "[-o gap open] [-e gap extend] [-g HIP workers] [-c CPU worker threads] "
                 "[-m blosum62|blosum50] [-batch]\n");
    std::exit(2);
}

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        const std::string f = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) usage();
            return argv[++i];
        };
        if (f == "-db") a.db = val();
        else if (f == "-q") a.query = val();
        else if (f == "-t") a.top = std::atoi(val());
        else if (f == "-o") a.open = std::atoi(val());
        else if (f == "-e") a.extend = std::atoi(val());
        else if (f == "-g") a.gpus = std::atoi(val());
        else if (f == "-m") a.matrix = val();
        else if (f == "-batch") a.batch = true;
        else if (f == "-c") a.cpus = std::atoi(val());
        else if (f == "-p") (void)val();   // Xeon Phi workers: none here
        else usage();
    }
    if (!a.db || !a.query || a.top < 1 || a.gpus < 0 || a.cpus < 0 || a.gpus + a.cpus == 0)
        usage();
    return a;
}

void print_hits(swhy_dbfile* db, const int32_t* score, const int64_t* index, int k) {
    std::printf("----------------top %d scores---------------\n", k);
    std::vector<char> t(4096);
    for (int i = 0; i < k; ++i) {
        uint64_t len = 0;
        swhy_dbfile_title(db, (uint64_t)index[i], t.data(), t.size(), &len);
        std::printf("score %d -- %s\n", score[i], t.data());
    }
    std::printf("--------------------------------------------\n");
}

int fail(const char* what, const char* msg) {
    std::fprintf(stderr, "swhybrid: %s: %s\n", what, msg);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const Args a = parse(argc, argv);
    int32_t matrix[32 * 32];
    if (!swhy::named_matrix(a.matrix, matrix)) return fail("matrix", a.matrix);
    swhy_fasta* fa = nullptr;
    if (swhy_fasta_open(a.query, &fa) != 0) return fail("query", swhy_last_error());
    uint64_t nq = 0;
    swhy_fasta_count(fa, &nq);
    if (nq == 0) return fail("query", "no FASTA record");
    std::vector<const uint8_t*> qs(nq);
    std::vector<int> qlens(nq);
    std::vector<const char*> qtitles(nq);
    for (uint64_t i = 0; i < nq; ++i) {
        uint64_t len = 0;
        swhy_fasta_record(fa, i, &qs[i], &len, &qtitles[i]);
        for (uint64_t j = 0; j < len; ++j)
            if (qs[i][j] == 0xFF) return fail(qtitles[i], "letter without a residue code");
        if (len == 0) return fail(qtitles[i], "empty query");
        qlens[i] = (int)len;
    }
    swhy_dbfile* db = nullptr;
    if (swhy_dbfile_open(a.db, &db) != 0) return fail("database", swhy_last_error());
    uint64_t records = 0, slots = 0, residues = 0, batches = 0;
    swhy_dbfile_info(db, &records, &slots, &residues, &batches);
    std::printf("Sequence Database size %2.2lfGB, %llu sequences\n",
                (double)residues / (1 << 30), (unsigned long long)records);
    std::vector<int32_t> ts(a.top);
    std::vector<int64_t> ti(a.top);
    if (nq == 1 && !a.batch) {
        std::printf("Query sequence length %d\n", qlens[0]);
        const swhy_pool_opts o{a.top, a.gpus, a.cpus, 0, 0, 0};
        swhy_pool_report rep{};
        if (swhy_dbfile_pool_search_ex(db, qs[0], qlens[0], matrix, a.open, a.extend, &o,
                                       ts.data(), ti.data(), nullptr, &rep) != 0)
            return fail("search", swhy_last_error());
        print_hits(db, ts.data(), ti.data(), a.top);
        std::printf("total calculation time: %lf, GCUPS %lf\n, total residue %llu\n",
                    rep.seconds, rep.gcups, (unsigned long long)residues);
        if (a.cpus > 0)
            std::printf("residues searched: HIP %llu, CPU %llu; rescored after the workers %llu\n",
                        (unsigned long long)rep.hip_residues, (unsigned long long)rep.cpu_residues,
                        (unsigned long long)rep.recalc);
    } else {
        const uint8_t* res = nullptr;
        const uint64_t* off = nullptr;
        const uint32_t* len = nullptr;
        swhy_dbfile_arrays(db, &res, &off, &len);
        swh_ctx* ctx = nullptr;
        swh_db* d = nullptr;
        if (a.gpus < 1) return fail("-batch", "needs a GPU (-g 1)");
        if (swh_ctx_create(0, matrix, a.open, a.extend, &ctx) != SWH_OK)
            return fail("device", swh_last_error());
        if (swh_db_create(ctx, res, off, len, slots, nullptr, &d) != SWH_OK)
            return fail("database upload", swh_last_error());
        std::vector<int32_t> scores(nq * slots);
        swh_stats st{};
        const auto t0 = std::chrono::steady_clock::now();
        if (swh_search_batch(ctx, d, qs.data(), qlens.data(), (int)nq, nullptr, scores.data(),
                             &st) != SWH_OK)
            return fail("search", swh_last_error());
        const double secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (uint64_t q = 0; q < nq; ++q) {
            std::printf("Query %s, length %d\n", qtitles[q], qlens[q]);
            swh_topk(scores.data() + q * slots, slots, a.top, ts.data(), ti.data());
            print_hits(db, ts.data(), ti.data(), a.top);
        }
        std::printf("total calculation time: %lf, GCUPS %lf\n, %llu queries, total residue %llu\n",
                    secs, st.cells / secs / 1e9, (unsigned long long)nq,
                    (unsigned long long)residues);
        swh_db_destroy(d);
        swh_ctx_destroy(ctx);
    }
    swhy_dbfile_close(db);
    swhy_fasta_close(fa);
    return 0;
}
