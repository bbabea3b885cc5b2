// host_stress.cpp -- drives the threaded host code of libswhybrid_host.so
// (no device): the DataPool and its fill threads, the CPU worker with its aux
// thread and scan team, the RecalcWorker list, ResultList merges, the
// packed-database scan, and the database files, from two caller threads at
// once.  Built with -fsanitize=thread and with -fsanitize=address,undefined
// (tests/test_sanitizers.py); every score is checked against the direct scan.
// Prints "ok" on success.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "Matrices.h"
#include "swhybrid_host.h"

namespace {

struct Db {
    std::vector<uint8_t> res;
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
};

Db random_db(uint64_t seed, int n, int max_len) {
    std::mt19937_64 rng(seed);
    Db d;
    for (int i = 0; i < n; ++i) {
        const uint32_t L = 1 + (uint32_t)(rng() % (uint64_t)max_len) / (1 + (uint32_t)(rng() % 4));
        d.off.push_back(d.res.size());
        d.len.push_back(L);
        for (uint32_t k = 0; k < L; ++k) d.res.push_back((uint8_t)(rng() % 20));
    }
    return d;
}

int fails = 0;
void check(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL %s: %s\n", what, swhy_last_error());
        ++fails;
    }
}

// The pipeline (DataPool -> CPU worker -> merged top-k) against the direct
// scan, with segments on and off.
void pipeline_case(uint64_t seed, const int32_t* m) {
    const Db d = random_db(seed, 2500, 1200);
    std::mt19937_64 rng(seed + 1);
    std::vector<uint8_t> q(200 + rng() % 900);
    for (auto& c : q) c = (uint8_t)(rng() % 20);
    std::vector<int32_t> direct(d.len.size()), direct_seg(d.len.size()), piped(d.len.size());
    double s1 = 0, s2 = 0;
    check(swhy_cpu_search(d.res.data(), d.off.data(), d.len.data(), d.len.size(), q.data(),
                          (int)q.size(), m, 10, 2, 3, -1, direct.data(), &s1, &s2) == 0,
          "cpu search");
    check(swhy_cpu_search(d.res.data(), d.off.data(), d.len.data(), d.len.size(), q.data(),
                          (int)q.size(), m, 10, 2, 2, 97, direct_seg.data(), &s1, &s2) == 0,
          "cpu search, segments");
    check(direct == direct_seg, "segmented scan differs");
    swhy_pool_opts o{10, 0, 3, 0, 60000};
    std::vector<int32_t> ts(10);
    std::vector<int64_t> ti(10);
    swhy_pool_report rep{};
    check(swhy_pool_search_ex(d.res.data(), d.off.data(), d.len.data(), d.len.size(), q.data(),
                              (int)q.size(), m, 10, 2, &o, ts.data(), ti.data(), piped.data(),
                              &rep) == 0,
          "pool search");
    check(piped == direct, "pipeline scores differ from the direct scan");
    check(rep.cpu_residues > 0 && rep.hip_residues == 0, "the CPU worker did the work");
}

void dbfile_case(const char* dir, int tag, const int32_t* m) {
    const std::string fa = std::string(dir) + "/db" + std::to_string(tag) + ".fa";
    const std::string base = std::string(dir) + "/db" + std::to_string(tag);
    const char* letters = "ARNDCQEGHILKMFPSTWYV";
    std::mt19937_64 rng(700 + tag);
    FILE* f = std::fopen(fa.c_str(), "w");
    for (int i = 0; i < 700; ++i) {
        std::fprintf(f, ">s%d some title\n", i);
        const int L = 1 + (int)(rng() % 700);
        for (int k = 0; k < L; ++k) std::fputc(letters[rng() % 20], f);
        std::fputc('\n', f);
    }
    std::fclose(f);
    uint64_t records = 0, slots = 0;
    check(swhy_dbfile_make(fa.c_str(), base.c_str(), 1 << 16, &records, &slots) == 0, "dbmaker");
    swhy_dbfile* h = nullptr;
    check(swhy_dbfile_open(base.c_str(), &h) == 0, "dbfile open");
    if (!h) return;
    const uint8_t* res = nullptr;
    const uint64_t* off = nullptr;
    const uint32_t* len = nullptr;
    check(swhy_dbfile_arrays(h, &res, &off, &len) == 0, "dbfile arrays");
    std::vector<uint8_t> q(333);
    for (auto& c : q) c = (uint8_t)(rng() % 20);
    std::vector<int32_t> direct(slots), piped(slots);
    double s1 = 0, s2 = 0;
    check(swhy_cpu_search(res, off, len, slots, q.data(), (int)q.size(), m, 10, 2, 2, 0,
                          direct.data(), &s1, &s2) == 0,
          "cpu search over the file");
    swhy_pool_opts o{5, 0, 2, 0, 20000};
    std::vector<int32_t> ts(5);
    std::vector<int64_t> ti(5);
    swhy_pool_report rep{};
    check(swhy_dbfile_pool_search_ex(h, q.data(), (int)q.size(), m, 10, 2, &o, ts.data(),
                                     ti.data(), piped.data(), &rep) == 0,
          "pool search over the file");
    check(piped == direct, "file pipeline scores differ");
    char title[64];
    uint64_t tl = 0;
    check(swhy_dbfile_title(h, ti[0], title, sizeof(title), &tl) == 0 && tl > 0, "title");
    swhy_dbfile_close(h);
}

}  // namespace

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : "/tmp";
    int32_t m[1024];
    if (!swhy::named_matrix("blosum62", m)) return 2;
    // two caller threads at once, each with its own pipelines and files
    std::thread a([&] {
        for (int r = 0; r < 3; ++r) pipeline_case(100 + r, m);
        dbfile_case(dir, 1, m);
    });
    std::thread b([&] {
        for (int r = 0; r < 3; ++r) pipeline_case(200 + r, m);
        dbfile_case(dir, 2, m);
    });
    a.join();
    b.join();
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}
