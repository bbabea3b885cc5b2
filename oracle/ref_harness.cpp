// ref_harness.cpp -- drives the reference's own CPU kernels, compiled from
// their sources under the reference checkout, for golden vectors and the CPU
// baseline.  TEST INFRASTRUCTURE ONLY; built into oracle/_ref/ by
// oracle/Makefile and never linked into the product.
//
// The harness restates only the reference's *callers*, with a correct layout:
//   - batch interleave dst[pos * LANES + lane], pad code 23, length rounded up
//     to a multiple of 4 (AVXWorker::packDB intent, src/AVXWorker.cpp:67-302;
//     src/Worker.cpp:143-166 for the 16-lane SSE layout);
//   - per-thread profile/qtable setup (src/AVXWorker.cpp:325-351);
//   - the OpenMP team around avxCompute / Compute (src/AVXWorker.cpp:397-427,
//     src/SSEWorker.cpp:120-144);
//   - overflowed lanes appended to a 23-terminated stream with vec_pos
//     (src/RecalcWorker.cpp:270-343) and rescored by avxCompute16 /
//     Compute16 (src/RecalcWorker.cpp:184-257).
// Query codes are padded to a multiple of 8 with 23 (src/DataPool.cpp:329-337).
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>
#include <pthread.h>
#include <mm_malloc.h>

// Kernel entry points of the reference (declared here with the signatures of
// src/AVXKernel.h:7-39 and src/SSEKernel.h:7-49).
void makeDupMatrix(short* dst16, char* dst8, int* src);
void avxCompute(char* sseQuery, int qlen, int gapoe, int gape, char* score_matrix,
                char* deviceDBSeq, int* map, size_t batchNum, int* sseResult,
                int* overflow_indices, volatile int& overflow_cnt,
                volatile int& globalCounter, pthread_mutex_t& idxMutex,
                pthread_mutex_t& cntMutex, const int SCORE_LIMIT_7, char* g_profile,
                char** g_qtable, char* externHF);
void avxCompute16(short* dup_score_matrix, char* dbSeq, int* vecPos, int numSeqs,
                  int* result, int start_idx, char* sseQuery, size_t qlen);
void initLocks();
void destroyLocks();
void Compute(char* sseQuery, int qlen, int gapoe, int gape, char* score_matrix,
             char* deviceDBSeq, int* map, size_t batchNum, int* sseResult,
             int* overflow_indices, volatile int& overflow_cnt,
             volatile int& globalCounter, const int SCORE_LIMIT_7);
void Compute16(short* score_matrix, char* dbSeq, int* vecPos, int numSeqs, int* result,
               int start_idx, char* sseQuery, size_t qlen);

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Packed {
    char* buf = nullptr;
    size_t bytes = 0;
    std::vector<int> map;           // (byte offset, padded length) per batch
    std::vector<int64_t> lane_seq;  // sequence id per (batch, lane), -1 for filler
};

// Sort by length (descending, ties by index) and interleave LANES sequences
// per batch.
Packed pack(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
            int64_t nseq, int lanes) {
    std::vector<int64_t> order(nseq);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return lengths[a] > lengths[b]; });
    Packed p;
    int64_t nb = (nseq + lanes - 1) / lanes;
    size_t total = 0;
    for (int64_t b = 0; b < nb; ++b) {
        uint32_t L = lengths[order[b * lanes]];
        L = (L + 3) & ~3u;
        if (L == 0) L = 4;
        total += (size_t)L * lanes;
    }
    p.bytes = total;
    p.buf = (char*)_mm_malloc(total + 64, 64);
    size_t pos = 0;
    p.lane_seq.assign(nb * lanes, -1);
    for (int64_t b = 0; b < nb; ++b) {
        uint32_t L = lengths[order[b * lanes]];
        L = (L + 3) & ~3u;
        if (L == 0) L = 4;
        p.map.push_back((int)pos);
        p.map.push_back((int)L);
        for (int lane = 0; lane < lanes; ++lane) {
            int64_t k = b * lanes + lane;
            const uint8_t* src = nullptr;
            uint32_t n = 0;
            if (k < nseq) {
                int64_t s = order[k];
                p.lane_seq[k] = s;
                src = residues + offsets[s];
                n = lengths[s];
            }
            for (uint32_t i = 0; i < L; ++i) {
                uint8_t c = i < n ? src[i] : 23;
                if (c > 23) c = 23;
                p.buf[pos + (size_t)i * lanes + lane] = (char)c;
            }
        }
        pos += (size_t)L * lanes;
    }
    return p;
}

// 23-terminated rescore stream (src/RecalcWorker.cpp:294-318).
void push_overflow(std::vector<char>& stream, std::vector<int>& vec_pos, const uint8_t* seq,
                   uint32_t len) {
    vec_pos.push_back((int)stream.size());
    for (uint32_t i = 0; i < len; ++i) stream.push_back((char)(seq[i] > 23 ? 23 : seq[i]));
    if (len == 0 || stream.back() != 23)
        for (int t = 0; t < 4; ++t) stream.push_back(23);
}

}  // namespace

extern "C" {

// path 0 = AVX2 (avxCompute + avxCompute16), 1 = SSE (Compute + Compute16).
// Returns 0 on success.  out_overflow (may be NULL) flags sequences that the
// 8-bit pass sent to the int16 rescore.
int ref_search(int path, const uint8_t* query, int qlen, const int32_t* matrix, int gap_open,
               int gap_extend, const uint8_t* residues, const uint64_t* offsets,
               const uint32_t* lengths, int64_t nseq, int32_t* out_scores,
               uint8_t* out_overflow, int nthreads, double* sec_pass8, double* sec_rescore) {
    if (qlen <= 0 || nseq < 0) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const int lanes = path == 0 ? 32 : 16;
    int qpad = (qlen + 7) & ~7;
    std::vector<char> q(qpad, 23);
    for (int j = 0; j < qlen; ++j) q[j] = (char)query[j];

    int pm[1024];
    int hi = 0;
    for (int i = 0; i < 1024; ++i) {
        pm[i] = matrix[i];
        hi = std::max(hi, matrix[i]);
    }
    const int limit7 = 128 - hi;  // src/Worker.cpp:62-70
    const int gapoe = gap_open + gap_extend;

    Packed p = pack(residues, offsets, lengths, nseq, lanes);
    int64_t nb = (int64_t)p.map.size() / 2;
    std::vector<int> res(nb * lanes + lanes, 0);
    std::vector<int> ovf_idx(nb * lanes + lanes, 0);
    volatile int ovf_cnt = 0;
    volatile int counter = -1;

    double t0 = now_s();
    if (path == 0) {
        char* dup8 = (char*)_mm_malloc(32 * 64, 32);
        short* dup16 = (short*)_mm_malloc(32 * 64 * sizeof(short), 32);
        makeDupMatrix(dup16, dup8, pm);
        char** qtable = (char**)_mm_malloc(sizeof(char*) * (size_t)nthreads * qpad, 32);
        char* profile = (char*)_mm_malloc((size_t)nthreads * 4096, 32);
        char* externHF = (char*)_mm_malloc(64, 32);
        for (int t = 0; t < nthreads; ++t)
            for (int j = 0; j < qpad; ++j)
                qtable[(size_t)t * qpad + j] = profile + (size_t)t * 4096 + 4 * 32 * q[j];
        pthread_mutex_t m1, m2;
        pthread_mutex_init(&m1, nullptr);
        pthread_mutex_init(&m2, nullptr);
        if (nb > 0) {
#pragma omp parallel num_threads(nthreads)
            avxCompute(q.data(), qpad, gapoe, gap_extend, dup8, p.buf, p.map.data(), nb,
                       res.data(), ovf_idx.data(), ovf_cnt, counter, m1, m2, limit7, profile,
                       qtable, externHF);
        }
        pthread_mutex_destroy(&m1);
        pthread_mutex_destroy(&m2);
        _mm_free(dup8);
        _mm_free(dup16);
        _mm_free(qtable);
        _mm_free(profile);
        _mm_free(externHF);
    } else {
        char sm[1024];
        for (int i = 0; i < 1024; ++i) sm[i] = (char)pm[i];
        initLocks();
        if (nb > 0) {
#pragma omp parallel num_threads(nthreads)
            Compute(q.data(), qpad, gapoe, gap_extend, sm, p.buf, p.map.data(), nb, res.data(),
                    ovf_idx.data(), ovf_cnt, counter, limit7);
        }
        destroyLocks();
    }
    double t1 = now_s();

    // Scores of the 8-bit pass, keyed back to the caller's sequence order.
    for (int64_t k = 0; k < nb * lanes; ++k) {
        int64_t s = p.lane_seq[k];
        if (s >= 0) {
            out_scores[s] = res[k];
            if (out_overflow) out_overflow[s] = 0;
        }
    }
    // Overflowed lanes -> int16 rescore stream, in index order (deterministic).
    std::vector<int64_t> ovf;
    for (int t = 0; t < ovf_cnt; ++t) {
        int64_t s = p.lane_seq[ovf_idx[t]];
        if (s >= 0) ovf.push_back(s);
    }
    std::sort(ovf.begin(), ovf.end());
    std::vector<char> stream;
    std::vector<int> vec_pos;
    for (int64_t s : ovf) push_overflow(stream, vec_pos, residues + offsets[s], lengths[s]);
    stream.resize(stream.size() + 64, 23);
    int n = (int)ovf.size();
    std::vector<int> rres(n > 0 ? n : 1, 0);
    short sm16[1024];
    for (int i = 0; i < 1024; ++i) sm16[i] = (short)pm[i];
    if (n > 0) {
        int nt = std::min(nthreads, std::max(1, n / 16));
        int chunk = (n + nt - 1) / nt;
#pragma omp parallel for num_threads(nt)
        for (int t = 0; t < nt; ++t) {
            int start = t * chunk;
            int cnt = std::min(chunk, n - start);
            if (cnt <= 0) continue;
            if (path == 0)
                avxCompute16(sm16, stream.data(), vec_pos.data(), cnt, rres.data(), start,
                             q.data(), (size_t)qpad);
            else
                Compute16(sm16, stream.data(), vec_pos.data(), cnt, rres.data(), start, q.data(),
                          (size_t)qpad);
        }
    }
    double t2 = now_s();
    for (int t = 0; t < n; ++t) {
        out_scores[ovf[t]] = rres[t];
        if (out_overflow) out_overflow[ovf[t]] = 1;
    }
    _mm_free(p.buf);
    if (sec_pass8) *sec_pass8 = t1 - t0;
    if (sec_rescore) *sec_rescore = t2 - t1;
    return 0;
}

}  // extern "C"
