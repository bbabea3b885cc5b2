// swh_capi.cpp -- host side of libswh.so: contexts, the resident database
// layout, query profiles, kernel launches and score copy-back.
#include "swh.h"
#include "swh_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

using namespace swh;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return fail(SWH_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                               \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

template <typename F>
void parallel_for(int64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> ts;
    int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        ts.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : ts) t.join();
}

// Device buffer that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

int check_scoring(const int32_t* matrix, int gap_open, int gap_extend) {
    if (!matrix) return fail(SWH_ERR_ARG, "matrix is NULL");
    for (int i = 0; i < 1024; ++i)
        if (matrix[i] < -128 || matrix[i] > 127)
            return fail(SWH_ERR_ARG, "matrix entry %d = %d outside [-128, 127]", i, matrix[i]);
    if (gap_open < 0 || gap_extend < 0 || gap_open + gap_extend > 32767)
        return fail(SWH_ERR_ARG, "invalid gap penalties open=%d extend=%d", gap_open,
                    gap_extend);
    return SWH_OK;
}

}  // namespace

struct swh_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    int32_t matrix[1024];
    int gap_open = 10, gap_extend = 2, max_score = 0;
    DevBuf border, counters, prof, scores, ovf;
    PinnedBuf h_prof, h_scores;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t border_budget = size_t(8) << 30;   // bytes of HBM for strip borders
};

struct swh_db {
    swh_ctx* ctx = nullptr;
    uint64_t nseq = 0, residues = 0, padded = 0;
    int n_pairs = 0;            // group-mode pair blocks, blocks [0, n_pairs)
    int n_blocks = 0;           // all blocks; lane blocks are [n_pairs, n_blocks)
    uint32_t max_rows_lane = 0, max_rows_pair = 0;
    uint64_t padded_pair = 0, padded_lane = 0;   // padded residues per layout
    uint64_t residues_pair = 0;                  // real residues in pair blocks
    std::vector<BlockDesc> blocks;
    DevBuf words, d_blocks, perm, sorted_len, sorted_blk;
};

extern "C" {

const char* swh_last_error(void) { return g_err.c_str(); }

int swh_version(void) { return 1; }

int swh_device_count(int* count) {
    if (!count) return fail(SWH_ERR_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return SWH_OK;
}

int swh_ctx_create(int device, const int32_t* matrix, int gap_open, int gap_extend,
                   swh_ctx** out) {
    if (!out) return fail(SWH_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_scoring(matrix, gap_open, gap_extend);
    if (rc) return rc;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(SWH_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= n) return fail(SWH_ERR_ARG, "device %d out of range", device);
    HIP_TRY(hipSetDevice(device));
    auto* c = new swh_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return fail(SWH_ERR_HIP, "hipGetDeviceProperties failed");
    }
    c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SWH_ERR_HIP, "hipStreamCreate failed");
    }
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    if (c->counters.reserve(64 * sizeof(uint32_t)) != hipSuccess) {
        swh_ctx_destroy(c);
        return fail(SWH_ERR_NOMEM, "counter allocation failed");
    }
    rc = swh_ctx_set_scoring(c, matrix, gap_open, gap_extend);
    if (rc) {
        swh_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return SWH_OK;
}

int swh_ctx_destroy(swh_ctx* c) {
    if (!c) return SWH_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->border.release();
    c->counters.release();
    c->prof.release();
    c->scores.release();
    c->ovf.release();
    c->h_prof.release();
    c->h_scores.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SWH_OK;
}

int swh_ctx_set_scoring(swh_ctx* c, const int32_t* matrix, int gap_open, int gap_extend) {
    if (!c) return fail(SWH_ERR_ARG, "ctx is NULL");
    int rc = check_scoring(matrix, gap_open, gap_extend);
    if (rc) return rc;
    std::memcpy(c->matrix, matrix, sizeof(c->matrix));
    c->gap_open = gap_open;
    c->gap_extend = gap_extend;
    int hi = 0;
    for (int i = 0; i < 1024; ++i) hi = std::max(hi, matrix[i]);
    c->max_score = hi;
    return SWH_OK;
}

int swh_db_create(swh_ctx* c, const uint8_t* residues, const uint64_t* offsets,
                  const uint32_t* lengths, uint64_t nseq, const swh_db_opts* opts,
                  swh_db** out) {
    if (!c || !out) return fail(SWH_ERR_ARG, "ctx/out is NULL");
    *out = nullptr;
    if (nseq > 0 && (!residues || !offsets || !lengths))
        return fail(SWH_ERR_ARG, "residues/offsets/lengths is NULL");
    if (nseq >= (uint64_t)kNoSeq / 2) return fail(SWH_ERR_ARG, "too many sequences");
    uint32_t group_min = (opts && opts->group_min_len) ? opts->group_min_len : 2048u;
    HIP_TRY(hipSetDevice(c->device));
    auto* db = new swh_db();
    db->ctx = c;
    db->nseq = nseq;
    // sort by length, longest first (ties keep caller order): the work queues
    // hand out the longest blocks first
    std::vector<uint32_t> order(nseq);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return lengths[a] > lengths[b]; });
    uint64_t total = 0;
    uint64_t n_long = 0;
    for (uint64_t i = 0; i < nseq; ++i) {
        total += lengths[i];
        if (lengths[i] >= group_min) ++n_long;
    }
    db->residues = total;
    const int np = (int)((n_long + 1) / 2);
    const uint64_t lane_first = std::min<uint64_t>(nseq, (uint64_t)np * 2);
    const int nl = (int)((nseq - lane_first + kLaneSeqs - 1) / kLaneSeqs);
    db->n_pairs = np;
    db->n_blocks = np + nl;
    db->blocks.resize(db->n_blocks);
    uint64_t woff = 0;
    for (int b = 0; b < np; ++b) {   // pair blocks: word (row/4)*2 + half
        uint32_t L = lengths[order[(size_t)b * 2]];
        uint32_t rows = std::max<uint32_t>(4, (L + 3) & ~3u);
        db->blocks[b] = BlockDesc{woff, rows, (uint32_t)(b * 2)};
        woff += (uint64_t)rows / 4 * 2;
        db->max_rows_pair = std::max(db->max_rows_pair, rows);
        db->padded_pair += (uint64_t)rows * 2;
    }
    for (int b = 0; b < nl; ++b) {   // lane blocks: word ((row/4)*64 + lane)*2 + half
        uint64_t first = lane_first + (uint64_t)b * kLaneSeqs;
        uint32_t L = lengths[order[first]];
        uint32_t rows = std::max<uint32_t>(4, (L + 3) & ~3u);
        db->blocks[np + b] = BlockDesc{woff, rows, (uint32_t)first};
        woff += (uint64_t)rows / 4 * kWave * 2;
        db->max_rows_lane = std::max(db->max_rows_lane, rows);
        db->padded_lane += (uint64_t)rows * kLaneSeqs;
    }
    db->padded = db->padded_pair + db->padded_lane;
    for (uint64_t k = 0; k < lane_first; ++k) db->residues_pair += lengths[order[k]];
    const uint64_t nslots = lane_first + (uint64_t)nl * kLaneSeqs + 2;
    std::vector<uint32_t> words(std::max<uint64_t>(woff, 1));
    std::vector<uint32_t> perm(nslots, kNoSeq);
    std::vector<uint32_t> slen(nslots, 0);
    std::vector<uint32_t> sblk(nslots, 0);
    auto put_seq = [&](uint64_t s, int b, uint32_t* w, uint32_t rows, size_t stride,
                       size_t col) {
        const uint8_t* src = nullptr;
        uint32_t n = 0;
        sblk[s] = (uint32_t)b;
        if (s < nseq) {
            uint32_t id = order[s];
            perm[s] = id;
            src = residues + offsets[id];
            n = lengths[id];
            slen[s] = n;
        }
        for (uint32_t k = 0; k < rows / 4; ++k) {
            uint32_t v = 0;
            for (int r = 0; r < 4; ++r) {
                uint32_t i = 4 * k + r;
                uint32_t code = i < n ? src[i] : 23u;
                if (code > 23) code = 23;
                v |= code << (8 * r);
            }
            w[(size_t)k * stride + col] = v;
        }
    };
    parallel_for(db->n_blocks, [&](int64_t b0, int64_t b1) {
        for (int64_t b = b0; b < b1; ++b) {
            const BlockDesc& bd = db->blocks[b];
            uint32_t* w = words.data() + bd.word_off;
            if (b < np) {
                for (int half = 0; half < 2; ++half)
                    put_seq(bd.first + half, (int)b, w, bd.rows, 2, half);
            } else {
                for (int half = 0; half < 2; ++half)
                    for (int lane = 0; lane < kWave; ++lane)
                        put_seq(bd.first + (uint64_t)half * kWave + lane, (int)b, w, bd.rows,
                                (size_t)kWave * 2, (size_t)lane * 2 + half);
            }
        }
    });
    auto up = [&](DevBuf& d, const void* h, size_t bytes) -> hipError_t {
        hipError_t e = d.reserve(std::max<size_t>(bytes, 16));
        if (e != hipSuccess) return e;
        return hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice);
    };
    hipError_t e = up(db->words, words.data(), words.size() * 4);
    if (e == hipSuccess && db->n_blocks > 0)
        e = up(db->d_blocks, db->blocks.data(), db->n_blocks * sizeof(BlockDesc));
    if (e == hipSuccess) e = up(db->perm, perm.data(), perm.size() * 4);
    if (e == hipSuccess) e = up(db->sorted_len, slen.data(), slen.size() * 4);
    if (e == hipSuccess) e = up(db->sorted_blk, sblk.data(), sblk.size() * 4);
    if (e != hipSuccess) {
        swh_db_destroy(db);
        return fail(e == hipErrorOutOfMemory ? SWH_ERR_NOMEM : SWH_ERR_HIP,
                    "database upload failed: %s", hipGetErrorString(e));
    }
    *out = db;
    return SWH_OK;
}

int swh_db_destroy(swh_db* db) {
    if (!db) return SWH_OK;
    if (db->ctx) (void)hipSetDevice(db->ctx->device);
    db->words.release();
    db->d_blocks.release();
    db->perm.release();
    db->sorted_len.release();
    db->sorted_blk.release();
    delete db;
    return SWH_OK;
}

int swh_db_info(const swh_db* db, uint64_t* nseq, uint64_t* residues, uint64_t* device_bytes) {
    if (!db) return fail(SWH_ERR_ARG, "db is NULL");
    if (nseq) *nseq = db->nseq;
    if (residues) *residues = db->residues;
    if (device_bytes)
        *device_bytes = db->words.bytes + db->d_blocks.bytes + db->perm.bytes +
                        db->sorted_len.bytes + db->sorted_blk.bytes;
    return SWH_OK;
}

static int pick_width(int qlen, int requested) {
    if (requested == 16 || requested == 24 || requested == 32) return requested;
    // smallest padding; ties go to the narrower strip (more waves per SIMD)
    int best = 16, best_pad = ((qlen + 15) / 16) * 16;
    for (int w : {24, 32}) {
        int pad = ((qlen + w - 1) / w) * w;
        if (pad < best_pad) {
            best = w;
            best_pad = pad;
        }
    }
    return best;
}

// Group mode shape: columns per lane W and lanes per group G.  Prefers the
// choice that leaves the fewest idle lane-columns, counting the per-step
// overhead of a lane (about 13 VALU ops beside the 9.5 per packed cell).
static void pick_group(int qlen, int req_w, int req_g, int* W_out, int* G_out) {
    double best = -1;
    int bw = 16, bg = 64;
    for (int W : {16, 8}) {
        if (req_w && req_w != W) continue;
        const int qp = ((qlen + W - 1) / W) * W;
        const int strips = qp / W;
        int G = req_g > 0 ? std::min(req_g, kWave) : std::min(strips, kWave);
        const int passes = (strips + G - 1) / G;
        const int gpw = kWave / G;
        const double lane_use = (double)strips / ((double)passes * G) * (double)(gpw * G) / kWave;
        const double eff = lane_use * (9.0 * W) / (9.5 * W + 13.0) * ((double)qlen / qp);
        if (eff > best) {
            best = eff;
            bw = W;
            bg = G;
        }
    }
    *W_out = bw;
    *G_out = bg;
}

int swh_search(swh_ctx* c, swh_db* db, const uint8_t* query, int qlen,
               const swh_search_opts* opts, int32_t* scores, swh_stats* stats) {
    const double t_start = now_ms();
    if (!c || !db || !query || (!scores && db->nseq > 0))
        return fail(SWH_ERR_ARG, "NULL argument");
    if (db->ctx != c) return fail(SWH_ERR_ARG, "database belongs to another context");
    if (qlen <= 0) return fail(SWH_ERR_ARG, "query length %d", qlen);
    for (int j = 0; j < qlen; ++j)
        if (query[j] > 23) return fail(SWH_ERR_ARG, "query code %d at %d > 23", query[j], j);
    swh_search_opts o{};
    if (opts) o = *opts;
    if (o.group_width && o.group_width != 8 && o.group_width != 16)
        return fail(SWH_ERR_ARG, "group_width must be 8 or 16");
    if (o.group_lanes < 0 || o.group_lanes > kWave)
        return fail(SWH_ERR_ARG, "group_lanes must be in 1..64");
    HIP_TRY(hipSetDevice(c->device));
    swh_stats st{};
    st.cells = db->residues * (uint64_t)qlen;
    st.cells_group = db->residues_pair * (uint64_t)qlen;
    if (db->nseq == 0) {
        st.total_ms = now_ms() - t_start;
        if (stats) *stats = st;
        return SWH_OK;
    }
    const int Wl = pick_width(qlen, o.strip_width);
    const int qpad_l = ((qlen + Wl - 1) / Wl) * Wl;
    int Wg = 16, G = 64;
    pick_group(qlen, o.group_width, o.group_lanes, &Wg, &G);
    const int qpad_g = ((qlen + Wg - 1) / Wg) * Wg;
    const int qpad_r = ((qlen + 7) / 8) * 8;
    const int pstride = std::max(qpad_l, qpad_g);
    {
        const int passes = (qpad_g / Wg + G - 1) / G;
        const int gpw = kWave / G;
        (void)gpw;
        st.padded_cells = db->padded_lane * (uint64_t)qpad_l +
                          db->padded_pair * (uint64_t)passes * G * Wg;
    }

    // query profile [24][pstride] int16, zero beyond qlen
    const size_t prof_bytes = (size_t)kProfRows * pstride * sizeof(int16_t);
    HIP_TRY(c->h_prof.reserve(prof_bytes));
    HIP_TRY(c->prof.reserve(prof_bytes));
    int16_t* hp = (int16_t*)c->h_prof.p;
    for (int r = 0; r < kProfRows; ++r)
        for (int j = 0; j < pstride; ++j)
            hp[(size_t)r * pstride + j] = j < qlen ? (int16_t)c->matrix[query[j] * 32 + r] : 0;
    HIP_TRY(hipMemcpyAsync(c->prof.p, hp, prof_bytes, hipMemcpyHostToDevice, c->stream));

    const size_t sc_bytes = db->nseq * sizeof(int32_t);
    HIP_TRY(c->scores.reserve(sc_bytes));
    HIP_TRY(c->ovf.reserve(db->nseq * sizeof(uint32_t)));
    uint32_t* ctr = (uint32_t*)c->counters.p;
    HIP_TRY(launch_reset(ctr, 64, c->stream));
    const int exact_thr = 65535 - c->max_score;
    const int thr = (o.overflow_threshold > 0) ? std::min(o.overflow_threshold, exact_thr)
                                               : exact_thr;
    const int gap_oe = c->gap_open + c->gap_extend;

    // border scratch: the two scans run one after the other on the stream
    const int n_lane = db->n_blocks - db->n_pairs;
    const int segw_l = std::min(qpad_l, kSegCols);
    const int ps_l = segw_l + kProfPad;
    const size_t lds_l = (size_t)kProfRows * ps_l * sizeof(int16_t);
    const int items_l = (n_lane + kWavesPerWG - 1) / kWavesPerWG;
    int grid_l = c->num_cus * lane_scan_max_wg_per_cu(Wl, lds_l);
    const uint64_t stride_l = (uint64_t)db->max_rows_lane * kWave;
    const uint64_t per_wg_l = qpad_l > Wl ? stride_l * sizeof(uint2) * kWavesPerWG : 0;
    if (per_wg_l) grid_l = (int)std::min<uint64_t>(grid_l, std::max<uint64_t>(1, c->border_budget / per_wg_l));
    grid_l = std::max(1, std::min(grid_l, items_l));

    const int gpw = kWave / G;
    const int pass_cols = G * Wg;
    const int ps_g = std::min(pass_cols, qpad_g) + kProfPad;
    const size_t lds_g = (size_t)kProfRows * ps_g * sizeof(int16_t);
    const int items_g = (db->n_pairs + kWavesPerWG * gpw - 1) / (kWavesPerWG * gpw);
    int grid_g = c->num_cus * group_scan_max_wg_per_cu(Wg, lds_g);
    const uint64_t stride_g = (uint64_t)gpw * db->max_rows_pair;
    const uint64_t per_wg_g = qpad_g > pass_cols ? stride_g * sizeof(uint2) * kWavesPerWG : 0;
    if (per_wg_g) grid_g = (int)std::min<uint64_t>(grid_g, std::max<uint64_t>(1, c->border_budget / per_wg_g));
    grid_g = std::max(1, std::min(grid_g, items_g));
    HIP_TRY(c->border.reserve(std::max<uint64_t>(
        {16, per_wg_l * (uint64_t)grid_l, per_wg_g * (uint64_t)grid_g})));

    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    if (db->n_pairs > 0) {
        GroupArgs g{};
        g.words = (const uint32_t*)db->words.p;
        g.blocks = (const BlockDesc*)db->d_blocks.p;
        g.perm = (const uint32_t*)db->perm.p;
        g.prof = (const int16_t*)c->prof.p;
        g.scores = (int32_t*)c->scores.p;
        g.ovf_list = (uint32_t*)c->ovf.p;
        g.ovf_count = ctr + 1;
        g.work_counter = ctr + 3;
        g.border = (uint2*)c->border.p;
        g.border_stride = stride_g;
        g.border_rows = db->max_rows_pair;
        g.blk_begin = 0;
        g.blk_end = db->n_pairs;
        g.group = G;
        g.n_items = items_g;
        g.qlen = qlen;
        g.qpad = qpad_g;
        g.ps = ps_g;
        g.pstride = pstride;
        g.gap_oe = gap_oe;
        g.gap_e = c->gap_extend;
        g.ovf_threshold = thr;
        HIP_TRY(launch_group_scan(Wg, g, grid_g, lds_g, c->stream));
        st.n_launches += 1;
    }
    HIP_TRY(hipEventRecord(c->ev[4], c->stream));
    if (n_lane > 0) {
        ScanArgs a{};
        a.words = (const uint32_t*)db->words.p;
        a.blocks = (const BlockDesc*)db->d_blocks.p;
        a.perm = (const uint32_t*)db->perm.p;
        a.prof = (const int16_t*)c->prof.p;
        a.scores = (int32_t*)c->scores.p;
        a.ovf_list = (uint32_t*)c->ovf.p;
        a.ovf_count = ctr + 1;
        a.work_counter = ctr + 0;
        a.border = (uint2*)c->border.p;
        a.border_stride = stride_l;
        a.blk_begin = db->n_pairs;
        a.blk_end = db->n_blocks;
        a.n_items = items_l;
        a.qlen = qlen;
        a.qpad = qpad_l;
        a.ps = ps_l;
        a.pstride = pstride;
        a.gap_oe = gap_oe;
        a.gap_e = c->gap_extend;
        a.ovf_threshold = thr;
        HIP_TRY(launch_lane_scan(Wl, a, grid_l, lds_l, c->stream));
        st.n_launches += 1;
    }
    HIP_TRY(hipEventRecord(c->ev[1], c->stream));

    // int32 rescore of the overflow list (the kernel reads the count itself)
    uint32_t* h_cnt = (uint32_t*)c->h_scores.p;
    HIP_TRY(c->h_scores.reserve(std::max<size_t>(sc_bytes, 64)));
    h_cnt = (uint32_t*)c->h_scores.p;
    HIP_TRY(hipMemcpyAsync(h_cnt, ctr + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t n_ovf = *h_cnt;
    st.n_overflow = n_ovf;
    if (n_ovf > 0) {
        RescoreArgs r{};
        r.words = (const uint32_t*)db->words.p;
        r.blocks = (const BlockDesc*)db->d_blocks.p;
        r.perm = (const uint32_t*)db->perm.p;
        r.sorted_len = (const uint32_t*)db->sorted_len.p;
        r.sorted_blk = (const uint32_t*)db->sorted_blk.p;
        r.prof = (const int16_t*)c->prof.p;
        r.scores = (int32_t*)c->scores.p;
        r.ovf_list = (const uint32_t*)c->ovf.p;
        r.ovf_count = ctr + 1;
        r.work_counter = ctr + 2;
        r.n_lane_blocks_first = db->n_pairs;
        r.qpad = qpad_r;
        r.ps = std::min(qpad_r, kSegCols) + kProfPad;
        r.pstride = pstride;
        r.gap_oe = gap_oe;
        r.gap_e = c->gap_extend;
        const size_t lds_r = (size_t)kProfRows * r.ps * sizeof(int16_t);
        const uint32_t lmax = std::max(db->max_rows_lane, db->max_rows_pair);
        const uint64_t rstride = (uint64_t)lmax * kWave;
        const uint64_t r_per_wg = rstride * sizeof(uint2) * kWavesPerWG;
        int rgrid = (int)std::min<uint64_t>((n_ovf + 255) / 256,
                                            (uint64_t)c->num_cus * rescore_max_wg_per_cu(lds_r));
        rgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>(rgrid, c->border_budget / r_per_wg));
        HIP_TRY(c->border.reserve(r_per_wg * rgrid));
        r.border = (uint2*)c->border.p;
        r.border_stride = rstride;
        HIP_TRY(hipEventRecord(c->ev[2], c->stream));
        HIP_TRY(launch_rescore(r, rgrid, lds_r, c->stream));
        HIP_TRY(hipEventRecord(c->ev[3], c->stream));
        st.n_launches += 1;
    }
    HIP_TRY(hipMemcpyAsync(c->h_scores.p, c->scores.p, sc_bytes, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(scores, c->h_scores.p, sc_bytes);
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    st.kernel_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[4]));
    st.group_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[4], c->ev[1]));
    st.lane_ms = ms;
    if (n_ovf > 0) {
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        st.rescore_ms = ms;
    }
    st.total_ms = now_ms() - t_start;
    if (stats) *stats = st;
    return SWH_OK;
}

int swh_search_batch(swh_ctx* c, swh_db* db, const uint8_t* const* queries, const int* qlens,
                     int nq, const swh_search_opts* opts, int32_t* scores, swh_stats* stats) {
    if (!c || !db || (nq > 0 && (!queries || !qlens || !scores)))
        return fail(SWH_ERR_ARG, "NULL argument");
    swh_stats total{};
    const double t0 = now_ms();
    for (int i = 0; i < nq; ++i) {
        swh_stats st{};
        int rc = swh_search(c, db, queries[i], qlens[i], opts, scores + (size_t)i * db->nseq, &st);
        if (rc) return rc;
        total.kernel_ms += st.kernel_ms;
        total.group_ms += st.group_ms;
        total.lane_ms += st.lane_ms;
        total.cells_group += st.cells_group;
        total.rescore_ms += st.rescore_ms;
        total.cells += st.cells;
        total.padded_cells += st.padded_cells;
        total.n_overflow += st.n_overflow;
        total.n_launches += st.n_launches;
    }
    total.total_ms = now_ms() - t0;
    if (stats) *stats = total;
    return SWH_OK;
}

int swh_topk(const int32_t* scores, uint64_t n, int k, int32_t* out_score, int64_t* out_index) {
    if (k < 0 || (k > 0 && (!out_score || !out_index)) || (n > 0 && !scores))
        return fail(SWH_ERR_ARG, "bad top-k arguments");
    // (score desc, index asc) among scores > 0; a min-heap of the current k best
    auto worse = [](const std::pair<int32_t, int64_t>& a, const std::pair<int32_t, int64_t>& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    };
    std::vector<std::pair<int32_t, int64_t>> heap;
    heap.reserve(k + 1);
    for (uint64_t i = 0; i < n && k > 0; ++i) {
        const int32_t s = scores[i];
        if (s <= 0) continue;
        if ((int)heap.size() < k) {
            heap.emplace_back(s, (int64_t)i);
            std::push_heap(heap.begin(), heap.end(), worse);
        } else if (s > heap.front().first) {
            std::pop_heap(heap.begin(), heap.end(), worse);
            heap.back() = {s, (int64_t)i};
            std::push_heap(heap.begin(), heap.end(), worse);
        }
    }
    std::sort(heap.begin(), heap.end(), worse);
    for (int t = 0; t < k; ++t) {
        out_score[t] = t < (int)heap.size() ? heap[t].first : 0;
        out_index[t] = t < (int)heap.size() ? heap[t].second : 0;
    }
    return SWH_OK;
}

}  // extern "C"
