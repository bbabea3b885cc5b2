// swh_capi.cpp -- host side of libswh.so: contexts, the resident database
// layout, query profiles, kernel launches and score copy-back.
#include "swh_host.h"

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace swhh {
thread_local std::string g_err;

int refuse_if_served(const swh_ctx* c, const char* what) {
    if (c->server.load() != nullptr)
        return fail(SWH_ERR_STATE, "%s: a server is attached to this context (destroy it first)",
                    what);
    return SWH_OK;
}
}  // namespace swhh
using namespace swhh;

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

int swh_pinned_alloc(size_t bytes, void** out) {
    if (!out) return fail(SWH_ERR_ARG, "out is NULL");
    *out = nullptr;
    hipError_t e = hipHostMalloc(out, std::max<size_t>(bytes, 1), hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(SWH_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return SWH_OK;
}

int swh_pinned_free(void* p) {
    if (p) HIP_TRY(hipHostFree(p));
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
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->upload, hipStreamNonBlocking) != hipSuccess) {
        swh_ctx_destroy(c);
        return fail(SWH_ERR_HIP, "hipStreamCreate failed");
    }
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    if (c->counters.reserve(64 * sizeof(uint32_t)) != hipSuccess) {
        swh_ctx_destroy(c);
        return fail(SWH_ERR_NOMEM, "counter allocation failed");
    }
    // the allocation-free search: this device's scratch array (looked up once
    // per device) and its own fixed page-locked descriptor staging
    hipError_t se = scratch_address(device, &c->scratch, &c->scratch_bytes);
    if (se != hipSuccess || c->na_h_desc.reserve(kNoAllocDescBytes) != hipSuccess) {
        swh_ctx_destroy(c);
        return fail(SWH_ERR_HIP, "scratch set-up failed: %s", hipGetErrorString(se));
    }
    c->na_h_desc.fixed = true;
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
    int rc = refuse_if_served(c, "swh_ctx_destroy");
    if (rc) return rc;
    if (const int n = c->live_dbs.load())
        return fail(SWH_ERR_STATE, "swh_ctx_destroy: %d database layout(s) made on this context "
                    "still exist (swh_db_destroy them first)", n);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->upload) (void)hipStreamSynchronize(c->upload);
    c->graph.release();
    c->cache.release();
    c->stage_raw.release();
    c->stage_soff.release();
    c->h_stage.release();
    free_retired(c);
    c->border.release();
    c->counters.release();
    c->prof.release();
    c->scores.release();
    c->ovf.release();
    c->h_prof.release();
    c->h_scores.release();
    c->desc.release();
    c->h_desc.release();
    c->h_ctr.release();
    c->na_h_desc.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);

    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->upload) (void)hipStreamDestroy(c->upload);
    delete c;
    return SWH_OK;
}

int swh_ctx_set_scoring(swh_ctx* c, const int32_t* matrix, int gap_open, int gap_extend) {
    if (!c) return fail(SWH_ERR_ARG, "ctx is NULL");
    int rc = check_scoring(matrix, gap_open, gap_extend);
    if (rc) return rc;
    if ((rc = refuse_if_served(c, "swh_ctx_set_scoring"))) return rc;
    // a captured graph holds the old gaps and overflow threshold in its
    // kernel arguments: it must not be replayed with the new scoring
    c->graph.release();
    std::memcpy(c->matrix, matrix, sizeof(c->matrix));
    c->gap_open = gap_open;
    c->gap_extend = gap_extend;
    int hi = 0;
    for (int i = 0; i < 1024; ++i) hi = std::max(hi, matrix[i]);
    c->max_score = hi;
    return SWH_OK;
}

// Default group-mode threshold: a lane block runs its sequences' rows through
// the whole query on one wave, so sequences longer than about 0.4 of a
// sequence slot's share of the database (total residues / (2 sequences x 256
// lanes x 2 workgroups x CUs)) go to group mode, where G lanes share a pair;
// at least 2048.  On a 5M-sequence, 1.75G-residue database this is 2,672, on
// a 2M-sequence, 4.3G-residue one with a tail to 35,000 residues 6,540 (the
// factor 0.4 measured best of 0.33, 0.4 and 0.5 over those two).
static uint32_t default_group_min_len(const swh_ctx* c, const uint32_t* lengths, uint64_t n) {
    const uint64_t cus = (uint64_t)std::max(c->num_cus, 1);
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) total += lengths[i];
    const uint64_t t = total / (cus * 2560u);
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(t, 2048u), 1u << 20);
}

// A database of fewer than 1,024 sequences per CU cannot give every workgroup
// slot (2 per CU, 512 sequences per lane item) one lane item of a query, so a
// single query leaves CUs idle.  Such a database also gets a second layout in
// which every sequence of at least 128 residues is a pair, split over G lanes,
// for launches with fewer lane items than slots (100k sequences, 29.9M
// residues, query length 256: 848 GCUPS on the lane layout, 5,882 on this
// one; 64 queries in one launch: 8,920 on the lane layout, 5,572 on this one;
// a 512 MB worker chunk of the 5M-sequence database's longest 500k sequences,
// 977 lane items, runs 6% faster on the lane layout).
constexpr uint64_t kSmallDbSeqsPerCu = 1024;
// Staging buffers up to this size stay with the context between uploads.
constexpr size_t kStageKeepBytes = size_t(2) << 30;
// Host-to-device copies of up to 16 KB are carried out by a copy kernel, which
// waits for CUs: beside a search that holds every CU (the worker pipeline
// builds chunk k + 1 while chunk k is searched) it starts only when that search
// ends, and every copy queued after it on the stream waits too.  Copies of
// 64 KB and more go to the copy engines and run beside the search (measured
// on MI355X: 4 and 16 KB waited 156 ms for a 370 ms kernel, 64 KB to 4 MB did
// not wait).  A database build therefore copies at least this much each time.
constexpr size_t kMinEngineCopy = size_t(64) << 10;
constexpr uint32_t kSmallDbGroupMinLen = 128;

// At least one lane item (4 lane blocks) per workgroup slot, 2 slots per CU.
static bool enough_lane_items(const swh_ctx* c, uint64_t lane_blocks, int nq) {
    const uint64_t items = (uint64_t)nq * ((lane_blocks + kWavesPerWG - 1) / kWavesPerWG);
    return items >= 2u * (uint64_t)std::max(c->num_cus, 1);
}

}  // extern "C"

// Whether p is page-locked host memory (hipHostMalloc, swh_pinned_alloc) that
// a device copy can write directly.
static bool is_pinned_host(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // pageable memory: not an error of this call
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

namespace swhh {

template <typename B>
static hipError_t grow_retiring(swh_ctx* c, B& b, size_t want, bool exact,
                                std::vector<B> swh_ctx::*list) {
    if (want <= b.bytes) return hipSuccess;
    if (b.fixed) return hipErrorOutOfMemory;
    const B old = b;
    b.p = nullptr;
    b.bytes = 0;
    const size_t n = exact ? want : std::max(want, old.bytes + old.bytes / 2);
    hipError_t e = b.reserve(n);
    if (e == hipErrorOutOfMemory) {   // the retired buffers may hold what is missing
        (void)hipGetLastError();
        free_retired(c);
        e = b.reserve(n);
    }
    if (e != hipSuccess) {
        b = old;
        return e;
    }
    if (old.p) {
        bool over;
        {
            std::lock_guard<std::mutex> g(c->retired_mu);
            (c->*list).push_back(old);
            c->retired_bytes += old.bytes;
            over = c->retired_bytes > swh_ctx::kRetiredMax;
        }
        if (over) free_retired(c);
    }
    return hipSuccess;
}

hipError_t grow_beside(swh_ctx* c, DevBuf& b, size_t want, bool exact) {
    return grow_retiring(c, b, want, exact, &swh_ctx::retired_dev);
}

hipError_t grow_beside(swh_ctx* c, PinnedBuf& b, size_t want, bool exact) {
    return grow_retiring(c, b, want, exact, &swh_ctx::retired_host);
}

void free_retired(swh_ctx* c) {
    std::lock_guard<std::mutex> g(c->retired_mu);
    for (DevBuf& b : c->retired_dev) b.release();
    for (PinnedBuf& b : c->retired_host) b.release();
    c->retired_dev.clear();
    c->retired_host.clear();
    c->retired_bytes = 0;
}

}  // namespace swhh

// One layout of the database: sequences of at least group_min residues in
// pair blocks, the rest in lane blocks.
static int create_layout(swh_ctx* c, const uint8_t* residues, const uint64_t* offsets,
                         const uint32_t* lengths, uint64_t nseq, uint32_t group_min,
                         swh_db** out) {
    *out = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    static std::atomic<uint64_t> next_uid{1};
    auto* db = new swh_db();
    db->ctx = c;
    c->live_dbs.fetch_add(1);
    db->uid = next_uid.fetch_add(1);
    db->nseq = nseq;
    db->group_min_len = group_min;
    // Residues go to the device as they are (the span the offsets cover, or
    // a compact copy when that span is mostly gaps) and are interleaved there
    // by k_pack: the host only plans the layout.  Page-locked ones (a worker's
    // chunk) are queued first, so the copy engine carries them while the host
    // plans (512 MB: 9.3 ms of copy beside about 10 ms of planning).
    uint64_t total = 0, lo = UINT64_MAX, hi = 0;
    for (uint64_t i = 0; i < nseq; ++i) {
        total += lengths[i];
        if (lengths[i] > 0) {
            lo = std::min(lo, offsets[i]);
            hi = std::max(hi, offsets[i] + lengths[i]);
        }
    }
    if (lo > hi) lo = hi = 0;
    const bool gappy = hi - lo > 2 * total + (uint64_t(64) << 20);
    const bool direct = !gappy && hi - lo >= kMinEngineCopy && is_pinned_host(residues + lo);
    std::lock_guard<std::mutex> stage_lock(c->stage_mu);
    DevBuf& raw = c->stage_raw;
    hipError_t e = hipSuccess;
    if (direct) {
        const uint64_t n = hi - lo;
        e = grow_beside(c, raw, n + n / 4);
        if (e == hipSuccess)
            e = hipMemcpyAsync(raw.p, residues + lo, n, hipMemcpyHostToDevice, c->upload);
    }
    // sort by length, longest first (ties keep caller order): the work queues
    // hand out the longest blocks first
    std::vector<uint32_t> order(nseq);
    std::iota(order.begin(), order.end(), 0u);
    // (a DataPool chunk arrives sorted already)
    if (!std::is_sorted(lengths, lengths + nseq, std::greater<uint32_t>()))
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return lengths[a] > lengths[b]; });
    uint64_t n_long = 0;
    for (uint64_t i = 0; i < nseq; ++i) n_long += lengths[i] >= group_min;
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
    db->blk_real.assign(db->n_blocks + 1, 0);
    db->blk_padded.assign(db->n_blocks + 1, 0);
    for (int b = 0; b < db->n_blocks; ++b) {
        const BlockDesc& bd = db->blocks[b];
        const uint64_t n = b < np ? 2 : kLaneSeqs;
        uint64_t real = 0;
        for (uint64_t k = bd.first; k < bd.first + n && k < nseq; ++k) real += lengths[order[k]];
        db->blk_real[b + 1] = db->blk_real[b] + real;
        db->blk_padded[b + 1] = db->blk_padded[b] + (uint64_t)bd.rows * n;
    }
    // The tables are written straight into the page-locked staging, each at
    // an offset padded to a copy of at least kMinEngineCopy bytes, and go up
    // from there (no pageable copies to fill and copy again: 2.1M sequences of
    // a 512 MB chunk took 20 ms that way).  The staging only grows, with
    // slack; its old buffer is kept (grow_beside), not freed.
    auto padded = [](size_t bytes) {
        return (std::max(bytes, kMinEngineCopy) + 4095) / 4096 * 4096;
    };
    const uint64_t nslots = lane_first + (uint64_t)nl * kLaneSeqs + 2;
    std::vector<uint8_t> compact;   // the residues, when their span is mostly gaps
    const uint8_t* stage = residues + lo;
    uint64_t stage_bytes = hi - lo;
    if (gappy) stage_bytes = total;
    // other residues: small pageable ones through the staging; large
    // pageable ones (a database built once) as a prefix the runtime stages
    // itself, in multiples of kMinEngineCopy, and a padded tail through the
    // staging
    const bool split = !direct && stage_bytes > (size_t(64) << 20);
    const size_t head = split ? stage_bytes / kMinEngineCopy * kMinEngineCopy : 0;
    const size_t tables[5] = {db->n_blocks * sizeof(BlockDesc), nslots * 4, nslots * 4,
                              nslots * 4, std::max<uint64_t>(nseq, 1) * 8};
    size_t region[6];   // offsets in the staging: blocks, perm, slen, sblk, soff, residue tail
    size_t h_need = 0;
    for (int t = 0; t < 5; ++t) {
        region[t] = h_need;
        h_need += padded(tables[t]);
    }
    region[5] = h_need;
    if (!direct) h_need += padded(stage_bytes - head);
    if (e == hipSuccess) e = grow_beside(c, c->h_stage, h_need + h_need / 4);
    char* const hs = (char*)c->h_stage.p;
    if (e == hipSuccess) {
        if (tables[0]) std::memcpy(hs + region[0], db->blocks.data(), tables[0]);
        uint32_t* perm = (uint32_t*)(hs + region[1]);
        uint32_t* slen = (uint32_t*)(hs + region[2]);
        uint32_t* sblk = (uint32_t*)(hs + region[3]);
        uint64_t* soff = (uint64_t*)(hs + region[4]);
        // slots past every block (the two spare ones at the end)
        const uint64_t covered = std::max<uint64_t>((uint64_t)np * 2,
                                                    lane_first + (uint64_t)nl * kLaneSeqs);
        for (uint64_t s2 = covered; s2 < nslots; ++s2) {
            perm[s2] = kNoSeq;
            slen[s2] = 0;
            sblk[s2] = 0;
        }
        parallel_for(db->n_blocks, [&](int64_t b0, int64_t b1) {
            for (int64_t b = b0; b < b1; ++b) {
                const BlockDesc& bd = db->blocks[b];
                const uint64_t n = b < np ? 2 : kLaneSeqs;
                for (uint64_t s2 = bd.first; s2 < bd.first + n; ++s2) {
                    sblk[s2] = (uint32_t)b;
                    const bool real = s2 < nseq;
                    perm[s2] = real ? order[s2] : kNoSeq;
                    slen[s2] = real ? lengths[order[s2]] : 0;
                }
            }
        });
        soff[0] = 0;
        if (gappy) {
            uint64_t at = 0;
            for (uint64_t i = 0; i < nseq; ++i) {
                soff[i] = at;
                at += lengths[i];
            }
            compact.resize(total);
            parallel_for((int64_t)nseq, [&](int64_t i0, int64_t i1) {
                for (int64_t i = i0; i < i1; ++i)
                    std::memcpy(compact.data() + soff[i], residues + offsets[i], lengths[i]);
            });
            stage = compact.data();
        } else {
            parallel_for((int64_t)nseq, [&](int64_t i0, int64_t i1) {
                for (int64_t i = i0; i < i1; ++i)
                    soff[i] = lengths[i] > 0 ? offsets[i] - lo : 0;
            });
        }
        if (!direct && stage_bytes > head)
            std::memcpy(hs + region[5], stage + head, stage_bytes - head);
    }
    // region t of the staging into d at offset at, padded
    auto up = [&](DevBuf& d, size_t at, int t, size_t bytes, bool cached) -> hipError_t {
        const size_t n = padded(bytes);
        hipError_t r = cached ? c->cache.take(d, at + n) : grow_beside(c, d, at + n + n / 4);
        if (r != hipSuccess) return r;
        return hipMemcpyAsync((char*)d.p + at, hs + region[t], n, hipMemcpyHostToDevice,
                              c->upload);
    };
    DevBuf& d_soff = c->stage_soff;
    db->dev_bytes = woff * 4 + tables[0] + tables[1] + tables[2] + tables[3];
    if (e == hipSuccess) e = c->cache.take(db->words, std::max<uint64_t>(woff, 1) * 4);
    if (e == hipSuccess) e = up(db->d_blocks, 0, 0, tables[0], true);
    if (e == hipSuccess) e = up(db->perm, 0, 1, tables[1], true);
    if (e == hipSuccess) e = up(db->sorted_len, 0, 2, tables[2], true);
    if (e == hipSuccess) e = up(db->sorted_blk, 0, 3, tables[3], true);
    if (e == hipSuccess) e = up(d_soff, 0, 4, tables[4], false);
    if (e == hipSuccess && !direct) {
        // sized once, as up() below sizes it: the prefix copy must land in the
        // buffer the tail goes to
        const size_t tail_n = padded(stage_bytes - head);
        e = grow_beside(c, raw, head + tail_n + tail_n / 4);
        if (e == hipSuccess && head)
            e = hipMemcpyAsync(raw.p, stage, head, hipMemcpyHostToDevice, c->upload);
        if (e == hipSuccess) e = up(raw, head, 5, stage_bytes - head, false);
    }
    if (e == hipSuccess && db->n_blocks > 0)
        e = launch_pack((const BlockDesc*)db->d_blocks.p, db->n_blocks, np,
                        (const uint32_t*)db->perm.p, (const uint32_t*)db->sorted_len.p,
                        (const uint8_t*)raw.p, (const uint64_t*)d_soff.p,
                        (uint32_t*)db->words.p, c->upload);
    // on failure too: no copy may still read the staging when this returns
    const hipError_t se = hipStreamSynchronize(c->upload);
    if (e == hipSuccess) e = se;
    if (raw.bytes > kStageKeepBytes) {   // a one-off large load: not kept
        raw.release();
        d_soff.release();
    }
    if (e != hipSuccess) {
        swh_db_destroy(db);
        return fail(e == hipErrorOutOfMemory ? SWH_ERR_NOMEM : SWH_ERR_HIP,
                    "database upload failed: %s", hipGetErrorString(e));
    }
    *out = db;
    return SWH_OK;
}

extern "C" {

int swh_db_create(swh_ctx* c, const uint8_t* residues, const uint64_t* offsets,
                  const uint32_t* lengths, uint64_t nseq, const swh_db_opts* opts,
                  swh_db** out) {
    if (!c || !out) return fail(SWH_ERR_ARG, "ctx/out is NULL");
    *out = nullptr;
    if (nseq > 0 && (!residues || !offsets || !lengths))
        return fail(SWH_ERR_ARG, "residues/offsets/lengths is NULL");
    if (nseq >= (uint64_t)kNoSeq / 2) return fail(SWH_ERR_ARG, "too many sequences");
    const bool auto_layout = !opts || !opts->group_min_len;
    uint32_t group_min = auto_layout ? default_group_min_len(c, lengths, nseq)
                                     : opts->group_min_len;
    bool second = auto_layout && nseq > 0 &&
                  nseq < (uint64_t)std::max(c->num_cus, 1) * kSmallDbSeqsPerCu &&
                  group_min > kSmallDbGroupMinLen;
    if (second && opts && opts->queries_per_launch > 0) {
        // the caller's launches all pick the same layout: build only that one
        uint64_t n_lane = 0;
        for (uint64_t i = 0; i < nseq; ++i) n_lane += lengths[i] < group_min;
        const uint64_t blocks = (n_lane + kLaneSeqs - 1) / kLaneSeqs;
        if (!enough_lane_items(c, blocks, (int)opts->queries_per_launch))
            group_min = kSmallDbGroupMinLen;
        second = false;
    }
    swh_db* db = nullptr;
    int rc = create_layout(c, residues, offsets, lengths, nseq, group_min, &db);
    if (rc) return rc;
    if (second) {
        rc = create_layout(c, residues, offsets, lengths, nseq, kSmallDbGroupMinLen, &db->alt);
        if (rc) {
            swh_db_destroy(db);
            return rc;
        }
    }
    *out = db;
    return SWH_OK;
}

int swh_db_destroy(swh_db* db) {
    if (!db) return SWH_OK;
    if (db->served.load())
        return fail(SWH_ERR_STATE, "swh_db_destroy: a server searches this database "
                                   "(destroy the server first)");
    if (db->alt) swh_db_destroy(db->alt);
    if (db->ctx) {   // the buffers go to the context's spares (see BufCache)
        (void)hipSetDevice(db->ctx->device);
        for (DevBuf* b : {&db->words, &db->d_blocks, &db->perm, &db->sorted_len, &db->sorted_blk})
            db->ctx->cache.give(*b);
        db->ctx->live_dbs.fetch_sub(1);
    }
    delete db;
    return SWH_OK;
}

int swh_db_info(const swh_db* db, uint64_t* nseq, uint64_t* residues, uint64_t* device_bytes) {
    if (!db) return fail(SWH_ERR_ARG, "db is NULL");
    if (nseq) *nseq = db->nseq;
    if (residues) *residues = db->residues;
    if (device_bytes) {
        *device_bytes = 0;
        for (const swh_db* d = db; d; d = d->alt) *device_bytes += d->dev_bytes;
    }
    return SWH_OK;
}

}  // extern "C"

namespace swhh {

// Real / padded residues of blocks [b0, b1).
uint64_t range_residues(const swh_db* db, int b0, int b1) {
    return b1 > b0 ? db->blk_real[b1] - db->blk_real[b0] : 0;
}
uint64_t range_padded(const swh_db* db, int b0, int b1) {
    return b1 > b0 ? db->blk_padded[b1] - db->blk_padded[b0] : 0;
}

// Group mode shape for one query: lanes per group G at column width W, in at
// most max_passes passes over the query.  G maximises the useful share of a
// wave's lane-steps: strips / (passes x G) of a pass, times (64 / G) x G / 64
// of the wave; ties go to fewer passes.  With max_passes = 1 that is the
// fewest passes (G = strips, at most 64).  G is at least min_lanes (when the
// query has that many strips), which bounds the groups per wave whose
// borders a multi-pass item keeps (group_min_lanes).
int group_lanes_for(int qlen, int W, int max_passes, int min_lanes) {
    const int strips = (qlen + W - 1) / W;
    const int g_hi = std::min(strips, kWave);
    const int g_lo = std::max({1, std::min((strips + max_passes - 1) / max_passes, g_hi),
                               std::min(min_lanes, g_hi)});
    int best = g_hi;
    double best_use = -1.0;
    for (int G = g_hi; G >= g_lo; --G) {
        const int passes = (strips + G - 1) / G;
        if (passes > max_passes) continue;
        const double use =
            (double)strips / ((double)passes * G) * (double)((kWave / G) * G) / kWave;
        if (use > best_use + 1e-9) {
            best_use = use;
            best = G;
        }
    }
    return best;
}

static double group_lane_use(int qlen, int W, int max_passes, int min_lanes) {
    const int strips = (qlen + W - 1) / W;
    const int G = group_lanes_for(qlen, W, max_passes, min_lanes);
    const int passes = (strips + G - 1) / G;
    return (double)strips / ((double)passes * G) * (double)((kWave / G) * G) / kWave;
}

// Passes a group item may make over its query at width W: a pair's item runs
// passes x (rows + G) steps of W columns on one wave, and the longest pair's
// must stay within 0.3 of the launch's cells per lane (launch cells / (2
// sequences x 512 lanes x CUs)), since longest-first ordering starts it
// first.  On the 5M-sequence database (pairs up to 35,000 rows) that allows
// one pass at query length 464 and two at 1000 (W = 24); on the 2M-sequence
// overflow-stress database, 11 at 5478.
// Fewest lanes per group such that every wave slot of a full grid (at most 3
// workgroups of 4 waves per CU) can keep a border of its longest pair for
// each of its 64 / G groups within the context's border budget, so that more
// passes never shrink the grid.
int group_min_lanes(const swh_ctx* c, uint32_t rows_max, uint64_t border_budget) {
    const uint64_t per_group = (uint64_t)std::max(c->num_cus, 1) * 3 * kWavesPerWG *
                               sizeof(uint2) * std::max<uint32_t>(rows_max, 1);
    const uint64_t m = std::max<uint64_t>(1, std::min<uint64_t>(kWave, border_budget / per_group));
    return (int)(kWave / (m + 1)) + 1;
}

// n elements in a fixed inline array when n <= N, else on the heap: the
// launch planning of a few queries allocates nothing.
template <typename T, size_t N>
struct SmallBuf {
    T local[N];
    std::vector<T> heap;
    T* p;
    explicit SmallBuf(size_t n) : p(local) {
        if (n > N) {
            heap.resize(n);
            p = heap.data();
        }
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

int group_pass_budget(const swh_ctx* c, uint64_t launch_cells, uint32_t rows_max, int W) {
    if (rows_max == 0) return kWave;
    const double per_lane = (double)launch_cells / (2.0 * 512.0 * std::max(c->num_cus, 1));
    const double item = ((double)rows_max + kWave) * W;
    return std::max(1, std::min(kWave, (int)(0.3 * per_lane / item)));
}

// Profile row stride of a query: a multiple of 96 covers every strip width
// (8, 16, 24, 32), so a profile built once serves any launch shape.
int profile_stride(int qlen) { return ((qlen + 95) / 96) * 96; }

// [24][pstride] int16 profile: s(q_j, r) for residue r, zero beyond qlen.
void build_profile(const int32_t* matrix, const uint8_t* q, int qlen, int16_t* out) {
    const int ps = profile_stride(qlen);
    for (int r = 0; r < kProfRows; ++r)
        for (int j = 0; j < ps; ++j)
            out[(size_t)r * ps + j] = j < qlen ? (int16_t)matrix[q[j] * 32 + r] : 0;
}

int check_opts(const swh_search_opts& o) {
    if (o.group_width && o.group_width != 8 && o.group_width != 16 && o.group_width != 24)
        return fail(SWH_ERR_ARG, "group_width must be 8, 16 or 24");
    if (o.group_width == 24 && o.strip_width && o.strip_width != 24)
        return fail(SWH_ERR_ARG, "group_width 24 runs with strip_width 24 only");
    if (o.group_lanes < 0 || o.group_lanes > kWave)
        return fail(SWH_ERR_ARG, "group_lanes must be in 1..64");
    if (o.strip_width && o.strip_width != 16 && o.strip_width != 24 && o.strip_width != 32)
        return fail(SWH_ERR_ARG, "strip_width must be 16, 24 or 32");
    if (o.profile_mode < SWH_PROFILE_AUTO || o.profile_mode > SWH_PROFILE_PAIR)
        return fail(SWH_ERR_ARG, "profile_mode must be 0, 1 or 2");
    if (o.profile_mode == SWH_PROFILE_PAIR && o.strip_width == 32)
        return fail(SWH_ERR_ARG, "the pair-table profile needs strip_width 16 or 24");
    return SWH_OK;
}

int check_query(const uint8_t* q, int qlen, int i) {
    if (!q) return fail(SWH_ERR_ARG, "query %d is NULL", i);
    if (qlen <= 0) return fail(SWH_ERR_ARG, "query %d length %d", i, qlen);
    for (int j = 0; j < qlen; ++j)
        if (q[j] > 23) return fail(SWH_ERR_ARG, "query %d code %d at %d > 23", i, q[j], j);
    return SWH_OK;
}

// Tuning-table row of a query length: the first bucket that holds it, else
// the last; -1 without a table.
static int tune_bucket(const swh_ctx* c, int qlen) {
    const int n = (int)c->tune.size();
    if (n == 0) return -1;
    for (int b = 0; b < n; ++b)
        if (qlen <= c->tune[b].qlen_max) return b;
    return n - 1;
}

// Lanes per group for one query: the caller's, else the tuned value of its
// bucket when that was measured at this group width (measured at the bucket's
// longest query, so at most one lane per strip of a shorter one), else the
// best lane use within the launch's pass budget (group_lanes_for).
int group_lanes(const swh_ctx* c, int qlen, int Wg, const swh_search_opts& o, int max_passes,
                int min_lanes) {
    if (o.group_lanes > 0) return std::min(o.group_lanes, kWave);
    const int b = tune_bucket(c, qlen);
    if (b >= 0 && c->tune[b].group_width == Wg && c->tune[b].group_lanes > 0)
        return std::min(c->tune[b].group_lanes, group_lanes_for(qlen, Wg, 1, 1));
    return group_lanes_for(qlen, Wg, max_passes, min_lanes);
}

// Lane columns a query costs at strip width W: it is computed to its length
// rounded up to 8, full strips of W and a narrower last strip, which pays a
// strip row's fixed work (table row loads, border, row maximum) over fewer
// columns: 8 columns cost 1.30 of a full strip's per column, 16 1.06, 24 1.02
// (from the instructions per row of each width's loop).
double lane_cols(int qlen, int W) {
    static const double kRemCost[4] = {0.0, 1.30, 1.06, 1.02};
    const int q8 = ((qlen + 7) / 8) * 8;
    const int full = q8 / W * W;
    const int rem = q8 - full;
    return full + rem * kRemCost[rem / 8];
}

// Built-in cost per lane column of each lane configuration (order of
// kLaneCfgs), measured on MI355X over a 5M-sequence database at query lengths
// 144-5478: the pair table at W=24 is 1.8-3.5% cheaper per column than at
// W=16 (fewer H/E round trips through LDS per cell); the single-residue
// profile costs 3-11% more.
static double builtin_cost(int k) {
    switch (k) {
        case 0: return 1.0;      // (16, pair)
        case 1: return 0.978;    // (24, pair)
        case 2: return 1.11;     // (16, single)
        case 3: return 1.05;     // (24, single)
        default: return 1.03;    // (32, single)
    }
}

// Time model of a group-mode launch, shared by the group width and pass
// choices (group_width_by_latency).  A pair is swept by one wave, P passes of
// rows + G - 1 steps; a step of W columns is priced at 105, 159 and 238 VALU
// instructions at W = 8, 16, 24 and about 6 cycles an instruction for a wave
// that shares its SIMD (the compiled rows with their wait states when the
// model was fit to measured launches; the rows now compile to 6% and 9% fewer
// issue slots at W = 16 and 24, and the constants stay with the fit).
// Throughput: the launch's cells at 9.4 T cells/s (per 256 CUs), the group
// share of them divided by the lanes' use and by what group mode reaches of
// that rate at each width (1 / 1.45, 1 / 1.55, 1 / 1.85 at W = 24, 16, 8).
struct GroupLaunchModel {
    double t_tp = 0;   // the launch's cells at the scan's full rate, seconds
    double gf = 0;     // share of the cells in group pairs
    uint32_t rows_max = 0;
    static double step(int w) { return w == 8 ? 105.0 : w == 16 ? 159.0 : 238.0; }
    static double k(int w) { return w == 8 ? 1.85 : w == 16 ? 1.55 : 1.45; }
    static double use(int qlen, int w, int G) {
        const int strips = (qlen + w - 1) / w;
        const int passes = (strips + G - 1) / G;
        return (double)strips / ((double)passes * G) * (double)((kWave / G) * G) / kWave;
    }
    double pass_s(int w, int G) const {
        return ((double)rows_max + G - 1) * step(w) * 6.0 / 2.4e9;
    }
    double throughput_s(int w, double lane_use) const {
        return t_tp * (k(w) * gf / lane_use + (1.0 - gf));
    }
};

static GroupLaunchModel group_model(const swh_ctx* c, const swh_db* db, uint64_t launch_cells) {
    GroupLaunchModel m;
    m.t_tp = (double)launch_cells / (9.4e12 * std::max(c->num_cus, 1) / 256.0);
    m.gf = db->residues ? (double)db->residues_pair / (double)db->residues : 0.0;
    m.rows_max = db->max_rows_pair;
    return m;
}

// Passes a query's group items may make: the launch's pass budget
// (group_pass_budget), unless the best group within it leaves more than a
// tenth of the wave's lanes idle; then as many as keep the longest pair's
// sweep within kSweepShare of the launch's predicted time.  On 100k sequences
// of up to 2,000 residues the budget allows one pass: 42 and 64 lanes a group
// at query length 1000 and 2005 (7.0 and 14.1 ms), against 21 lanes in 2 and
// 4 passes (4.4 and 8.9 ms, the longest pair's sweep 55% of the launch).  A
// worker's first chunk of the 5M-sequence database, its 117,771 longest
// sequences, all in pairs of up to 34,988 rows, at query length 5478: 64
// lanes in 4 passes; 21 lanes in 11, a sweep as long as the predicted launch,
// made the pipeline's kernels 11% slower.  On the whole 5M-sequence and the
// overflow-stress databases the budget's groups use 94% of the lanes or more.
constexpr double kSweepShare = 0.6;

int group_passes(const swh_ctx* c, const swh_db* db, uint64_t launch_cells, int qlen, int W,
                 int max_passes, int min_lanes) {
    const int G = group_lanes_for(qlen, W, max_passes, min_lanes);
    const double u = GroupLaunchModel::use(qlen, W, G);
    if (u >= 0.9 || db->max_rows_pair == 0) return max_passes;
    const GroupLaunchModel m = group_model(c, db, launch_cells);
    const double t_launch = m.throughput_s(W, 0.95);
    const int p = (int)(kSweepShare * t_launch / m.pass_s(W, kWave));
    return std::max(max_passes, std::min(kWave, p));
}

// Group width by the launch model (GroupLaunchModel): the larger of the
// throughput time and the longest pair's sweep.  Another width replaces W
// when its time is 5% below W's.  Measured on 100k sequences of up to 2,000
// residues, 90% of the cells in group pairs (ms at W = 8 / 16 / 24): query 50
// 0.53 / 0.81 / 1.05, 100 0.74 / 0.81 / 1.07, 144 0.95 / 0.82 / 1.11, 200
// 1.42 / 1.19 / 1.15, 256 1.47 / 1.22 / 1.37, 464 3.16 / 2.47 / 2.25.  On the
// 5M-sequence database (7% of the cells in pairs, pairs of up to 34,988 rows)
// the widths' times differ by less than 5% from query length 144 up, and
// from 200 up the width chosen by lane use stays: a 21 ms sweep at W = 24 is
// within kSweepShare of those launches.
static int group_width_by_latency(const swh_ctx* c, const int* qlens, int nq, const swh_db* db,
                                  int W) {
    if (db->n_pairs == 0 || db->max_rows_pair == 0 || db->residues == 0) return W;
    uint64_t cells = 0;
    int qmax = 0;
    for (int i = 0; i < nq; ++i) {
        cells += (uint64_t)qlens[i] * db->residues;
        qmax = std::max(qmax, qlens[i]);
    }
    const GroupLaunchModel m = group_model(c, db, cells);
    const int min_lanes = group_min_lanes(c, db->max_rows_pair, c->border_budget);
    struct Est {
        double tp, sweep;
        double t() const { return std::max(tp, sweep); }
        bool risky() const { return sweep > kSweepShare * tp; }
    };
    auto predict = [&](int w) {
        const int budget = group_pass_budget(c, cells, db->max_rows_pair, w);
        const int passes_max = group_passes(c, db, cells, qmax, w, budget, min_lanes);
        const int G = group_lanes_for(qmax, w, passes_max, min_lanes);
        const int strips = (qmax + w - 1) / w;
        const int passes = (strips + G - 1) / G;
        return Est{m.throughput_s(w, GroupLaunchModel::use(qmax, w, G)), passes * m.pass_s(w, G)};
    };
    const int widths[3] = {24, 16, 8};
    Est e[3];
    int bi = -1;
    for (int i = 0; i < 3; ++i) {
        e[i] = predict(widths[i]);
        if (widths[i] == W) bi = i;
    }
    if (bi < 0) return W;
    double t_best = e[bi].t() * 0.95;
    for (int i = 0; i < 3; ++i) {
        if (i == bi || widths[i] == 24) continue;   // 24 only ever stays, never replaces
        if (e[i].t() < t_best) {
            bi = i;
            t_best = e[i].t();
        }
    }
    // A sweep above kSweepShare of the throughput time leaves the launch's
    // end to one wave and its SIMD neighbours: at query length 144 on the
    // 5M-sequence database (a 21 ms sweep at W = 24 in a 27 ms launch) the
    // launch took 26.7-35 ms at W = 24, a steady 26.7 ms at W = 16.  A width
    // whose sweep fits the share replaces it when predicted within 5%.
    if (e[bi].risky()) {
        const double t_cap = e[bi].t() * 1.05;
        int alt = -1;
        for (int i = 0; i < 3; ++i) {
            if (i == bi || e[i].risky() || e[i].t() > t_cap) continue;
            if (widths[i] == 24 && W != 24) continue;
            if (alt < 0 || e[i].t() < e[alt].t()) alt = i;
        }
        if (alt >= 0) bi = alt;
    }
    return widths[bi];
}

// Launch shape shared by the queries of a launch.  Lane scan: the width and
// profile placement with the least predicted time, summed over the queries as
// cost per lane column x lane columns (lane_cols).  Costs come from the tuning
// table when one is installed, else from builtin_cost.  Group scan: the
// width of the bucket holding most of the batch's columns, else 16 or 24 by
// lane use (24 is instantiated with 24-column lanes only; 8 columns per lane
// lose 2% on the 5M-sequence database at query length 144), unless the
// launch is bound by its longest pair (group_width_by_latency).
Shape choose_shape(const swh_ctx* c, const int* qlens, int nq, const swh_search_opts& o,
                   const swh_db* db, uint64_t db_residues) {
    Shape sh{16, o.group_width, true};
    double best = 1e300;
    for (int k = 0; k < SWH_TUNE_LANE_CONFIGS; ++k) {
        const LaneCfg& L = kLaneCfgs[k];
        if (o.strip_width && o.strip_width != L.width) continue;
        if (o.group_width == 24 && L.width != 24) continue;   // built with 24-column lanes only
        if (o.profile_mode == SWH_PROFILE_PAIR && !L.pair) continue;
        if (o.profile_mode == SWH_PROFILE_SINGLE && L.pair) continue;
        double t = 0;
        for (int i = 0; i < nq; ++i) {
            const int b = tune_bucket(c, qlens[i]);
            double w = builtin_cost(k);
            if (b >= 0) {
                const float m = c->tune[b].lane_cost[k];
                w = m > 0 ? m : 1e30;   // not measured: never preferred
            }
            t += lane_cols(qlens[i], L.width) * w;
        }
        if (t < best) {
            best = t;
            sh.Wl = L.width;
            sh.pair = L.pair;
        }
    }
    if (!sh.Wg) {
        // 16, or 24 next to 24-column lanes where its lane use, weighted by its
        // 10% faster step (measured at equal lane use), covers more columns
        sh.Wg = 16;
        if (sh.Wl == 24) {
            uint64_t cells = 0;
            for (int i = 0; i < nq; ++i) cells += (uint64_t)qlens[i] * db_residues;
            const int p16 = group_pass_budget(c, cells, db->max_rows_pair, 16);
            const int p24 = group_pass_budget(c, cells, db->max_rows_pair, 24);
            const int gmin = group_min_lanes(c, db->max_rows_pair, c->border_budget);
            double u16 = 0, u24 = 0;
            for (int i = 0; i < nq; ++i) {
                u16 += qlens[i] * group_lane_use(qlens[i], 16, p16, gmin);
                u24 += qlens[i] * group_lane_use(qlens[i], 24, p24, gmin) * 1.10;
            }
            if (u24 > u16) sh.Wg = 24;
        }
        if (!c->tune.empty()) {
            double cols[SWH_TUNE_MAX] = {};
            const int nb = std::min((int)c->tune.size(), SWH_TUNE_MAX);
            for (int i = 0; i < nq; ++i) cols[tune_bucket(c, qlens[i])] += qlens[i];
            const int b = (int)(std::max_element(cols, cols + nb) - cols);
            if (c->tune[b].group_width > 0) sh.Wg = c->tune[b].group_width;   // 0: built-in
        }
        if (sh.Wg == 24 && sh.Wl != 24) sh.Wg = 16;   // a bucket tuned with other lanes
        // a tuned width was measured on a throughput-bound launch; a launch on
        // another database may be bound by its longest pair instead
        sh.Wg = group_width_by_latency(c, qlens, nq, db, sh.Wg);
    }
    return sh;
}

// Work items longest first across every task (LPT order): a merge of the
// per-task lists, each already longest first because blocks are sorted by
// length.  Duration model in VALU ops per lane: a lane item runs its first
// (longest) block's rows through every strip at 9.5 ops per packed cell plus
// 13 per strip row; a group item runs its first pair's rows (plus the lane
// skew of group_step_rows rows per lane) through every pass at 9.5 ops per
// cell plus 17 per row.
void order_items(const swh_db* db, const Shape& sh, const QueryDesc* lane_d,
                 const QueryDesc* group_d, int nt, uint2* out) {
    auto item_cost = [&](int list, int k) -> double {
        if (list < nt) {
            const QueryDesc& Gd = group_d[list];
            const int p = Gd.blk_begin + k * kWavesPerWG * (kWave / Gd.group);
            const int npass = (Gd.qpad / sh.Wg + Gd.group - 1) / Gd.group;
            const int nr = group_step_rows(sh.Wl, sh.Wg);
            return (double)npass * (db->blocks[p].rows + nr * (Gd.group - 1)) *
                   (9.5 * sh.Wg + 17.0);
        }
        const QueryDesc& L = lane_d[list - nt];
        const int b = L.blk_begin + k * kWavesPerWG;
        const int strips = (L.qpad + sh.Wl - 1) / sh.Wl;
        return (double)db->blocks[b].rows * (9.5 * L.qpad + 13.0 * strips);
    };
    struct Head {
        double cost;
        int list, k;
        bool operator<(const Head& h) const {   // max-heap on cost, then list order
            return cost != h.cost ? cost < h.cost : list > h.list;
        }
    };
    SmallBuf<Head, 16> heap(2 * (size_t)nt);
    size_t nh = 0;
    for (int l = 0; l < 2 * nt; ++l)
        if ((l < nt ? group_d[l].n_items : lane_d[l - nt].n_items) > 0)
            heap[nh++] = Head{item_cost(l, 0), l, 0};
    std::make_heap(heap.p, heap.p + nh);
    for (int n = 0; nh > 0; ++n) {
        std::pop_heap(heap.p, heap.p + nh);
        Head& h = heap[nh - 1];
        out[n] = h.list < nt ? make_uint2((uint32_t)h.list, (uint32_t)h.k)
                             : make_uint2((uint32_t)(h.list - nt) | kLaneItem, (uint32_t)h.k);
        const int n_list = h.list < nt ? group_d[h.list].n_items : lane_d[h.list - nt].n_items;
        if (++h.k < n_list) {
            h.cost = item_cost(h.list, h.k);
            std::push_heap(heap.p, heap.p + nh);
        } else {
            --nh;
        }
    }
}

// Enqueue, without waiting: counter reset, one scan launch for every task and
// the int32 rescore of their overflow lists.  counters needs 16 + nt words;
// counters[16 + i] ends up holding task i's overflow count.
int enqueue_tasks(swh_ctx* c, swh_db* db, const Task* t, int nt, const Shape& sh,
                  const swh_search_opts& o, uint32_t* counters, const LaunchRes& lr,
                  swh_stats* st) {
    // exact below both the saturation of an add (65535 - max(matrix)) and the
    // fp16 ordering of the three-way max
    // a range outside its block kind would make the scan read pair blocks as
    // lane blocks (or the reverse), past their rows' border slots
    for (int i = 0; i < nt; ++i)
        if (t[i].pair_begin < 0 || t[i].pair_begin > t[i].pair_end || t[i].pair_end > db->n_pairs ||
            t[i].lane_begin < db->n_pairs || t[i].lane_begin > t[i].lane_end ||
            t[i].lane_end > db->n_blocks)
            return fail(SWH_ERR_ARG, "task %d: pair blocks [%d, %d) / lane blocks [%d, %d) outside "
                        "[0, %d) / [%d, %d)", i, t[i].pair_begin, t[i].pair_end, t[i].lane_begin,
                        t[i].lane_end, db->n_pairs, db->n_pairs, db->n_blocks);
    const int exact_thr = std::min(65535 - c->max_score, kScoreLimitF16);
    const int thr = (o.overflow_threshold > 0) ? std::min(o.overflow_threshold, exact_thr)
                                               : exact_thr;
    const int gap_oe = c->gap_open + c->gap_extend;
    SmallBuf<QueryDesc, 24> hd(3 * (size_t)nt);
    uint64_t launch_cells = 0;
    for (int i = 0; i < nt; ++i)
        launch_cells += (uint64_t)t[i].qlen * (range_residues(db, t[i].pair_begin, t[i].pair_end) +
                                               range_residues(db, t[i].lane_begin, t[i].lane_end));
    const int max_passes = group_pass_budget(c, launch_cells, db->max_rows_pair, sh.Wg);
    const int min_lanes = group_min_lanes(c, db->max_rows_pair, lr.border_budget);
    int items_l = 0, items_g = 0, max_qpad_l = 0, max_pass_g = 0, max_gpw = 1, max_pass_r = 0;
    int max_gpw_r = 0;   // lane groups per wave of the multi-pass rescores (0: none)
    bool lane_border = false, group_border = false;
    for (int i = 0; i < nt; ++i) {
        const Task& T = t[i];
        QueryDesc base{};
        base.prof = T.prof;
        base.scores = T.scores;
        base.ovf_list = T.ovf;
        base.ovf_count = counters + 16 + i;
        base.qlen = T.qlen;
        base.pstride = profile_stride(T.qlen);
        base.ovf_threshold = thr;
        QueryDesc& L = hd[i];
        L = base;
        L.qpad = ((T.qlen + 7) / 8) * 8;   // the last strip may be narrower than Wl
        L.group = 1;
        L.blk_begin = T.lane_begin;
        L.blk_end = T.lane_end;
        L.n_items = (std::max(0, T.lane_end - T.lane_begin) + kWavesPerWG - 1) / kWavesPerWG;
        items_l += L.n_items;
        max_qpad_l = std::max(max_qpad_l, L.qpad);
        lane_border |= L.n_items > 0 && L.qpad > sh.Wl;
        QueryDesc& Gd = hd[nt + i];
        Gd = base;
        Gd.qpad = ((T.qlen + sh.Wg - 1) / sh.Wg) * sh.Wg;
        Gd.group = group_lanes(c, T.qlen, sh.Wg, o,
                               group_passes(c, db, launch_cells, T.qlen, sh.Wg, max_passes,
                                            min_lanes),
                               min_lanes);
        Gd.blk_begin = T.pair_begin;
        Gd.blk_end = T.pair_end;
        const int per_item = kWavesPerWG * (kWave / Gd.group);
        Gd.n_items = (std::max(0, T.pair_end - T.pair_begin) + per_item - 1) / per_item;
        items_g += Gd.n_items;
        if (Gd.n_items > 0) {
            max_pass_g = std::max(max_pass_g, std::min(Gd.group * sh.Wg, Gd.qpad));
            // only queries swept in more than one pass keep a border per pair
            if (Gd.qpad > Gd.group * sh.Wg) {
                group_border = true;
                max_gpw = std::max(max_gpw, kWave / Gd.group);
            }
        }
        QueryDesc& Rd = hd[2 * nt + i];
        Rd = base;
        Rd.qpad = ((T.qlen + kRescoreWidth - 1) / kRescoreWidth) * kRescoreWidth;
        Rd.group = std::min(kWave, Rd.qpad / kRescoreWidth);
        max_pass_r = std::max(max_pass_r, Rd.group * kRescoreWidth);
        if (Rd.qpad > Rd.group * kRescoreWidth) max_gpw_r = std::max(max_gpw_r, kWave / Rd.group);
        st->cells += (uint64_t)T.qlen *
                     (range_residues(db, T.pair_begin, T.pair_end) +
                      range_residues(db, T.lane_begin, T.lane_end));
        st->cells_group += (uint64_t)T.qlen * range_residues(db, T.pair_begin, T.pair_end);
        const int passes = (Gd.qpad / sh.Wg + Gd.group - 1) / Gd.group;
        st->padded_cells +=
            range_padded(db, T.lane_begin, T.lane_end) * (uint64_t)L.qpad +
            range_padded(db, T.pair_begin, T.pair_end) * (uint64_t)passes * Gd.group * sh.Wg;
    }
    const int n_items = items_g + items_l;
    const size_t desc_bytes = 3 * (size_t)nt * sizeof(QueryDesc);
    const size_t total_bytes = desc_bytes + (size_t)n_items * sizeof(uint2);
    if ((lr.h_desc->fixed && total_bytes > lr.h_desc->bytes) ||
        (lr.desc->fixed && total_bytes > lr.desc->bytes))
        return fail(SWH_ERR_CAPACITY, "%zu bytes of work descriptors, above the fixed %zu",
                    total_bytes, std::min(lr.h_desc->bytes, lr.desc->bytes));
    HIP_TRY(grow_beside(c, *lr.h_desc, total_bytes));
    HIP_TRY(grow_beside(c, *lr.desc, total_bytes));
    std::memcpy(lr.h_desc->p, hd.p, desc_bytes);
    order_items(db, sh, &hd[0], &hd[nt], nt, (uint2*)((char*)lr.h_desc->p + desc_bytes));
    *lr.desc_used = total_bytes;
    HIP_TRY(hipMemcpyAsync(lr.desc->p, lr.h_desc->p, total_bytes, hipMemcpyHostToDevice,
                           lr.stream));
    const QueryDesc* d_lane = (const QueryDesc*)lr.desc->p;
    const QueryDesc* d_group = d_lane + nt;
    const QueryDesc* d_res = d_lane + 2 * nt;
    const uint2* d_items = (const uint2*)((const char*)lr.desc->p + desc_bytes);
    HIP_TRY(launch_reset(counters, 16 + nt, lr.stream));

    // scan: every item of every task, one launch
    const int ps_l = std::min(std::max(max_qpad_l, 8), kSegCols) + kProfPad;
    const int ps_g = max_pass_g + kProfPad;
    const size_t lds = std::max(lane_lds_bytes(sh.Wl, sh.pair, ps_l),
                                (size_t)kProfRows * ps_g * sizeof(int16_t));
    const uint64_t stride_l = lane_border ? (uint64_t)db->max_rows_lane * kWave : 0;
    const uint64_t stride_g = group_border ? (uint64_t)max_gpw * db->max_rows_pair : 0;
    const uint64_t stride = std::max(stride_l, stride_g);
    const uint64_t per_wg = stride * sizeof(uint2) * kWavesPerWG;
    const int items = n_items;
    int grid = c->num_cus * scan_max_wg_per_cu(sh.Wl, sh.Wg, sh.pair, lds);
    if (per_wg)
        grid = (int)std::min<uint64_t>(grid, std::max<uint64_t>(1, lr.border_budget / per_wg));
    grid = std::max(1, std::min(grid, items));
    RescoreLaunch rl = plan_rescore(c, db, max_pass_r, max_gpw_r, lr.border_budget, d_res, nt,
                                    counters + 2);
    const uint64_t border_need =
        std::max<uint64_t>({16, per_wg * (uint64_t)grid, rl.per_wg * (uint64_t)rl.grid});
    if (lr.border->fixed && border_need > lr.border->bytes)
        return fail(SWH_ERR_CAPACITY, "strip borders need %llu bytes for one workgroup, above "
                    "the fixed %zu", (unsigned long long)border_need, lr.border->bytes);
    HIP_TRY(grow_beside(c, *lr.border, border_need, true));   // within the border budget

    ScanArgs a{};
    a.words = (const uint32_t*)db->words.p;
    a.blocks = (const BlockDesc*)db->d_blocks.p;
    a.perm = (const uint32_t*)db->perm.p;
    a.group_q = d_group;
    a.lane_q = d_lane;
    a.desc_stride = 1;
    a.items = d_items;
    a.work_counter = counters + 0;
    a.border = (uint2*)lr.border->p;
    a.border_stride = stride;
    a.border_rows = db->max_rows_pair;
    a.n_queries = nt;
    a.n_items = n_items;
    a.ps_lane = ps_l;
    a.ps_group = ps_g;
    a.gap_oe = gap_oe;
    a.gap_e = c->gap_extend;
    if (items > 0) {
        HIP_TRY(launch_scan(sh.Wl, sh.Wg, sh.pair, a, grid, lds, lr.stream));
        st->n_launches += 1;
    }
    rl.args.border = (uint2*)lr.border->p;
    HIP_TRY(launch_rescore(rl.args, rl.grid, rl.lds, lr.stream));
    st->n_launches += 1;
    return SWH_OK;
}

// The int32 rescore of nq queries' overflow lists: a lane group per
// sequence, one pass's profile columns in LDS; only queries swept in more
// than one pass keep a border per group, and the grid is capped so those
// borders fit 1 GiB.  max_pass_r: the widest pass (group x 8 columns);
// max_gpw_r: lane groups per wave of the multi-pass queries (0: none).  The
// caller sets args.border once the border buffer holds per_wg x grid bytes.
RescoreLaunch plan_rescore(const swh_ctx* c, const swh_db* db, int max_pass_r, int max_gpw_r,
                           uint64_t border_budget, const QueryDesc* d_queries, int nq,
                           uint32_t* work_counter) {
    RescoreLaunch rl{};
    const int ps_r = max_pass_r + kProfPad;
    rl.lds = (size_t)kProfRows * ps_r * sizeof(int16_t);
    const uint32_t lmax = std::max(db->max_rows_lane, db->max_rows_pair);
    const uint64_t rstride = (uint64_t)max_gpw_r * lmax;
    rl.per_wg = rstride * sizeof(uint2) * kWavesPerWG;
    rl.grid = c->num_cus * rescore_max_wg_per_cu(rl.lds);
    const uint64_t rbudget = std::min<uint64_t>(size_t(1) << 30, border_budget);
    if (rl.per_wg)
        rl.grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(rl.grid, rbudget / rl.per_wg));
    RescoreArgs& r = rl.args;
    r.words = (const uint32_t*)db->words.p;
    r.blocks = (const BlockDesc*)db->d_blocks.p;
    r.perm = (const uint32_t*)db->perm.p;
    r.sorted_len = (const uint32_t*)db->sorted_len.p;
    r.sorted_blk = (const uint32_t*)db->sorted_blk.p;
    r.queries = d_queries;
    r.work_counter = work_counter;
    r.border_stride = rstride;
    r.border_rows = lmax;
    r.n_queries = nq;
    r.n_lane_blocks_first = db->n_pairs;
    r.ps = ps_r;
    r.gap_oe = c->gap_open + c->gap_extend;
    r.gap_e = c->gap_extend;
    return rl;
}

// Search every query over the whole database in one launch sequence.
// scores_host gets nq rows of nseq: copied there straight from the device
// when it is page-locked, else through the context's pinned staging (up to 64
// MB) and a host copy.
// The layout a launch of nq queries runs on: the small-database pair layout
// when the lane layout would give fewer lane items than workgroup slots.
swh_db* pick_layout(const swh_ctx* c, swh_db* db, int nq) {
    if (!db->alt) return db;
    return enough_lane_items(c, (uint64_t)(db->n_blocks - db->n_pairs), nq) ? db : db->alt;
}

int run_queries(swh_ctx* c, swh_db* db_in, const uint8_t* const* qs, const int* qlens,
                       int nq, const swh_search_opts& o, int32_t* scores_host,
                       swh_stats* stats) {
    swh_db* db = pick_layout(c, db_in, nq);
    const double t_start = now_ms();
    swh_stats st{};
    int rc = check_opts(o);
    if (rc) return rc;
    for (int i = 0; i < nq; ++i)
        if ((rc = check_query(qs[i], qlens[i], i))) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (db->nseq == 0 || nq == 0) {
        st.total_ms = now_ms() - t_start;
        if (stats) *stats = st;
        return SWH_OK;
    }
    const Shape sh = choose_shape(c, qlens, nq, o, db, db->residues);
    std::vector<size_t> prof_off(nq);
    size_t prof_elems = 0;
    for (int i = 0; i < nq; ++i) {
        prof_off[i] = prof_elems;
        prof_elems += (size_t)kProfRows * profile_stride(qlens[i]);
    }
    const size_t nseq = db->nseq;
    const size_t score_bytes = (size_t)nq * nseq * sizeof(int32_t);
    // (a database build may be running beside this search: no buffer is freed)
    HIP_TRY(grow_beside(c, c->h_prof, prof_elems * sizeof(int16_t)));
    HIP_TRY(grow_beside(c, c->prof, prof_elems * sizeof(int16_t)));
    HIP_TRY(grow_beside(c, c->scores, score_bytes));
    HIP_TRY(grow_beside(c, c->ovf, (size_t)nq * nseq * sizeof(uint32_t)));
    HIP_TRY(grow_beside(c, c->counters, (16 + (size_t)nq) * sizeof(uint32_t)));
    HIP_TRY(grow_beside(c, c->h_ctr, (16 + (size_t)nq) * sizeof(uint32_t)));
    const bool direct_out = is_pinned_host(scores_host);
    const bool pinned_out = !direct_out && score_bytes <= (size_t(64) << 20);
    if (pinned_out) HIP_TRY(grow_beside(c, c->h_scores, score_bytes));
    int16_t* hp = (int16_t*)c->h_prof.p;
    std::vector<Task> tasks(nq);
    for (int i = 0; i < nq; ++i) {
        build_profile(c->matrix, qs[i], qlens[i], hp + prof_off[i]);
        Task& T = tasks[i];
        T.prof = (const int16_t*)c->prof.p + prof_off[i];
        T.qlen = qlens[i];
        T.scores = (int32_t*)c->scores.p + (size_t)i * nseq;
        T.ovf = (uint32_t*)c->ovf.p + (size_t)i * nseq;
        T.pair_begin = 0;
        T.pair_end = db->n_pairs;
        T.lane_begin = db->n_pairs;
        T.lane_end = db->n_blocks;
    }
    uint32_t* ctr = (uint32_t*)c->counters.p;
    // the launch sequence after the profile upload: scan and rescore, then
    // the copy-back
    auto enqueue_copyback = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(c->h_ctr.p, ctr, (16 + (size_t)nq) * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, c->stream));
        if (pinned_out || direct_out)
            HIP_TRY(hipMemcpyAsync(direct_out ? (void*)scores_host : c->h_scores.p, c->scores.p,
                                   score_bytes, hipMemcpyDeviceToHost, c->stream));
        return SWH_OK;
    };
    auto enqueue_rest = [&](swh_stats* pst) -> int {
        int r2 = enqueue_tasks(c, db, tasks.data(), nq, sh, o, ctr, ctx_res(c), pst);
        return r2 ? r2 : enqueue_copyback();
    };
    std::vector<long long> key;
    if (o.use_graph) {
        // the database by id: a new one may reuse a destroyed one's address
        key = {(long long)db->uid, nq, sh.Wl, sh.Wg, sh.pair, (long long)c->tune_gen,
               o.group_lanes, o.overflow_threshold, c->gap_open, c->gap_extend,
               c->max_score,
               pinned_out, (long long)(uintptr_t)(direct_out ? scores_host : nullptr),
               (long long)(uintptr_t)c->prof.p, (long long)(uintptr_t)c->scores.p,
               (long long)(uintptr_t)c->ovf.p, (long long)(uintptr_t)c->counters.p,
               (long long)(uintptr_t)c->h_prof.p, (long long)(uintptr_t)c->h_ctr.p,
               (long long)(uintptr_t)c->h_scores.p};
        key.insert(key.end(), qlens, qlens + nq);
    }
    const bool replay = o.use_graph && c->graph.exec && c->graph.key == key &&
                        c->graph.desc == c->desc.p && c->graph.border == c->border.p &&
                        c->graph.h_desc == c->h_desc.p;
    if (replay) {
        // graph: profile upload, scan, rescore and copy-back in one submission;
        // calls in between may have reused the pinned descriptor staging
        st = c->graph.st;
        std::memcpy(c->h_desc.p, c->graph.desc_bytes.data(), c->graph.desc_bytes.size());
        HIP_TRY(hipEventRecord(c->ev[0], c->stream));
        HIP_TRY(hipGraphLaunch(c->graph.exec, c->stream));
        HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    } else {
        HIP_TRY(hipMemcpyAsync(c->prof.p, hp, prof_elems * sizeof(int16_t),
                               hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipEventRecord(c->ev[0], c->stream));
        rc = enqueue_tasks(c, db, tasks.data(), nq, sh, o, ctr, ctx_res(c), &st);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(c->ev[1], c->stream));
        // The copy-back is queued once the launches have ended, not behind
        // them: a copy waiting for a running launch holds up the copy engine,
        // and with it the uploads of a database that another thread builds
        // meanwhile (the worker pipeline builds chunk k + 1 while chunk k is
        // searched; measured on MI355X, its uploads then ran only after the
        // scan, and with this wait beside it).
        HIP_TRY(hipEventSynchronize(c->ev[1]));
        rc = enqueue_copyback();
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (o.use_graph && !replay) {
        // capture the same sequence (buffers are sized now) for the next call
        // with this shape; the profile upload node re-reads the pinned staging
        c->graph.release();
        hipGraph_t g = nullptr;
        swh_stats cap{};
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipMemcpyAsync(c->prof.p, hp, prof_elems * sizeof(int16_t),
                                       hipMemcpyHostToDevice, c->stream);
        int r2 = ce == hipSuccess ? enqueue_rest(&cap) : SWH_ERR_HIP;
        hipError_t ee = hipStreamEndCapture(c->stream, &g);
        if (ce != hipSuccess || r2 != SWH_OK || ee != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return fail(SWH_ERR_HIP, "graph capture failed");
        }
        hipError_t ie = hipGraphInstantiate(&c->graph.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            c->graph.exec = nullptr;
            return fail(SWH_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
        }
        c->graph.key = key;
        c->graph.st = cap;
        c->graph.desc = c->desc.p;
        c->graph.border = c->border.p;
        c->graph.h_desc = c->h_desc.p;
        const char* hb = (const char*)c->h_desc.p;
        c->graph.desc_bytes.assign(hb, hb + c->desc_used);
    }
    if (pinned_out) std::memcpy(scores_host, c->h_scores.p, score_bytes);
    else if (!direct_out)
        HIP_TRY(hipMemcpy(scores_host, c->scores.p, score_bytes, hipMemcpyDeviceToHost));
    const uint32_t* hc = (const uint32_t*)c->h_ctr.p;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    st.kernel_ms = ms;   // scan + rescore (+ copies when replayed from a graph)
    st.n_overflow = 0;
    for (int i = 0; i < nq; ++i) st.n_overflow += hc[16 + i];
    st.total_ms = now_ms() - t_start;
    if (stats) *stats = st;
    return SWH_OK;
}

}  // namespace swhh

extern "C" {

int swh_search(swh_ctx* c, swh_db* db, const uint8_t* query, int qlen,
               const swh_search_opts* opts, int32_t* scores, swh_stats* stats) {
    if (!c || !db || !query || (!scores && db->nseq > 0))
        return fail(SWH_ERR_ARG, "NULL argument");
    if (db->ctx != c) return fail(SWH_ERR_ARG, "database belongs to another context");
    int rc = refuse_if_served(c, "swh_search");
    if (rc) return rc;
    swh_search_opts o{};
    if (opts) o = *opts;
    const uint8_t* qs[1] = {query};
    return run_queries(c, db, qs, &qlen, 1, o, scores, stats);
}

int swh_search_batch(swh_ctx* c, swh_db* db, const uint8_t* const* queries, const int* qlens,
                     int nq, const swh_search_opts* opts, int32_t* scores, swh_stats* stats) {
    if (!c || !db || nq < 0 || (nq > 0 && (!queries || !qlens || (!scores && db->nseq > 0))))
        return fail(SWH_ERR_ARG, "NULL argument");
    if (db->ctx != c) return fail(SWH_ERR_ARG, "database belongs to another context");
    int rc = refuse_if_served(c, "swh_search_batch");
    if (rc) return rc;
    swh_search_opts o{};
    if (opts) o = *opts;
    return run_queries(c, db, queries, qlens, nq, o, scores, stats);
}

int swh_topk(const int32_t* scores, uint64_t n, int k, int32_t* out_score, int64_t* out_index) {
    if (k < 0 || (k > 0 && (!out_score || !out_index)) || (n > 0 && !scores))
        return fail(SWH_ERR_ARG, "bad top-k arguments");
    // (score desc, index asc) among scores > 0: a total order, so the best k
    // of the union of the best k of each part are the best k of the whole
    using Entry = std::pair<int32_t, int64_t>;
    auto worse = [](const Entry& a, const Entry& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    };
    // the best k of [lo, hi) as a min-heap of the current k best: once it is
    // full a score must beat its weakest entry, and blocks of 64 whose
    // maximum cannot are skipped
    auto part = [&](uint64_t lo, uint64_t hi, std::vector<Entry>& heap) {
        heap.clear();
        heap.reserve(k + 1);
        int32_t floor = 0;   // scores <= floor cannot enter
        constexpr uint64_t kBlock = 64;
        for (uint64_t b = lo; b < hi && k > 0; b += kBlock) {
            const uint64_t e = std::min(hi, b + kBlock);
            int32_t m = scores[b];
            for (uint64_t i = b + 1; i < e; ++i) m = std::max(m, scores[i]);
            if (m <= floor) continue;
            for (uint64_t i = b; i < e; ++i) {
                const int32_t s = scores[i];
                if (s <= floor) continue;
                if ((int)heap.size() < k) {
                    heap.emplace_back(s, (int64_t)i);
                    std::push_heap(heap.begin(), heap.end(), worse);
                } else {
                    std::pop_heap(heap.begin(), heap.end(), worse);
                    heap.back() = {s, (int64_t)i};
                    std::push_heap(heap.begin(), heap.end(), worse);
                }
                if ((int)heap.size() == k) floor = heap.front().first;
            }
        }
    };
    std::vector<Entry> best;
    // large score vectors (a search over millions of sequences) in 16 parts
    constexpr uint64_t kPartMin = uint64_t(1) << 18;
    const uint64_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint64_t parts = std::min<uint64_t>(hw, std::max<uint64_t>(1, n / kPartMin));
    if (parts <= 1 || k == 0) {
        part(0, n, best);
    } else {
        std::vector<std::vector<Entry>> heaps(parts);
        std::vector<std::thread> team;
        for (uint64_t p = 1; p < parts; ++p)
            team.emplace_back([&, p] { part(n * p / parts, n * (p + 1) / parts, heaps[p]); });
        part(0, n / parts, heaps[0]);
        for (auto& t : team) t.join();
        for (auto& h : heaps) best.insert(best.end(), h.begin(), h.end());
    }
    std::sort(best.begin(), best.end(), worse);
    for (int t = 0; t < k; ++t) {
        out_score[t] = t < (int)best.size() ? best[t].first : 0;
        out_index[t] = t < (int)best.size() ? best[t].second : 0;
    }
    return SWH_OK;
}

}  // extern "C"
