// swh_noalloc.cpp -- swh_search_noalloc: the one-query search of swh_search
// for callers that cannot allocate (code under graph capture, tight serving
// loops, harnesses that forbid allocation).  Every device buffer of the call
// (counters, matrix, query, profile, descriptor tables, the device score
// vector, the overflow list and the strip borders) is a part of the static
// scratch array of swh_kernels.hip, whose address swh_ctx_create looked up on
// the context's device; the descriptor tables are staged through the fixed
// page-locked area the context allocated when it was created.  The call makes
// no hipMalloc, hipFree, hipHostMalloc or other allocating call.
#include "swh_host.h"

using namespace swhh;

namespace {

constexpr int kMaxDevices = 64;
// One allocation-free search at a time per device: every call on a device
// uses the same array.
std::mutex g_scratch_mu[kMaxDevices];

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Waits for the stream when it goes out of scope, so no kernel of this call
// still reads or writes the scratch array once its lock is released (on the
// error paths too).
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

}  // namespace

extern "C" int swh_search_noalloc(swh_ctx* c, swh_db* db_in, const uint8_t* query, int qlen,
                                  const swh_search_opts* opts, int32_t* scores,
                                  swh_stats* stats) {
    if (!c || !db_in || !query || (!scores && db_in->nseq > 0))
        return fail(SWH_ERR_ARG, "NULL argument");
    if (db_in->ctx != c) return fail(SWH_ERR_ARG, "database belongs to another context");
    int rc = refuse_if_served(c, "swh_search_noalloc");
    if (rc) return rc;
    swh_search_opts o{};
    if (opts) o = *opts;
    if (o.use_graph) return fail(SWH_ERR_ARG, "swh_search_noalloc does not capture graphs");
    if ((rc = check_opts(o))) return rc;
    if ((rc = check_query(query, qlen, 0))) return rc;
    if (qlen > kNoAllocMaxQlen)
        return fail(SWH_ERR_CAPACITY, "query length %d above the allocation-free maximum %d", qlen,
                    kNoAllocMaxQlen);
    if (!c->scratch || c->device < 0 || c->device >= kMaxDevices)
        return fail(SWH_ERR_STATE, "the context has no scratch array");
    const double t_start = now_ms();
    swh_stats st{};
    HIP_TRY(hipSetDevice(c->device));
    swh_db* db = pick_layout(c, db_in, 1);
    const size_t nseq = db->nseq;
    if (nseq == 0) {
        st.total_ms = now_ms() - t_start;
        if (stats) *stats = st;
        return SWH_OK;
    }
    // the scratch array, front to back; the strip borders take the rest
    const int ps = profile_stride(qlen);
    size_t off = 0;
    auto carve = [&](size_t n) {
        const size_t at = off;
        off = align256(off + n);
        return at;
    };
    const size_t o_ctr = carve(64 * sizeof(uint32_t));   // 16 launch counters + 1 overflow count
    const size_t o_mat = carve(1024 * sizeof(int32_t));
    const size_t o_query = carve((size_t)qlen);
    const size_t o_prof = carve((size_t)kProfRows * ps * sizeof(int16_t));
    const size_t o_desc = carve(kNoAllocDescBytes);
    const size_t o_scores = carve(nseq * sizeof(int32_t));
    const size_t o_ovf = carve(nseq * sizeof(uint32_t));
    if (off >= c->scratch_bytes)
        return fail(SWH_ERR_CAPACITY, "a database of %zu sequences needs %zu bytes of scratch "
                    "before its strip borders; the allocation-free maximum is %zu", nseq, off,
                    c->scratch_bytes);

    std::lock_guard<std::mutex> lock(g_scratch_mu[c->device]);
    StreamDrain drain{c->stream};   // destroyed before the lock is released
    uint8_t* base = static_cast<uint8_t*>(c->scratch);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(base + o_ctr);
    int32_t* d_scores = reinterpret_cast<int32_t*>(base + o_scores);
    int16_t* d_prof = reinterpret_cast<int16_t*>(base + o_prof);
    HIP_TRY(hipMemcpyAsync(base + o_mat, c->matrix, sizeof(c->matrix), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(base + o_query, query, (size_t)qlen, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(launch_build_profile(reinterpret_cast<const int32_t*>(base + o_mat), base + o_query,
                                 qlen, ps, d_prof, c->stream));
    Task T{};
    T.prof = d_prof;
    T.qlen = qlen;
    T.scores = d_scores;
    T.ovf = reinterpret_cast<uint32_t*>(base + o_ovf);
    T.pair_begin = 0;
    T.pair_end = db->n_pairs;
    T.lane_begin = db->n_pairs;
    T.lane_end = db->n_blocks;
    const Shape sh = choose_shape(c, &qlen, 1, o, db, db->residues);
    DevBuf desc = DevBuf::view(base + o_desc, kNoAllocDescBytes);
    DevBuf border = DevBuf::view(base + off, c->scratch_bytes - off);
    size_t desc_used = 0;
    const LaunchRes lr{c->stream, &desc, &c->na_h_desc, &border, &desc_used, border.bytes};
    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    if ((rc = enqueue_tasks(c, db, &T, 1, sh, o, ctr, lr, &st))) return rc;
    HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    uint32_t hctr[17] = {};
    HIP_TRY(hipMemcpyAsync(hctr, ctr, sizeof(hctr), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(scores, d_scores, nseq * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    st.kernel_ms = ms;
    st.n_overflow = hctr[16];
    st.total_ms = now_ms() - t_start;
    if (stats) *stats = st;
    return SWH_OK;
}
