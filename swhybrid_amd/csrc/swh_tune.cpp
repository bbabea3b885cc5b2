// swh_tune.cpp -- the autotuner: launch shapes per query-length bucket,
// measured on a length-stratified sample of the database.
//
// Three choices are tuned, all of them layout or schedule only (the scores do
// not depend on any of them):
//   * lanes per sequence: the database length from which sequences are stored
//     as pairs and searched by a group of lanes (swh_db_opts.group_min_len),
//     and per query bucket the group size G;
//   * profile placement for the lane scan: residue-pair table or single
//     per-residue rows, both in LDS;
//   * tile depth: query columns held in registers per lane (lane scan 16, 24
//     or 32; group scan 8 or 16).
// Every candidate is timed as a whole multi-query launch on the sample
// (HIP events around scan and rescore), so interactions inside the fused
// persistent kernel -- group items running beside lane items -- are measured,
// not modelled.
#include "swh_host.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace swhh;

namespace {

// Deterministic random query residues over the 20 standard codes.
std::vector<uint8_t> tune_query(int len, uint64_t seed) {
    std::vector<uint8_t> q(len);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    static const uint8_t kStd[20] = {0, 2, 3, 4, 5, 6, 7, 8, 9, 10,
                                     11, 12, 13, 14, 15, 16, 17, 18, 19, 21};
    for (int i = 0; i < len; ++i) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        q[i] = kStd[(x >> 32) % 20];
    }
    return q;
}

struct Sample {
    const uint8_t* residues;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> lengths;
};

struct DbHolder {
    swh_db* db = nullptr;
    ~DbHolder() { swh_db_destroy(db); }
};

}  // namespace

// A fixed group width and size must beat the launch-time choice by this
// factor on the sample to be kept (64 batched queries on the 5M-sequence
// database: fixed per-bucket choices from a 1/8 sample ran 1% slower).
constexpr double kFixedGroupMargin = 0.98;
// Configurations timed within this factor of the fastest are timed again.
constexpr double kRetimeWithin = 1.2;

extern "C" {

int swh_set_tuning(swh_ctx* c, const swh_tune_entry* e, int n) {
    if (!c) return fail(SWH_ERR_ARG, "ctx is NULL");
    int rc = refuse_if_served(c, "swh_set_tuning");
    if (rc) return rc;
    if (n < 0 || n > SWH_TUNE_MAX) return fail(SWH_ERR_ARG, "n must be in 0..%d", SWH_TUNE_MAX);
    if (n > 0 && !e) return fail(SWH_ERR_ARG, "entries is NULL");
    for (int i = 0; i < n; ++i) {
        const swh_tune_entry& t = e[i];
        if (t.qlen_max <= 0 || (i > 0 && t.qlen_max <= e[i - 1].qlen_max))
            return fail(SWH_ERR_ARG, "entry %d: qlen_max must be positive and ascending", i);
        if (t.strip_width != 16 && t.strip_width != 24 && t.strip_width != 32)
            return fail(SWH_ERR_ARG, "entry %d: strip_width %d", i, t.strip_width);
        if (t.profile_mode != SWH_PROFILE_SINGLE && t.profile_mode != SWH_PROFILE_PAIR)
            return fail(SWH_ERR_ARG, "entry %d: profile_mode %d", i, t.profile_mode);
        if (t.profile_mode == SWH_PROFILE_PAIR && t.strip_width == 32)
            return fail(SWH_ERR_ARG, "entry %d: pair table needs strip width 16 or 24", i);
        if (t.group_lanes < 0 || t.group_lanes > kWave)
            return fail(SWH_ERR_ARG, "entry %d: group_lanes %d", i, t.group_lanes);
        if (t.group_width != 0 && t.group_width != 8 && t.group_width != 16 &&
            !(t.group_width == 24 && t.strip_width == 24))
            return fail(SWH_ERR_ARG, "entry %d: group_width %d", i, t.group_width);
        if (t.group_width == 0 && t.group_lanes != 0)
            return fail(SWH_ERR_ARG, "entry %d: group_lanes needs a group_width", i);
        bool any = false;
        for (float v : t.lane_cost) {
            if (!(v >= 0.0f)) return fail(SWH_ERR_ARG, "entry %d: negative lane cost", i);
            any |= v > 0.0f;
        }
        if (!any) return fail(SWH_ERR_ARG, "entry %d: no lane configuration measured", i);
    }
    c->tune.assign(e, e + n);
    c->tune_gen += 1;
    return SWH_OK;
}

int swh_get_tuning(const swh_ctx* c, swh_tune_entry* e, int cap, int* n) {
    if (!c || !n) return fail(SWH_ERR_ARG, "ctx/n is NULL");
    *n = (int)c->tune.size();
    if (cap > 0 && e) std::memcpy(e, c->tune.data(), sizeof(swh_tune_entry) * std::min(cap, *n));
    return SWH_OK;
}

int swh_autotune(swh_ctx* c, const uint8_t* residues, const uint64_t* offsets,
                 const uint32_t* lengths, uint64_t nseq, int sample_every,
                 const int* bucket_qlens, int nb, int qpb, int reps, swh_tune_result* out) {
    if (!c || !out) return fail(SWH_ERR_ARG, "ctx/out is NULL");
    if (int rc0 = refuse_if_served(c, "swh_autotune")) return rc0;
    if (nseq == 0 || !residues || !offsets || !lengths)
        return fail(SWH_ERR_ARG, "the database sample is empty");
    if (!bucket_qlens || nb < 1 || nb > SWH_TUNE_MAX)
        return fail(SWH_ERR_ARG, "n_buckets must be in 1..%d", SWH_TUNE_MAX);
    for (int b = 0; b < nb; ++b)
        if (bucket_qlens[b] <= 0 || bucket_qlens[b] > (1 << 20) ||
            (b > 0 && bucket_qlens[b] <= bucket_qlens[b - 1]))
            return fail(SWH_ERR_ARG, "bucket_qlens must be positive and ascending");
    if (sample_every < 0 || qpb < 0 || reps < 0) return fail(SWH_ERR_ARG, "negative count");
    if (sample_every == 0) sample_every = 8;
    if (qpb == 0) qpb = 4;
    if (reps == 0) reps = 1;
    const double t_start = now_ms();
    std::memset(out, 0, sizeof(*out));

    // length-stratified sample: every k-th sequence in length order
    std::vector<uint32_t> order(nseq);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return lengths[a] > lengths[b]; });
    Sample smp{residues, {}, {}};
    for (uint64_t i = 0; i < nseq; i += (uint64_t)sample_every) {
        smp.offsets.push_back(offsets[order[i]]);
        smp.lengths.push_back(lengths[order[i]]);
    }
    const uint64_t ns = smp.lengths.size();
    out->sample_seqs = ns;
    for (uint32_t L : smp.lengths) out->sample_residues += L;

    // queries: qpb per bucket, each of the bucket's length
    std::vector<std::vector<uint8_t>> qstore;
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < qpb; ++k)
            qstore.push_back(tune_query(bucket_qlens[b], (uint64_t)b * 131 + k + 1));
    std::vector<const uint8_t*> qptr(qstore.size());
    std::vector<int> qlen(qstore.size());
    for (size_t i = 0; i < qstore.size(); ++i) {
        qptr[i] = qstore[i].data();
        qlen[i] = (int)qstore[i].size();
    }
    std::vector<int32_t> scores((size_t)nb * qpb * ns);

    // tune with the built-in shape choice, restore the caller's table on failure
    std::vector<swh_tune_entry> saved = c->tune;
    c->tune.clear();
    c->tune_gen += 1;
    auto restore = [&](int rc) {
        if (rc != SWH_OK) {
            c->tune = saved;
            c->tune_gen += 1;
        }
        return rc;
    };
    // best-of-reps launch time (scan + rescore) of queries [q0, q0+n)
    auto timed = [&](swh_db* db, int q0, int n, const swh_search_opts& o, double* ms) -> int {
        *ms = 1e300;
        for (int r = 0; r < reps; ++r) {
            swh_stats st{};
            int rc = run_queries(c, db, qptr.data() + q0, qlen.data() + q0, n, o, scores.data(),
                                 &st);
            if (rc) return rc;
            *ms = std::min(*ms, st.kernel_ms);
        }
        return SWH_OK;
    };
    auto make_db = [&](uint32_t min_len, DbHolder& h) -> int {
        swh_db_opts dopt{min_len};
        return swh_db_create(c, smp.residues, smp.offsets.data(), smp.lengths.data(), ns, &dopt,
                             &h.db);
    };

    // 1. lanes per sequence, database side: which lengths go to group mode.
    // Every query of every bucket in one launch, so that the launch is long
    // next to its longest item (a lane block of the longest lane-mode
    // sequences) and the timing reflects throughput, as in a batched call;
    // thresholds that give the same layout as a shorter one are skipped.
    const uint32_t kThresholds[] = {1024u, 2048u, 4096u, 8192u, 0xffffffffu};
    uint32_t best_min = 0xffffffffu;
    double best_ms = 1e300;
    uint64_t prev_long = ~0ull;
    bool warmed = false;
    const int nall = (int)qptr.size();
    for (uint32_t thr : kThresholds) {
        uint64_t n_long = 0;
        for (uint32_t L : smp.lengths) n_long += L >= thr;
        if (n_long == prev_long) continue;
        prev_long = n_long;
        DbHolder h;
        int rc = make_db(thr, h);
        if (rc) return restore(rc);
        swh_search_opts o{};
        if (!warmed) {   // the first launch of the process pays code-object loading
            swh_stats st{};
            rc = run_queries(c, h.db, qptr.data(), qlen.data(), qpb, o, scores.data(), &st);
            if (rc) return restore(rc);
            warmed = true;
        }
        double ms;
        if ((rc = timed(h.db, 0, nall, o, &ms))) return restore(rc);
        if (ms < best_ms) {
            best_ms = ms;
            best_min = thr;
        }
    }
    out->group_min_len = best_min;

    // 2. per bucket, on the chosen layout: lane-scan width and profile
    // placement, then group size and group width with that lane shape
    // (qpb queries of the bucket's length per launch)
    DbHolder h;
    int rc = make_db(best_min, h);
    if (rc) return restore(rc);
    const bool has_pairs = h.db->n_pairs > 0;
    for (int b = 0; b < nb; ++b) {
        swh_tune_entry& E = out->entries[b];
        const int L = bucket_qlens[b];
        E.qlen_max = L;
        const int q0 = b * qpb;
        // A launch on the sample now and then runs 12-18% slow (never fast):
        // every configuration within kRetimeWithin of the fastest is timed
        // again and keeps its faster time, so one slow launch of the best
        // configuration does not hand its bucket to another (over three
        // tunings of the 5M-sequence database, one bucket in each had gone
        // to a configuration 4-10% slower).
        auto lane_opts = [](int k) {
            swh_search_opts o{};
            o.strip_width = kLaneCfgs[k].width;
            o.profile_mode = kLaneCfgs[k].pair ? SWH_PROFILE_PAIR : SWH_PROFILE_SINGLE;
            return o;
        };
        double ms_k[SWH_TUNE_LANE_CONFIGS];
        double fastest = 1e300;
        for (int k = 0; k < SWH_TUNE_LANE_CONFIGS; ++k) {
            if ((rc = timed(h.db, q0, qpb, lane_opts(k), &ms_k[k]))) return restore(rc);
            fastest = std::min(fastest, ms_k[k]);
        }
        for (int k = 0; k < SWH_TUNE_LANE_CONFIGS; ++k) {
            if (ms_k[k] > fastest * kRetimeWithin) continue;
            double again;
            if ((rc = timed(h.db, q0, qpb, lane_opts(k), &again))) return restore(rc);
            ms_k[k] = std::min(ms_k[k], again);
        }
        double best = 1e300;
        for (int k = 0; k < SWH_TUNE_LANE_CONFIGS; ++k) {
            const int w = kLaneCfgs[k].width;
            E.lane_cost[k] = (float)(ms_k[k] / ((double)qpb * lane_cols(L, w)));
            if (ms_k[k] < best) {
                best = ms_k[k];
                E.strip_width = w;
                E.profile_mode = lane_opts(k).profile_mode;
            }
        }
        // group mode: the launch-time choice (width 0 and lanes 0 in the table)
        // unless a fixed width and size is measurably faster on the sample;
        // the launch-time rule follows the real launch's size, a fixed one
        // is tied to the sample's
        E.group_lanes = 0;
        E.group_width = 0;
        if (!has_pairs) continue;
        double base_ms;   // the launch-time group choice
        {
            swh_search_opts o{};
            o.strip_width = E.strip_width;
            o.profile_mode = E.profile_mode;
            if ((rc = timed(h.db, q0, qpb, o, &base_ms))) return restore(rc);
            best = base_ms * kFixedGroupMargin;
        }
        std::vector<int> widths = {8, 16};
        if (E.strip_width == 24) widths.push_back(24);
        for (int wg : widths) {
            const int strips = (L + wg - 1) / wg;
            std::vector<int> cand = {group_lanes_for(L, wg, 1, 1)};
            for (int g : {8, 16, 32})
                if (g < strips) cand.push_back(g);
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            for (int g : cand) {
                swh_search_opts o{};
                o.strip_width = E.strip_width;
                o.profile_mode = E.profile_mode;
                o.group_width = wg;
                o.group_lanes = g;
                double ms;
                if ((rc = timed(h.db, q0, qpb, o, &ms))) return restore(rc);
                if (ms < best) {
                    best = ms;
                    E.group_lanes = g;
                    E.group_width = wg;
                }
            }
        }
        if (E.group_width) {
            // a fixed choice won: time both again (a slow launch of the
            // launch-time choice must not be what made it win)
            swh_search_opts o{};
            o.strip_width = E.strip_width;
            o.profile_mode = E.profile_mode;
            double again;
            if ((rc = timed(h.db, q0, qpb, o, &again))) return restore(rc);
            base_ms = std::min(base_ms, again);
            o.group_width = E.group_width;
            o.group_lanes = E.group_lanes;
            if ((rc = timed(h.db, q0, qpb, o, &again))) return restore(rc);
            if (std::min(best, again) >= base_ms * kFixedGroupMargin) {
                E.group_lanes = 0;
                E.group_width = 0;
            }
        }
    }
    out->n_entries = nb;
    rc = swh_set_tuning(c, out->entries, nb);
    if (rc) return restore(rc);
    out->seconds = (now_ms() - t_start) / 1e3;
    return SWH_OK;
}

}  // extern "C"
