// CpuKernel.cpp -- compiled with -mavx2; callers check avx2_available().
#include "CpuKernel.h"

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <thread>

namespace swhy {
namespace cpu {

bool avx2_available() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2");
}

namespace {

constexpr int kRows = 4;        // database rows per profile block
constexpr int kSymbols = 24;    // query residue codes 0..23

// Biased int8: score s is held as s - 128, so the saturating floor -128 is the
// local-alignment zero; scores up to 255 are representable.
struct I8 {
    static constexpr int kLanes = 32;
    static constexpr int kBias = 128;
    static constexpr int kTop = 255;
    static __m256i adds(__m256i a, __m256i b) { return _mm256_adds_epi8(a, b); }
    static __m256i subs(__m256i a, __m256i b) { return _mm256_subs_epi8(a, b); }
    static __m256i max(__m256i a, __m256i b) { return _mm256_max_epi8(a, b); }
    static __m256i set1(int v) { return _mm256_set1_epi8((char)v); }
    // s(a, d) for the 32 database codes of one row
    static __m256i lookup(const __m128i* tab, const uint8_t* codes) {
        const __m256i c = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(codes));
        const __m256i lo = _mm256_broadcastsi128_si256(tab[0]);
        const __m256i hi = _mm256_broadcastsi128_si256(tab[1]);
        const __m256i upper = _mm256_cmpgt_epi8(c, _mm256_set1_epi8(15));
        return _mm256_blendv_epi8(_mm256_shuffle_epi8(lo, c), _mm256_shuffle_epi8(hi, c), upper);
    }
    static void scores(__m256i s, int32_t* out) {
        alignas(32) int8_t t[32];
        _mm256_store_si256(reinterpret_cast<__m256i*>(t), s);
        for (int l = 0; l < 32; ++l) out[l] = (int)t[l] + kBias;
    }
};

// Biased int16: s - 32768, scores up to 65535.
struct I16 {
    static constexpr int kLanes = 16;
    static constexpr int kBias = 32768;
    static constexpr int kTop = 65535;
    static __m256i adds(__m256i a, __m256i b) { return _mm256_adds_epi16(a, b); }
    static __m256i subs(__m256i a, __m256i b) { return _mm256_subs_epi16(a, b); }
    static __m256i max(__m256i a, __m256i b) { return _mm256_max_epi16(a, b); }
    static __m256i set1(int v) { return _mm256_set1_epi16((short)v); }
    static __m256i lookup(const __m128i* tab, const uint8_t* codes) {
        const __m128i c = _mm_loadu_si128(reinterpret_cast<const __m128i*>(codes));
        const __m128i upper = _mm_cmpgt_epi8(c, _mm_set1_epi8(15));
        const __m128i s8 =
            _mm_blendv_epi8(_mm_shuffle_epi8(tab[0], c), _mm_shuffle_epi8(tab[1], c), upper);
        return _mm256_cvtepi8_epi16(s8);
    }
    static void scores(__m256i s, int32_t* out) {
        alignas(32) int16_t t[16];
        _mm256_store_si256(reinterpret_cast<__m256i*>(t), s);
        for (int l = 0; l < 16; ++l) out[l] = (int)t[l] + kBias;
    }
};

struct Tables {
    __m128i row[kSymbols][2];   // matrix row of each query code, database codes 0..31
    int present[kSymbols];      // query codes that occur
    int n_present = 0;
    int max_score = 0;
};

void make_tables(const Scoring& sc, const uint8_t* q, int qlen, Tables& t) {
    bool seen[kSymbols] = {};
    for (int j = 0; j < qlen; ++j) seen[q[j]] = true;
    t.n_present = 0;
    for (int a = 0; a < kSymbols; ++a) {
        alignas(16) int8_t r[32];
        for (int d = 0; d < 32; ++d) r[d] = (int8_t)sc.matrix[a * 32 + d];
        t.row[a][0] = _mm_load_si128(reinterpret_cast<const __m128i*>(r));
        t.row[a][1] = _mm_load_si128(reinterpret_cast<const __m128i*>(r + 16));
        if (seen[a]) t.present[t.n_present++] = a;
    }
    t.max_score = 0;
    for (int i = 0; i < 32 * 32; ++i) t.max_score = std::max(t.max_score, sc.matrix[i]);
}

__m256i* align32(std::vector<uint8_t>& v) {
    return reinterpret_cast<__m256i*>((reinterpret_cast<uintptr_t>(v.data()) + 31) &
                                      ~uintptr_t(31));
}

// The profile rows of every query column, in s.qp.
const void* const* column_rows(const uint8_t* q, int m, const __m256i* prof, Scratch& s) {
    s.qp.resize((size_t)m);
    for (int j = 0; j < m; ++j) s.qp[j] = prof + q[j] * kRows;
    return s.qp.data();
}

// Lane l of the group holds sequence idx[l] (fewer than kLanes: pad lanes);
// inter[row * kLanes + l] = its code at row (23 beyond its end or above 23)
// for rows [0, rows).
template <class T>
void interleave(const uint8_t* const* seqs, const uint32_t* lens, uint32_t rows,
                const uint32_t* idx, int n, uint8_t* inter) {
    const size_t total = (size_t)rows * T::kLanes;
    std::memset(inter, 23, total);
    for (int l = 0; l < n; ++l) {
        const uint8_t* src = seqs[idx[l]];
        const uint32_t len = std::min(lens[idx[l]], rows);
        uint8_t* o = inter + l;
        for (uint32_t r = 0; r < len; ++r) o[(size_t)r * T::kLanes] = src[r];
    }
    const __m256i top = _mm256_set1_epi8(23);
    for (size_t i = 0; i < total; i += 32) {   // total is a multiple of 64
        __m256i* p = reinterpret_cast<__m256i*>(inter + i);
        _mm256_storeu_si256(p, _mm256_min_epu8(_mm256_loadu_si256(p), top));
    }
}

// One lane group over every row, the query in segments of at most seg
// columns (the reference's multipass for long queries, src/AVXKernel.cpp:
// 291-406,493-552): a segment sweeps every row, 4 rows per pass through its
// columns, with H and E of the previous row block per column in col (2 x seg
// vectors, so the column state stays in the core's cache however long the
// query is).  Each 4-row block starts from the (H, F) its rows reached at the
// previous segment's last column and leaves its own in edge (2 vectors per
// row); the H of the row above the block at that column is carried from the
// block before.  Per cell: 1 add, 3 sub, 5 max (the reference's count).
// qp[j]: the 4 profile rows of query column j (prof + 4 * q[j]), set once per
// thread and query.
template <class T>
__m256i scan_group(const Tables& tb, const void* const* qp, int m, const uint8_t* inter,
                   uint32_t len, int gap_oe, int gap_e, int seg, __m256i* col, __m256i* prof,
                   __m256i* edge) {
    const __m256i Z = T::set1(-T::kBias);
    const __m256i OE = T::set1(gap_oe);
    const __m256i GE = T::set1(gap_e);
    __m256i S = Z;
    for (int c0 = 0; c0 < m; c0 += seg) {
        const int w = std::min(seg, m - c0);
        const bool first = c0 == 0, last = c0 + w == m;
        const void* const* qs = qp + c0;
        for (int j = 0; j < w; ++j) {
            col[2 * j] = Z;       // H of the row above the block
            col[2 * j + 1] = Z;   // E of the block's first row
        }
        __m256i carried = Z;   // H(i0 - 1, c0 - 1)
        for (uint32_t i0 = 0; i0 < len; i0 += kRows) {
            for (int r = 0; r < kRows; ++r) {
                const uint8_t* codes = inter + (size_t)(i0 + r) * T::kLanes;
                for (int k = 0; k < tb.n_present; ++k) {
                    const int a = tb.present[k];
                    prof[a * kRows + r] = T::lookup(tb.row[a], codes);
                }
            }
            __m256i F[kRows], Hl[kRows];
            __m256i diag_top = Z;
            if (first) {
                for (int r = 0; r < kRows; ++r) F[r] = Hl[r] = Z;
            } else {
                const __m256i* in = edge + 2 * (size_t)i0;
                for (int r = 0; r < kRows; ++r) {
                    Hl[r] = in[2 * r];
                    F[r] = in[2 * r + 1];
                }
                diag_top = carried;
                carried = Hl[kRows - 1];
            }
            for (int j = 0; j < w; ++j) {
                const __m256i* P = static_cast<const __m256i*>(qs[j]);
                const __m256i hup = col[2 * j];
                __m256i e = col[2 * j + 1];
                __m256i d = diag_top;
                for (int r = 0; r < kRows; ++r) {
                    __m256i h = T::adds(d, P[r]);
                    h = T::max(h, e);
                    h = T::max(h, F[r]);
                    S = T::max(S, h);
                    const __m256i hq = T::subs(h, OE);
                    F[r] = T::max(T::subs(F[r], GE), hq);
                    e = T::max(T::subs(e, GE), hq);
                    d = Hl[r];
                    Hl[r] = h;
                }
                diag_top = hup;
                col[2 * j] = Hl[kRows - 1];
                col[2 * j + 1] = e;
            }
            if (!last) {
                __m256i* out = edge + 2 * (size_t)i0;
                for (int r = 0; r < kRows; ++r) {
                    out[2 * r] = Hl[r];       // H(i, c0 + w - 1)
                    out[2 * r + 1] = F[r];    // F(i, c0 + w)
                }
            }
        }
    }
    return S;
}

template <class T>
void run_groups(const Tables& tb, const uint8_t* q, int m, const uint8_t* const* seqs,
                const uint32_t* lens, const uint32_t* idx, size_t n, int gap_oe, int gap_e,
                int seg, int32_t* score_of_idx_pos, Scratch& s) {
    s.colstate.resize((size_t)std::min(m, seg) * 2 * sizeof(__m256i) + 32);
    s.prof.resize(kSymbols * kRows * sizeof(__m256i) + 32);
    __m256i* col = align32(s.colstate);
    __m256i* prof = align32(s.prof);
    const void* const* qp = column_rows(q, m, prof, s);
    for (size_t g = 0; g < n; g += T::kLanes) {
        const int cnt = (int)std::min<size_t>(T::kLanes, n - g);
        uint32_t rows = 0;
        for (int l = 0; l < cnt; ++l) rows = std::max(rows, lens[idx[g + l]]);
        rows = std::max<uint32_t>(kRows, (rows + kRows - 1) / kRows * kRows);
        if (s.inter.size() < (size_t)rows * T::kLanes) s.inter.resize((size_t)rows * T::kLanes);
        interleave<T>(seqs, lens, rows, idx + g, cnt, s.inter.data());
        __m256i* edge = nullptr;
        if (seg < m) {
            if (s.edge.size() < (size_t)rows * 2 * sizeof(__m256i) + 32)
                s.edge.resize((size_t)rows * 2 * sizeof(__m256i) + 32);
            edge = align32(s.edge);
        }
        const __m256i S =
            scan_group<T>(tb, qp, m, s.inter.data(), rows, gap_oe, gap_e, seg, col, prof, edge);
        int32_t lanes[T::kLanes];
        T::scores(S, lanes);
        for (int l = 0; l < cnt; ++l) score_of_idx_pos[g + l] = lanes[l];
    }
}

}  // namespace

int query_segment(const Scoring& sc, int qlen) {
    const int seg = sc.query_segment == 0 ? kDefaultQuerySegment : sc.query_segment;
    return seg < 0 || seg >= qlen ? qlen : seg;
}

void score_seqs(const Scoring& sc, const uint8_t* query, int qlen, const uint8_t* const* seqs,
                const uint32_t* lens, uint32_t n, int32_t* out, Scratch& s) {
    if (n == 0) return;
    if (qlen <= 0) {
        std::fill(out, out + n, 0);
        return;
    }
    Tables tb;
    make_tables(sc, query, qlen, tb);
    const int oe = sc.gap_open + sc.gap_extend, ge = sc.gap_extend;
    const int seg = query_segment(sc, qlen);
    s.all.resize(n);
    for (uint32_t k = 0; k < n; ++k) s.all[k] = k;
    s.ovf.clear();
    if (oe <= 127) {
        // int8: exact while no add saturated, i.e. while the score stays below
        // 256 - max(matrix)
        s.s8.resize(n);
        run_groups<I8>(tb, query, qlen, seqs, lens, s.all.data(), n, oe, ge, seg, s.s8.data(), s);
        for (uint32_t k = 0; k < n; ++k) {
            if (s.s8[k] >= I8::kTop + 1 - tb.max_score) s.ovf.push_back(k);
            else out[k] = s.s8[k];
        }
    } else {
        s.ovf = s.all;
    }
    if (s.ovf.empty()) return;
    s.ovf_scores.resize(s.ovf.size());
    run_groups<I16>(tb, query, qlen, seqs, lens, s.ovf.data(), s.ovf.size(), oe, ge, seg,
                    s.ovf_scores.data(), s);
    for (size_t i = 0; i < s.ovf.size(); ++i) {
        const int v = s.ovf_scores[i];
        out[s.ovf[i]] = v >= I16::kTop + 1 - tb.max_score ? -1 : v;
    }
}

void score_team(const Scoring& sc, const uint8_t* query, int qlen, const uint8_t* const* seqs,
                const uint32_t* lens, size_t n, int32_t* out, int threads) {
    if (n == 0) return;
    std::vector<uint32_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return lens[a] > lens[b]; });
    // pieces of whole 32-lane groups of about 256 KB of residues, longest first
    struct Piece {
        size_t first, n;
    };
    std::vector<Piece> pieces;
    for (size_t i = 0; i < n;) {
        size_t k = 0, bytes = 0;
        while (i + k < n && (k == 0 || k % 32 != 0 || bytes < (size_t(256) << 10))) {
            bytes += lens[order[i + k]];
            ++k;
        }
        pieces.push_back(Piece{i, k});
        i += k;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        Scratch scratch;
        std::vector<int32_t> sc_out;
        std::vector<const uint8_t*> pp;
        std::vector<uint32_t> pl;
        for (size_t p = next++; p < pieces.size(); p = next++) {
            const Piece& pc = pieces[p];
            pp.resize(pc.n);
            pl.resize(pc.n);
            for (size_t k = 0; k < pc.n; ++k) {
                pp[k] = seqs[order[pc.first + k]];
                pl[k] = lens[order[pc.first + k]];
            }
            sc_out.resize(pc.n);
            score_seqs(sc, query, qlen, pp.data(), pl.data(), (uint32_t)pc.n, sc_out.data(),
                       scratch);
            for (size_t k = 0; k < pc.n; ++k) out[order[pc.first + k]] = sc_out[k];
        }
    };
    std::vector<std::thread> team;
    for (int t = 1; t < threads; ++t) team.emplace_back(work);
    work();
    for (auto& t : team) t.join();
}

void pack_db(const uint8_t* const* seqs, const uint32_t* lens, size_t n, PackedDb& db,
             int threads) {
    db.seqs = seqs;
    db.lens = lens;
    db.n = n;
    std::vector<uint32_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return lens[a] > lens[b]; });
    const size_t ng = (n + I8::kLanes - 1) / I8::kLanes;
    db.group_off.assign(ng, 0);
    db.group_rows.assign(ng, 0);
    db.lane_seq.assign(ng * I8::kLanes, ~0u);
    uint64_t off = 0;
    for (size_t g = 0; g < ng; ++g) {
        const uint32_t L = lens[order[g * I8::kLanes]];
        const uint32_t rows = std::max<uint32_t>(kRows, (L + kRows - 1) / kRows * kRows);
        db.group_off[g] = off;
        db.group_rows[g] = rows;
        off += (uint64_t)rows * I8::kLanes;
        for (int l = 0; l < I8::kLanes && g * I8::kLanes + l < n; ++l)
            db.lane_seq[g * I8::kLanes + l] = order[g * I8::kLanes + l];
    }
    db.data.resize(off);
    std::atomic<size_t> next{0};
    auto work = [&]() {
        uint32_t idx[I8::kLanes];
        for (size_t g = next++; g < ng; g = next++) {
            int cnt = 0;
            while (cnt < I8::kLanes && db.lane_seq[g * I8::kLanes + cnt] != ~0u) {
                idx[cnt] = db.lane_seq[g * I8::kLanes + cnt];
                ++cnt;
            }
            interleave<I8>(seqs, lens, db.group_rows[g], idx, cnt, db.data.data() + db.group_off[g]);
        }
    };
    std::vector<std::thread> team;
    for (int t = 1; t < threads; ++t) team.emplace_back(work);
    work();
    for (auto& t : team) t.join();
}

void scan_db(const Scoring& sc, const uint8_t* query, int qlen, const PackedDb& db, int32_t* out,
             int threads) {
    if (db.n == 0) return;
    if (qlen <= 0) {
        std::fill(out, out + db.n, 0);
        return;
    }
    Tables tb;
    make_tables(sc, query, qlen, tb);
    const int oe = sc.gap_open + sc.gap_extend, ge = sc.gap_extend;
    const int seg = query_segment(sc, qlen);
    const size_t ng = db.group_rows.size();
    std::vector<uint32_t> ovf;
    std::mutex ovf_mu;
    if (oe <= 127) {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            Scratch s;
            s.colstate.resize((size_t)seg * 2 * sizeof(__m256i) + 32);
            s.prof.resize(kSymbols * kRows * sizeof(__m256i) + 32);
            __m256i* col = align32(s.colstate);
            __m256i* prof = align32(s.prof);
            const void* const* qp = column_rows(query, qlen, prof, s);
            std::vector<uint32_t> mine;
            for (size_t g = next++; g < ng; g = next++) {
                __m256i* edge = nullptr;
                if (seg < qlen) {
                    const size_t want = (size_t)db.group_rows[g] * 2 * sizeof(__m256i) + 32;
                    if (s.edge.size() < want) s.edge.resize(want);
                    edge = align32(s.edge);
                }
                const __m256i S = scan_group<I8>(tb, qp, qlen, db.data.data() + db.group_off[g],
                                                 db.group_rows[g], oe, ge, seg, col, prof, edge);
                int32_t lanes[I8::kLanes];
                I8::scores(S, lanes);
                for (int l = 0; l < I8::kLanes; ++l) {
                    const uint32_t k = db.lane_seq[g * I8::kLanes + l];
                    if (k == ~0u) break;
                    // exact while no add saturated: below 256 - max(matrix)
                    if (lanes[l] >= I8::kTop + 1 - tb.max_score) mine.push_back(k);
                    else out[k] = lanes[l];
                }
            }
            std::lock_guard<std::mutex> g(ovf_mu);
            ovf.insert(ovf.end(), mine.begin(), mine.end());
        };
        std::vector<std::thread> team;
        for (int t = 1; t < threads; ++t) team.emplace_back(work);
        work();
        for (auto& t : team) t.join();
    } else {
        ovf.resize(db.n);
        for (size_t k = 0; k < db.n; ++k) ovf[k] = (uint32_t)k;
    }
    if (ovf.empty()) return;
    // int16 for the lanes that may have saturated, longest first
    std::sort(ovf.begin(), ovf.end(), [&](uint32_t a, uint32_t b) {
        return db.lens[a] != db.lens[b] ? db.lens[a] > db.lens[b] : a < b;
    });
    std::atomic<size_t> next{0};
    auto work16 = [&]() {
        Scratch s;
        std::vector<int32_t> r;
        for (size_t p = next++ * I16::kLanes; p < ovf.size(); p = next++ * I16::kLanes) {
            const size_t cnt = std::min<size_t>(I16::kLanes, ovf.size() - p);
            r.resize(cnt);
            run_groups<I16>(tb, query, qlen, db.seqs, db.lens, ovf.data() + p, cnt, oe, ge, seg,
                            r.data(), s);
            for (size_t i = 0; i < cnt; ++i)
                out[ovf[p + i]] = r[i] >= I16::kTop + 1 - tb.max_score ? -1 : r[i];
        }
    };
    std::vector<std::thread> team;
    for (int t = 1; t < threads; ++t) team.emplace_back(work16);
    work16();
    for (auto& t : team) t.join();
}

void score_batch(const Scoring& sc, const uint8_t* query, int qlen, const uint8_t* seqs,
                 uint32_t n, uint32_t len, int32_t* out, Scratch& s) {
    std::vector<const uint8_t*> ptrs(n);
    std::vector<uint32_t> lens(n, len);
    for (uint32_t k = 0; k < n; ++k) ptrs[k] = seqs + (size_t)k * len;
    score_seqs(sc, query, qlen, ptrs.data(), lens.data(), n, out, s);
}

}  // namespace cpu
}  // namespace swhy
