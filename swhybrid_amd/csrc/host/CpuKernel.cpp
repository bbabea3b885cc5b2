// CpuKernel.cpp -- compiled with -mavx2; callers check avx2_available().
#include "CpuKernel.h"

#include <immintrin.h>

#include <algorithm>
#include <cstring>

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

// One lane group over every row: H and E of the previous row block per query
// column in colstate, 4 rows per pass through the query.  Per cell: 1 add,
// 3 sub, 5 max (the reference's count).
template <class T>
__m256i scan_group(const Tables& tb, const uint8_t* q, int m, const uint8_t* inter, uint32_t len,
                   int gap_oe, int gap_e, __m256i* col, __m256i* prof) {
    const __m256i Z = T::set1(-T::kBias);
    const __m256i OE = T::set1(gap_oe);
    const __m256i GE = T::set1(gap_e);
    for (int j = 0; j < m; ++j) {
        col[2 * j] = Z;       // H of the row above the block
        col[2 * j + 1] = Z;   // E of the block's first row
    }
    __m256i S = Z;
    for (uint32_t i0 = 0; i0 < len; i0 += kRows) {
        for (int r = 0; r < kRows; ++r) {
            const uint8_t* codes = inter + (size_t)(i0 + r) * T::kLanes;
            for (int k = 0; k < tb.n_present; ++k) {
                const int a = tb.present[k];
                prof[a * kRows + r] = T::lookup(tb.row[a], codes);
            }
        }
        __m256i F[kRows], Hl[kRows];
        for (int r = 0; r < kRows; ++r) F[r] = Hl[r] = Z;
        __m256i diag_top = Z;
        for (int j = 0; j < m; ++j) {
            const __m256i* P = prof + q[j] * kRows;
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
    }
    return S;
}

template <class T>
void run_groups(const Tables& tb, const uint8_t* q, int m, const uint8_t* const* seqs,
                const uint32_t* lens, const uint32_t* idx, size_t n, int gap_oe, int gap_e,
                int32_t* score_of_idx_pos, Scratch& s) {
    s.colstate.resize((size_t)m * 2 * sizeof(__m256i) + 32);
    s.prof.resize(kSymbols * kRows * sizeof(__m256i) + 32);
    auto align = [](std::vector<uint8_t>& v) {
        return reinterpret_cast<__m256i*>(
            (reinterpret_cast<uintptr_t>(v.data()) + 31) & ~uintptr_t(31));
    };
    __m256i* col = align(s.colstate);
    __m256i* prof = align(s.prof);
    for (size_t g = 0; g < n; g += T::kLanes) {
        const int cnt = (int)std::min<size_t>(T::kLanes, n - g);
        uint32_t rows = 0;
        for (int l = 0; l < cnt; ++l) rows = std::max(rows, lens[idx[g + l]]);
        rows = std::max<uint32_t>(kRows, (rows + kRows - 1) / kRows * kRows);
        if (s.inter.size() < (size_t)rows * T::kLanes) s.inter.resize((size_t)rows * T::kLanes);
        interleave<T>(seqs, lens, rows, idx + g, cnt, s.inter.data());
        const __m256i S = scan_group<T>(tb, q, m, s.inter.data(), rows, gap_oe, gap_e, col, prof);
        int32_t lanes[T::kLanes];
        T::scores(S, lanes);
        for (int l = 0; l < cnt; ++l) score_of_idx_pos[g + l] = lanes[l];
    }
}

}  // namespace

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
    s.all.resize(n);
    for (uint32_t k = 0; k < n; ++k) s.all[k] = k;
    s.ovf.clear();
    if (oe <= 127) {
        // int8: exact while no add saturated, i.e. while the score stays below
        // 256 - max(matrix)
        s.s8.resize(n);
        run_groups<I8>(tb, query, qlen, seqs, lens, s.all.data(), n, oe, ge, s.s8.data(), s);
        for (uint32_t k = 0; k < n; ++k) {
            if (s.s8[k] >= I8::kTop + 1 - tb.max_score) s.ovf.push_back(k);
            else out[k] = s.s8[k];
        }
    } else {
        s.ovf = s.all;
    }
    if (s.ovf.empty()) return;
    s.ovf_scores.resize(s.ovf.size());
    run_groups<I16>(tb, query, qlen, seqs, lens, s.ovf.data(), s.ovf.size(), oe, ge,
                    s.ovf_scores.data(), s);
    for (size_t i = 0; i < s.ovf.size(); ++i) {
        const int v = s.ovf_scores[i];
        out[s.ovf[i]] = v >= I16::kTop + 1 - tb.max_score ? -1 : v;
    }
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
