// swh_kernels.hip -- gfx950 kernels of the Smith-Waterman database search.
//
// Inter-sequence affine-gap local alignment (the recurrence of the reference's
// ONE_CELL_UPDATE, src/AVXKernel.cpp:196-218):
//   H(i,j) = max(0, H(i-1,j-1) + s(q_j, d_i), E(i,j), F(i,j))
//   E(i,j) = max(E(i-1,j) - ext, H(i-1,j) - (open+ext))      (gap along the db)
//   F(i,j) = max(F(i,j-1) - ext, H(i,j-1) - (open+ext))      (gap along the query)
//   score  = max H
// Database sequences run along the rows i, the query along the columns j.
//
// Scan kernel (int16): every lane carries TWO database sequences, one in each
// 16-bit half of a VGPR, and all arithmetic is packed and saturating
// (v_pk_add_i16 clamp / v_pk_sub_i16 clamp / v_pk_max_i16, and for the
// three-way maxima v_pk_minimum3_f16 on the same bit patterns, see vmax3).
// Values are biased so that 0x8000 encodes score 0: the saturating floor at -32768 is the
// local-alignment zero floor, exactly as the reference's 0x80 / 0x8000 bias
// (src/AVXKernel.cpp:443, :903).  A lane keeps a strip of W query columns in
// registers (H and E per column) and walks down its sequences' rows; the H/F
// values on the strip's right edge go to a per-wave border array in HBM and
// feed the next strip.  The profile comes from LDS: either a residue-pair
// table built per strip (a row reads its pair's W dwords directly), or the
// query profile (24 x qlen int16) staged per 1344-column segment, whose two
// residue rows a row interleaves with one v_perm_b32 per column.  Sequences
// too long for one lane are swept by groups of lanes (group_item), and a
// query's last strip is as narrow as its length rounded up to 8 allows.
//
// A sequence whose int16 score reaches the overflow threshold (default 31744,
// below which no add can have saturated and the three-way max orders) is
// appended to an overflow list and rescored exactly by the int32 kernel, on
// the device.
#include "swh_internal.h"

#include <mutex>
#include <type_traits>

namespace swh {

typedef short s2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s2 adds(s2 a, s2 b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ s2 subs(s2 a, s2 b) { return __builtin_elementwise_sub_sat(a, b); }
__device__ __forceinline__ s2 vmax(s2 a, s2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ int adds(int a, int b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ int subs(int a, int b) { return __builtin_elementwise_sub_sat(a, b); }
__device__ __forceinline__ int vmax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ uint32_t bits(s2 v) { return __builtin_bit_cast(uint32_t, v); }

// Three-way max of biased int16 scores in one instruction.  The bit pattern
// 0x8000 + s of a score s is, read as an fp16, the negative number -f(s) with
// f increasing in s while s < 0x7C00 (subnormals included: the kernels run
// with fp16 denormals enabled), so the IEEE minimum of the patterns is the
// pattern of the largest score: v_pk_minimum3_f16 does two packed max ops.
// Patterns of scores from 31744 on are -inf or NaN and do not order; the scan
// sends every sequence whose score reaches kScoreLimitF16 to the int32
// rescore, and such a pattern, once formed, makes the score reach it.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s2 vmax3(s2 a, s2 b, s2 c) {
    const h2 x = __builtin_bit_cast(h2, a), y = __builtin_bit_cast(h2, b),
             z = __builtin_bit_cast(h2, c);
    return __builtin_bit_cast(s2, __builtin_elementwise_minimum(
                                      __builtin_elementwise_minimum(x, y), z));
}
__device__ __forceinline__ int vmax3(int a, int b, int c) { return vmax(vmax(a, b), c); }
__device__ __forceinline__ s2 as_s2(uint32_t v) { return __builtin_bit_cast(s2, v); }

constexpr uint32_t kZero16x2 = 0x80008000u;   // biased zero in both halves
constexpr int kZero32 = INT32_MIN;             // biased zero, int32 rescore

// Copy profile rows 0..23, columns [seg0, seg0+segw) into LDS rows of stride ps.
__device__ __forceinline__ void stage_profile(short* lds, const int16_t* __restrict__ prof,
                                              int pstride, int ps, int seg0, int segw) {
    const int per_row = segw >> 3;
    const int total = kProfRows * per_row;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int r = t / per_row;
        const int c8 = t - r * per_row;
        const uint4 v = *reinterpret_cast<const uint4*>(prof + (size_t)r * pstride + seg0 + c8 * 8);
        *reinterpret_cast<uint4*>(lds + r * ps + c8 * 8) = v;
    }
}

// One row of a strip: W cells, packed pairs.  P[c] = (s(q_c, a), s(q_c, b)).
// h holds H(i-1,c) on entry and H(i,c) on exit; e holds E(i,c) on entry and
// E(i+1,c) on exit, so H(i,c) - (open+ext) is used where it is made and never
// kept across rows.
template <int W, typename T>
__device__ __forceinline__ void row_cells(T (&h)[W], T (&e)[W], const T (&P)[W], T diag0, T hl,
                                          T fl, T Q, T R, T& fout) {
    T X = adds(diag0, P[0]);      // H(i-1, c0-1) + s
    T f = fl;                      // F(i, c0-1)
    T hqp = subs(hl, Q);           // H(i, c0-1) - (open+ext)
#pragma unroll
    for (int c = 0; c < W; ++c) {
        f = vmax(subs(f, R), hqp);                 // F(i, c)
        T Xn = X;
        if (c + 1 < W) Xn = adds(h[c], P[c + 1]);  // diagonal term of column c+1
        const T hn = vmax3(X, e[c], f);            // H(i, c); zero floor by saturation
        const T hq = subs(hn, Q);
        e[c] = vmax(subs(e[c], R), hq);            // E(i+1, c)
        h[c] = hn;
        hqp = hq;
        X = Xn;
    }
    fout = f;
}

// The running score folded with the maximum of a row: a tree of three-way
// maxima over the W values and S (no state carried between rows).
template <int W, typename T>
__device__ __forceinline__ T row_max_into(T S, const T (&h)[W]) {
    T m[W + 1];
#pragma unroll
    for (int c = 0; c < W; ++c) m[c] = h[c];
    m[W] = S;
    int n = W + 1;
#pragma unroll
    for (int level = 0; level < 8; ++level) {
        if (n == 1) break;
        int k = 0;
#pragma unroll
        for (int i = 0; i < W + 1; i += 3) {
            if (i >= n) break;
            if (i + 2 < n) m[k++] = vmax3(m[i], m[i + 1], m[i + 2]);
            else if (i + 1 < n) m[k++] = vmax(m[i], m[i + 1]);
            else m[k++] = m[i];
        }
        n = k;
    }
    return m[0];
}

// Raw profile rows of the lane's two residues (W int16 each), and their
// interleave into (s(q_c, a), s(q_c, b)) pairs with one v_perm_b32 per column.
template <int W>
struct RawProfile {
    uint4 va[W / 8];
    uint4 vb[W / 8];
};

template <int W>
__device__ __forceinline__ void load_profile(RawProfile<W>& rp, const short* lds, int ps, int a,
                                             int b, int jj0) {
    const uint4* pa = reinterpret_cast<const uint4*>(lds + __umul24(a, ps) + jj0);
    const uint4* pb = reinterpret_cast<const uint4*>(lds + __umul24(b, ps) + jj0);
#pragma unroll
    for (int k = 0; k < W / 8; ++k) {
        rp.va[k] = pa[k];
        rp.vb[k] = pb[k];
    }
}

// The same rows from a lane's first staged column (col0) and the row stride in
// bytes: one multiply-add per residue forms the LDS address.
template <int W>
__device__ __forceinline__ void load_profile_at(RawProfile<W>& rp, const char* col0,
                                                uint32_t row_bytes, uint32_t a, uint32_t b) {
    const uint4* pa = reinterpret_cast<const uint4*>(col0 + __umul24(a, row_bytes));
    const uint4* pb = reinterpret_cast<const uint4*>(col0 + __umul24(b, row_bytes));
#pragma unroll
    for (int k = 0; k < W / 8; ++k) {
        rp.va[k] = pa[k];
        rp.vb[k] = pb[k];
    }
}

template <int W>
__device__ __forceinline__ void interleave_profile(s2 (&P)[W], const RawProfile<W>& rp) {
#pragma unroll
    for (int k = 0; k < W / 8; ++k) {
        const uint4 va = rp.va[k];
        const uint4 vb = rp.vb[k];
        P[8 * k + 0] = as_s2(__builtin_amdgcn_perm(vb.x, va.x, 0x05040100u));
        P[8 * k + 1] = as_s2(__builtin_amdgcn_perm(vb.x, va.x, 0x07060302u));
        P[8 * k + 2] = as_s2(__builtin_amdgcn_perm(vb.y, va.y, 0x05040100u));
        P[8 * k + 3] = as_s2(__builtin_amdgcn_perm(vb.y, va.y, 0x07060302u));
        P[8 * k + 4] = as_s2(__builtin_amdgcn_perm(vb.z, va.z, 0x05040100u));
        P[8 * k + 5] = as_s2(__builtin_amdgcn_perm(vb.z, va.z, 0x07060302u));
        P[8 * k + 6] = as_s2(__builtin_amdgcn_perm(vb.w, va.w, 0x05040100u));
        P[8 * k + 7] = as_s2(__builtin_amdgcn_perm(vb.w, va.w, 0x07060302u));
    }
}

// Pair table: for the W columns of one strip, the interleaved profile pairs
// (s(q_c, a), s(q_c, b)) of every residue pair (a, b), 576 rows of W dwords
// padded to an odd number of 16-byte chunks (rows spread over LDS banks).
// Built by the whole workgroup from the global profile once per strip; a lane
// then reads its row's pairs directly, with no per-column interleave.
constexpr int kPairs = kProfRows * kProfRows;

template <int W>
struct PairTab {
    static constexpr int stride = ((W / 4) % 2 == 0) ? W + 4 : W;   // dwords per pair row
    static constexpr size_t bytes = (size_t)kPairs * stride * sizeof(uint32_t);
};

template <int W>
__device__ __forceinline__ void build_pair_table(uint32_t* tab, const int16_t* __restrict__ prof,
                                                 int pstride, int c0) {
    for (int p = threadIdx.x; p < kPairs; p += blockDim.x) {
        const int a = p / kProfRows;
        const int b = p - a * kProfRows;
        const uint4* ga = reinterpret_cast<const uint4*>(prof + (size_t)a * pstride + c0);
        const uint4* gb = reinterpret_cast<const uint4*>(prof + (size_t)b * pstride + c0);
        uint4* out = reinterpret_cast<uint4*>(tab + (size_t)p * PairTab<W>::stride);
#pragma unroll
        for (int k = 0; k < W / 8; ++k) {
            const uint4 va = ga[k];
            const uint4 vb = gb[k];
            out[2 * k] = make_uint4(__builtin_amdgcn_perm(vb.x, va.x, 0x05040100u),
                                    __builtin_amdgcn_perm(vb.x, va.x, 0x07060302u),
                                    __builtin_amdgcn_perm(vb.y, va.y, 0x05040100u),
                                    __builtin_amdgcn_perm(vb.y, va.y, 0x07060302u));
            out[2 * k + 1] = make_uint4(__builtin_amdgcn_perm(vb.z, va.z, 0x05040100u),
                                        __builtin_amdgcn_perm(vb.z, va.z, 0x07060302u),
                                        __builtin_amdgcn_perm(vb.w, va.w, 0x05040100u),
                                        __builtin_amdgcn_perm(vb.w, va.w, 0x07060302u));
        }
    }
}

// Profile sources for lane_strip: the staged single profile (two residue rows
// and a v_perm_b32 per column) or the pair table (one row, no interleave).
template <int W>
struct SingleSource {
    using Raw = RawProfile<W>;
    const short* lds;
    int ps, jj0;
    __device__ __forceinline__ void load(Raw& r, int a, int b) const {
        load_profile<W>(r, lds, ps, a, b, jj0);
    }
    __device__ __forceinline__ void pairs(s2 (&P)[W], const Raw& r) const {
        interleave_profile<W>(P, r);
    }
};

template <int W>
struct PairSource {
    struct Raw {
        uint4 v[W / 4];
    };
    const uint32_t* tab;
    __device__ __forceinline__ void load(Raw& r, int a, int b) const {
        const uint4* p = reinterpret_cast<const uint4*>(
            tab + __umul24(__umul24(a, kProfRows) + b, PairTab<W>::stride));
#pragma unroll
        for (int k = 0; k < W / 4; ++k) r.v[k] = p[k];
    }
    __device__ __forceinline__ void pairs(s2 (&P)[W], const Raw& r) const {
#pragma unroll
        for (int k = 0; k < W / 4; ++k) {
            P[4 * k + 0] = as_s2(r.v[k].x);
            P[4 * k + 1] = as_s2(r.v[k].y);
            P[4 * k + 2] = as_s2(r.v[k].z);
            P[4 * k + 3] = as_s2(r.v[k].w);
        }
    }
};

// One strip of W columns over every row of a lane-mode block.  IN / OUT: the
// strip reads / writes the border array (compile-time, so that a 4-row tile is
// a single basic block and the compiler keeps each row's work in place).
template <int W, bool IN, bool OUT, typename Src>
__device__ __forceinline__ void lane_strip(const uint2* __restrict__ wds, int quads,
                                           uint2* __restrict__ bord, const Src& src, s2 Q, s2 R,
                                           s2& S, int lane) {
    const s2 Z = as_s2(kZero16x2);
    s2 h[W], e[W];
#pragma unroll
    for (int c = 0; c < W; ++c) {
        h[c] = Z;
        e[c] = Z;
    }
    s2 diag0 = Z;
    const uint2 zz = make_uint2(kZero16x2, kZero16x2);
    // three register sets of (word, borders): tile k computes from one while
    // the next tile's set is in flight and tile k + 2's loads are issued
    // (quads >= 1: a block has at least 4 rows)
    uint2 wa = wds[lane];
    uint2 ba[4], bb[4], bc[4];
    const int k1 = max(min(1, quads - 1), 0);
    uint2 wb = wds[(size_t)k1 * kWave + lane], wc = wb;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        ba[r] = IN ? bord[r * kWave + lane] : zz;
        bb[r] = IN ? bord[((size_t)k1 * 4 + r) * kWave + lane] : zz;
        bc[r] = bb[r];
    }
    typename Src::Raw rp;
    src.load(rp, wa.x & 0xff, wa.y & 0xff);
    // tile k: its word w and borders b, the next tile's word wn (in flight),
    // and the set (wf, bf) that receives tile k + 2
    auto tile = [&](int k, const uint2& w, const uint2 (&b)[4], const uint2& wn, uint2& wf,
                    uint2 (&bf)[4]) {
        // prefetch tile k + 2 (the last tiles reload the last one: no branch)
        const int kf = min(k + 2, quads - 1);
        wf = wds[(size_t)kf * kWave + lane];
        if (IN) {
            const uint2* bp = bord + (size_t)kf * 4 * kWave + lane;
#pragma unroll
            for (int r = 0; r < 4; ++r) bf[r] = bp[r * kWave];
        }
        uint2* bo = bord + (size_t)k * 4 * kWave + lane;
        // the profile of row r+1 is loaded while row r computes: the barrier
        // after the loads keeps the scheduler from sinking them to the end of
        // the row (where the next row would wait for them), the one at the end
        // of the row from hoisting every load of the tile at once
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const typename Src::Raw cur = rp;
            const uint2 nw = r < 3 ? w : wn;
            const int sh = 8 * ((r + 1) & 3);
            src.load(rp, (nw.x >> sh) & 0xff, (nw.y >> sh) & 0xff);
            __builtin_amdgcn_sched_barrier(0);
            s2 P[W];
            src.pairs(P, cur);
            const s2 hl = as_s2(b[r].x);
            const s2 fl = as_s2(b[r].y);
            s2 fout;
            row_cells<W, s2>(h, e, P, diag0, hl, fl, Q, R, fout);
            S = row_max_into<W, s2>(S, h);
            if (OUT) bo[r * kWave] = make_uint2(bits(h[W - 1]), bits(fout));
            diag0 = hl;
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // unrolled three times with the sets rotating roles, so no copies move the
    // prefetched values (rotating them made every tile wait for its loads)
    int k = 0;
    for (; k + 2 < quads; k += 3) {
        tile(k, wa, ba, wb, wc, bc);
        tile(k + 1, wb, bb, wc, wa, ba);
        tile(k + 2, wc, bc, wa, wb, bb);
    }
    if (k < quads) tile(k, wa, ba, wb, wc, bc);
    if (k + 1 < quads) tile(k + 1, wb, bb, wc, wa, ba);
}

template <int W, typename Src>
__device__ __forceinline__ void lane_strip_any(const uint2* wds, int quads, uint2* bord,
                                               const Src& src, bool has_in, bool has_out, s2 Q,
                                               s2 R, s2& S, int lane) {
    if (has_in) {
        if (has_out) lane_strip<W, true, true>(wds, quads, bord, src, Q, R, S, lane);
        else lane_strip<W, true, false>(wds, quads, bord, src, Q, R, S, lane);
    } else {
        if (has_out) lane_strip<W, false, true>(wds, quads, bord, src, Q, R, S, lane);
        else lane_strip<W, false, false>(wds, quads, bord, src, Q, R, S, lane);
    }
}

// Score of sorted sequence s, or its overflow-list entry (and *ovf = 1, an
// LDS flag of the workgroup's item).
__device__ __forceinline__ void emit(const QueryDesc& Qd, const uint32_t* perm, uint32_t s,
                                     int sc, int* ovf) {
    const uint32_t idx = perm[s];
    if (idx == kNoSeq) return;
    if (sc >= Qd.ovf_threshold) {
        const uint32_t slot = atomicAdd(Qd.ovf_count, 1u);
        Qd.ovf_list[slot] = s;
        Qd.scores[idx] = -1;
        *ovf = 1;
    } else {
        Qd.scores[idx] = sc;
    }
}

// Calls f with std::integral_constant<int, w> for a strip of wl columns: the
// kernel's width W, or for the query's last strip a narrower one (wl is a
// multiple of 8 below W), so a query is computed to its length rounded to 8
// rather than to W.
template <int W, typename F>
__device__ __forceinline__ void with_strip_width(int wl, F&& f) {
    if (wl == W) {
        f(std::integral_constant<int, W>{});
    } else {
        if constexpr (W > 24) {
            if (wl == 24) {
                f(std::integral_constant<int, 24>{});
                return;
            }
        }
        if constexpr (W > 16) {
            if (wl == 16) {
                f(std::integral_constant<int, 16>{});
                return;
            }
        }
        f(std::integral_constant<int, 8>{});
    }
}

// One lane-mode work item: 4 consecutive lane blocks of query q, one per wave.
// PAIR (W = 16 or 24): per strip the workgroup builds the pair table in LDS;
// otherwise the single profile is staged per 1344-column segment and each row
// interleaves two residue rows (a 32-wide pair table would not leave room for
// two workgroups per CU, so W = 32 always stages the single profile).  The
// last strip is qpad - c0 wide (qpad = qlen rounded up to 8).
template <int W, bool PAIR>
__device__ __forceinline__ void lane_item(const ScanArgs& A, const QueryDesc& Qd, int q,
                                          int local, short* lds, uint2* bord, int& loaded_q,
                                          int& loaded_seg, int* ovf) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const s2 Q = {(short)A.gap_oe, (short)A.gap_oe};
    const s2 R = {(short)A.gap_e, (short)A.gap_e};
    const int blk = __builtin_amdgcn_readfirstlane(Qd.blk_begin + local * kWavesPerWG + wave);
    const bool active = blk < Qd.blk_end;
    BlockDesc bd = {0, 0, 0};
    if (active) bd = A.blocks[blk];
    const uint2* wds = reinterpret_cast<const uint2*>(A.words + bd.word_off);
    const int quads = (int)(bd.rows >> 2);
    s2 S = as_s2(kZero16x2);
    if constexpr (!PAIR) {
        const int nseg = (Qd.qpad + kSegCols - 1) / kSegCols;
        for (int seg = 0; seg < nseg; ++seg) {
            const int seg0 = seg * kSegCols;
            const int segw = min(kSegCols, Qd.qpad - seg0);
            if (q != loaded_q || seg != loaded_seg) {
                __syncthreads();
                stage_profile(lds, Qd.prof, Qd.pstride, A.ps_lane, seg0, segw);
                __syncthreads();
                loaded_q = q;
                loaded_seg = seg;
            }
            if (active) {
                for (int c0 = seg0; c0 < seg0 + segw; c0 += W) {
                    const int wl = min(W, seg0 + segw - c0);
                    with_strip_width<W>(wl, [&](auto wc) {
                        constexpr int w = decltype(wc)::value;
                        const SingleSource<w> src{lds, A.ps_lane, c0 - seg0};
                        lane_strip_any<w>(wds, quads, bord, src, c0 > 0, c0 + w < Qd.qpad, Q,
                                          R, S, lane);
                    });
                }
            }
        }
    } else {
        uint32_t* tab = reinterpret_cast<uint32_t*>(lds);
        for (int c0 = 0; c0 < Qd.qpad; c0 += W) {
            const int wl = min(W, Qd.qpad - c0);
            with_strip_width<W>(wl, [&](auto wc) {
                constexpr int w = decltype(wc)::value;
                __syncthreads();
                build_pair_table<w>(tab, Qd.prof, Qd.pstride, c0);
                __syncthreads();
                if (active) {
                    const PairSource<w> src{tab};
                    lane_strip_any<w>(wds, quads, bord, src, c0 > 0, c0 + w < Qd.qpad, Q, R, S,
                                      lane);
                }
            });
        }
        loaded_q = -1;   // the single-profile staging area was overwritten
    }
    if (active) {
        emit(Qd, A.perm, bd.first + lane, (int)S.x + 32768, ovf);
        emit(Qd, A.perm, bd.first + kWave + lane, (int)S.y + 32768, ovf);
    }
}

// ---------------------------------------------------------------------------
// Group mode for long database sequences.  G lanes share one sequence pair:
// lane g of a group keeps query columns [pass*G*W + g*W, +W) in registers and
// at step t works on rows NR(t-g) .. NR(t-g)+NR-1 (NR = group_step_rows, one
// or two); the H and F of its last column in each row reach lane g+1 one step
// later through a wave-wide DPP shift (wave_shr:1), so the pair is swept as a
// pipeline along the database with NR rows of skew per lane.  The
// first lane of a group reads the previous pass's right edge from the border
// array, the last active lane writes it for the next pass.  Lane 0 of the
// wave reads 0 from the shift (bound_ctrl), a value no lane uses: the first
// lane of every group takes the border instead, so no old value is set up.
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}

template <int WL, int WG>
constexpr int kGroupStepRows = group_step_rows(WL, WG);

template <int W, int NR>
__device__ __forceinline__ void group_item(const ScanArgs& A, const QueryDesc& Qd, int local,
                                           short* lds, uint2* slot, uint32_t (*s_red)[kWave],
                                           int* ovf) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const s2 Q = {(short)A.gap_oe, (short)A.gap_oe};
    const s2 R = {(short)A.gap_e, (short)A.gap_e};
    const s2 Z = as_s2(kZero16x2);
    const int G = Qd.group;
    const int gpw = kWave / G;
    const int gi = lane / G;
    const int gl = lane - gi * G;
    const bool in_group = gi < gpw;
    uint2* bord = slot + (size_t)(in_group ? gi : 0) * A.border_rows;
    const int p = Qd.blk_begin + (local * kWavesPerWG + wave) * gpw + gi;
    const bool has_pair = in_group && p < Qd.blk_end;
    BlockDesc bd = {0, 0, 0};
    if (has_pair) bd = A.blocks[p];
    uint32_t rmax = bd.rows;
    for (int off = 32; off > 0; off >>= 1) rmax = max(rmax, (uint32_t)__shfl_xor((int)rmax, off));
    rmax = __builtin_amdgcn_readfirstlane(rmax);   // equal in every lane: keep loops scalar
    const uint2* pw = reinterpret_cast<const uint2*>(A.words + bd.word_off);
    const int pass_cols = G * W;
    const int npass = (Qd.qpad + pass_cols - 1) / pass_cols;
    s2 S = Z;
    for (int pass = 0; pass < npass; ++pass) {
        const int col_base = pass * pass_cols;
        const int pcols = min(pass_cols, Qd.qpad - col_base);
        const int nl = pcols / W;
        __syncthreads();
        stage_profile(lds, Qd.prof, Qd.pstride, A.ps_group, col_base, pcols);
        __syncthreads();
        const bool has_in = pass > 0;
        const bool has_out = pass + 1 < npass;
        const bool on = has_pair && gl < nl;
        const int jj0 = on ? gl * W : 0;   // idle lanes load staged columns too
        s2 h[W], e[W];
#pragma unroll
        for (int c = 0; c < W; ++c) {
            h[c] = Z;
            e[c] = Z;
        }
        s2 diag0 = Z;
        // NR rows per step: at step t lane g works on rows NR(t-g) .. NR(t-g)+NR-1
        // (block rows are a multiple of 4, so all or none are in range); with
        // two, the second row's columns interleave with the first's in the
        // schedule, which the F chain of a single row leaves idle
        uint32_t ch[NR], cf[NR];   // this lane's last (H, F) of each row of its step
#pragma unroll
        for (int j = 0; j < NR; ++j) ch[j] = cf[j] = kZero16x2;
        // the last lane of a group writes its edge for the next pass; pend: the
        // step before ran rows whose edge is not written yet
        const bool edge_out = has_out && gl == nl - 1;
        bool pend = false;
        constexpr int U = 4 / NR;   // steps per 4 rows
        const int steps = rmax > 0 ? (int)(rmax / NR) + nl - 1 : 0;
        // residue words and border entries are loaded 4 rows ahead; rows
        // outside the pair read a clamped word (any code is a valid LDS row)
        const bool reads_border = has_in && gl == 0 && has_pair;
        const int rows = (int)bd.rows;
        const int last_row = max(rows - 1, 0);
        auto word_at = [&](int t) { return pw[min(max(NR * (t - gl), 0), last_row) >> 2]; };
        uint2 wa[U], wb[U], ba[U][NR], bb[U][NR];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            wa[u] = word_at(u);
            wb[u] = wa[u];
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                ba[u][j] = make_uint2(kZero16x2, kZero16x2);
                if (reads_border && NR * u + j < rows) ba[u][j] = bord[NR * u + j];
                bb[u][j] = ba[u][j];
            }
        }
        // a step's first profile row is loaded while the step before it
        // computes, its other row at the start of the step.  Steps run in
        // units of U starting at multiples of U (NR x U = 4 rows), so row
        // r0 + k of step t0 + u sits at byte (NR u + k - NR g) & 3 of its word:
        // four shifts per lane, fixed for the pass.
        static_assert(NR * U == 4, "a unit of steps covers one residue word");
        const char* col0 = reinterpret_cast<const char*>(lds + jj0);
        const uint32_t row_bytes = (uint32_t)A.ps_group * sizeof(short);
        uint32_t shv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) shv[k] = 8u * (uint32_t)((k - NR * gl) & 3);
        RawProfile<W> rp;
        load_profile_at<W>(rp, col0, row_bytes, (wa[0].x >> shv[0]) & 0xff,
                           (wa[0].y >> shv[0]) & 0xff);
        // U steps from (wc, bc) while the next U steps' words and borders load
        // into (wn, bn); unrolled twice with the two sets swapping roles
        auto unit = [&](int t0, const uint2 (&wc)[U], const uint2 (&bc)[U][NR], uint2 (&wn)[U],
                        uint2 (&bn)[U][NR]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int tn = t0 + U + u;
                wn[u] = word_at(tn);
#pragma unroll
                for (int j = 0; j < NR; ++j)
                    if (reads_border && NR * tn + j < rows) bn[u][j] = bord[NR * tn + j];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int r0 = NR * (t0 + u - gl);
                const RawProfile<W> cur = rp;
                std::conditional_t<NR == 2, RawProfile<W>, char> rq;   // the second row
                if constexpr (NR == 2) {
                    const uint32_t sh = shv[(NR * u + 1) & 3];
                    load_profile_at<W>(rq, col0, row_bytes, (wc[u].x >> sh) & 0xff,
                                       (wc[u].y >> sh) & 0xff);
                }
                const uint2 nw = u + 1 < U ? wc[u + 1] : wn[0];
                const uint32_t sh = shv[(NR * u + NR) & 3];
                load_profile_at<W>(rp, col0, row_bytes, (nw.x >> sh) & 0xff, (nw.y >> sh) & 0xff);
                __builtin_amdgcn_sched_barrier(0);   // issue the loads before the step's work
                // every lane takes part in the shifts (no lane may be masked off here)
                uint32_t hl[NR], fl[NR];
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const uint32_t hx = wave_shr1(ch[j]);
                    const uint32_t fx = wave_shr1(cf[j]);
                    hl[j] = gl == 0 ? bc[u][j].x : hx;
                    fl[j] = gl == 0 ? bc[u][j].y : fx;
                }
                // the last lane's edge of the step before goes to the border
                // array here, ahead of this step's rows, so that the rows, their
                // E updates and maxima stay one block for the scheduler
                if (pend) {
#pragma unroll
                    for (int j = 0; j < NR; ++j) bord[r0 - NR + j] = make_uint2(ch[j], cf[j]);
                }
                const bool run = on && (uint32_t)r0 < (uint32_t)rows;   // 0 <= r0 < rows
                if (run) {
                    auto row = [&](const RawProfile<W>& raw, int j) {
                        s2 P[W];
                        interleave_profile<W>(P, raw);
                        s2 fout;
                        row_cells<W, s2>(h, e, P, diag0, as_s2(hl[j]), as_s2(fl[j]), Q, R, fout);
                        S = row_max_into<W, s2>(S, h);
                        ch[j] = bits(h[W - 1]);
                        cf[j] = bits(fout);
                        diag0 = as_s2(hl[j]);
                    };
                    row(cur, 0);
                    if constexpr (NR == 2) row(rq, 1);
                }
                pend = run && edge_out;
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        int t_end = 0;   // steps run
        for (int t0 = 0; t0 < steps; t0 += 2 * U) {
            unit(t0, wa, ba, wb, bb);
            t_end = t0 + U;
            if (t0 + U >= steps) break;   // a uniform branch: no tail copy of the code
            unit(t0 + U, wb, bb, wa, ba);
            t_end = t0 + 2 * U;
        }
        if (pend) {   // the last step's edge
            const int r = NR * (t_end - 1 - gl);
#pragma unroll
            for (int j = 0; j < NR; ++j) bord[r + j] = make_uint2(ch[j], cf[j]);
        }
    }
    s_red[wave][lane] = bits(S);
    __syncthreads();
    if (has_pair && gl == 0) {
        s2 m = S;
        for (int k = 1; k < G; ++k) m = vmax(m, as_s2(s_red[wave][lane + k]));
        emit(Qd, A.perm, bd.first, (int)m.x + 32768, ovf);
        emit(Qd, A.perm, bd.first + 1, (int)m.y + 32768, ovf);
    }
}

// The search kernel: one persistent launch, one device work queue.  The host
// orders the items longest first across every query of the launch (group
// items of long pairs and lane items alike), so long work starts first and
// short blocks fill in around it.
__device__ __forceinline__ bool next_scan_item(const ScanArgs& A, int* s_buf, int& local,
                                               int& q, bool& is_group) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const int it = (int)atomicAdd(A.work_counter, 1u);
        uint2 v = make_uint2(0u, 0xffffffffu);
        if (it < A.n_items) v = A.items[it];
        s_buf[0] = (int)v.y;
        s_buf[1] = (int)(v.x & ~kLaneItem);
        s_buf[2] = (v.x & kLaneItem) ? 1 : 0;
    }
    __syncthreads();
    // one value for the whole workgroup: kept in scalar registers, so item
    // descriptors load through the scalar cache and loops over them are scalar
    local = __builtin_amdgcn_readfirstlane(s_buf[0]);
    q = __builtin_amdgcn_readfirstlane(s_buf[1]);
    is_group = __builtin_amdgcn_readfirstlane(s_buf[2]) == 0;
    return local >= 0;
}

__device__ __forceinline__ uint32_t sys_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint64_t sys_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Live queue (LiveQueue): thread 0 draws a ticket and waits, sleeping between
// polls, until the host has published it, closed the queue or the idle limit
// has passed.  The published count and the closed flag are one word, so one
// read sees both as the host last wrote them.
__device__ __forceinline__ bool next_live_item(const ScanArgs& A, int* s_buf, int& local, int& q,
                                               int& kind, uint32_t& ticket) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = atomicAdd(A.work_counter, 1u);
        uint2 v = make_uint2(0u, 0xffffffffu);
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        int nap = 1;   // s_sleep 127 (about 3.4 us) per unit
        for (;;) {
            const uint64_t c = sys_load(A.live.ctl);
            if (t < (uint32_t)c) {
                const uint64_t e = sys_load(reinterpret_cast<const uint64_t*>(
                    A.live.ring + (t & (A.live.cap - 1))));
                v = make_uint2((uint32_t)e, (uint32_t)(e >> 32));
                break;
            }
            if ((c >> 32) != 0) break;   // closed, and t was not published
            if (__builtin_amdgcn_s_memrealtime() - t0 > A.live.idle_ticks) {
                __hip_atomic_store(A.live.status, 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            for (int k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(127);
            nap = min(2 * nap, 32);
        }
        s_buf[0] = (int)v.y;
        s_buf[1] = (int)(v.x & ~(kLaneItem | kPrepItem));
        s_buf[2] = (v.x & kPrepItem) ? 2 : (v.x & kLaneItem) ? 1 : 0;
        s_buf[3] = 0;   // the item's overflow flag (emit sets it)
        s_buf[4] = (int)t;
    }
    __syncthreads();
    local = __builtin_amdgcn_readfirstlane(s_buf[0]);
    q = __builtin_amdgcn_readfirstlane(s_buf[1]);
    kind = __builtin_amdgcn_readfirstlane(s_buf[2]);
    ticket = (uint32_t)__builtin_amdgcn_readfirstlane(s_buf[4]);
    return local >= 0;
}

// Prep item of task q: its descriptor block and profile from the page-locked
// arena to the device arena, 16 bytes a thread per step (the profile's size,
// 24 x pstride int16 with pstride a multiple of 96, is a multiple of 16).
__device__ __forceinline__ void prep_item(const ScanArgs& A, int q) {
    const size_t doff = (size_t)q * kLiveDescStride * sizeof(QueryDesc);
    const QueryDesc* hd = reinterpret_cast<const QueryDesc*>(A.live.host_arena + doff);
    const size_t poff = (size_t)(reinterpret_cast<const char*>(hd->prof) - A.live.dev_arena);
    const size_t pbytes = (size_t)kProfRows * hd->pstride * sizeof(int16_t);
    const uint4* src = reinterpret_cast<const uint4*>(A.live.host_arena + poff);
    uint4* dst = reinterpret_cast<uint4*>(A.live.dev_arena + poff);
    const int n = (int)(pbytes / sizeof(uint4));
    constexpr int kUnroll = 4;
    for (int i = threadIdx.x; i < n; i += kUnroll * 256) {
        uint4 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            if (i + u * 256 < n) v[u] = src[i + u * 256];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            if (i + u * 256 < n) dst[i + u * 256] = v[u];
    }
    constexpr int kDescVecs = kLiveDescStride * sizeof(QueryDesc) / sizeof(uint4);
    if (threadIdx.x < kDescVecs)
        reinterpret_cast<uint4*>(A.live.dev_arena + doff)[threadIdx.x] =
            reinterpret_cast<const uint4*>(A.live.host_arena + doff)[threadIdx.x];
}

// End of a live item: every wave's scores and overflow entries are written
// back to memory (a system-scope release: the L2 is not coherent with the
// copy engine that reads them), then thread 0 reports the ticket done.
__device__ __forceinline__ void finish_live_item(const ScanArgs& A, const int* s_buf,
                                                 uint32_t ticket) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t v = ((ticket + 1u) << 1) | (s_buf[3] ? 1u : 0u);
        __hip_atomic_store(A.live.done + (ticket & (A.live.cap - 1)), v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// WL = 16: at most 168 VGPRs, so that three workgroups (12 waves) share a CU
// as its LDS allows; wider strips run two.  LIVE: items come from the live
// queue (A.live) instead of the launch's item list.
template <int WL, int WG, bool PAIR, bool LIVE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WL == 16 ? 3 : 2)))
void k_scan(ScanArgs A) {
    static_assert(!PAIR || WL == 16 || WL == 24, "pair table only for W = 16, 24");
    extern __shared__ __align__(16) short lds[];
    __shared__ int s_buf[5];
    __shared__ uint32_t s_red[kWavesPerWG][kWave];
    const int wave = threadIdx.x >> 6;
    uint2* slot = A.border + (size_t)(blockIdx.x * kWavesPerWG + wave) * A.border_stride;
    int loaded_q = -1, loaded_seg = -1;
    int local, q;
    bool is_group;
    int kind = 0;   // live: 0 group, 1 lane, 2 prep item
    uint32_t ticket = 0;
    while (LIVE ? next_live_item(A, s_buf, local, q, kind, ticket)
                : next_scan_item(A, s_buf, local, q, is_group)) {
        const size_t dq = (size_t)q * A.desc_stride;
        if constexpr (LIVE) is_group = kind == 0;
        if (LIVE && kind == 2) {
            prep_item(A, q);
        } else if (is_group) {
            const QueryDesc Qd = A.group_q[dq];
            group_item<WG, kGroupStepRows<WL, WG>>(A, Qd, local, lds, slot, s_red, s_buf + 3);
            loaded_q = -1;   // the group item restaged the profile
        } else {
            const QueryDesc Qd = A.lane_q[dq];
            lane_item<WL, PAIR>(A, Qd, q, local, lds, slot, loaded_q, loaded_seg, s_buf + 3);
        }
        if constexpr (LIVE) finish_live_item(A, s_buf, ticket);
    }
}

#define SWH_SCAN_CASES(X)                                                                  \
    X(16, 8, true) X(16, 16, true) X(24, 8, true) X(24, 16, true) X(16, 8, false)          \
    X(16, 16, false) X(24, 8, false) X(24, 16, false) X(32, 8, false) X(32, 16, false)     \
    X(24, 24, true) X(24, 24, false)

template <typename K>
static int occ(K kern, size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, 256, lds) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

template <int WL, int WG, bool PAIR, bool LIVE>
static hipError_t launch_scan_w(const ScanArgs& a, int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_scan<WL, WG, PAIR, LIVE>), dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

#ifdef SWH_LIVE_ONLY
// The live-queue instantiations of the scan (swh_kernels_live.hip compiles
// this file with SWH_LIVE_ONLY, so both halves build in parallel).
hipError_t launch_scan_live(int wl, int wg, bool pair, const ScanArgs& a, int grid, size_t lds,
                            hipStream_t s) {
#define X(L, G, P) \
    if (wl == L && wg == G && pair == P) return launch_scan_w<L, G, P, true>(a, grid, lds, s);
    SWH_SCAN_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

int scan_live_max_wg_per_cu(int wl, int wg, bool pair, size_t lds) {
#define X(L, G, P) \
    if (wl == L && wg == G && pair == P) return occ(k_scan<L, G, P, true>, lds);
    SWH_SCAN_CASES(X)
#undef X
    return 1;
}
#else

// ---------------------------------------------------------------------------
// int32 rescore of the overflowed sequences, in the pipeline of the group
// items: G lanes share one sequence (G = the query's rescore group, up to 64),
// lane g holds query columns [pass*G*8 + g*8, +8) and at step t works on row
// t-g; the (H, F) of its last column reaches lane g+1 through the wave-wide
// DPP shift.  32-bit saturating arithmetic biased by INT32_MIN.  A sequence
// of length n costs ceil(qpad / (8G)) passes of n + G - 1 steps on one wave,
// so a few long overflowed sequences no longer run their whole matrix on one
// lane each.
template <int W>
__device__ __forceinline__ void single_profile(int (&P)[W], const short* lds, int ps, int a,
                                               int jj0) {
    const short* pa = lds + __umul24(a, ps) + jj0;
#pragma unroll
    for (int c = 0; c < W; ++c) P[c] = pa[c];
}

// Rescore work items: kWavesPerWG * (64 / G) overflow-list entries of one
// query (one per lane group).  Thread 0 walks the (short) query table to map
// a claimed item to its query, so no host round trip sizes this launch.
__device__ __forceinline__ bool next_rescore_item(const RescoreArgs& A, int* s_buf, int& local,
                                                  int& q) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int it = (int)atomicAdd(A.work_counter, 1u);
        int found = -1;
        for (int j = 0; j < A.n_queries; ++j) {
            const uint32_t per_item = kWavesPerWG * (kWave / A.queries[j].group);
            const int n = (int)((*A.queries[j].ovf_count + per_item - 1) / per_item);
            if (it < n) {
                found = j;
                break;
            }
            it -= n;
        }
        s_buf[0] = found >= 0 ? it : -1;
        s_buf[1] = found;
    }
    __syncthreads();
    local = s_buf[0];
    q = s_buf[1];
    return local >= 0;
}

__global__ __launch_bounds__(256) void k_rescore(RescoreArgs A) {
    constexpr int W = kRescoreWidth;
    extern __shared__ __align__(16) short lds[];
    __shared__ int s_buf[2];
    __shared__ int s_red[kWavesPerWG][kWave];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    uint2* slot = A.border + (size_t)(blockIdx.x * kWavesPerWG + wave) * A.border_stride;
    int local, q;
    while (next_rescore_item(A, s_buf, local, q)) {
        const QueryDesc Qd = A.queries[q];
        const int n = (int)*Qd.ovf_count;
        const int G = Qd.group;
        const int gpw = kWave / G;
        const int gi = lane / G;
        const int gl = lane - gi * G;
        const bool in_group = gi < gpw;
        const int entry = (local * kWavesPerWG + wave) * gpw + gi;
        const bool valid = in_group && entry < n;
        uint2* bord = slot + (size_t)(in_group ? gi : 0) * A.border_rows;
        uint32_t s = 0, len = 0;
        uint64_t base = 0;
        uint32_t stride = 0;
        if (valid) {
            s = Qd.ovf_list[entry];
            len = A.sorted_len[s];
            const uint32_t blk = A.sorted_blk[s];
            const BlockDesc bd = A.blocks[blk];
            const uint32_t pos = s - bd.first;
            if ((int)blk >= A.n_lane_blocks_first) {
                base = bd.word_off + (uint64_t)(pos & 63) * 2 + (pos >> 6);
                stride = 2 * kWave;
            } else {
                base = bd.word_off + pos;
                stride = 2;
            }
        }
        // rows of the wave = its longest sequence (equal in every lane: scalar loops)
        uint32_t rmax = len;
        for (int off = 32; off > 0; off >>= 1) rmax = max(rmax, (uint32_t)__shfl_xor((int)rmax, off));
        rmax = __builtin_amdgcn_readfirstlane(rmax);
        const int rows = (int)len;
        const int pass_cols = G * W;
        const int npass = (Qd.qpad + pass_cols - 1) / pass_cols;
        int S = kZero32;
        for (int pass = 0; pass < npass; ++pass) {
            const int col_base = pass * pass_cols;
            const int pcols = min(pass_cols, Qd.qpad - col_base);
            const int nl = pcols / W;
            __syncthreads();
            stage_profile(lds, Qd.prof, Qd.pstride, A.ps, col_base, pcols);
            __syncthreads();
            const bool has_in = pass > 0;
            const bool has_out = pass + 1 < npass;
            const bool on = valid && gl < nl;
            const int jj0 = gl * W;
            int h[W], e[W];
#pragma unroll
            for (int c = 0; c < W; ++c) {
                h[c] = kZero32;
                e[c] = kZero32;
            }
            int diag0 = kZero32;
            int ch = kZero32, cf = kZero32;   // this lane's last (H, F)
            const int steps = rmax > 0 ? (int)rmax + nl - 1 : 0;
            // residue words and border entries are loaded 4 steps ahead
            const bool reads_border = has_in && gl == 0 && valid;
            uint32_t wn[4];
            uint2 bn[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int row = u - gl;
                wn[u] = 0u;
                bn[u] = make_uint2((uint32_t)kZero32, (uint32_t)kZero32);
                if (on && row >= 0 && row < rows) wn[u] = A.words[base + (uint64_t)(row >> 2) * stride];
                if (reads_border && u < rows) bn[u] = bord[u];
            }
            for (int t0 = 0; t0 < steps; t0 += 4) {
                uint32_t wc[4];
                uint2 bc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    wc[u] = wn[u];
                    bc[u] = bn[u];
                    const int tn = t0 + 4 + u;
                    const int row = tn - gl;
                    if (on && row >= 0 && row < rows)
                        wn[u] = A.words[base + (uint64_t)(row >> 2) * stride];
                    if (reads_border && tn < rows) bn[u] = bord[tn];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int row = t0 + u - gl;
                    // every lane takes part in the shift (no lane may be masked off here)
                    const int hx = (int)wave_shr1((uint32_t)ch);
                    const int fx = (int)wave_shr1((uint32_t)cf);
                    const int hl_u = gl == 0 ? (int)bc[u].x : hx;
                    const int fl_u = gl == 0 ? (int)bc[u].y : fx;
                    if (on && row >= 0 && row < rows) {
                        const int a = (wc[u] >> (8 * (row & 3))) & 0xff;
                        int P[W];
                        single_profile<W>(P, lds, A.ps, a, jj0);
                        int fout;
                        row_cells<W, int>(h, e, P, diag0, hl_u, fl_u, A.gap_oe, A.gap_e, fout);
                        S = row_max_into<W, int>(S, h);
                        ch = h[W - 1];
                        cf = fout;
                        diag0 = hl_u;
                        if (has_out && gl == nl - 1) bord[row] = make_uint2((uint32_t)ch, (uint32_t)cf);
                    }
                }
            }
        }
        s_red[wave][lane] = S;
        __syncthreads();
        if (valid && gl == 0) {
            int m = S;
            for (int k = 1; k < G; ++k) m = vmax(m, s_red[wave][lane + k]);
            const uint32_t idx = A.perm[s];
            if (idx != kNoSeq) Qd.scores[idx] = (int)((int64_t)m - (int64_t)kZero32);
        }
    }
}

// Database packing: block b's words from the residues of its sequences.  Lane
// blocks: word ((row/4)*64 + lane)*2 + half holds rows 4k..4k+3 of sorted
// sequence first + half*64 + lane; pair blocks: word (row/4)*2 + half.  Codes
// above 23 and rows beyond a sequence's end become 23.  Byte work: stores are
// coalesced, loads are 4-byte reads of 64 different sequences per wave.
__global__ __launch_bounds__(256) void k_pack(const BlockDesc* __restrict__ blocks, int n_pairs,
                                              const uint32_t* __restrict__ perm,
                                              const uint32_t* __restrict__ slen,
                                              const uint8_t* __restrict__ raw,
                                              const uint64_t* __restrict__ soff,
                                              uint32_t* __restrict__ words) {
    const int b = blockIdx.x;
    const BlockDesc bd = blocks[b];
    const bool pair = b < n_pairs;
    const uint32_t per4 = pair ? 2u : (uint32_t)kLaneSeqs;
    const uint64_t nw = (uint64_t)(bd.rows / 4) * per4;
    for (uint64_t w = threadIdx.x; w < nw; w += blockDim.x) {
        const uint32_t row4 = (uint32_t)(w / per4);
        const uint32_t r = (uint32_t)(w - (uint64_t)row4 * per4);
        const uint32_t s = pair ? bd.first + r : bd.first + (r & 1u) * kWave + (r >> 1);
        uint32_t v = 0x17171717u;
        const uint32_t id = perm[s];
        if (id != kNoSeq) {
            const uint32_t len = slen[s];
            const uint32_t p0 = row4 * 4;
            if (p0 < len) {
                const uint8_t* src = raw + soff[id] + p0;
                v = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    uint32_t c = p0 + k < len ? src[k] : 23u;
                    v |= min(c, 23u) << (8 * k);
                }
            }
        }
        words[bd.word_off + w] = v;
    }
}

__global__ void k_reset(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Query profile on the device: out[r * pstride + j] = matrix[q_j * 32 + r] for
// j < qlen, 0 beyond (the host's build_profile, for callers that stage the
// query in device memory).
__global__ void k_build_profile(const int32_t* __restrict__ matrix,
                                const uint8_t* __restrict__ query, int qlen, int pstride,
                                int16_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kProfRows * pstride) return;
    const int r = i / pstride;
    const int j = i - r * pstride;
    out[i] = j < qlen ? (int16_t)matrix[query[j] * 32 + r] : (int16_t)0;
}

// The scratch array of the allocation-free search (swh_search_noalloc).  In
// host code its name is only a stand-in for the device symbol: the memory is
// reached through the address hipGetSymbolAddress returns for each device.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

hipError_t scratch_address(int device, void** out, size_t* bytes) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;   // the caller makes it current
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *out = addr[device];
    *bytes = kScratchBytes;
    return hipSuccess;
}

// ---------------------------------------------------------------------------
hipError_t launch_reset(uint32_t* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_reset, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_build_profile(const int32_t* matrix, const uint8_t* query, int qlen,
                                int pstride, int16_t* out, hipStream_t s) {
    const int n = kProfRows * pstride;
    hipLaunchKernelGGL(k_build_profile, dim3((n + 255) / 256), dim3(256), 0, s, matrix, query,
                       qlen, pstride, out);
    return hipGetLastError();
}

hipError_t launch_pack(const BlockDesc* blocks, int n_blocks, int n_pairs, const uint32_t* perm,
                       const uint32_t* sorted_len, const uint8_t* raw, const uint64_t* soff,
                       uint32_t* words, hipStream_t s) {
    hipLaunchKernelGGL(k_pack, dim3(n_blocks), dim3(256), 0, s, blocks, n_pairs, perm,
                       sorted_len, raw, soff, words);
    return hipGetLastError();
}

hipError_t launch_scan(int wl, int wg, bool pair, const ScanArgs& a, int grid, size_t lds,
                       hipStream_t s) {
#define X(L, G, P) \
    if (wl == L && wg == G && pair == P) return launch_scan_w<L, G, P, false>(a, grid, lds, s);
    SWH_SCAN_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_rescore(const RescoreArgs& a, int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_rescore, dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

int scan_max_wg_per_cu(int wl, int wg, bool pair, size_t lds) {
#define X(L, G, P) \
    if (wl == L && wg == G && pair == P) return occ(k_scan<L, G, P, false>, lds);
    SWH_SCAN_CASES(X)
#undef X
    return 1;
}

int rescore_max_wg_per_cu(size_t lds) { return occ(k_rescore, lds); }

size_t lane_lds_bytes(int width, bool pair, int ps_lane) {
    if (pair && width == 16) return PairTab<16>::bytes;
    if (pair && width == 24) return PairTab<24>::bytes;
    return (size_t)kProfRows * ps_lane * sizeof(int16_t);
}
#endif  // SWH_LIVE_ONLY

}  // namespace swh
