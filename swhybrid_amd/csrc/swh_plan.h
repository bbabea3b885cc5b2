// swh_plan.h -- the serving loop's budgeted-round planning, as a pure
// function of the database's block table (no device state), so that its
// block ranges can be checked on the host.
//
// Blocks: pair blocks [0, n_pairs), then lane blocks [n_pairs, n_blocks);
// padded[b] is the prefix sum of padded residues over blocks [0, b)
// (n_blocks + 1 entries).  A request's cursors are pair_pos in
// [0, n_pairs] and lane_pos in [n_pairs, n_blocks]; a slice is
// [pair_pos, pe) of pair blocks and [lane_pos, le) of lane blocks.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace swhh {

struct PlanRequest {
    int qlen;
    int pair_pos, lane_pos;
    double waited_ms;   // since it was submitted
};

struct PlanSlice {
    int request;        // index into the requests
    int pair_begin, pair_end, lane_begin, lane_end;
    bool last;          // the slice reaches the end of both kinds
};

// Padded cells request r still has to scan.
inline double plan_remaining(const uint64_t* padded, int n_pairs, int n_blocks,
                             const PlanRequest& r) {
    return (double)r.qlen * (double)((padded[n_pairs] - padded[r.pair_pos]) +
                                     (padded[n_blocks] - padded[r.lane_pos]));
}

// First block b in [b0, b1] with padded(b0, b) >= want (b1 if none).
inline int plan_advance(const uint64_t* padded, int b0, int b1, double want) {
    const uint64_t target = padded[b0] + (uint64_t)std::max(0.0, std::ceil(want));
    const uint64_t* it = std::lower_bound(padded + b0, padded + b1 + 1, target);
    return it == padded + b1 + 1 ? b1 : (int)(it - padded);
}

// One budgeted round: about budget_cells padded cells, handed out shortest
// remaining work first -- the key is a request's remaining cells less
// wait_weight x the cells the device could have scanned (rate, per ms) while
// it has been waiting, so long requests are not starved while short ones
// keep arriving.
// A request given g of its remaining W cells takes g / W of its remaining
// pair work and of its remaining lane work.  Each request of the round gets
// at least min_round_lane_blocks / n lane blocks (n: requests in the round),
// so that a round has enough lane items to balance its tail.  Requests with
// nothing left get no slice.  Every slice advances at least one block.
inline std::vector<PlanSlice> plan_budgeted(const uint64_t* padded, int n_pairs, int n_blocks,
                                            const std::vector<PlanRequest>& reqs,
                                            double budget_cells, double rate,
                                            int min_round_lane_blocks, double wait_weight) {
    struct Cand {
        double key, rem;
        int i;
    };
    std::vector<Cand> cand;
    for (int i = 0; i < (int)reqs.size(); ++i) {
        const double rem = plan_remaining(padded, n_pairs, n_blocks, reqs[i]);
        if (reqs[i].pair_pos < n_pairs || reqs[i].lane_pos < n_blocks)
            cand.push_back(Cand{rem - wait_weight * reqs[i].waited_ms * rate, rem, i});
    }
    std::stable_sort(cand.begin(), cand.end(),
                     [](const Cand& a, const Cand& b) { return a.key < b.key; });
    double left = std::max(1.0, budget_cells);
    std::vector<std::pair<const Cand*, double>> given;
    for (const Cand& k : cand) {
        if (left <= 0) break;
        const double g = std::min(k.rem, left);
        given.emplace_back(&k, g);
        left -= g;
    }
    const int min_lane =
        std::max(1, min_round_lane_blocks / std::max(1, (int)given.size()));
    std::vector<PlanSlice> out;
    for (const auto& gv : given) {
        const PlanRequest& r = reqs[gv.first->i];
        const double f = gv.first->rem > 0 ? gv.second / gv.first->rem : 1.0;
        int pe = n_pairs, le = n_blocks;
        if (f < 0.999) {
            pe = plan_advance(padded, r.pair_pos, n_pairs,
                              f * (double)(padded[n_pairs] - padded[r.pair_pos]));
            le = plan_advance(padded, r.lane_pos, n_blocks,
                              f * (double)(padded[n_blocks] - padded[r.lane_pos]));
            le = std::min(n_blocks, std::max(le, r.lane_pos + min_lane));
            if (pe == r.pair_pos && le == r.lane_pos) {   // progress: one block
                if (pe < n_pairs) ++pe;
                else ++le;
            }
        }
        out.push_back(PlanSlice{gv.first->i, r.pair_pos, pe, r.lane_pos, le,
                                pe == n_pairs && le == n_blocks});
    }
    return out;
}

}  // namespace swhh
