// serving_plan_check.cpp -- host-only check of the serving loop's budgeted
// round planning (swhybrid_amd/csrc/swh_plan.h).  Random block tables and
// arrival patterns are planned round after round until every request is
// done; every slice must stay inside its block kind, start at the request's
// cursor, advance it, and the slices of a request must cover each of its
// blocks exactly once.  Prints "ok <rounds> <slices>" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "swh_plan.h"

using namespace swhh;

#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::printf("FAIL %s: ", #cond);              \
            std::printf(__VA_ARGS__);                     \
            std::printf("\n");                            \
            std::exit(1);                                 \
        }                                                 \
    } while (0)

int main(int argc, char** argv) {
    const int trials = argc > 1 ? std::atoi(argv[1]) : 2000;
    std::mt19937_64 rng(12345);
    long total_rounds = 0, total_slices = 0;
    for (int trial = 0; trial < trials; ++trial) {
        const int n_pairs = (int)(rng() % 50);
        const int n_blocks = n_pairs + (int)(rng() % 3000);
        std::vector<uint64_t> padded(n_blocks + 1, 0);
        for (int b = 0; b < n_blocks; ++b) {   // longest first within each kind
            const uint64_t rows = b < n_pairs ? 40000 - 700 * (uint64_t)b
                                              : 4 + (uint64_t)(rng() % 2000) / (1 + b / 200);
            padded[b + 1] = padded[b] + rows * (b < n_pairs ? 2 : 128);
        }
        const int nreq = 1 + (int)(rng() % 12);
        std::vector<PlanRequest> reqs(nreq);
        std::vector<int> arrive(nreq);
        std::vector<std::vector<int>> covered(nreq, std::vector<int>(n_blocks, 0));
        std::vector<bool> done(nreq, false);
        for (int i = 0; i < nreq; ++i) {
            reqs[i] = PlanRequest{1 + (int)(rng() % 6000), 0, n_pairs, 0.0};
            arrive[i] = (int)(rng() % 40);   // round in which it is admitted
            if (n_blocks == 0) done[i] = true;
        }
        const double rate = 1e6 * (1 + rng() % 10000);
        const double budget = rate * (0.5 + (double)(rng() % 100));
        const int min_lane = 16 * (1 + (int)(rng() % 300));
        const double wait_weight = (double)(rng() % 5) * 0.5;
        int round = 0;
        for (;; ++round) {
            CHECK(round < 200000, "trial %d does not terminate", trial);
            std::vector<PlanRequest> active;
            std::vector<int> who;
            bool pending = false;
            for (int i = 0; i < nreq; ++i) {
                if (done[i]) continue;
                pending = true;
                if (arrive[i] > round) continue;
                PlanRequest r = reqs[i];
                r.waited_ms = (double)(round - arrive[i]) * 3.0;
                active.push_back(r);
                who.push_back(i);
            }
            if (!pending) break;
            if (active.empty()) continue;
            const std::vector<PlanSlice> plan =
                plan_budgeted(padded.data(), n_pairs, n_blocks, active, budget, rate, min_lane,
                              wait_weight);
            CHECK(!plan.empty(), "trial %d round %d: no slice for %zu requests", trial, round,
                  active.size());
            for (const PlanSlice& p : plan) {
                CHECK(p.request >= 0 && p.request < (int)active.size(), "bad index");
                const int i = who[p.request];
                PlanRequest& r = reqs[i];
                CHECK(p.pair_begin == r.pair_pos && p.lane_begin == r.lane_pos,
                      "trial %d: slice does not start at the cursor", trial);
                CHECK(0 <= p.pair_begin && p.pair_begin <= p.pair_end && p.pair_end <= n_pairs,
                      "trial %d: pair range [%d, %d) outside [0, %d)", trial, p.pair_begin,
                      p.pair_end, n_pairs);
                CHECK(n_pairs <= p.lane_begin && p.lane_begin <= p.lane_end &&
                          p.lane_end <= n_blocks,
                      "trial %d: lane range [%d, %d) outside [%d, %d)", trial, p.lane_begin,
                      p.lane_end, n_pairs, n_blocks);
                CHECK(p.pair_end > p.pair_begin || p.lane_end > p.lane_begin,
                      "trial %d: a slice without progress", trial);
                CHECK(p.last == (p.pair_end == n_pairs && p.lane_end == n_blocks),
                      "trial %d: wrong last flag", trial);
                for (int b = p.pair_begin; b < p.pair_end; ++b) covered[i][b] += 1;
                for (int b = p.lane_begin; b < p.lane_end; ++b) covered[i][b] += 1;
                r.pair_pos = p.pair_end;
                r.lane_pos = p.lane_end;
                if (p.last) done[i] = true;
                ++total_slices;
            }
        }
        total_rounds += round;
        for (int i = 0; i < nreq; ++i)
            for (int b = 0; b < n_blocks; ++b)
                CHECK(covered[i][b] == 1, "trial %d request %d: block %d covered %d times", trial,
                      i, b, covered[i][b]);
    }
    std::printf("ok %ld %ld\n", total_rounds, total_slices);
    return 0;
}
