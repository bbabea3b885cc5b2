// ResultList.h -- fixed-size top-k list.
//
// Semantics of the reference's ResultList<int> (src/ResultList.h:18-92): the
// list always holds k entries, starting as k (0, 0) results; an accepted push
// enters before the first entry that ranks below it and the last entry drops
// out.  Entries with equal scores are ordered by index, ascending, and a push
// is rejected unless it ranks above the last entry.  For pushes in increasing
// index order that is the reference's rule (reject a score <= the bottom
// score, first-pushed order among ties); it also keeps merges of several
// workers' lists, whose indices interleave, independent of thread timing (a
// tie with the bottom entry and a smaller index enters).
#pragma once
#include <cstddef>
#include <cstdio>
#include <vector>

template <class T>
struct Result {
    Result() : score(0), idx(0) {}
    Result(T s, size_t i) : score(s), idx(i) {}
    T score;
    size_t idx;
};

template <class T>
class ResultList {
  public:
    explicit ResultList(size_t k) : size_(k), bottom_(0), results_(k) {}

    const std::vector<Result<T>>& getResult() const { return results_; }
    size_t size() const { return size_; }

    void push(const T* scores, size_t n, size_t baseIdx) {
        for (size_t i = 0; i < n; ++i)
            if (scores[i] >= bottom_) push(scores[i], baseIdx + i);
    }

    void mergeList(const ResultList<T>& other) {
        for (const auto& r : other.results_) push(r.score, r.idx);
    }

    void push(T score, size_t idx) {
        if (size_ == 0 || score <= 0 || !ranks_below(results_[size_ - 1], score, idx)) return;
        size_t pos = 0;
        while (pos < size_ && !ranks_below(results_[pos], score, idx)) ++pos;
        if (pos == size_) return;
        for (size_t t = size_ - 1; t > pos; --t) results_[t] = results_[t - 1];
        results_[pos] = Result<T>(score, idx);
        bottom_ = results_[size_ - 1].score;
    }

    void print() const {
        for (const auto& r : results_) std::printf("score %d id %zu\n", (int)r.score, r.idx);
    }

  private:
    // an entry ranks below (score, idx) when its score is lower, or equal with a
    // larger index; filler (0, 0) entries rank below every positive score
    static bool ranks_below(const Result<T>& e, T score, size_t idx) {
        return e.score < score || (e.score == score && e.idx > idx);
    }
    size_t size_;
    T bottom_;
    std::vector<Result<T>> results_;
};
