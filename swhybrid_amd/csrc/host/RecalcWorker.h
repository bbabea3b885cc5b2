// RecalcWorker.h -- collects sequences a worker could not score exactly and
// rescores them after the workers finish.
//
// Interface of the reference's RecalcWorker (src/RecalcWorker.h:27-100):
// pushOverflowSeq(seq, len, idx) under a mutex, launch() after every worker
// has joined, getResult() merged last.  The reference rescored on the CPU in
// int16 with gaps fixed at 12/2 (src/AVXKernel.cpp:914-915); this one runs the
// exact int32-capable GPU search of libswh on the collected sequences with the
// search's own scoring.  HIPWorker never pushes here (it rescores overflowed
// lanes on the device); the CPU worker pushes what overflows its int16 pass.
#pragma once
#include <cstdint>
#include <mutex>
#include <vector>

#include "Param.h"
#include "ResultList.h"

class RecalcWorker {
  public:
    RecalcWorker(Param* params, int device);
    ~RecalcWorker();
    void setQuery(const char* query, size_t qlen);
    void pushOverflowSeq(const char* seq, size_t len, size_t idx);
    void launch();
    size_t getNum() const { return idx_.size(); }
    // after launch(): pool index and exact score of every rescored sequence
    const std::vector<size_t>& indices() const { return idx_; }
    const std::vector<int32_t>& scores() const { return scores_; }
    ResultList<int>& getResult() { return result_; }
    double getStartTime() const { return t0_; }
    double getStopTime() const { return t1_; }

  private:
    Param* params_;
    int device_;
    std::vector<uint8_t> query_;
    std::vector<uint8_t> residues_;
    std::vector<uint64_t> offsets_;
    std::vector<uint32_t> lengths_;
    std::vector<size_t> idx_;
    std::vector<int32_t> scores_;
    ResultList<int> result_;
    std::mutex mu_;
    double t0_ = 0, t1_ = 0;
};
