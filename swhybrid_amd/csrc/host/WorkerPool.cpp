#include "WorkerPool.h"

#include <algorithm>
#include <thread>

WorkerPool::WorkerPool(DataPool* loader, Param* params, size_t chunkBytes, size_t cpuChunkBytes,
                       size_t firstChunkBytes)
    : loader_(loader), params_(params), resultList_(params->topNum) {
    recalcWorker_ = new RecalcWorker(params, 0);
    recalcWorker_->setQuery(loader->getQuery(), loader->getQueryLen());
    for (int i = 0; i < params->hipNum; ++i) {
        workerList_.push_back(new PoolHIPWorker(loader, params, recalcWorker_, i, chunkBytes,
                                                firstChunkBytes));
        isCpu_.push_back(false);
    }
    if (params->cpuNum > 0) {
        workerList_.push_back(new CPUWorker(loader, params, recalcWorker_, params->cpuNum,
                                            cpuChunkBytes ? cpuChunkBytes : (size_t(4) << 20)));
        isCpu_.push_back(true);
        // the last device chunk's worth of work is left to the device workers
        if (params->hipNum > 0) loader->setCpuReserve(chunkBytes * (size_t)params->hipNum);
        // chunk copies leave the CPU worker's threads their cores (16-thread
        // fills beside a 4-thread CPU worker made the whole call 3% slower
        // than the device worker alone at query length 1000)
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        loader->setFillThreads(std::max(2u, hw - std::min(hw, (unsigned)params->cpuNum)));
    }
}

WorkerPool::~WorkerPool() {
    for (Worker* w : workerList_) delete w;
    delete recalcWorker_;
}

int WorkerPool::run(int* scoreSink) {
    if (workerList_.empty()) {
        error_ = "no computing device";
        return -1;
    }
    for (Worker* w : workerList_)
        if (scoreSink) w->setScoreSink(scoreSink);
    for (Worker* w : workerList_) w->startThread();
    for (Worker* w : workerList_) w->waitForThread();
    for (Worker* w : workerList_)
        if (!w->lastError().empty()) {
            error_ = w->lastError();
            return -1;
        }
    try {
        recalcWorker_->launch();
    } catch (const char* msg) {
        error_ = msg;
        return -1;
    }
    double start = 1e300, stop = 0;
    for (Worker* w : workerList_) {
        if (w->getStartTime() > 0) start = std::min(start, w->getStartTime());
        stop = std::max(stop, w->getStopTime());
    }
    if (recalcWorker_->getNum() > 0) stop = std::max(stop, recalcWorker_->getStopTime());
    seconds_ = stop > start ? stop - start : 0;
    gcups_ = seconds_ > 0
                 ? (double)loader_->getTotalResidue() * loader_->getQueryLen() / seconds_ / 1e9
                 : 0;
    for (Worker* w : workerList_) resultList_.mergeList(w->getResult());
    resultList_.mergeList(recalcWorker_->getResult());
    if (scoreSink)
        for (size_t i = 0; i < recalcWorker_->indices().size(); ++i)
            scoreSink[recalcWorker_->indices()[i]] = recalcWorker_->scores()[i];
    return 0;
}

uint64_t WorkerPool::residues(bool cpu) const {
    uint64_t r = 0;
    for (size_t i = 0; i < workerList_.size(); ++i)
        if (isCpu_[i] == cpu) r += workerList_[i]->getResidueCount();
    return r;
}

double WorkerPool::busySeconds(bool cpu) const {
    double s = 0;
    for (size_t i = 0; i < workerList_.size(); ++i)
        if (isCpu_[i] == cpu) s += workerList_[i]->busySeconds();
    return s;
}

Worker::Phases WorkerPool::devicePhases() const {
    Worker::Phases p;
    for (size_t i = 0; i < workerList_.size(); ++i)
        if (!isCpu_[i]) {
            const Worker::Phases& w = workerList_[i]->phases();
            p.first_fill += w.first_fill;
            p.aux_fill += w.aux_fill;
            p.search += w.search;
            p.aux_wait += w.aux_wait;
            p.chunks += w.chunks;
        }
    return p;
}

double WorkerPool::devicePackSeconds() const {
    double s = 0;
    for (size_t i = 0; i < workerList_.size(); ++i)
        if (!isCpu_[i]) s += static_cast<const HIPWorker*>(workerList_[i])->packSeconds();
    return s;
}

double WorkerPool::kernelSeconds() const {
    double s = 0;
    for (size_t i = 0; i < workerList_.size(); ++i)
        if (!isCpu_[i]) s += workerList_[i]->busySeconds();
    return s;
}
