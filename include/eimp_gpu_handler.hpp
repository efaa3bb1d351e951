// eimp_gpu_handler.hpp -- C++ drop-in for the reference's GPUHandler
// (src/GPUHandler.h:47-90, src/GPUHandler.cpp) over the C ABI in eimp.h.
//
// The reference's producer runs Target::prepPhase (findBestHaps, findMonos,
// findCallLocs) on the host under OpenMP and hands only the condensed
// reference + PBWT to the GPU (src/GPUHandler.cpp:119-136).  This handler's
// producer runs all of it on the device: one eimp_engine_run_batch per block
// of targets (top-K, monos, split/call sites, condensed reference, forward and
// backward compressed PBWT), then copies each target's cPBWT rows, count0 and
// prepPhase results into the queue item.  The consumer phases each target
// from its cPBWT with the beam search of eimp.h (eimp_phase_target), as
// processCondensedRef does with PBWT(cpbwt, count0, K, 4*Kwords, Mp) +
// Target::phaseExtPBWT (src/GPUHandler.cpp:259-266).
//
// Kept from the reference: the two member functions with their queue /
// termination / thread-index arguments, initialize(K, Nrefhapscorr,
// NrefhapsAD) returning the block size, one producer thread per GPU engine,
// the queue item types.  The cPBWT queue item carries count0, which the
// reference's handler writes and reads (src/GPUHandler.cpp:214,259) but its
// PBWTPhaser::cPBWTQueue_type lacks (src/PBWTPhaser.h:47-52).  TBB is not
// needed: BoundedQueue provides the members of tbb::concurrent_bounded_queue
// that the reference calls (push, pop, abort, set_capacity).  VCFData and
// Target are replaced by the plain structs below, holding what the handler
// reads of them.  Errors throw std::runtime_error.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "eimp.h"
#include "eimp_gpu_engine.hpp"

namespace eimp_cpp {

// Thrown by BoundedQueue::push / pop after abort() (tbb::user_abort).
struct QueueAborted : std::exception {
    const char *what() const noexcept override { return "queue aborted"; }
};

// Blocking bounded MPMC queue with the members of tbb::concurrent_bounded_queue
// that PBWTPhaser and GPUHandler use.
template <class T>
class BoundedQueue {
public:
    void set_capacity(size_t c) {
        std::lock_guard<std::mutex> lk(mu_);
        cap_ = c ? c : SIZE_MAX;
    }
    void push(T v) {
        std::unique_lock<std::mutex> lk(mu_);
        not_full_.wait(lk, [&] { return aborted_ || q_.size() < cap_; });
        if (aborted_) throw QueueAborted();
        q_.push_back(std::move(v));
        not_empty_.notify_one();
    }
    void pop(T &v) {
        std::unique_lock<std::mutex> lk(mu_);
        not_empty_.wait(lk, [&] { return aborted_ || !q_.empty(); });
        if (aborted_) throw QueueAborted();
        v = std::move(q_.front());
        q_.pop_front();
        not_full_.notify_one();
    }
    void abort() {
        std::lock_guard<std::mutex> lk(mu_);
        aborted_ = true;
        not_empty_.notify_all();
        not_full_.notify_all();
    }
    size_t size() const {
        std::lock_guard<std::mutex> lk(mu_);
        return q_.size();
    }

private:
    mutable std::mutex mu_;
    std::condition_variable not_empty_, not_full_;
    std::deque<T> q_;
    size_t cap_ = SIZE_MAX;
    bool aborted_ = false;
};

// What the handler reads of VCFData: the site-major reference rows it uploads
// (getReferenceT_const()[0].getData() / getCapacity(), src/GPUHandler.cpp:69),
// the number of sites (getNSNPs, :41) and, for the device prepPhase, the
// genetic positions (src/VCFData.cpp:1189), optionally the hap-major rows
// (getReference, src/VCFData.h:203; the engine transposes the site-major
// rows without them) and the haploid flags of the reference samples
// (getHaploidsRef, src/Target.cpp:804-806,835).
struct PanelData {
    size_t M = 0;
    size_t refTCapacity = 0;             // bytes per site-major row
    const uint64_t *refTData = nullptr;  // M rows, contiguous
    const float *cM = nullptr;           // M genetic positions
    const uint64_t *hapData = nullptr;   // optional: NrefhapsAD rows of hapStride words
    size_t hapStride = 0;
    const uint8_t *haploidRef = nullptr; // optional: NrefhapsAD/2 flags
    float cMmaxSplit = 0.125f;           // src/Args.cpp:290
};

// One target: what Target::prepPhase reads (the genotypes of its window and
// the sample to skip, src/Target.cpp:95-163, 791-795) and what prepPhase and
// phaseExtPBWT leave in it.
struct Target {
    size_t idx = 0;                  // Target::getIdx
    const uint64_t *gt = nullptr;    // GenotypeVector words of the window
    uint64_t s0 = 0, W = 0;          // window [s0, s0 + W) of the panel's sites
    int64_t ignoreSample = -1;       // reference sample skipped in top-K, -1 none
    int32_t startPhase = 0;          // VCFData::getStartPhase
    const float *af = nullptr;       // allele frequencies (imputing missing genotypes)
    // prepPhase results (device)
    uint64_t K = 0, S = 0, ncalls = 0;
    std::vector<uint32_t> besthaps;      // getBestHaps
    std::vector<uint32_t> splitsites;    // getSplitSites (window-relative)
    std::vector<float> cMsplitDiffs;     // S + 1
    // phaseExtPBWT results (consumer)
    std::vector<uint64_t> mat, pat;      // phased haplotypes, ceil(W/64) words each
    std::vector<float> dosageMat, dosagePat, callConf;
    float totalconf = 0.f;               // getTotalConfidence
    uint64_t ncallsites = 0;             // getNCallSites
    int32_t nmissing = 0, nmonhets = 0;
};

// PBWTPhaser's queue types (src/PBWTPhaser.h:43-57); cPBWT items carry count0.
using targetQueue_type = std::shared_ptr<std::vector<Target *>>;
struct cPBWTQueue_type {
    Target *target = nullptr;
    uint32_t *cpbwt = nullptr;           // 2*Mp rows of 4*Kwords uint32 (malloc'd; the consumer frees)
    std::vector<int> *count0 = nullptr;  // Mp zero counts (new'd; the consumer deletes)
};
struct confidence_type {
    size_t id;
    float totalconf;
    size_t ncalls;
};

class GPUHandler {
public:
    // GPUHandler(Hybridsys&, const VCFData&) (src/GPUHandler.cpp:34-54): one
    // engine per listed device; nTargets as VCFData::getNTarget.
    GPUHandler(const std::vector<int> &devices, const PanelData &panel, size_t nTargets)
        : panel_(panel), totalTargets_(nTargets) {
        if (devices.empty()) throw std::runtime_error("GPUHandler: no devices");
        if (!panel.refTData || !panel.cM || !panel.M)
            throw std::runtime_error("GPUHandler: panel rows and positions are required");
        for (int d : devices) engines_.emplace_back(d);
        eimp_phase_default_params(&params_);
    }

    // src/GPUHandler.cpp:57-75: initializes every engine, returns the block size
    size_t initialize(size_t K, size_t Nrefhapscorr, size_t NrefhapsAD) {
        K_ = K;
        Kwords_ = (K + 511) / 512 * 8;
        Nrefhapscorr_ = Nrefhapscorr;
        NrefhapsAD_ = NrefhapsAD;
        for (auto &eng : engines_) {
            eng.initialize(K_, Kwords_, panel_.M, Nrefhapscorr_, NrefhapsAD_, panel_.refTCapacity,
                           panel_.refTData);
            eng.setPanelExtras(panel_.cM, panel_.hapData, panel_.hapStride, panel_.haploidRef);
        }
        blockSize_ = engines_[0].getBlockSize();
        totalBlocks_ = (totalTargets_ + blockSize_ - 1) / blockSize_;
        blocksRemaining_ = totalBlocks_;
        targetsRemaining_ = totalTargets_;
        return blockSize_;
    }

    // Phasing parameters of the consumer (eimp_phase_default_params gives the
    // reference's defaults).
    eimp_phase_params &phaseParams() { return params_; }

    // src/GPUHandler.cpp:78-186, with prepPhase on the device.
    void preProcessGPUOnly(BoundedQueue<targetQueue_type> &inqueue,
                           BoundedQueue<cPBWTQueue_type> &outqueue,
                           std::atomic_bool &termination_request, int threadIndex) {
        GPUEngine &eng = engines_.at((size_t)threadIndex);
        while (!termination_request) {
            if (blocksRemaining_.fetch_sub(1) <= 0) break;
            targetQueue_type block;
            try {
                inqueue.pop(block);
            } catch (QueueAborted &) {
                break;
            }
            const size_t n = block->size();
            std::vector<eimp_target> in(n);
            for (size_t t = 0; t < n; t++) {
                const Target *tg = block->at(t);
                in[t] = eimp_target{tg->gt, tg->s0, tg->W, K_, tg->ignoreSample};
            }
            // the reference's layout: rows of 4*Kwords uint32 (src/GPUHandler.cpp:127,259)
            const std::vector<eimp_target_result> res = eng.runBatch(in, panel_.cMmaxSplit, 0u);
            for (size_t t = 0; t < n; t++) {
                Target *tg = block->at(t);
                const eimp_target_result &r = res[t];
                tg->K = r.K;
                tg->S = r.S;
                tg->ncalls = r.ncalls;
                tg->besthaps.assign(r.K, 0);
                tg->splitsites.assign(r.S, 0);
                tg->cMsplitDiffs.assign(r.S ? r.S + 1 : 0, 0.f);
                const size_t rs = 4 * ((r.K + 511) / 512 * 8);
                cPBWTQueue_type item;
                item.target = tg;
                item.cpbwt = static_cast<uint32_t *>(std::malloc(2 * r.Mp * rs * sizeof(uint32_t)));
                item.count0 = new std::vector<int>(r.Mp);
                if (!item.cpbwt) {
                    delete item.count0;
                    throw std::runtime_error("GPUHandler: host cPBWT allocation failed");
                }
                eng.copyTarget(t, tg->besthaps.data(), nullptr,
                               tg->S ? tg->splitsites.data() : nullptr, nullptr,
                               tg->S ? tg->cMsplitDiffs.data() : nullptr,
                               reinterpret_cast<int32_t *>(item.count0->data()), item.cpbwt);
                outqueue.push(std::move(item));
            }
        }
    }

    // src/GPUHandler.cpp:189-275: phases each target from its cPBWT and
    // reports (id, total confidence, call sites); frees the item's buffers,
    // as the reference's PBWT destructor does (src/PBWT.h:49-52).
    void processCondensedRef(BoundedQueue<cPBWTQueue_type> &inqueue,
                             BoundedQueue<confidence_type> &outqueue,
                             std::atomic_bool &termination_request, int /*threadIndex*/) {
        while (!termination_request) {
            if (targetsRemaining_.fetch_sub(1) <= 0) break;
            cPBWTQueue_type item;
            try {
                inqueue.pop(item);
            } catch (QueueAborted &) {
                break;
            }
            Target *tg = item.target;
            const uint32_t Mp = (uint32_t)item.count0->size();
            const uint32_t rs = (uint32_t)(4 * ((tg->K + 511) / 512 * 8));
            const size_t words = (tg->W + 63) / 64;
            tg->mat.assign(words, 0);
            tg->pat.assign(words, 0);
            tg->dosageMat.assign(tg->W, 0.f);
            tg->dosagePat.assign(tg->W, 0.f);
            tg->callConf.assign(tg->ncalls ? tg->ncalls : 1, 0.f);
            eimp_phase_input pin{item.cpbwt, reinterpret_cast<const int32_t *>(item.count0->data()),
                                 (uint32_t)tg->K, rs, Mp, (uint32_t)tg->W, tg->gt,
                                 tg->splitsites.data(), tg->cMsplitDiffs.data(), tg->af,
                                 tg->startPhase, 0};
            eimp_phase_output pout{};
            pout.mat = tg->mat.data();
            pout.pat = tg->pat.data();
            pout.dosage_mat = tg->dosageMat.data();
            pout.dosage_pat = tg->dosagePat.data();
            pout.call_conf = tg->callConf.data();
            const int rc = eimp_phase_target(&params_, &pin, &pout);
            std::free(item.cpbwt);
            delete item.count0;
            check(rc, "GPUHandler::processCondensedRef");
            tg->totalconf = pout.total_conf;
            tg->ncallsites = pout.ncallsites;
            tg->nmissing = pout.nmissing;
            tg->nmonhets = pout.nmonhets;
            outqueue.push(confidence_type{tg->idx, tg->totalconf, (size_t)tg->ncallsites});
        }
    }

    size_t getBlockSize() const { return blockSize_; }
    GPUEngine &engine(size_t i) { return engines_.at(i); }

private:
    PanelData panel_;
    std::vector<GPUEngine> engines_;
    eimp_phase_params params_{};
    size_t K_ = 0, Kwords_ = 0, Nrefhapscorr_ = 0, NrefhapsAD_ = 0;
    size_t totalTargets_ = 0, blockSize_ = 0, totalBlocks_ = 0;
    std::atomic<long> blocksRemaining_{0}, targetsRemaining_{0};
};

}  // namespace eimp_cpp
