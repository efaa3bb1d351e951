// eimp_gpu_engine.hpp -- C++ drop-in for the reference's GPUEngine
// (src/GPUEngine.h:41-98) over the C ABI in eimp.h.
//
// Same constructor and member signatures as the reference class, so
// GPUHandler (src/GPUHandler.cpp) compiles against it unchanged:
//   GPUEngine(int index);
//   void initialize(K, Kwords, M, Nrefhapscorr, NrefhapsAD, refTCapacity, refTData);
//   size_t getBlockSize();
//   void runCrefPBWTKernel(tgtData, splitdata, besthapdata, crefTOffsets, cPBWT, count0s);
//   void free();
//   int getCUDAIndex() const;
// Errors throw std::runtime_error (the reference throws thrust::system_error,
// src/GPUKernels.h:29-38).  Beyond the reference: prepPhase on the device
// (setPanelExtras + runBatch) and copies of per-target outputs.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "eimp.h"

namespace eimp_cpp {

using data_type = uint64_t;

inline void check(int rc, const char *what) {
    if (rc != EIMP_OK)
        throw std::runtime_error(std::string(what) + ": " + eimp_last_error() + " (status " +
                                 std::to_string(rc) + ")");
}

class GPUEngine {
public:
    GPUEngine() = delete;
    explicit GPUEngine(int index) : index(index) {
        eimp_engine *e = nullptr;
        check(eimp_engine_create(index, &e), "GPUEngine");
        h.reset(e);
    }
    GPUEngine(GPUEngine &&) = default;
    GPUEngine &operator=(GPUEngine &&) = default;
    ~GPUEngine() = default;

    // src/GPUEngine.h:50
    void initialize(size_t K_, size_t Kwords_, size_t M_, size_t Nrefhapscorr_, size_t NrefhapsAD_,
                    size_t refTCapacity, const data_type *refTData) {
        check(eimp_engine_initialize(h.get(), K_, Kwords_, M_, Nrefhapscorr_, NrefhapsAD_,
                                     refTCapacity, refTData),
              "GPUEngine::initialize");
        K = K_;
        Kwords = Kwords_;
        check(eimp_engine_block_size(h.get(), &blocksize), "GPUEngine::getBlockSize");
    }

    // src/GPUEngine.h:53
    size_t getBlockSize() { return blocksize; }

    // src/GPUEngine.h:55-61 -- count0s[t] is resized to Mp as the reference's
    // caller allocates it (src/GPUHandler.cpp:135)
    void runCrefPBWTKernel(const std::vector<const data_type *> &tgtData,
                           const std::vector<const data_type *> &splitdata,
                           const std::vector<const data_type *> &besthapdata,
                           const std::vector<uint32_t> &crefTOffsets, std::vector<uint32_t *> &cPBWT,
                           std::vector<std::vector<int> *> &count0s) {
        const size_t n = tgtData.size();
        if (splitdata.size() != n || besthapdata.size() != n || cPBWT.size() != n ||
            count0s.size() != n || crefTOffsets.size() != n + 1)
            throw std::runtime_error("GPUEngine::runCrefPBWTKernel: inconsistent block sizes");
        std::vector<int32_t *> c0(n);
        for (size_t t = 0; t < n; t++) {
            count0s[t]->resize(crefTOffsets[t + 1] - crefTOffsets[t]);
            c0[t] = reinterpret_cast<int32_t *>(count0s[t]->data());
        }
        check(eimp_engine_run_cref_pbwt(h.get(), n, tgtData.data(), splitdata.data(),
                                        besthapdata.data(), crefTOffsets.data(), cPBWT.data(),
                                        c0.data()),
              "GPUEngine::runCrefPBWTKernel");
    }

    // src/GPUEngine.h:64
    void free() { check(eimp_engine_free(h.get()), "GPUEngine::free"); }

    // src/GPUEngine.h:66
    int getCUDAIndex() const { return index; }

    // --- device-side prepPhase (no counterpart in the reference engine)
    void setPanelExtras(const float *cM, const uint64_t *hapData = nullptr, size_t hapStride = 0,
                        const uint8_t *haploidRef = nullptr) {
        check(eimp_engine_set_panel_extras(h.get(), cM, hapData, hapStride, haploidRef),
              "GPUEngine::setPanelExtras");
    }
    std::vector<eimp_target_result> runBatch(const std::vector<eimp_target> &targets,
                                             float cMmaxSplit = 0.125f,
                                             uint32_t flags = EIMP_BATCH_DIGEST) {
        std::vector<eimp_target_result> r(targets.size());
        check(eimp_engine_run_batch(h.get(), targets.size(), targets.data(), cMmaxSplit, flags,
                                    r.data()),
              "GPUEngine::runBatch");
        return r;
    }
    void copyTarget(size_t t, uint32_t *besthaps, uint64_t *bhflags, uint32_t *splitsites,
                    uint64_t *splits, float *cMdiffs, int32_t *count0, uint32_t *pbwt) {
        check(eimp_engine_copy_target(h.get(), t, besthaps, bhflags, splitsites, splits, cMdiffs,
                                      count0, pbwt, nullptr),
              "GPUEngine::copyTarget");
    }
    eimp_engine *handle() { return h.get(); }

private:
    struct Deleter {
        void operator()(eimp_engine *e) const { eimp_engine_destroy(e); }
    };
    std::unique_ptr<eimp_engine, Deleter> h;
    int index;
    size_t K = 0, Kwords = 0, blocksize = 0;
};

}  // namespace eimp_cpp
