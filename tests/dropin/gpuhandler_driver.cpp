// gpuhandler_driver.cpp -- the reference's GPUHandler::preProcessGPUOnly flow
// (src/GPUHandler.cpp:78-227) driving the drop-in eimp_cpp::GPUEngine:
// host prepPhase per target (here: the oracle), crefTOffsets prefix sums,
// host cPBWT buffers of 8*Mp*Kwords uint32 and count0 vectors, one
// runCrefPBWTKernel call per block.  Every target's result is compared with
// the oracle's compressed PBWT converted to the drop-in layout.
//
// TEST INFRASTRUCTURE: links the oracle and the generator.
// usage: gpuhandler_driver N M F T K seed
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <vector>

#include "../../include/eimp_gpu_engine.hpp"
#include "../../oracle/eimp_oracle.h"

extern "C" {
struct synth_params {
    uint64_t seed, N, M, F;
    double beta_a, maf_lo, maf_hi;
    int64_t bp_per_site;
    double switch_per_cM, flip_rate;
};
int synth_panel(const synth_params *, uint64_t T, uint64_t *hap, uint64_t hs, uint64_t *site,
                uint64_t ss, uint64_t *tgt, float *cM, int32_t *bp, int nthreads);
int synth_target_gt(const uint64_t *h1, const uint64_t *h2, uint64_t s0, uint64_t W, double miss,
                    uint64_t seed, uint64_t idx, uint64_t *gt);
}

static int run(int argc, char **argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s N M F T K seed\n", argv[0]);
        return 2;
    }
    const uint64_t N = std::strtoull(argv[1], 0, 10), M = std::strtoull(argv[2], 0, 10);
    const uint64_t F = std::strtoull(argv[3], 0, 10), T = std::strtoull(argv[4], 0, 10);
    const uint64_t Karg = std::strtoull(argv[5], 0, 10), seed = std::strtoull(argv[6], 0, 10);
    const uint64_t hs = eo_unit_words(M), ss = eo_unit_words(N);
    std::vector<uint64_t> hap(N * hs), site(M * ss), tgt(2 * T * hs);
    std::vector<float> cM(M);
    synth_params sp{seed, N, M, F, 0.5, 0.005, 0.5, 1000, 1.0, 1e-3};
    if (synth_panel(&sp, T, hap.data(), hs, site.data(), ss, tgt.data(), cM.data(), nullptr, 0))
        return 3;
    const uint64_t K = Karg < N ? Karg : N;  // localK = min(Ncorr, K)
    const uint64_t Kwords = eo_unit_words(K);

    // GPUHandler::initialize (src/GPUHandler.cpp:62-75)
    eimp_cpp::GPUEngine engine(0);
    engine.initialize(K, Kwords, M, N, N, ss * 8, site.data());
    const size_t blocksize = engine.getBlockSize();

    eo_panel p{N, N, M, hap.data(), hs, site.data(), ss, cM.data(), nullptr, -1};
    int bad = 0;
    for (uint64_t b0 = 0; b0 < T; b0 += blocksize) {
        const uint64_t nb = (T - b0 < blocksize) ? T - b0 : blocksize;
        std::vector<std::vector<uint64_t>> gts(nb);
        std::vector<eo_result> res(nb);
        std::vector<const uint64_t *> tgtData(nb), splitData(nb), besthapData(nb);
        std::vector<uint32_t> crefTOffsets(nb + 1, 0);
        std::vector<uint32_t *> cPBWTData(nb);
        std::vector<std::vector<int> *> count0s(nb);
        for (uint64_t t = 0; t < nb; t++) {  // prepPhase per target (:119-136)
            const uint64_t g = b0 + t;
            gts[t].assign(2 * eo_unit_words(M), 0);
            synth_target_gt(&tgt[2 * g * hs], &tgt[(2 * g + 1) * hs], 0, M, 0.01, seed, g,
                            gts[t].data());
            eo_target et{gts[t].data(), 0, M, K, 0.125f};
            if (eo_process_target(&p, &et, 1, &res[t])) return 4;
            tgtData[t] = gts[t].data();
            splitData[t] = res[t].splits;
            besthapData[t] = res[t].flags;
            crefTOffsets[t + 1] = (uint32_t)res[t].Mp;
            cPBWTData[t] = (uint32_t *)std::malloc(8 * res[t].Mp * Kwords * sizeof(uint32_t));
            count0s[t] = new std::vector<int>(res[t].Mp);
        }
        uint32_t cur = 0;  // prefix sums (:139-143)
        for (auto &o : crefTOffsets) {
            cur += o;
            o = cur;
        }
        engine.runCrefPBWTKernel(tgtData, splitData, besthapData, crefTOffsets, cPBWTData, count0s);
        for (uint64_t t = 0; t < nb; t++) {
            const uint64_t Mp = res[t].Mp;
            std::vector<uint32_t> exp(2 * Mp * eo_gpu_row_words(K));
            eo_pbwt_to_gpu_layout(res[t].pbwt, K, Mp, exp.data());
            bool ok = std::memcmp(exp.data(), cPBWTData[t], exp.size() * 4) == 0;
            for (uint64_t r = 0; r < Mp; r++) ok = ok && (*count0s[t])[r] == res[t].count0[r];
            if (!ok) {
                std::printf("target %llu differs\n", (unsigned long long)(b0 + t));
                bad++;
            }
            std::free(cPBWTData[t]);
            delete count0s[t];
            eo_result_free(&res[t]);
        }
    }
    engine.free();
    std::printf("%s: %llu targets, K=%llu, %d differing\n", bad ? "FAIL" : "OK",
                (unsigned long long)T, (unsigned long long)K, bad);
    return bad ? 1 : 0;
}

int main(int argc, char **argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception &ex) {
        std::printf("ERROR: %s\n", ex.what());
        return 1;
    }
}
