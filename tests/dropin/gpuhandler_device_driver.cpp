// gpuhandler_device_driver.cpp -- the reference's PBWTPhaser::phaseGPU flow
// (src/PBWTPhaser.cpp:373-533) over the drop-in eimp_cpp::GPUHandler whose
// producer runs prepPhase on the device: blocks of targets into the provide
// queue, one producer thread (preProcessGPUOnly), consumer threads
// (processCondensedRef) reporting confidences.  Between the two stages this
// driver takes every cPBWT queue item, compares its rows, count0 and the
// target's prepPhase results with the oracle, and passes it on.  After the
// run each target's phasing equals the phasing of the oracle's own rows.
//
// TEST INFRASTRUCTURE: links the oracle and the generator.
// usage: gpuhandler_device_driver N M F T K seed consumers haploid_every
//   haploid_every > 0: every that-many-th reference sample is haploid
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

#include "../../include/eimp_gpu_handler.hpp"
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

using namespace eimp_cpp;

static int run(int argc, char **argv) {
    if (argc != 9) {
        std::fprintf(stderr, "usage: %s N M F T K seed consumers haploid_every\n", argv[0]);
        return 2;
    }
    const uint64_t N = std::strtoull(argv[1], 0, 10), M = std::strtoull(argv[2], 0, 10);
    const uint64_t F = std::strtoull(argv[3], 0, 10), T = std::strtoull(argv[4], 0, 10);
    const uint64_t Karg = std::strtoull(argv[5], 0, 10), seed = std::strtoull(argv[6], 0, 10);
    const int ncons = std::atoi(argv[7]);
    const uint64_t hapEvery = std::strtoull(argv[8], 0, 10);
    const uint64_t hs = eo_unit_words(M), ss = eo_unit_words(N);
    std::vector<uint64_t> hap(N * hs), site(M * ss), tgt(2 * T * hs);
    std::vector<float> cM(M);
    synth_params sp{seed, N, M, F, 0.5, 0.005, 0.5, 1000, 1.0, 1e-3};
    if (synth_panel(&sp, T, hap.data(), hs, site.data(), ss, tgt.data(), cM.data(), nullptr, 0))
        return 3;
    std::vector<uint8_t> haploid(N / 2, 0);
    uint64_t nhaploid = 0;
    if (hapEvery)
        for (uint64_t s = 0; s < N / 2; s += hapEvery) haploid[s] = 1, nhaploid++;
    const uint64_t Ncorr = N - nhaploid;
    const uint64_t K = Karg < Ncorr ? Karg : Ncorr;  // localK = min(Ncorr, K)

    // targets: whole-panel windows, 1% missing
    std::vector<std::vector<uint64_t>> gts(T);
    std::vector<Target> targets(T);
    for (uint64_t t = 0; t < T; t++) {
        gts[t].assign(2 * eo_unit_words(M), 0);
        synth_target_gt(&tgt[2 * t * hs], &tgt[(2 * t + 1) * hs], 0, M, 0.01, seed, t, gts[t].data());
        targets[t].idx = t;
        targets[t].gt = gts[t].data();
        targets[t].s0 = 0;
        targets[t].W = M;
    }

    PanelData panel;
    panel.M = M;
    panel.refTCapacity = ss * 8;
    panel.refTData = site.data();
    panel.cM = cM.data();
    panel.haploidRef = hapEvery ? haploid.data() : nullptr;
    GPUHandler handler({0}, panel, T);
    const size_t blocksize = handler.initialize(K, Ncorr, N);

    BoundedQueue<targetQueue_type> provideQueue;
    BoundedQueue<cPBWTQueue_type> cpbwtQueue, checkedQueue;
    BoundedQueue<confidence_type> confidenceQueue;
    confidenceQueue.set_capacity(T);
    std::atomic_bool term(false);

    std::thread producer([&] { handler.preProcessGPUOnly(provideQueue, cpbwtQueue, term, 0); });
    std::vector<std::thread> consumers;
    for (int c = 0; c < ncons; c++)
        consumers.emplace_back(
            [&, c] { handler.processCondensedRef(checkedQueue, confidenceQueue, term, c); });

    // fill the provide queue in blocks of the engine's block size (src/PBWTPhaser.cpp:503-525)
    for (uint64_t b0 = 0; b0 < T; b0 += blocksize) {
        auto block = std::make_shared<std::vector<Target *>>();
        for (uint64_t t = b0; t < T && t < b0 + blocksize; t++) block->push_back(&targets[t]);
        provideQueue.push(block);
    }

    eo_panel p{N, Ncorr, M, hap.data(), hs, site.data(), ss, cM.data(),
               hapEvery ? haploid.data() : nullptr, -1};
    const uint64_t KW = eo_pbwt_row_words(K), RS = eo_gpu_row_words(K);
    std::vector<std::vector<uint32_t>> refRows(T);
    std::vector<std::vector<int32_t>> refC0(T);
    int bad = 0;
    for (uint64_t i = 0; i < T; i++) {
        cPBWTQueue_type item;
        cpbwtQueue.pop(item);
        const Target &tg = *item.target;
        eo_result r;
        eo_target et{tg.gt, tg.s0, tg.W, K, 0.125f};
        if (eo_process_target(&p, &et, 1, &r)) return 4;
        bool ok = tg.K == r.K && tg.S == r.S && tg.ncalls == r.ncalls &&
                  item.count0->size() == r.Mp &&
                  std::memcmp(tg.besthaps.data(), r.besthaps, r.K * 4) == 0 &&
                  (r.S == 0 || std::memcmp(tg.splitsites.data(), r.splitsites, r.S * 4) == 0) &&
                  (r.S == 0 || std::memcmp(tg.cMsplitDiffs.data(), r.cMdiffs, (r.S + 1) * 4) == 0);
        for (uint64_t m = 0; ok && m < r.Mp; m++) ok = (*item.count0)[m] == r.count0[m];
        for (uint64_t row = 0; ok && row < 2 * r.Mp; row++) {
            const uint32_t *g = item.cpbwt + row * RS;
            ok = std::memcmp(g, r.pbwt + row * KW, KW * 4) == 0;
            for (uint64_t w = KW; ok && w < RS; w++) ok = g[w] == 0;
        }
        if (!ok) {
            std::printf("target %zu: cPBWT item differs from the oracle\n", tg.idx);
            bad++;
        }
        refRows[tg.idx].assign(r.pbwt, r.pbwt + 2 * r.Mp * KW);
        refC0[tg.idx].assign(r.count0, r.count0 + r.Mp);
        eo_result_free(&r);
        checkedQueue.push(std::move(item));
    }
    producer.join();
    std::vector<bool> seen(T, false);
    for (uint64_t i = 0; i < T; i++) {
        confidence_type c;
        confidenceQueue.pop(c);
        if (c.id >= T || seen[c.id]) {
            std::printf("confidence for an unknown or repeated target %zu\n", c.id);
            bad++;
            continue;
        }
        seen[c.id] = true;
        if (c.totalconf != targets[c.id].totalconf || c.ncalls != targets[c.id].ncallsites) bad++;
    }
    for (auto &th : consumers) th.join();

    // each target phased from the oracle's rows (CPU layout) gives the same result
    eimp_phase_params P;
    eimp_phase_default_params(&P);
    for (uint64_t t = 0; t < T; t++) {
        const Target &tg = targets[t];
        const size_t words = (M + 63) / 64;
        std::vector<uint64_t> mat(words), pat(words);
        std::vector<float> dm(M), dp(M), cc(tg.ncalls ? tg.ncalls : 1);
        eimp_phase_input in{refRows[t].data(), refC0[t].data(), (uint32_t)K, (uint32_t)KW,
                            (uint32_t)refC0[t].size(), (uint32_t)M, tg.gt, tg.splitsites.data(),
                            tg.cMsplitDiffs.data(), nullptr, 0, 0};
        eimp_phase_output out{};
        out.mat = mat.data();
        out.pat = pat.data();
        out.dosage_mat = dm.data();
        out.dosage_pat = dp.data();
        out.call_conf = cc.data();
        if (eimp_phase_target(&P, &in, &out) != EIMP_OK) return 5;
        const bool ok = mat == tg.mat && pat == tg.pat && out.total_conf == tg.totalconf &&
                        out.ncallsites == tg.ncallsites &&
                        std::memcmp(dm.data(), tg.dosageMat.data(), M * 4) == 0 &&
                        std::memcmp(dp.data(), tg.dosagePat.data(), M * 4) == 0;
        if (!ok) {
            std::printf("target %llu: phasing differs\n", (unsigned long long)t);
            bad++;
        }
    }
    std::printf("%s: %llu targets, K=%llu, Ncorr=%llu, block %zu, %d consumers, %d differing\n",
                bad ? "FAIL" : "OK", (unsigned long long)T, (unsigned long long)K,
                (unsigned long long)Ncorr, blocksize, ncons, bad);
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
