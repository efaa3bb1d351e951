// eimp_internal.h -- types shared by the device kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eimp {

// Elements of one condensed-reference block: 16 consecutive cref rows are
// stored transposed, one uint16 per conditioning haplotype (bit j = row
// 16*b + j).  The PBWT kernel carries these bits with each haplotype index
// so that a step needs no gather.
constexpr int kBlockRows = 16;

// Excluded haplotype (ignored target sample or second hap of a haploid
// reference sample) in the distance array: never selected.
constexpr uint32_t kExcluded = 0xFFFFFFFFu;

// Per-target flags
enum : uint32_t {
    TF_SELECT = 1u,    // K < Ncorr: top-K selection by (nerr, idx)
    TF_IDENTITY = 2u,  // bestHaps == 0..N_ad-1 (whole panel, no exclusions)
};

// Per-target record in device memory.  Filled by the host in two stages:
// the prep fields before phase A, the build fields (Mp, offsets) after the
// split counts are known.
struct TgtDev {
    // inputs
    const uint64_t *gt;    // window-relative genotype words (is0, is2 interleaved)
    uint64_t s0;           // window start (panel site)
    uint32_t W;            // window sites
    uint32_t K;            // conditioning haplotypes (<= Ncorr)
    int32_t ignore;        // sample whose haplotypes are skipped, -1 none
    uint32_t flags;        // TF_*
    float cMmaxSplit;
    uint32_t pad0;
    // phase A (prep) outputs / scratch
    uint32_t *nerr;        // N_ad   (TF_SELECT only)
    uint32_t *bh;          // K      bestHaps ascending
    uint64_t *bhflags;     // unit_words(N_ad)
    const uint64_t *mand;  // all-ones-at-site bits (window or panel coordinates)
    const uint64_t *mor;   // any-one-at-site bits
    uint64_t *mand_w;      // window-relative accumulators written by k_monos (TF_IDENTITY off)
    uint64_t *mor_w;
    uint64_t mono_off;     // bit offset of window site 0 in mand/mor
    uint32_t *splitsites;  // W
    uint64_t *splits;      // unit_words(W)
    float *cMdiffs;        // W + 1
    uint32_t *stats;       // [0] S, [1] ncalls, [2] first het, [3] reserved
    // phase B (build)
    uint16_t *cref;        // nblk x Kpad, transposed 16-row blocks
    int32_t *count0;       // Mp
    uint32_t *out;         // 2 x Mp x RS uint32 (forward rows, then backward rows)
    uint64_t *digest;      // [0] pbwt words, [1] count0
    uint32_t Mp;
    uint32_t Kpad;         // K rounded up to 64
    uint32_t RS;           // output row stride in uint32
    uint32_t S;
};

// cref work item: target t, blocks [b0, b1)
struct CrefItem {
    uint32_t t, b0, b1, pad;
};

// PBWT work item: target t, direction (0 fwd, 1 bck)
struct PbwtItem {
    uint32_t t, dir;
};

// Panel in device memory (resident)
struct PanelDev {
    const uint64_t *hap;   // N_ad rows x hap_stride words (bit m = site m)
    const uint64_t *site;  // M rows x site_stride words (bit h = hap h)
    const float *cM;       // M
    const uint8_t *haploid; // N_ad/2 flags or null
    const uint64_t *pand;  // panel-level all-ones bits per site
    const uint64_t *por;   // panel-level any-one bits per site
    uint64_t hap_stride, site_stride;
    uint32_t N_ad, N_corr, M, pad;
};

// 32-bit finaliser of MurmurHash3 (Appleby): bijective avalanche mix used for
// the output digest.  The CPU oracle computes the same sums.
__host__ __device__ inline uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
__host__ __device__ inline uint32_t dig_word(uint64_t g, uint32_t v) {
    return fmix32(v + (uint32_t)g * 0x9E3779B1u);
}
__host__ __device__ inline uint32_t dig_count0(uint32_t r, int32_t c0) {
    return fmix32((uint32_t)c0 + r * 0x7FEB352Du + 0x68E31DA4u);
}

}  // namespace eimp
