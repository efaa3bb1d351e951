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
    TF_RAW = 4u,       // output rows hold data words only (raw layout)
};

// Output row layouts and their stride in uint32 per stored row
enum : uint32_t {
    LAYOUT_GPU = 0u,      // 4*Kwords: (data, cumulative zeros) pairs, tail pairs zero
    LAYOUT_COMPACT = 1u,  // 2*ceil(K/32) pairs, the CPU layout
    LAYOUT_RAW = 2u,      // roundToMultiple(K,512)/32 data words, tail words all ones
};
__host__ __device__ inline uint32_t row_stride(uint32_t K, uint32_t layout) {
    return layout == LAYOUT_COMPACT ? 2 * ((K + 31) / 32)
         : layout == LAYOUT_RAW     ? (K + 511) / 512 * 16
                                    : 4 * ((K + 511) / 512 * 8);
}

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
    uint32_t *stats;       // [0] S, [1] ncalls, [2] deferred (stream contexts), [3] best haps set
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

// Device-side header of a batch run by a stream context (device-driven
// path: no host synchronisation between the phases).
struct BatchDev {
    uint32_t n;           // targets in this submission (written by the host)
    uint32_t ncref;       // cref blocks queued by k_plan: subset (gather) targets
    uint32_t npbwt;       // PBWT items queued by k_plan
    uint32_t nfit;        // targets whose outputs fit the pools
    uint32_t cref_ctr;    // work-queue counters (zeroed by k_plan)
    uint32_t pbwt_ctr;
    uint32_t ncref_id;    // cref blocks of whole-panel (identity) targets
    uint32_t cref_id_ctr;
    // cref queue order per kind ([0] subset, [1] identity): 0 target-major
    // (prefix sums in blockpref), else block-major over the nrr targets listed
    // in blockpref (item i = block i / nrr of target i % nrr)
    uint32_t rr[2];
    uint32_t nrr[2];
};

// 16-row blocks per work item of the subset cref kernel (256 splits)
constexpr uint32_t kCrefGroup = 128;

// cref work item: target t, blocks [b0, b1)
struct CrefItem {
    uint32_t t, b0, b1, pad;
};

// PBWT work item: target t, direction (0 fwd, 1 bck)
struct PbwtItem {
    uint32_t t, dir;
};

// One item of the large-K PBWT kernel (order in global memory): one
// direction of one target, or the panel PBWT.
struct BigItem {
    const uint64_t *rows;    // row r of this item's rows (bit k = haplotype k) at rows + r * rstride
    uint64_t rstride;        // uint64 words per row
    uint32_t *out;           // 2 x Mp (or Mp) rows of RS uint32, this direction's first row
    int32_t *count0;         // Mp zero counts (forward items), else null
    int32_t *absperm;        // (Mp + 1) x K absolute orders (panel PBWT) or null
    uint32_t *ord0, *ord1;   // Kpad indexes each
    uint64_t *digest;        // [0] += digest of the rows, or null
    uint64_t gbase;          // global index of this direction's first logical word (digest)
    uint32_t K, Mp, RS, rev, raw, pad;
};

// Output pool and queues of a stream context, for k_plan.  Each target gets
// one contiguous region of the pool: cref blocks, count0, PBWT rows.
struct PlanArgs {
    char *pool;
    uint64_t pool_cap;
    uint64_t *dig_base;    // 2 words per target
    uint32_t *blockpref;     // nmax + 1 prefix sums of cref blocks (gather targets)
    uint32_t *blockpref_id;  // nmax + 1 prefix sums of cref blocks (identity targets)
    PbwtItem *pitems;        // 2 * nmax
    uint32_t layout;         // LAYOUT_*
    uint32_t digest;
    uint32_t ident_ok;       // identity targets take the transpose kernel
    uint32_t pad;
};

// Bytes of a target's region in a stream context's pool (each part rounded
// to 256 bytes): cref blocks, count0, forward + backward PBWT rows.
__host__ __device__ inline uint64_t region_cref_bytes(uint32_t Mp, uint32_t Kpad) {
    return ((uint64_t)((Mp + 15) / 16) * Kpad * 2 + 255) / 256 * 256;
}
__host__ __device__ inline uint64_t region_c0_bytes(uint32_t Mp) {
    return ((uint64_t)Mp * 4 + 255) / 256 * 256;
}
__host__ __device__ inline uint64_t region_out_bytes(uint32_t Mp, uint32_t RS) {
    return (2ull * Mp * RS * 4 + 255) / 256 * 256;
}

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

// Output digest (the CPU oracle computes the same sums): over the PBWT words,
// sum of rotl32(v ^ g * 0x9E3779B1, g mod 32), g the word's global index;
// over count0, sum of fmix32(c0 + r * 0x7FEB352D + 0x68E31DA4).  Each term is
// a bijection of its word.
// 32-bit finaliser of MurmurHash3 (Appleby).
__host__ __device__ inline uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
constexpr uint32_t kDigKey = 0x9E3779B1u;
__host__ __device__ inline uint32_t dig_word(uint64_t g, uint32_t v) {
    const uint32_t x = v ^ ((uint32_t)g * kDigKey);
    const uint32_t r = (uint32_t)(g & 31);
    return r ? (x << r) | (x >> (32 - r)) : x;
}
__host__ __device__ inline uint32_t dig_count0(uint32_t r, int32_t c0) {
    return fmix32((uint32_t)c0 + r * 0x7FEB352Du + 0x68E31DA4u);
}

}  // namespace eimp
