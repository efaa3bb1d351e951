/*
 * eimp_oracle.h -- CPU restatement of EagleImp's phasing pre-stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (eagleimp_amd/) never links or calls it.
 *
 * Each function restates one piece of the reference CPU path and cites the
 * reference file:line it follows.  Bit layouts follow the reference exactly:
 *   - bit vectors are LSB-first uint64 words, padded to 512-bit units
 *     (src/Datatypes.h:70 UNITWORDS=8, src/Datatypes.h:335-339);
 *   - genotype vectors interleave an "is0" and an "is2" word per 64 sites,
 *     Het = (0,0), Miss = (1,1) (src/Datatypes.cpp:27-56);
 *   - the compressed PBWT holds ceil(K/32) (data, cumulative-zero) uint32 pairs
 *     per row, forward rows 0..Mp-1 then backward rows (src/PBWT.cpp:207-285).
 *
 * Pinning (how each function is checked):
 *   eo_compare_with_preinit  golden vectors from the reference Datatypes.cpp
 *                            (tests/golden) + closed form
 *   eo_find_best_haps        brute-force sort of (nerr, idx) + KATs (ties,
 *                            K = Ncorr identity branch, haploid/ignore skips)
 *   eo_find_monos            brute force ("all chosen rows equal")
 *   eo_find_call_locs        KATs (mono first het, missing, no het, no call,
 *                            trailing-hom truncation, float cM edges)
 *   eo_condensed_ref         brute-force per-segment OR + gather; count0 invariant
 *   eo_build_pbwt            golden vectors from the reference PBWT.cpp
 *                            (tests/golden) + Durbin's reversed-prefix sort
 *   eo_map_interval          golden vectors from the reference PBWT.cpp
 */
#ifndef EIMP_ORACLE_H_
#define EIMP_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Panel description (reference: VCFData getters src/VCFData.h:203-222). */
typedef struct eo_panel {
    uint64_t N_ad;            /* NrefhapsAllDiploid: haplotypes incl. haploid-as-diploid */
    uint64_t N_corr;          /* NrefhapsCorrected: haplotypes really available */
    uint64_t M;               /* panel sites */
    const uint64_t *hap;      /* hap-major rows (src/VCFData.h:203): N_ad rows */
    uint64_t hap_stride;      /* uint64 words per hap-major row, >= ceil(M/64) */
    const uint64_t *site;     /* site-major rows (src/VCFData.h:205): M rows */
    uint64_t site_stride;     /* uint64 words per site-major row, >= ceil(N_ad/64) */
    const float *cM;          /* M genetic positions as float (src/VCFData.cpp:1189) */
    const uint8_t *haploid_ref; /* N_ad/2 flags (src/VCFData.h getHaploidsRef) or NULL */
    int64_t ignore_sample;    /* sample skipped in top-K (src/Target.cpp:791-795), -1 none */
} eo_panel;

/* One target: a window [s0, s0+W) of the panel's sites (the reference's chunk
 * semantics, src/VCFData.cpp:648-830), its genotypes over that window and K. */
typedef struct eo_target {
    const uint64_t *gt;       /* window-relative GenotypeVector words (is0, is2 interleaved) */
    uint64_t s0;
    uint64_t W;
    uint64_t K;               /* must be <= N_corr (localK = min(Ncorr, K), src/PBWTPhaser.cpp:683) */
    float cMmaxSplit;         /* src/Args.cpp:290 default 0.125f */
} eo_target;

/* Sizes (src/utils.h:86-92 roundToMultiple). */
uint64_t eo_round_up(uint64_t n, uint64_t mult);
uint64_t eo_unit_words(uint64_t nbits);          /* roundToMultiple(nbits,512)/64 */
uint64_t eo_pbwt_row_words(uint64_t K);          /* 2*ceil(K/32)  (src/PBWT.cpp:36) */
uint64_t eo_gpu_row_words(uint64_t K);           /* 4*roundToMultiple(K,512)/64 (src/GPUHandler.cpp:64,259) */

/* a3: src/Datatypes.cpp:107-116 */
uint64_t eo_compare_with_preinit(const uint64_t *gt, uint64_t gt_words,
                                 const uint64_t *hap, uint64_t hap_words);

/* a4: src/Target.cpp:779-875.  besthaps[K] ascending, flags has
 * eo_unit_words(N_ad) words (zeroed here). */
void eo_find_best_haps(const eo_panel *p, const eo_target *t,
                       uint32_t *besthaps, uint64_t *flags);

/* a5: src/Target.cpp:878-897.  monos has eo_unit_words(W) words. */
void eo_find_monos(const eo_panel *p, const eo_target *t,
                   const uint32_t *besthaps, uint64_t *monos);

/* a6: src/Target.cpp:900-950.  splitsites/callLocs sized >= W, cMdiffs >= W+1,
 * splits has eo_unit_words(W) words.  Returns S (number of split sites). */
uint64_t eo_find_call_locs(const eo_panel *p, const eo_target *t,
                           const uint64_t *monos, uint64_t *splits,
                           uint32_t *splitsites, uint32_t *callLocs,
                           uint64_t *ncalls, float *cMdiffs);

/* a7: src/Target.cpp:953-1059.  cref has Mp = 2S+1 rows of
 * eo_unit_words(K) words (zeroed here), count0 has Mp ints. */
void eo_condensed_ref(const eo_panel *p, const eo_target *t,
                      const uint32_t *besthaps, const uint32_t *splitsites,
                      uint64_t S, uint64_t *cref, int32_t *count0);

/* a9: src/PBWT.cpp:33-51,131-153,207-285.  cref rows of cref_stride words,
 * out has 2*Mp rows of eo_pbwt_row_words(K) uint32 (forward then backward). */
/* Panel PBWT for imputation (PBWT(refT, true), src/PBWT.cpp:53-71,131-153):
 * forward only over M site rows (row_stride uint64 words, bit h = haplotype h)
 * of K haplotypes; out M rows of eo_pbwt_row_words(K) uint32, count0 M,
 * absperm (M+1) x K (row m = order before site m) or NULL. */
void eo_build_ref_pbwt(const uint64_t *rows, uint64_t row_stride, uint64_t K, uint64_t M,
                       uint32_t *out, int32_t *count0, int32_t *absperm);

void eo_build_pbwt(const uint64_t *cref, uint64_t cref_stride, uint64_t K,
                   uint64_t Mp, const int32_t *count0, uint32_t *out);

/* a10: CPU layout -> drop-in GPU layout (stride eo_gpu_row_words(K), tail
 * pairs zero), src/GPUHandler.cpp:127,259 and src/GPUEngine.cu:182-211. */
void eo_pbwt_to_gpu_layout(const uint32_t *cpu, uint64_t K, uint64_t Mp, uint32_t *gpu);

/* a11: src/PBWT.cpp:288-382.  Maps interval [l, r] (inclusive, l > r empty)
 * of row m (reverse selects the backward rows).  out4 = {l0, r0, l1, r1}. */
void eo_map_interval(const uint32_t *pbwt, uint64_t K, uint64_t Mp,
                     const int32_t *count0, int reverse, uint64_t m,
                     int32_t l, int32_t r, int32_t *out4);
int32_t eo_decode_rel_perm(const uint32_t *row, int32_t gcnt0, int32_t from);

/* Digest of the logical cPBWT content and of count0 (the sink comparison
 * used by the device engine; same definition on both sides). */
uint64_t eo_mix64(uint64_t z);
void eo_digest(const uint32_t *pbwt, uint64_t row_stride, uint64_t K, uint64_t Mp,
               const int32_t *count0, uint64_t *dig_pbwt, uint64_t *dig_count0);

/* Whole target (a4..a9 sequenced as src/Target.cpp:95-163,251-282).
 * Allocates internally; results are returned through the eo_result which
 * must be released with eo_result_free. */
typedef struct eo_result {
    uint64_t K, W, S, Mp, ncalls;
    uint32_t *besthaps;   /* K */
    uint64_t *flags;      /* eo_unit_words(N_ad) */
    uint64_t *monos;      /* eo_unit_words(W) */
    uint64_t *splits;     /* eo_unit_words(W) */
    uint32_t *splitsites; /* S */
    uint32_t *callLocs;   /* ncalls */
    float *cMdiffs;       /* S+1 (0 if S == 0) */
    uint64_t *cref;       /* Mp x eo_unit_words(K) */
    int32_t *count0;      /* Mp */
    uint32_t *pbwt;       /* 2*Mp x eo_pbwt_row_words(K) */
    uint64_t dig_pbwt, dig_count0;
} eo_result;

int eo_process_target(const eo_panel *p, const eo_target *t, int keep_arrays, eo_result *res);
void eo_result_free(eo_result *res);

/* Batch over targets with OpenMP schedule(dynamic,1), as
 * src/PBWTPhaser.cpp:699-700.  Writes S, Mp and digests per target.
 * nthreads <= 0 uses the OpenMP default. */
int eo_process_batch(const eo_panel *p, uint64_t ntargets, const eo_target *targets,
                     int nthreads, uint64_t *S, uint64_t *Mp,
                     uint64_t *dig_pbwt, uint64_t *dig_count0);

#ifdef __cplusplus
}
#endif
#endif
