/*
 * eimp_oracle.c -- CPU restatement of EagleImp's phasing pre-stage
 * (top-K conditioning haplotypes, monos, split/call sites, condensed
 * reference and compressed forward/backward PBWT).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device engine and the timed
 * CPU baseline.  Nothing in the product path links it.
 *
 * Every function names the reference lines it restates.  See eimp_oracle.h
 * for the pinning of each function.
 */
#include "eimp_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define UNITBITS 512u /* UNITWORDS(8) * 64, src/Datatypes.h:70 */

uint64_t eo_round_up(uint64_t n, uint64_t mult) {
    /* src/utils.h:86-92 roundToMultiple */
    return (n % mult) ? n + mult - (n % mult) : n;
}

uint64_t eo_unit_words(uint64_t nbits) { return eo_round_up(nbits, UNITBITS) / 64; }
uint64_t eo_pbwt_row_words(uint64_t K) { return 2 * ((K + 31) / 32); }
uint64_t eo_gpu_row_words(uint64_t K) { return 4 * (eo_round_up(K, UNITBITS) / 64); }

static inline int bit_at(const uint64_t *v, uint64_t pos) {
    /* BooleanVector::at, src/Datatypes.h:207-212 */
    return (int)((v[pos >> 6] >> (pos & 63)) & 1u);
}
static inline void bit_set(uint64_t *v, uint64_t pos) {
    /* BooleanVector::setWithPreInit, src/Datatypes.h:180-185 */
    v[pos >> 6] |= 1ull << (pos & 63);
}

/* Genotype codes as src/Datatypes.h:81-86 */
enum { G_HOMREF = 0, G_HET = 1, G_HOMALT = 2, G_MISS = 3 };

static inline int gt_at(const uint64_t *gt, uint64_t pos) {
    /* GenotypeVector::at, src/Datatypes.cpp:79-90 */
    uint64_t w = pos >> 6, m = 1ull << (pos & 63);
    int is0 = (gt[2 * w] & m) != 0, is2 = (gt[2 * w + 1] & m) != 0;
    if (is0 && is2) return G_MISS;
    if (is0) return G_HOMREF;
    if (is2) return G_HOMALT;
    return G_HET;
}

/* ---------------------------------------------------------------------- */
/* Window view: the reference runs the same algorithm on a chunk of sites;  */
/* a window [s0, s0+W) is that chunk.  Hap-major rows are sliced so that    */
/* window site 0 is bit 0 (the chunk's own BooleanVectors,                  */
/* src/VCFData.cpp:691-704), site-major rows and cM are plain offsets.      */
/* ---------------------------------------------------------------------- */
typedef struct {
    const uint64_t *hap;
    uint64_t hap_stride;
    uint64_t *owned;
    const uint64_t *site;
    uint64_t site_stride;
    const float *cM;
    uint64_t W;
    uint64_t hw; /* eo_unit_words(W) */
} eo_window;

static int window_open(const eo_panel *p, const eo_target *t, eo_window *w) {
    w->W = t->W;
    w->hw = eo_unit_words(t->W);
    w->site = p->site + t->s0 * p->site_stride;
    w->site_stride = p->site_stride;
    w->cM = p->cM + t->s0;
    w->owned = NULL;
    if (t->s0 == 0 && t->W == p->M) {
        w->hap = p->hap;
        w->hap_stride = p->hap_stride;
        return 0;
    }
    uint64_t hw = w->hw;
    uint64_t nbuf = p->N_ad * hw;
    uint64_t *buf = (uint64_t *)calloc(nbuf > 0 ? nbuf : 1, sizeof(uint64_t));
    if (!buf) return -1;
    uint64_t nfull = t->W / 64, rem = t->W % 64;
    uint64_t sh = t->s0 & 63, base = t->s0 >> 6;
    for (uint64_t h = 0; h < p->N_ad; h++) {
        const uint64_t *src = p->hap + h * p->hap_stride + base;
        uint64_t *dst = buf + h * hw;
        uint64_t nw = nfull + (rem ? 1 : 0);
        for (uint64_t i = 0; i < nw; i++) {
            uint64_t lo = src[i] >> sh;
            /* the high part exists only if the window continues past this source word */
            uint64_t hi = 0;
            if (sh && (t->s0 + 64 * i + (64 - sh)) < t->s0 + t->W) hi = src[i + 1] << (64 - sh);
            dst[i] = lo | hi;
        }
        if (rem) dst[nfull] &= (1ull << rem) - 1;
    }
    w->owned = buf;
    w->hap = buf;
    w->hap_stride = hw;
    return 0;
}

static void window_close(eo_window *w) { free(w->owned); w->owned = NULL; }

/* ---------------------------------------------------------------------- */
/* a3: GenotypeVector::compare_withPreInit, src/Datatypes.cpp:107-116       */
/* ---------------------------------------------------------------------- */
uint64_t eo_compare_with_preinit(const uint64_t *gt, uint64_t gt_words,
                                 const uint64_t *hap, uint64_t hap_words) {
    uint64_t nerr = 0;
    uint64_t limit = gt_words / 2 < hap_words ? gt_words / 2 : hap_words;
    for (uint64_t i = 0; i < limit; i++) {
        uint64_t tmp = (gt[2 * i] & hap[i]) | (gt[2 * i + 1] & ~hap[i]);
        nerr += (uint64_t)__builtin_popcountll(tmp);
    }
    return nerr;
}

/* ---------------------------------------------------------------------- */
/* a4: Target::findBestHaps, src/Target.cpp:779-875                        */
/* ---------------------------------------------------------------------- */
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static void best_haps_impl(const eo_panel *p, const eo_window *w, const eo_target *t,
                           uint32_t *bh, uint64_t *flags) {
    uint64_t K = t->K;
    memset(flags, 0, eo_unit_words(p->N_ad) * sizeof(uint64_t));
    int64_t ignore = p->ignore_sample;
    if (K >= p->N_corr) { /* take all references (:798-826) */
        if (p->N_ad != p->N_corr) {
            uint64_t n = 0;
            for (uint64_t nr = 0; nr < p->N_ad / 2; nr++) {
                if ((int64_t)nr == ignore) continue;
                bh[n++] = (uint32_t)(2 * nr);
                if (!(p->haploid_ref && p->haploid_ref[nr])) bh[n++] = (uint32_t)(2 * nr + 1);
            }
        } else {
            for (uint64_t nr = 0; nr < p->N_ad; nr++) bh[nr] = (uint32_t)nr;
        }
        for (uint64_t k = 0; k < K; k++) bit_set(flags, bh[k]);
        return;
    }
    /* K < Ncorr: score, sort pair(nerr, idx), take K, re-sort by idx (:827-853).
     * pair<size_t,size_t> ordering == ordering of (nerr << 32 | idx) since
     * nerr <= W < 2^32 and idx < 2^32. */
    uint64_t gtw = 2 * w->hw;
    uint64_t *keys = (uint64_t *)malloc((p->N_corr ? p->N_corr : 1) * sizeof(uint64_t));
    uint64_t n = 0;
    for (uint64_t nr = 0; nr < p->N_ad / 2; nr++) {
        if ((int64_t)nr == ignore) continue;
        uint64_t h = 2 * nr;
        uint64_t e = eo_compare_with_preinit(t->gt, gtw, w->hap + h * w->hap_stride, w->hw);
        keys[n++] = (e << 32) | h;
        if (!(p->haploid_ref && p->haploid_ref[nr])) {
            h = 2 * nr + 1;
            e = eo_compare_with_preinit(t->gt, gtw, w->hap + h * w->hap_stride, w->hw);
            keys[n++] = (e << 32) | h;
        }
    }
    qsort(keys, n, sizeof(uint64_t), cmp_u64);
    for (uint64_t k = 0; k < K; k++) bh[k] = (uint32_t)(keys[k] & 0xffffffffu);
    qsort(bh, K, sizeof(uint32_t), cmp_u32);
    for (uint64_t k = 0; k < K; k++) bit_set(flags, bh[k]);
    free(keys);
}

void eo_find_best_haps(const eo_panel *p, const eo_target *t, uint32_t *bh, uint64_t *flags) {
    eo_window w;
    if (window_open(p, t, &w)) return;
    best_haps_impl(p, &w, t, bh, flags);
    window_close(&w);
}

/* ---------------------------------------------------------------------- */
/* a5: Target::findMonos, src/Target.cpp:878-897 (init per                  */
/* BooleanVector::setInitV, src/Datatypes.h:350-362)                        */
/* ---------------------------------------------------------------------- */
static void monos_impl(const eo_window *w, const eo_target *t, const uint32_t *bh, uint64_t *monos) {
    uint64_t datasize = w->hw;
    memset(monos, 0xff, datasize * sizeof(uint64_t));
    if (w->W % UNITBITS) {
        uint64_t mask = (1ull << (w->W % 64)) - 1ull;
        uint64_t idx = datasize - 8 + (w->W % UNITBITS) / 64;
        monos[idx] = mask;
        for (uint64_t i = idx + 1; i < datasize; i++) monos[i] = 0;
    }
    for (uint64_t curr = 1; curr < t->K; curr++) {
        const uint64_t *last = w->hap + (uint64_t)bh[curr - 1] * w->hap_stride;
        const uint64_t *cur = w->hap + (uint64_t)bh[curr] * w->hap_stride;
        for (uint64_t i = 0; i < datasize; i++) monos[i] &= ~(last[i] ^ cur[i]);
    }
}

void eo_find_monos(const eo_panel *p, const eo_target *t, const uint32_t *bh, uint64_t *monos) {
    eo_window w;
    if (window_open(p, t, &w)) return;
    monos_impl(&w, t, bh, monos);
    window_close(&w);
}

/* ---------------------------------------------------------------------- */
/* a6: Target::findCallLocs, src/Target.cpp:900-950                         */
/* ---------------------------------------------------------------------- */
static uint64_t call_locs_impl(const eo_window *w, const eo_target *t, const uint64_t *monos,
                               uint64_t *splits, uint32_t *splitsites, uint32_t *callLocs,
                               uint64_t *ncalls_out, float *cMdiffs) {
    uint64_t W = w->W;
    memset(splits, 0, w->hw * sizeof(uint64_t));
    const float *cM = w->cM;
    float last = cM[0];
    uint64_t S = 0, nc = 0, m;
    /* the first het site starts the phasing, it may be mono (:907-913) */
    for (m = 0; m < W; m++)
        if (gt_at(t->gt, m) == G_HET) break;
    for (; m < W; m++) {
        int g = gt_at(t->gt, m);
        if (bit_at(monos, m) || g == G_MISS) continue;
        float d = cM[m] - last; /* float arithmetic, fp_type = float */
        if (g == G_HET || d > t->cMmaxSplit) {
            if (g == G_HET) callLocs[nc++] = (uint32_t)S;
            splitsites[S++] = (uint32_t)m;
            bit_set(splits, m);
            last = cM[m];
        }
    }
    *ncalls_out = 0;
    if (S == 0) return 0;                      /* :927-929 */
    if (nc == 0) {                             /* :931-935 */
        memset(splits, 0, w->hw * sizeof(uint64_t));
        return 0;
    }
    S = (uint64_t)callLocs[nc - 1] + 1;        /* :939 truncate after last call */
    {                                          /* :940 BooleanVector::clearAfter */
        uint64_t pos = splitsites[S - 1] + 1;
        uint64_t idx = pos / 64;
        splits[idx] &= (1ull << (pos % 64)) - 1ull;
        for (idx++; idx < w->hw; idx++) splits[idx] = 0;
    }
    *ncalls_out = nc;
    /* :943-948 cMsplitDiffs */
    cMdiffs[0] = cM[splitsites[0]] - cM[0];
    for (uint64_t i = 1; i < S; i++) cMdiffs[i] = cM[splitsites[i]] - cM[splitsites[i - 1]];
    cMdiffs[S] = cM[W - 1] - cM[splitsites[S - 1]];
    return S;
}

uint64_t eo_find_call_locs(const eo_panel *p, const eo_target *t, const uint64_t *monos,
                           uint64_t *splits, uint32_t *splitsites, uint32_t *callLocs,
                           uint64_t *ncalls, float *cMdiffs) {
    eo_window w;
    w.W = t->W;
    w.hw = eo_unit_words(t->W);
    w.cM = p->cM + t->s0;
    return call_locs_impl(&w, t, monos, splits, splitsites, callLocs, ncalls, cMdiffs);
}

/* ---------------------------------------------------------------------- */
/* a7: Target::condensedReferenceT + mergeInconsistencies,                  */
/* src/Target.cpp:953-1059                                                  */
/* ---------------------------------------------------------------------- */
static void cref_impl(const eo_panel *p, const eo_window *w, const eo_target *t,
                      const uint32_t *bh, const uint32_t *splitsites, uint64_t S,
                      uint64_t *cref, int32_t *count0) {
    uint64_t K = t->K, Mp = 2 * S + 1;
    uint64_t crefw = eo_unit_words(K);
    memset(cref, 0, Mp * crefw * sizeof(uint64_t));
    uint64_t incw = eo_unit_words(p->N_ad);
    uint64_t *inc = (uint64_t *)calloc(incw, sizeof(uint64_t));
    uint64_t nopad = (p->N_ad + 63) / 64;
    uint64_t padmask = (p->N_ad % 64) ? ((1ull << (p->N_ad % 64)) - 1) : ~0ull;
    int full = (K == p->N_ad);
    uint64_t row = 0, si = 0;
    uint64_t nextsplit = S ? splitsites[si++] : ~0ull;

#define EMIT_FROM(src, is_inc)                                                     \
    do {                                                                           \
        int count1 = 0;                                                            \
        uint64_t *dst = cref + row * crefw;                                        \
        if (full) {                                                                \
            if (is_inc) inc[nopad - 1] &= padmask;                                 \
            memcpy(dst, (src), nopad * sizeof(uint64_t));                          \
            for (uint64_t q = 0; q < nopad; q++) count1 += __builtin_popcountll((src)[q]); \
        } else {                                                                   \
            for (uint64_t k = 0; k < K; k++)                                       \
                if (bit_at((src), bh[k])) { bit_set(dst, k); count1++; }           \
        }                                                                          \
        count0[row] = (int32_t)K - count1;                                         \
        row++;                                                                     \
    } while (0)

    for (uint64_t m = 0; m < w->W; m++) {
        const uint64_t *srow = w->site + m * w->site_stride;
        if (m == nextsplit) {
            EMIT_FROM(inc, 1);                                    /* :962-977 */
            memset(inc, 0, incw * sizeof(uint64_t));              /* :983 */
            EMIT_FROM(srow, 0);                                   /* :986-1003 */
            nextsplit = si < S ? splitsites[si++] : ~0ull;        /* :1007 */
        } else {
            int g = gt_at(t->gt, m);                              /* :1040-1059 */
            if (g == G_HOMREF) {
                for (uint64_t q = 0; q < nopad; q++) inc[q] |= srow[q];
            } else if (g == G_HOMALT) {
                for (uint64_t q = 0; q < nopad; q++) inc[q] |= ~srow[q];
            }
        }
    }
    EMIT_FROM(inc, 1);                                            /* :1021-1035 */
#undef EMIT_FROM
    free(inc);
}

void eo_condensed_ref(const eo_panel *p, const eo_target *t, const uint32_t *bh,
                      const uint32_t *splitsites, uint64_t S, uint64_t *cref, int32_t *count0) {
    eo_window w;
    w.W = t->W;
    w.hw = eo_unit_words(t->W);
    w.site = p->site + t->s0 * p->site_stride;
    w.site_stride = p->site_stride;
    w.cM = p->cM + t->s0;
    cref_impl(p, &w, t, bh, splitsites, S, cref, count0);
}

/* ---------------------------------------------------------------------- */
/* a9: PBWT::buildComprPBWT, src/PBWT.cpp:207-285, driven as PBWT::init     */
/* (:131-153) with the precomputed count0 (:33-51).                         */
/* ---------------------------------------------------------------------- */
void eo_build_pbwt(const uint64_t *cref, uint64_t cref_stride, uint64_t K, uint64_t Mp,
                   const int32_t *count0, uint32_t *out) {
    uint64_t KW = eo_pbwt_row_words(K);
    int32_t *perm = (int32_t *)malloc((K ? K : 1) * sizeof(int32_t));
    int32_t *nperm = (int32_t *)malloc((K ? K : 1) * sizeof(int32_t));
    for (int reverse = 0; reverse < 2; reverse++) {
        uint32_t *compr = out + (reverse ? Mp * KW : 0);
        for (uint64_t i = 0; i < K; i++) perm[i] = (int32_t)i; /* :212-215 */
        for (uint64_t midx = 0; midx < Mp; midx++) {
            uint32_t *cp = compr;
            uint64_t m = reverse ? Mp - 1 - midx : midx;
            const uint64_t *x = cref + m * cref_stride;
            uint64_t p0 = 0, p1 = (uint64_t)count0[m];
            uint32_t data = 0, mask = 1;
            for (uint64_t i = 0; i < K; i++) {
                int32_t e = perm[i];
                if (!bit_at(x, (uint64_t)e)) {
                    nperm[p0++] = e;
                } else {
                    data |= mask;
                    nperm[p1++] = e;
                }
                mask <<= 1;
                if (mask == 0) {
                    *cp++ = data;
                    *cp++ = (uint32_t)p0;
                    data = 0;
                    mask = 1;
                }
            }
            if (K % 32) { /* 1-padding of the last word (:261-267) */
                mask = ~(mask - 1);
                *cp++ = data | mask;
                *cp++ = (uint32_t)p0;
            }
            compr += KW;
            int32_t *tmp = perm; perm = nperm; nperm = tmp;
        }
    }
    free(perm);
    free(nperm);
}

/* ---------------------------------------------------------------------- */
/* Panel PBWT for imputation: PBWT(refT, true), src/PBWT.cpp:53-71,131-153 */
/* and buildComprPBWT (:207-285) with refPBWT set: forward only over M site */
/* rows of K haplotypes, count0 computed per site (haveCount0 false,        */
/* :228,275-276), the absolute permutation before every site kept (:218-222, */
/* absperm row m; row M = after the last site).                            */
/* ---------------------------------------------------------------------- */
void eo_build_ref_pbwt(const uint64_t *rows, uint64_t row_stride, uint64_t K, uint64_t M,
                       uint32_t *out, int32_t *count0, int32_t *absperm) {
    uint64_t KW = eo_pbwt_row_words(K);
    int32_t *perm = (int32_t *)malloc((K ? K : 1) * sizeof(int32_t));
    int32_t *nperm = (int32_t *)malloc((K ? K : 1) * sizeof(int32_t));
    for (uint64_t i = 0; i < K; i++) perm[i] = (int32_t)i;
    if (absperm) memcpy(absperm, perm, K * sizeof(int32_t));
    for (uint64_t m = 0; m < M; m++) {
        const uint64_t *x = rows + m * row_stride;
        uint64_t c0 = 0;
        for (uint64_t i = 0; i < K; i++) c0 += !bit_at(x, i);
        uint32_t *cp = out + m * KW;
        uint64_t p0 = 0, p1 = c0;
        uint32_t data = 0, mask = 1;
        for (uint64_t i = 0; i < K; i++) {
            int32_t e = perm[i];
            if (!bit_at(x, (uint64_t)e)) {
                nperm[p0++] = e;
            } else {
                data |= mask;
                nperm[p1++] = e;
            }
            mask <<= 1;
            if (mask == 0) {
                *cp++ = data;
                *cp++ = (uint32_t)p0;
                data = 0;
                mask = 1;
            }
        }
        if (K % 32) {
            mask = ~(mask - 1);
            *cp++ = data | mask;
            *cp++ = (uint32_t)p0;
        }
        count0[m] = (int32_t)p0;
        int32_t *tmp = perm; perm = nperm; nperm = tmp;
        if (absperm) memcpy(absperm + (m + 1) * K, perm, K * sizeof(int32_t));
    }
    free(perm);
    free(nperm);
}

/* ---------------------------------------------------------------------- */
/* a10: drop-in GPU layout (src/GPUHandler.cpp:127,259; src/GPUEngine.cu:160,182-211) */
/* ---------------------------------------------------------------------- */
void eo_pbwt_to_gpu_layout(const uint32_t *cpu, uint64_t K, uint64_t Mp, uint32_t *gpu) {
    uint64_t KW = eo_pbwt_row_words(K), GW = eo_gpu_row_words(K);
    for (uint64_t q = 0; q < 2 * Mp; q++) {
        memcpy(gpu + q * GW, cpu + q * KW, KW * sizeof(uint32_t));
        memset(gpu + q * GW + KW, 0, (GW - KW) * sizeof(uint32_t));
    }
}

/* ---------------------------------------------------------------------- */
/* a11: PBWT::decodeRelPerm / mapInterval, src/PBWT.cpp:288-382             */
/* ---------------------------------------------------------------------- */
int32_t eo_decode_rel_perm(const uint32_t *row, int32_t gcnt0, int32_t from) {
    uint32_t off = row[2 * (from / 32) + 1];
    uint32_t field = row[2 * (from / 32)];
    uint32_t mask = 1u << (from % 32);
    uint32_t bit = field & mask;
    mask--;
    field |= mask;
    uint32_t fcnt0 = off - (uint32_t)__builtin_popcount(~field);
    if (bit) return gcnt0 + (from - (int32_t)fcnt0);
    return (int32_t)fcnt0;
}

void eo_map_interval(const uint32_t *pbwt, uint64_t K, uint64_t Mp, const int32_t *count0,
                     int reverse, uint64_t m, int32_t fl, int32_t fr, int32_t *o) {
    uint64_t KW = eo_pbwt_row_words(K);
    const uint32_t *row = pbwt + (reverse ? (2 * Mp - 1 - m) : m) * KW;
    int32_t cnt0 = count0[m];
    /* empty defaults: (0, -1) as PBWTInterval() */
    o[0] = 0; o[1] = -1; o[2] = 0; o[3] = -1;
    if (fl > fr) return;
    if (fl == fr) {
        int32_t i = eo_decode_rel_perm(row, cnt0, fl);
        if (i < cnt0) { o[0] = i; o[1] = i; }
        else { o[2] = i; o[3] = i; }
        return;
    }
    int32_t l0 = eo_decode_rel_perm(row, cnt0, fl);
    int32_t l1 = fl - l0 + cnt0;
    if (l0 >= cnt0) { int32_t tt = l0; l0 = l1; l1 = tt; }
    int32_t r0 = eo_decode_rel_perm(row, cnt0, fr);
    int32_t r1 = fr - r0 + cnt0 - 1;
    if (r0 >= cnt0) { int32_t tt = r0; r0 = r1; r1 = tt; }
    o[0] = l0; o[1] = r0; o[2] = l1; o[3] = r1;
}

/* ---------------------------------------------------------------------- */
/* Digest of the logical output (both sides compute the same sums): the sum  */
/* over every logical word v of rotl32(v ^ g * 0x9E3779B1, g mod 32), g the  */
/* word's index in (direction, stored row, word) order, and over count0 of   */
/* fmix32(c0 + r * 0x7FEB352D + 0x68E31DA4).  Each term is a bijection of    */
/* its word, so any single differing word changes the sum.                  */
/* ---------------------------------------------------------------------- */
static inline uint32_t dig_term(uint64_t g, uint32_t v) {
    const uint32_t x = v ^ ((uint32_t)g * 0x9E3779B1u);
    const uint32_t r = (uint32_t)(g & 31);
    return r ? (x << r) | (x >> (32 - r)) : x;
}
static inline uint32_t fmix32(uint32_t h) {
    /* MurmurHash3 32-bit finaliser (Appleby) */
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

uint64_t eo_mix64(uint64_t z) {
    /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void eo_digest(const uint32_t *pbwt, uint64_t row_stride, uint64_t K, uint64_t Mp,
               const int32_t *count0, uint64_t *dp, uint64_t *dc) {
    uint64_t KW = eo_pbwt_row_words(K);
    uint64_t s = 0;
    for (uint64_t q = 0; q < 2 * Mp; q++)
        for (uint64_t w = 0; w < KW; w++) {
            uint64_t g = q * KW + w;
            s += dig_term(g, pbwt[q * row_stride + w]);
        }
    uint64_t c = 0;
    for (uint64_t r = 0; r < Mp; r++)
        c += fmix32((uint32_t)count0[r] + (uint32_t)r * 0x7FEB352Du + 0x68E31DA4u);
    *dp = s;
    *dc = c;
}

/* ---------------------------------------------------------------------- */
/* a8: Target::prepPhase + Target::phase up to the PBWT,                    */
/* src/Target.cpp:95-163,251-282                                            */
/* ---------------------------------------------------------------------- */
void eo_result_free(eo_result *r) {
    free(r->besthaps); free(r->flags); free(r->monos); free(r->splits);
    free(r->splitsites); free(r->callLocs); free(r->cMdiffs);
    free(r->cref); free(r->count0); free(r->pbwt);
    memset(r, 0, sizeof(*r));
}

int eo_process_target(const eo_panel *p, const eo_target *t, int keep, eo_result *r) {
    memset(r, 0, sizeof(*r));
    if (t->K < 1 || t->K > p->N_corr || t->W < 1 || t->s0 + t->W > p->M) return -1;
    eo_window w;
    if (window_open(p, t, &w)) return -2;
    uint64_t K = t->K, W = t->W;
    r->K = K;
    r->W = W;
    r->besthaps = (uint32_t *)malloc(K * sizeof(uint32_t));
    r->flags = (uint64_t *)calloc(eo_unit_words(p->N_ad), sizeof(uint64_t));
    r->monos = (uint64_t *)calloc(w.hw, sizeof(uint64_t));
    r->splits = (uint64_t *)calloc(w.hw, sizeof(uint64_t));
    r->splitsites = (uint32_t *)malloc(W * sizeof(uint32_t));
    r->callLocs = (uint32_t *)malloc(W * sizeof(uint32_t));
    r->cMdiffs = (float *)malloc((W + 1) * sizeof(float));
    best_haps_impl(p, &w, t, r->besthaps, r->flags);
    monos_impl(&w, t, r->besthaps, r->monos);
    r->S = call_locs_impl(&w, t, r->monos, r->splits, r->splitsites, r->callLocs,
                          &r->ncalls, r->cMdiffs);
    r->Mp = 2 * r->S + 1;
    uint64_t crefw = eo_unit_words(K);
    r->cref = (uint64_t *)malloc(r->Mp * crefw * sizeof(uint64_t));
    r->count0 = (int32_t *)malloc(r->Mp * sizeof(int32_t));
    cref_impl(p, &w, t, r->besthaps, r->splitsites, r->S, r->cref, r->count0);
    window_close(&w);
    uint64_t KW = eo_pbwt_row_words(K);
    r->pbwt = (uint32_t *)malloc(2 * r->Mp * KW * sizeof(uint32_t));
    eo_build_pbwt(r->cref, crefw, K, r->Mp, r->count0, r->pbwt);
    eo_digest(r->pbwt, KW, K, r->Mp, r->count0, &r->dig_pbwt, &r->dig_count0);
    if (!keep) {
        uint64_t S = r->S, Mp = r->Mp, nc = r->ncalls, dp = r->dig_pbwt, dc = r->dig_count0;
        eo_result_free(r);
        r->K = K; r->W = W; r->S = S; r->Mp = Mp; r->ncalls = nc;
        r->dig_pbwt = dp; r->dig_count0 = dc;
    }
    return 0;
}

int eo_process_batch(const eo_panel *p, uint64_t n, const eo_target *ts, int nthreads,
                     uint64_t *S, uint64_t *Mp, uint64_t *dp, uint64_t *dc) {
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        eo_result r;
        if (eo_process_target(p, &ts[i], 0, &r)) { err |= 1; continue; }
        S[i] = r.S; Mp[i] = r.Mp; dp[i] = r.dig_pbwt; dc[i] = r.dig_count0;
    }
    return err ? -1 : 0;
}
