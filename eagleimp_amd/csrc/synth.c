/*
 * synth.c -- deterministic synthetic reference panels and targets.
 *
 * Panel model: F founder haplotypes with per-site minor-allele frequency
 * p ~ Beta(a, a) clipped to [maf_lo, maf_hi] (the minor allele, 0 or 1, is
 * drawn per site); every panel and target haplotype is a mosaic of founders
 * with switch points from a Poisson process along the genetic map and
 * independent allele flips.  Sites are equally spaced in base pairs and the
 * genetic map is linear; cM values are computed as EagleImp's
 * MapInterpolater::interp does for a map without records (sentinel at
 * bp 1e9 with rate 1 cM/Mb and 1000 cM, src/MapInterpolater.cpp:108-130) and
 * stored as float (src/VCFData.cpp:1189).
 *
 * Every random stream is derived from (seed, tag, index) so results do not
 * depend on the number of threads.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SYNTH_API __attribute__((visibility("default")))

enum { TAG_FOUNDER = 1, TAG_PANEL = 2, TAG_TARGET = 3, TAG_MISSING = 4, TAG_TPARAM = 5,
       TAG_ARRIVAL = 6 };

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t sm64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed, uint64_t tag, uint64_t idx) {
    uint64_t x = seed ^ (tag * 0x9E3779B97F4A7C15ull) ^ (idx * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 4; i++) r->s[i] = sm64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(rng_t *r) { /* xoshiro256** (Blackman, Vigna) */
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}
static inline double rng_u01(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
static inline uint64_t rng_below(rng_t *r, uint64_t n) {
    /* unbiased: rejection on the top of the range */
    if (n <= 1) return 0;
    uint64_t lim = UINT64_MAX - UINT64_MAX % n;
    uint64_t x;
    do { x = rng_next(r); } while (x >= lim);
    return x % n;
}
static inline double rng_exp(rng_t *r) { return -log(1.0 - rng_u01(r)); }
static double rng_normal(rng_t *r) { /* Box-Muller, one value */
    double u1 = 1.0 - rng_u01(r), u2 = rng_u01(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
static double rng_gamma(rng_t *r, double a) { /* Marsaglia & Tsang 2000 */
    if (a < 1.0) {
        double g = rng_gamma(r, a + 1.0);
        return g * pow(1.0 - rng_u01(r), 1.0 / a);
    }
    double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (;;) {
        double x = rng_normal(r), v = 1.0 + c * x;
        if (v <= 0) continue;
        v = v * v * v;
        double u = 1.0 - rng_u01(r);
        if (log(u) < 0.5 * x * x + d - d * v + d * log(v)) return d * v;
    }
}
static double rng_beta(rng_t *r, double a, double b) {
    double x = rng_gamma(r, a), y = rng_gamma(r, b);
    return x / (x + y);
}
/* number of trials until the next success with probability p (>= 1) */
static inline uint64_t rng_geom(rng_t *r, double p) {
    if (p <= 0) return UINT64_MAX;
    if (p >= 1) return 1;
    double u = 1.0 - rng_u01(r);
    double k = floor(log(u) / log1p(-p));
    if (k > 1e18) return UINT64_MAX;
    return (uint64_t)k + 1;
}

typedef struct synth_params {
    uint64_t seed;
    uint64_t N;            /* panel haplotypes */
    uint64_t M;            /* sites */
    uint64_t F;            /* founders */
    double beta_a;         /* MAF ~ Beta(a, a) */
    double maf_lo, maf_hi; /* clip */
    int64_t bp_per_site;   /* spacing in bp; site m sits at bp (m+1)*spacing */
    double switch_per_cM;  /* mosaic switch rate */
    double flip_rate;      /* per-allele flip probability */
} synth_params;

SYNTH_API void synth_cm(const synth_params *sp, float *cM, int32_t *bp) {
    for (uint64_t m = 0; m < sp->M; m++) {
        int32_t b = (int32_t)((int64_t)(m + 1) * sp->bp_per_site);
        if (bp) bp[m] = b;
        /* ubGen + 1e-6 * ubRate * (bp - ubBp) with the end sentinel */
        cM[m] = (float)(1000.0 + 1e-6 * 1.0 * (double)(b - 1000000000));
    }
}

static void make_founders(const synth_params *sp, uint64_t *fd, uint64_t fw, int nthreads) {
    uint64_t M = sp->M, F = sp->F;
    uint64_t nwords = (M + 63) / 64;
    memset(fd, 0, F * fw * sizeof(uint64_t));
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t w = 0; w < (int64_t)nwords; w++) {
        for (uint64_t m = (uint64_t)w * 64; m < (uint64_t)w * 64 + 64 && m < M; m++) {
            rng_t r;
            rng_seed(&r, sp->seed, TAG_FOUNDER, m);
            double p = rng_beta(&r, sp->beta_a, sp->beta_a);
            if (p < sp->maf_lo) p = sp->maf_lo;
            if (p > sp->maf_hi) p = sp->maf_hi;
            int minor = (int)(rng_next(&r) >> 63);
            uint64_t bit = 1ull << (m & 63);
            for (uint64_t f = 0; f < F; f++) {
                int a = rng_u01(&r) < p ? minor : !minor;
                if (a) fd[f * fw + (uint64_t)w] |= bit;
            }
        }
    }
}

/* copy bits [a, b) of src into dst (same bit positions) */
static inline void copy_bits(uint64_t *dst, const uint64_t *src, uint64_t a, uint64_t b) {
    while (a < b) {
        uint64_t w = a >> 6, lo = a & 63;
        uint64_t hi = (b - (w << 6)) >= 64 ? 64 : (b - (w << 6));
        uint64_t mask = (hi == 64 ? ~0ull : ((1ull << hi) - 1)) & ~((1ull << lo) - 1);
        dst[w] = (dst[w] & ~mask) | (src[w] & mask);
        a = (w << 6) + hi;
    }
}

static void make_mosaic(const synth_params *sp, const uint64_t *fd, uint64_t fw, const float *cM,
                        uint64_t tag, uint64_t idx, uint64_t *out) {
    rng_t r;
    rng_seed(&r, sp->seed, tag, idx);
    uint64_t M = sp->M;
    uint64_t f = rng_below(&r, sp->F);
    double rate = sp->switch_per_cM > 0 ? sp->switch_per_cM : 1e-300;
    double next = (double)cM[0] + rng_exp(&r) / rate;
    uint64_t start = 0;
    while (start < M) {
        /* first site at or beyond the next switch point */
        uint64_t lo = start, hi = M;
        while (lo < hi) {
            uint64_t mid = (lo + hi) / 2;
            if ((double)cM[mid] >= next) hi = mid; else lo = mid + 1;
        }
        uint64_t end = lo;
        if (end > start) copy_bits(out, fd + f * fw, start, end);
        if (end >= M) break;
        f = rng_below(&r, sp->F);
        next += rng_exp(&r) / rate;
        start = end;
    }
    /* allele flips */
    uint64_t pos = rng_geom(&r, sp->flip_rate) - 1;
    while (pos < M) {
        out[pos >> 6] ^= 1ull << (pos & 63);
        uint64_t g = rng_geom(&r, sp->flip_rate);
        if (g > M) break;
        pos += g;
    }
}

/* 64x64 bit-matrix transpose (Hacker's Delight 7-3, 64-bit form) in that
 * book's MSB-first bit numbering. */
static void transpose64(uint64_t a[64]) {
    uint64_t m = 0x00000000FFFFFFFFull;
    for (int j = 32; j != 0; j >>= 1, m ^= (m << j)) {
        for (int k = 0; k < 64; k = ((k | j) + 1) & ~j) {
            uint64_t t = (a[k] ^ (a[k | j] >> j)) & m;
            a[k] ^= t;
            a[k | j] ^= t << j;
        }
    }
}

/* hap-major (N rows, bit m = site m) -> site-major (M rows, bit h = hap h) */
SYNTH_API int synth_transpose(const uint64_t *hap, uint64_t N, uint64_t hap_stride, uint64_t M,
                              uint64_t *site, uint64_t site_stride, int nthreads) {
    uint64_t nhb = (N + 63) / 64, nsw = (M + 63) / 64;
    (void)nthreads;
    for (uint64_t m = 0; m < M; m++) memset(site + m * site_stride, 0, site_stride * sizeof(uint64_t));
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t hb = 0; hb < (int64_t)nhb; hb++) {
        for (int64_t sw = 0; sw < (int64_t)nsw; sw++) {
            uint64_t a[64];
            /* transpose64 flips about the anti-diagonal for LSB-first bits;
             * reversing the row order on both sides makes it a transpose */
            for (int i = 0; i < 64; i++) {
                uint64_t h = (uint64_t)hb * 64 + (uint64_t)i;
                a[63 - i] = h < N ? hap[h * hap_stride + (uint64_t)sw] : 0;
            }
            transpose64(a);
            for (int i = 0; i < 64; i++) {
                uint64_t m = (uint64_t)sw * 64 + (uint64_t)i;
                if (m < M) site[m * site_stride + (uint64_t)hb] = a[63 - i];
            }
        }
    }
    return 0;
}

/* Panel (N mosaics, stream TAG_PANEL) and 2T held-out target haplotypes
 * (stream TAG_TARGET).  hap / tgt_hap rows have hap_stride words (zeroed
 * here).  site may be NULL. */
SYNTH_API int synth_panel(const synth_params *sp, uint64_t T, uint64_t *hap, uint64_t hap_stride,
                          uint64_t *site, uint64_t site_stride, uint64_t *tgt_hap, float *cM,
                          int32_t *bp, int nthreads) {
    if (!sp || sp->F < 1 || sp->M < 1 || hap_stride < (sp->M + 63) / 64) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    synth_cm(sp, cM, bp);
    uint64_t fw = (sp->M + 63) / 64;
    uint64_t *fd = (uint64_t *)malloc(sp->F * fw * sizeof(uint64_t));
    if (!fd) return -2;
    make_founders(sp, fd, fw, nthreads);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t h = 0; h < (int64_t)sp->N; h++) {
        uint64_t *row = hap + (uint64_t)h * hap_stride;
        memset(row, 0, hap_stride * sizeof(uint64_t));
        make_mosaic(sp, fd, fw, cM, TAG_PANEL, (uint64_t)h, row);
    }
    if (tgt_hap) {
#pragma omp parallel for schedule(dynamic, 8)
        for (int64_t j = 0; j < (int64_t)(2 * T); j++) {
            uint64_t *row = tgt_hap + (uint64_t)j * hap_stride;
            memset(row, 0, hap_stride * sizeof(uint64_t));
            make_mosaic(sp, fd, fw, cM, TAG_TARGET, (uint64_t)j, row);
        }
    }
    free(fd);
    if (site) synth_transpose(hap, sp->N, hap_stride, sp->M, site, site_stride, nthreads);
    return 0;
}

static inline uint64_t get_word(const uint64_t *row, uint64_t bitoff, uint64_t nbits_left) {
    uint64_t w = bitoff >> 6, sh = bitoff & 63;
    uint64_t v = row[w] >> sh;
    if (sh && nbits_left > 64 - sh) v |= row[w + 1] << (64 - sh);
    if (nbits_left < 64) v &= (1ull << nbits_left) - 1;
    return v;
}

/* Target genotypes over the window [s0, s0+W) from two haplotypes, coded as
 * EagleImp's GenotypeVector (is0 / is2 words interleaved, Het = (0,0),
 * Miss = (1,1); src/Datatypes.cpp:27-56).  gt has 2*ceil(W/512)*8 words. */
SYNTH_API int synth_target_gt(const uint64_t *h1, const uint64_t *h2, uint64_t s0, uint64_t W,
                              double miss, uint64_t seed, uint64_t idx, uint64_t *gt) {
    uint64_t units = (W + 511) / 512;
    memset(gt, 0, 2 * units * 8 * sizeof(uint64_t));
    uint64_t nw = (W + 63) / 64;
    for (uint64_t i = 0; i < nw; i++) {
        uint64_t left = W - 64 * i;
        uint64_t a = get_word(h1, s0 + 64 * i, left), b = get_word(h2, s0 + 64 * i, left);
        uint64_t valid = left >= 64 ? ~0ull : ((1ull << left) - 1);
        gt[2 * i] = ~(a | b) & valid;  /* HomRef */
        gt[2 * i + 1] = (a & b) & valid; /* HomAlt */
    }
    rng_t r;
    rng_seed(&r, seed, TAG_MISSING, idx);
    uint64_t pos = rng_geom(&r, miss) - 1;
    while (pos < W) {
        gt[2 * (pos >> 6)] |= 1ull << (pos & 63);
        gt[2 * (pos >> 6) + 1] |= 1ull << (pos & 63);
        uint64_t g = rng_geom(&r, miss);
        if (g > W) break;
        pos += g;
    }
    return 0;
}

/* Per-target window, K and missing rate: W log-uniform on [wmin, wmax]
 * (clipped to M), s0 uniform, K uniform from kset, miss uniform on
 * [miss_lo, miss_hi]. */
SYNTH_API int synth_target_params(uint64_t seed, uint64_t T, uint64_t M, uint64_t wmin, uint64_t wmax,
                                  const uint32_t *kset, uint64_t nk, double miss_lo, double miss_hi,
                                  uint64_t *s0, uint64_t *W, uint32_t *K, double *miss) {
    for (uint64_t t = 0; t < T; t++) {
        rng_t r;
        rng_seed(&r, seed, TAG_TPARAM, t);
        uint64_t w;
        if (wmin >= wmax) w = wmin;
        else {
            double lw = log((double)wmin) + rng_u01(&r) * (log((double)wmax) - log((double)wmin));
            w = (uint64_t)llround(exp(lw));
        }
        if (w < 1) w = 1;
        if (w > M) w = M;
        W[t] = w;
        s0[t] = rng_below(&r, M - w + 1);
        K[t] = nk ? kset[rng_below(&r, nk)] : 0;
        miss[t] = miss_lo + rng_u01(&r) * (miss_hi - miss_lo);
    }
    return 0;
}

/* Serving workload: n request batches arriving as a Poisson process of the
 * given rate (batches/s), sizes log-uniform on [bmin, bmax], each batch a run
 * of consecutive targets of a pool of npool targets starting at first[i]
 * (mod npool).  Stream TAG_ARRIVAL, one substream per request. */
SYNTH_API int synth_arrivals(uint64_t seed, uint64_t n, double rate, uint64_t bmin, uint64_t bmax,
                             uint64_t npool, double *t, uint32_t *size, uint32_t *first) {
    if (rate <= 0 || bmin < 1 || bmax < bmin || npool < 1) return -1;
    double now = 0;
    for (uint64_t i = 0; i < n; i++) {
        rng_t r;
        rng_seed(&r, seed, TAG_ARRIVAL, i);
        now += rng_exp(&r) / rate;
        t[i] = now;
        uint64_t b = bmin;
        if (bmax > bmin) {
            double lb = log((double)bmin) + rng_u01(&r) * (log((double)bmax) - log((double)bmin));
            b = (uint64_t)llround(exp(lb));
        }
        if (b < bmin) b = bmin;
        if (b > bmax) b = bmax;
        size[i] = (uint32_t)b;
        first[i] = (uint32_t)rng_below(&r, npool);
    }
    return 0;
}
