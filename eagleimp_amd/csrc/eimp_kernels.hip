// eimp_kernels.hip -- gfx950 kernels of the phasing pre-stage.
//
// Everything here is integer/bitwise work; wave64 ballots and LDS do the
// per-site and per-step partitions, HBM is streamed with coalesced loads.
//
//   k_site_to_hap    site-major -> hap-major panel (64x64 bit tiles via LDS)
//   k_panel_mono     per-site "all ones" / "any one" bits of the whole panel
//   k_distance       nerr(h) = sum popcount((is0 & h) | (is2 & ~h))
//                    (GenotypeVector::compare_withPreInit, src/Datatypes.cpp:107-116)
//   k_select         exact K smallest (nerr, h), ascending h
//                    (Target::findBestHaps, src/Target.cpp:779-875)
//   k_monos          AND / OR of the chosen rows per window word
//                    (Target::findMonos, src/Target.cpp:878-897)
//   k_splits         split / call sites, one wave per target
//                    (Target::findCallLocs, src/Target.cpp:900-950)
//   k_bits_to_list   bitmap -> ascending index list (drop-in inputs)
//   k_cref           condensed reference in 16-row transposed blocks
//                    (Target::condensedReferenceT, src/Target.cpp:953-1059)
//   k_pbwt           compressed forward / backward PBWT
//                    (PBWT::buildComprPBWT, src/PBWT.cpp:207-285)
#include <map>
#include <mutex>
#include <type_traits>

#include "eimp_internal.h"

namespace eimp {

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Pointers read out of device structs are generic, so accesses through them
// compile to FLAT instructions, which count in the LDS counter as well: the
// next s_waitcnt for an LDS result then also waits for a FLAT store to be
// acknowledged by memory.  Inside loops that mix LDS traffic with global
// stores the pointers are cast to the global address space first.
template <class T>
using gptr_t = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr_t<T> gmem(T *p) {
    return (gptr_t<T>)p;
}
// Arrays that no kernel running at the same time writes, read at wave-uniform
// addresses through the constant address space: scalar loads.
template <class T>
using cptr_t = const __attribute__((address_space(4))) T *;
template <class T>
__device__ __forceinline__ cptr_t<T> cmem(const T *p) {
    return (cptr_t<T>)p;
}
// LDS (address space 3) pointers: 32-bit byte addresses
template <class T>
using lds_ptr_t = __attribute__((address_space(3))) T *;
// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// clang vector types: assignable through global-address-space pointers
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o));
    return v;
}
// inclusive scan over the wave
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    uint32_t l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o);
        if (l >= (uint32_t)o) v += u;
    }
    return v;
}

// bits [bitoff, bitoff+64) of a bit row with `words` readable words
__device__ __forceinline__ uint64_t funnel(const uint64_t *__restrict__ row, uint64_t bitoff,
                                           uint64_t words) {
    uint64_t w = bitoff >> 6;
    uint32_t sh = (uint32_t)(bitoff & 63);
    uint64_t lo = w < words ? row[w] : 0ull;
    if (sh == 0) return lo;
    uint64_t hi = (w + 1) < words ? row[w + 1] : 0ull;
    return (lo >> sh) | (hi << (64 - sh));
}

__device__ __forceinline__ bool excluded_hap(const PanelDev &P, int32_t ignore, uint32_t h) {
    uint32_t nr = h >> 1;
    if ((int32_t)nr == ignore) return true;
    return (h & 1u) && P.haploid && P.haploid[nr];
}

// ---------------------------------------------------------------------------
// panel set-up
// ---------------------------------------------------------------------------
// grid (ceil(hap_stride/4), ceil(N_ad/64)), block 256: one 64-hap x 256-site tile
__global__ __launch_bounds__(256) void k_site_to_hap(PanelDev P, uint64_t *__restrict__ hap) {
    __shared__ uint64_t tile[256];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, l = lane_id();
    const uint64_t hb = blockIdx.y;
    const uint64_t sw0 = (uint64_t)blockIdx.x * 4;
    const uint64_t m = sw0 * 64 + tid;
    tile[tid] = (m < P.M && hb < P.site_stride) ? P.site[m * P.site_stride + hb] : 0ull;
    __syncthreads();
    uint64_t out = 0;
#pragma unroll 8
    for (int j = 0; j < 64; j++) out |= ((tile[wave * 64 + j] >> l) & 1ull) << j;
    const uint64_t h = hb * 64 + l, sw = sw0 + wave;
    if (h < P.N_ad && sw < P.hap_stride) hap[h * P.hap_stride + sw] = out;
}

// grid ceil(M/4), block 256: one wave per site
__global__ __launch_bounds__(256) void k_panel_mono(PanelDev P, uint64_t *__restrict__ pand,
                                                   uint64_t *__restrict__ por) {
    const uint32_t wave = threadIdx.x >> 6, l = lane_id();
    const uint64_t m = (uint64_t)blockIdx.x * 4 + wave;
    if (m >= P.M) return;
    const uint64_t nw = (P.N_ad + 63) / 64;
    uint32_t c = 0;
    for (uint64_t w = l; w < nw; w += 64) c += __popcll(P.site[m * P.site_stride + w]);
    c = wave_sum(c);
    if (l == 0) {
        unsigned long long bit = 1ull << (m & 63);
        if (c == P.N_ad) atomicOr((unsigned long long *)&pand[m >> 6], bit);
        if (c > 0) atomicOr((unsigned long long *)&por[m >> 6], bit);
    }
}

// ---------------------------------------------------------------------------
// a3: distances for groups of DIST_G targets.  grid (ceil(N_ad/16),
// ceil(n/DIST_G)), block 256: 16 haplotypes (4 per wave) x DIST_G targets per
// workgroup.  Over the union of the group's windows, 64 panel words at a time,
// the group's genotype words are staged in LDS shifted to panel alignment
// (sites outside a target's window zero); each lane then reads one word of its
// 4 haplotype rows once for all targets of the group, so a panel word is read
// once per group rather than once per target.  nerr = sum of
// popcount((is0 & h) | (is2 & ~h)) over the window (compare_withPreInit,
// src/Datatypes.cpp:107-116).
// ---------------------------------------------------------------------------
constexpr int DIST_G = 8;

// genotype words (is0, is2) of T for panel word p, bits of sites outside
// [s0, s0 + W) zero
__device__ __forceinline__ void gt_panel_word(const TgtDev &T, uint64_t p, uint64_t &a,
                                              uint64_t &b) {
    const uint64_t nq = ((uint64_t)T.W + 63) / 64;  // window words with data
    const int64_t d = (int64_t)(64 * p) - (int64_t)T.s0;  // window bit of panel bit 0
    auto at = [&](int64_t q, int k) -> uint64_t {
        return (q >= 0 && (uint64_t)q < nq) ? T.gt[2 * q + k] : 0ull;
    };
    if (d >= 0) {
        const int64_t q = d >> 6;
        const uint32_t sh = (uint32_t)(d & 63);
        a = at(q, 0);
        b = at(q, 1);
        if (sh) {
            a = (a >> sh) | (at(q + 1, 0) << (64 - sh));
            b = (b >> sh) | (at(q + 1, 1) << (64 - sh));
        }
    } else if (d > -64) {
        const uint32_t sh = (uint32_t)(-d);
        a = at(0, 0) << sh;
        b = at(0, 1) << sh;
    } else {
        a = b = 0ull;
    }
    // panel bits at or beyond the window end
    const int64_t end = (int64_t)T.s0 + (int64_t)T.W - (int64_t)(64 * p);
    const uint64_t keep = end >= 64 ? ~0ull : (end <= 0 ? 0ull : ((1ull << end) - 1));
    a &= keep;
    b &= keep;
}

__global__ __launch_bounds__(256) void k_distance(PanelDev P, const TgtDev *__restrict__ tg,
                                                 const uint32_t *__restrict__ tlist, uint32_t nlist,
                                                 const uint32_t *__restrict__ d_n) {
    // tlist: targets that select (nlist of them); without it every target of the
    // batch (size *d_n), those that do not select are skipped
    __shared__ uint64_t sg[DIST_G][64][2];
    const uint32_t n = tlist ? nlist : *d_n;
    const uint32_t j0 = blockIdx.y * DIST_G;
    if (j0 >= n) return;
    const TgtDev *Tj[DIST_G];
    bool act[DIST_G];
    uint64_t plo = ~0ull, phi = 0;
#pragma unroll
    for (int j = 0; j < DIST_G; j++) {
        const uint32_t q = j0 + j;
        act[j] = q < n;
        Tj[j] = act[j] ? &tg[tlist ? tlist[q] : q] : tg;
        act[j] = act[j] && (Tj[j]->flags & TF_SELECT);
        if (act[j]) {
            plo = min(plo, Tj[j]->s0 / 64);
            phi = max(phi, (Tj[j]->s0 + Tj[j]->W + 63) / 64);
        }
    }
    if (phi == 0) return;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, l = lane_id();
    const uint32_t hb = blockIdx.x * 16 + wave * 4;
    uint32_t e[4][DIST_G];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < DIST_G; j++) e[i][j] = 0;
    for (uint64_t c0 = plo; c0 < phi; c0 += 64) {
        for (uint32_t q = tid; q < DIST_G * 64; q += 256) {
            const uint32_t j = q / 64, pi = q % 64;
            uint64_t a = 0, b = 0;
#pragma unroll
            for (int jj = 0; jj < DIST_G; jj++)  // static indexing of Tj / act
                if (jj == (int)j && act[jj] && c0 + pi < phi) gt_panel_word(*Tj[jj], c0 + pi, a, b);
            sg[j][pi][0] = a;
            sg[j][pi][1] = b;
        }
        __syncthreads();
        const uint64_t p = c0 + l;
        uint64_t hv[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
            hv[i] = (hb + i < P.N_ad && p < phi) ? P.hap[(uint64_t)(hb + i) * P.hap_stride + p] : 0ull;
#pragma unroll
        for (int j = 0; j < DIST_G; j++) {
            if (!act[j]) continue;
            const uint64_t g0 = sg[j][l][0], g2 = sg[j][l][1];
#pragma unroll
            for (int i = 0; i < 4; i++) e[i][j] += (uint32_t)__popcll((g0 & hv[i]) | (g2 & ~hv[i]));
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < DIST_G; j++) {
        if (!act[j]) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t h = hb + i;
            const uint32_t s = wave_sum(e[i][j]);
            if (l == 0 && h < P.N_ad)
                Tj[j]->nerr[h] = excluded_hap(P, Tj[j]->ignore, h) ? kExcluded : s;
        }
    }
}

// ---------------------------------------------------------------------------
// a4: exact top-K selection / take-all.  grid n, block SEL_NT.  Workgroups of
// 256 threads fit on CUs beside a PBWT workgroup, so a context's selection can
// run while another context's PBWT holds the device.
// ---------------------------------------------------------------------------
constexpr int SEL_NT = 256;
constexpr int SEL_NW = SEL_NT / 64;
constexpr int SEL_BINS = 1024;                    // 10-bit digits
constexpr int SEL_BPT = SEL_BINS / SEL_NT;        // bins per thread in the bin scan

// block-wide exclusive scan of one value per thread; also returns the total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_w, uint32_t &total) {
    const uint32_t wave = threadIdx.x >> 6, l = lane_id();
    uint32_t inc = wave_scan_incl(v);
    if (l == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SEL_NW; w++) {
        uint32_t x = s_w[w];
        tot += x;
        if (w < (int)wave) before += x;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
}

__global__ __launch_bounds__(SEL_NT) void k_select(PanelDev P, const TgtDev *__restrict__ tg,
                                                   const uint32_t *__restrict__ d_n) {
    if (d_n && blockIdx.x >= *d_n) return;
    const TgtDev &T = tg[blockIdx.x];
    __shared__ uint32_t hist[SEL_BINS];
    __shared__ uint32_t s_w[SEL_NW];
    __shared__ uint32_t s_bin, s_before;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, l = lane_id();
    const uint32_t N = P.N_ad;
    const bool select = (T.flags & TF_SELECT) != 0;
    uint32_t thr = 0, need = 0;
    if (select) {
        const uint32_t *__restrict__ nerr = T.nerr;
        uint32_t mx = 0;
        for (uint32_t h = tid; h < N; h += SEL_NT) {
            uint32_t v = nerr[h];
            if (v != kExcluded) mx = max(mx, v);
        }
        mx = wave_max(mx);
        if (l == 0) s_w[wave] = mx;
        __syncthreads();
        mx = 0;
        for (int w = 0; w < SEL_NW; w++) mx = max(mx, s_w[w]);
        __syncthreads();
        int hi = mx ? 32 - __clz(mx) : 0;
        uint32_t k = T.K;  // rank (1-based) of the threshold value
        uint64_t prefix = 0;
        while (hi > 0) {
            const int lo = hi > 10 ? hi - 10 : 0;
            const uint32_t nb = 1u << (hi - lo);
            for (uint32_t b = tid; b < nb; b += SEL_NT) hist[b] = 0;
            __syncthreads();
            for (uint32_t h = tid; h < N; h += SEL_NT) {
                uint32_t v = nerr[h];
                if (v != kExcluded && ((uint64_t)v >> hi) == (prefix >> hi))
                    atomicAdd(&hist[(v >> lo) & (nb - 1)], 1u);
            }
            __syncthreads();
            // thread tid scans bins SEL_BPT*tid .. SEL_BPT*tid + SEL_BPT-1
            uint32_t cb[SEL_BPT], c = 0, total;
#pragma unroll
            for (int q = 0; q < SEL_BPT; q++) {
                const uint32_t bin = SEL_BPT * tid + (uint32_t)q;
                cb[q] = bin < nb ? hist[bin] : 0u;
                c += cb[q];
            }
            uint32_t ex = block_excl_scan(c, s_w, total);
#pragma unroll
            for (int q = 0; q < SEL_BPT; q++) {
                const uint32_t bin = SEL_BPT * tid + (uint32_t)q;
                if (bin < nb && ex < k && k <= ex + cb[q]) {
                    s_bin = bin;
                    s_before = ex;
                }
                ex += cb[q];
            }
            __syncthreads();
            k -= s_before;
            prefix |= (uint64_t)s_bin << lo;
            hi = lo;
            __syncthreads();
        }
        thr = (uint32_t)prefix;
        need = k;
    }
    // compaction in index order: ties at the threshold value go to the lowest indices
    const uint32_t nwords = ((N + 511) / 512) * 8;
    uint32_t ties_before = 0, sel_before = 0;
    for (uint32_t base = 0; base < nwords * 64; base += SEL_NT) {
        const uint32_t h = base + tid;
        bool sel, tie = false;
        if (h >= N) {
            sel = false;
        } else if (select) {
            uint32_t v = T.nerr[h];
            tie = (v == thr);
            sel = v < thr;
        } else {
            sel = !excluded_hap(P, T.ignore, h);
        }
        if (select) {
            uint32_t tcount;
            uint32_t trank = ties_before + block_excl_scan(tie ? 1u : 0u, s_w, tcount);
            if (tie && trank < need) sel = true;
            ties_before += tcount;
        }
        uint32_t scount;
        const uint32_t pos = sel_before + block_excl_scan(sel ? 1u : 0u, s_w, scount);
        // exactly K when Ncorr is consistent; the bitmap always matches bh
        sel = sel && pos < T.K;
        if (sel) T.bh[pos] = h;
        const uint64_t sm = __ballot(sel);
        const uint32_t word = base / 64 + wave;
        if (l == 0 && word < nwords) T.bhflags[word] = sm;
        sel_before += scount;
    }
    if (tid == 0) T.stats[3] = sel_before;
}

// ---------------------------------------------------------------------------
// a5: monos for targets that select a subset.  grid (ceil(Ww/64), n), block
// MONO_NT: the waves split the K chosen rows, lanes cover 64 window words
// (256 threads: they fit beside a PBWT workgroup, as the selection does).
// ---------------------------------------------------------------------------
constexpr int MONO_NT = 256;
constexpr int MONO_NW = MONO_NT / 64;

__global__ __launch_bounds__(MONO_NT) void k_monos(PanelDev P, const TgtDev *__restrict__ tg,
                                                   const uint32_t *__restrict__ tlist,
                                                   const uint32_t *__restrict__ d_n) {
    // tlist: targets that need monos; without it every target of the batch (size *d_n)
    if (!tlist && blockIdx.y >= *d_n) return;
    const TgtDev &T = tg[tlist ? tlist[blockIdx.y] : blockIdx.y];
    if (T.flags & TF_IDENTITY) return;
    __shared__ uint64_t s_and[MONO_NW][64], s_or[MONO_NW][64];
    const uint32_t wave = threadIdx.x >> 6, l = lane_id();
    const uint64_t Ww = (T.W + 63) / 64;
    const uint64_t i = (uint64_t)blockIdx.x * 64 + l;
    const bool valid = i < Ww;
    uint64_t a = ~0ull, o = 0ull;
    const uint64_t bitoff = T.s0 + 64 * i;
    if (valid) {
        uint32_t k = wave;
        for (; k + 3 * MONO_NW < T.K; k += 4 * MONO_NW) {
            uint64_t v0 = funnel(P.hap + (uint64_t)T.bh[k] * P.hap_stride, bitoff, P.hap_stride);
            uint64_t v1 = funnel(P.hap + (uint64_t)T.bh[k + MONO_NW] * P.hap_stride, bitoff,
                                 P.hap_stride);
            uint64_t v2 = funnel(P.hap + (uint64_t)T.bh[k + 2 * MONO_NW] * P.hap_stride, bitoff,
                                 P.hap_stride);
            uint64_t v3 = funnel(P.hap + (uint64_t)T.bh[k + 3 * MONO_NW] * P.hap_stride, bitoff,
                                 P.hap_stride);
            a &= v0 & v1 & v2 & v3;
            o |= v0 | v1 | v2 | v3;
        }
        for (; k < T.K; k += MONO_NW) {
            uint64_t v = funnel(P.hap + (uint64_t)T.bh[k] * P.hap_stride, bitoff, P.hap_stride);
            a &= v;
            o |= v;
        }
    }
    s_and[wave][l] = a;
    s_or[wave][l] = o;
    __syncthreads();
    if (wave == 0 && valid) {
        for (int w = 1; w < MONO_NW; w++) {
            a &= s_and[w][l];
            o |= s_or[w][l];
        }
        T.mand_w[i] = a;
        T.mor_w[i] = o;
    }
}

// ---------------------------------------------------------------------------
// a6: split and call sites.  grid n, block 64 (one wave per target).
// Sequential over 64-site chunks; within a chunk het candidates are always
// splits and hom candidates are resolved lowest-first against the cM of the
// previous split (see the comment in the loop).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_splits(PanelDev P, const TgtDev *__restrict__ tg,
                                              const uint32_t *__restrict__ d_n) {
    if (d_n && blockIdx.x >= *d_n) return;
    const TgtDev &T = tg[blockIdx.x];
    const uint32_t l = lane_id();
    const uint32_t W = T.W;
    const uint32_t Ww = (W + 63) / 64;
    const float *__restrict__ cMw = P.cM + T.s0;
    const float thr = T.cMmaxSplit;
    const uint64_t lt = lanemask_lt();
    const uint64_t mono_words = T.mand == T.mand_w ? (uint64_t)((W + 511) / 512) * 8
                                                   : (uint64_t)((P.M + 511) / 512) * 8;
    float Lprev = cMw[0];
    bool started = false;
    uint32_t S = 0, ncalls = 0;
    int64_t lastcall = -1;      // index (in splitsites) of the last het split
    float lastcall_cm = 0.f;
    uint32_t lastcall_word = 0;
    uint64_t lastcall_mask = 0;
    for (uint32_t c = 0; c < Ww; c++) {
        const uint32_t m = c * 64 + l;
        const bool valid = m < W;
        const uint64_t g0 = T.gt[2 * c], g2 = T.gt[2 * c + 1];
        const bool is0 = (g0 >> l) & 1ull, is2 = (g2 >> l) & 1ull;
        const bool het = valid && !is0 && !is2;
        const bool miss = is0 && is2;
        const uint64_t hetm = __ballot(het);
        bool elig = true;
        if (!started) {
            if (hetm == 0ull) continue;  // the phasing starts at the first het (:907-913)
            started = true;
            elig = l >= (uint32_t)(__ffsll((unsigned long long)hetm) - 1);
        }
        const uint64_t ma = funnel(T.mand, T.mono_off + 64ull * c, mono_words);
        const uint64_t mo = funnel(T.mor, T.mono_off + 64ull * c, mono_words);
        const bool mono = ((ma >> l) & 1ull) || !((mo >> l) & 1ull);
        const bool cand = valid && elig && !mono && !miss;
        const uint64_t hetc = __ballot(cand && het);
        const uint64_t homc = __ballot(cand && !het);
        const float cm = valid ? cMw[m] : 0.f;
        uint64_t splitm = hetc, pending = homc;
        // Resolve hom candidates.  All splits fixed so far are true splits and
        // L (cM of the previous split) only grows as splits are added, so a
        // pending candidate failing now fails for good, and the lowest one
        // passing now has exactly the true L: accept it and repeat.
        while (pending) {
            const uint64_t below = splitm & lt;
            const int src = below ? 63 - __clzll((long long)below) : (int)l;
            const float Ls = __shfl(cm, src);
            const float L = below ? Ls : Lprev;
            const float d = cm - L;
            const bool pass = ((pending >> l) & 1ull) && (d > thr);
            const uint64_t passm = __ballot(pass);
            pending &= passm;
            if (!passm) break;
            const uint64_t p = passm & (~passm + 1);  // lowest passing lane
            splitm |= p;
            pending &= ~p;
        }
        if (splitm) {
            // emit split sites and their cM differences (src/Target.cpp:943-947)
            const uint64_t below = splitm & lt;
            const int src = below ? 63 - __clzll((long long)below) : (int)l;
            const float Ls = __shfl(cm, src);
            const float L = below ? Ls : Lprev;
            if ((splitm >> l) & 1ull) {
                const uint32_t idx = S + (uint32_t)__popcll(below);
                T.splitsites[idx] = m;
                T.cMdiffs[idx] = cm - L;
            }
            if (hetc) {
                const int top = 63 - __clzll((long long)hetc);
                const uint64_t upto = top == 63 ? ~0ull : ((1ull << (top + 1)) - 1);
                lastcall = (int64_t)S + __popcll(splitm & upto) - 1;
                lastcall_cm = __shfl(cm, top);
                lastcall_word = c;
                lastcall_mask = splitm & upto;
                ncalls += (uint32_t)__popcll(hetc);
            }
            S += (uint32_t)__popcll(splitm);
            Lprev = __shfl(cm, 63 - __clzll((long long)splitm));
            if (l == 0) T.splits[c] = splitm;
        } else if (l == 0) {
            T.splits[c] = 0ull;
        }
    }
    // no split, or no call: S = 0 (:927-935); else truncate after the last call (:939-940)
    uint32_t Sfin = (S == 0 || ncalls == 0) ? 0u : (uint32_t)(lastcall + 1);
    const uint32_t nwords = ((W + 511) / 512) * 8;
    const uint32_t keep_to = Sfin ? lastcall_word : 0u;
    // words below keep_to hold the loop's values (k_splits_clear zeroed the
    // words the loop skipped before the first het)
    for (uint32_t w = l; w < nwords; w += 64) {
        if (Sfin == 0 || w > keep_to) T.splits[w] = 0ull;
        else if (w == keep_to) T.splits[w] = lastcall_mask;
    }
    if (l == 0) {
        if (Sfin) T.cMdiffs[Sfin] = cMw[W - 1] - lastcall_cm;
        T.stats[0] = Sfin;
        T.stats[1] = Sfin ? ncalls : 0u;
    }
}

// Words skipped before the first het must be zero in the split bitmap: the
// loop `continue`s there without writing.  This pass clears them (grid n,
// block 64) -- it runs before k_splits.
__global__ __launch_bounds__(64) void k_splits_clear(const TgtDev *__restrict__ tg,
                                                    const uint32_t *__restrict__ d_n) {
    if (d_n && blockIdx.x >= *d_n) return;
    const TgtDev &T = tg[blockIdx.x];
    const uint32_t nwords = ((T.W + 511) / 512) * 8;
    for (uint32_t w = lane_id(); w < nwords; w += 64) T.splits[w] = 0ull;
}

// ---------------------------------------------------------------------------
// bitmap -> ascending index list (drop-in path: best-haps flags, split
// bitmap).  grid n, block SEL_NT.  which: 0 best haps, 1 split sites.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SEL_NT) void k_bits_to_list(const TgtDev *__restrict__ tg,
                                                        uint32_t nbits_bh, int which) {
    const TgtDev &T = tg[blockIdx.x];
    __shared__ uint32_t s_w[SEL_NW];
    const uint32_t tid = threadIdx.x;
    const uint64_t *bits = which == 0 ? T.bhflags : T.splits;
    uint32_t *out = which == 0 ? T.bh : T.splitsites;
    const uint32_t nbits = which == 0 ? nbits_bh : T.W;
    const uint32_t cap = which == 0 ? T.K : T.W;  // list capacity; the host rejects a mismatch
    uint32_t before = 0;
    for (uint32_t base = 0; base < nbits; base += SEL_NT) {
        const uint32_t i = base + tid;
        const bool set = i < nbits && ((bits[i >> 6] >> (i & 63)) & 1ull);
        uint32_t total;
        uint32_t pos = before + block_excl_scan(set ? 1u : 0u, s_w, total);
        if (set && pos < cap) out[pos] = i;
        before += total;
    }
    if (tid == 0 && which == 1) T.stats[0] = before;
    if (tid == 0 && which == 0) T.stats[3] = before;
}

// ---------------------------------------------------------------------------
// a7: condensed reference, in 16-row blocks (segments 8b..8b+7): row 2s is the
// OR of the inconsistencies of segment s (mergeInconsistencies,
// src/Target.cpp:1040-1059), row 2s+1 the haplotypes at split s (:986-1003);
// one uint16 per haplotype holds a block's 16 rows.  Two kernels:
//  - whole-panel targets (all N haplotypes in panel order): one work item =
//    one block, its rows built in LDS from the site-major panel (every thread
//    streaming two words through all the block's site rows, 8 independent
//    loads in flight) and transposed in registers;
//  - the others: one work item = CREF_G blocks, from the haplotype-major
//    panel, reading only the K chosen rows.
// count0 (zeros per row) is written by the forward PBWT pass.
// ---------------------------------------------------------------------------
constexpr int CREF_NT = 256;
constexpr uint32_t CREF_G = kCrefGroup;  // blocks per work item of the subset kernel
constexpr uint32_t CREF_TW = 512;  // words of the haplotype axis per tile (16 rows: 64 KB)

__host__ __device__ inline uint32_t cref_tile_words(uint32_t nw) {
    return nw < CREF_TW ? nw : CREF_TW;
}

// The rows of block b over words [w0, w0 + tw) into srow[row * tw + (w - w0)]
// (row 2(s-8b)+1 only for s < S).
__device__ __forceinline__ void cref_fill(const PanelDev &P, const TgtDev &T, uint32_t b,
                                          uint32_t w0, uint32_t tw, uint32_t nw, uint64_t *srow) {
    const uint64_t *__restrict__ site = P.site + T.s0 * P.site_stride;
    const uint64_t ss = P.site_stride;
    const uint32_t S = T.S, W = T.W;
    const gptr_t<const uint32_t> splitsites = gmem((const uint32_t *)T.splitsites);
    const gptr_t<const uint64_t> splits = gmem((const uint64_t *)T.splits);
    const gptr_t<const uint64_t> gt = gmem((const uint64_t *)T.gt);
    const uint32_t sa = 8 * b, sb = min(8 * b + 8, S + 1);
    const uint64_t lastmask = (P.N_ad & 63) ? ((1ull << (P.N_ad & 63)) - 1) : ~0ull;
    // sites of the block: after split sa-1 up to split sb-1 (or the window end)
    const uint32_t m0 = sa == 0 ? 0u : splitsites[sa - 1] + 1;
    const uint32_t m1 = sb <= S ? splitsites[sb - 1] + 1 : W;
    for (uint32_t wi = threadIdx.x; wi < tw; wi += 2 * CREF_NT) {
        const uint32_t wj = wi + CREF_NT;  // this thread's second word
        const bool two = wj < tw;
        const uint32_t w = w0 + wi, w2 = w0 + (two ? wj : wi);
        const uint64_t mask0 = (w == nw - 1) ? lastmask : ~0ull;
        const uint64_t mask1 = (w2 == nw - 1) ? lastmask : ~0ull;
        uint32_t s = sa;
        uint64_t a0 = 0, a1 = 0;
        // one pass over the sites; per group of 4 sites the site-row words, the
        // genotype bits and the split bits are independent loads
        for (uint32_t mb = m0; mb < m1; mb += 4) {
            uint64_t v0[4], v1[4];
            uint32_t sp[4], is0[4], is2[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t m = mb + u;
                const bool in = m < m1;
                const uint32_t mw = in ? m : m0;
                v0[u] = in ? site[(uint64_t)m * ss + w] : 0ull;
                v1[u] = in ? site[(uint64_t)m * ss + w2] : 0ull;
                sp[u] = (uint32_t)(splits[mw >> 6] >> (mw & 63)) & 1u;
                is0[u] = (uint32_t)(gt[2 * (mw >> 6)] >> (mw & 63)) & 1u;
                is2[u] = (uint32_t)(gt[2 * (mw >> 6) + 1] >> (mw & 63)) & 1u;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t m = mb + u;
                if (m >= m1) break;
                if (sp[u]) {  // split: close segment s, its haplotypes are row 2s+1
                    const uint32_t r0 = 2 * (s - sa);
                    srow[(uint64_t)r0 * tw + wi] = a0 & mask0;
                    srow[(uint64_t)(r0 + 1) * tw + wi] = v0[u];
                    if (two) {
                        srow[(uint64_t)r0 * tw + wj] = a1 & mask1;
                        srow[(uint64_t)(r0 + 1) * tw + wj] = v1[u];
                    }
                    a0 = a1 = 0;
                    s++;
                } else {
                    // HomRef: the ones are inconsistent; HomAlt: the zeros; het/missing: none
                    const uint64_t flip = is2[u] ? ~0ull : 0ull;
                    const uint64_t keep = (is0[u] ^ is2[u]) ? ~0ull : 0ull;
                    a0 |= (v0[u] ^ flip) & keep;
                    a1 |= (v1[u] ^ flip) & keep;
                }
            }
        }
        if (sb == S + 1) {  // the last segment has no split row
            const uint32_t r0 = 2 * (S - sa);
            srow[(uint64_t)r0 * tw + wi] = a0 & mask0;
            if (two) srow[(uint64_t)r0 * tw + wj] = a1 & mask1;
        }
    }
}

// Blocks [b0, b1) of a subset target from the haplotype-major panel.  Thread
// x takes the chosen haplotypes k = x, x + CREF_NT, ... and streams each one's
// row over the sites of the blocks, 32 window sites per step:
//   inc = (v & HomRef) | (~v & HomAlt)  (HomRef = is0 & ~is2, HomAlt =
//                                        is2 & ~is0; split sites and sites
//                                        outside the range masked out)
//   C   = carries of ~splits + inc + cin
// A split kills the carry chain, any other site propagates it and an
// inconsistency generates it, so bit p of C says whether the segment ending
// at split p has an inconsistency (row 2s, mergeInconsistencies,
// src/Target.cpp:1040-1059) and bit p of v is row 2s+1 (src/Target.cpp:986-1003);
// the carry out of bit 31 goes into the next word, and after the last word it
// is the final segment's row 2S.  The split positions are the same for every
// haplotype: a wave-uniform loop over them packs two bits per split into the
// block's uint16.  Only the K chosen rows are read (the site-major layout
// would need all N bits of every site).  Haplotypes K..Kpad-1 get all-ones
// values (the PBWT kernel's padding); count0 comes from the PBWT kernel.
__device__ __forceinline__ void cref_range_hap(const PanelDev &P, const TgtDev &T, uint32_t b0,
                                               uint32_t b1) {
    const uint32_t tid = threadIdx.x;
    const uint32_t K = T.K, S = T.S, Kpad = T.Kpad, W = T.W;
    const bool ident = (T.flags & TF_IDENTITY) != 0;
    // the target's arrays, written by earlier kernels and read-only here
    const cptr_t<uint32_t> splitsites = cmem((const uint32_t *)T.splitsites);
    const cptr_t<uint32_t> sp32 = cmem((const uint32_t *)T.splits);
    const cptr_t<uint32_t> gt32 = cmem((const uint32_t *)T.gt);
    const gptr_t<const uint32_t> bh = gmem((const uint32_t *)T.bh);
    const gptr_t<uint16_t> dst = gmem(T.cref);
    const uint32_t sa = 8 * b0, sb = min(8 * b1, S + 1);
    const bool last = sb == S + 1;  // the range ends with the final segment (no split row)
    // window sites [m0, m1): after split sa-1, up to split sb-1 or the window end
    const uint32_t m0 = sa == 0 ? 0u : splitsites[sa - 1] + 1;
    const uint32_t m1 = last ? W : splitsites[sb - 1] + 1;
    const uint32_t i0 = m0 >> 5, i1 = (m1 + 31) >> 5;  // 32-site window words
    for (uint32_t k = K + tid; k < Kpad; k += CREF_NT)
        for (uint32_t b = b0; b < b1; b++) dst[(uint64_t)b * Kpad + k] = 0xFFFFu;
    // window word i of a row = bits sh.. of row words q0 + i, q0 + i + 1
    const uint64_t rowlen = 2 * P.hap_stride;  // uint32 words per haplotype row
    const uint64_t q0 = T.s0 >> 5;
    const uint32_t sh = (uint32_t)T.s0 & 31;
    for (uint32_t kb = 0; kb < K; kb += CREF_NT) {
        const uint32_t k = kb + tid;
        const bool act = k < K;
        const uint32_t h = act ? (ident ? k : bh[k]) : 0u;
        const gptr_t<const uint32_t> row = gmem((const uint32_t *)P.hap) + (uint64_t)h * rowlen;
        uint32_t cin = 0, acc = 0, j2 = 0, b = b0;  // j2: bit position of the next pair
        // The row is read in chunks of 16 words (64 bytes) aligned in the row
        // (rows start on 64-byte boundaries), each as four 16-byte loads, two
        // chunks ahead of the words being processed: a lane asks for both
        // halves of a 128-byte line close together, so fewer lines are
        // fetched again after the cache has dropped them.  Window word i is
        // bits sh.. of row words q0 + i and q0 + i + 1.
        auto chunk = [&](uint64_t R, uint32_t (&c)[16]) {
            if (R + 16 <= rowlen) {
                const gptr_t<const u32x4> r4 = (gptr_t<const u32x4>)(row + R);
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const u32x4 x = r4[u];
                    c[4 * u] = x[0];
                    c[4 * u + 1] = x[1];
                    c[4 * u + 2] = x[2];
                    c[4 * u + 3] = x[3];
                }
            } else {
#pragma unroll
                for (int u = 0; u < 16; u++) c[u] = R + u < rowlen ? row[R + u] : 0u;
            }
        };
        const uint64_t R0 = (q0 + i0) & ~15ull;  // the first chunk
        uint32_t cur[16], nx[16], nx2[16];
        chunk(R0, nx);
        chunk(R0 + 16, nx2);
        for (uint64_t R = R0; R < q0 + i1; R += 16) {
#pragma unroll
            for (int u = 0; u < 16; u++) {
                cur[u] = nx[u];
                nx[u] = nx2[u];
            }
            chunk(R + 32, nx2);
            uint32_t w[17];
#pragma unroll
            for (int u = 0; u < 16; u++) w[u] = cur[u];
            w[16] = nx[0];
            // window word of the chunk's first word (below i0 for the first
            // chunk: those words are skipped)
            const int64_t ib = (int64_t)R - (int64_t)q0;
#pragma unroll
            for (int u = 0; u < 16; u++) {
                if (ib + u < (int64_t)i0) continue;
                const uint32_t i = (uint32_t)(ib + u);
                if (i >= i1) break;
                const uint32_t v = __builtin_amdgcn_alignbit(w[u + 1], w[u], sh);
                uint32_t rng = ~0u;
                if (i == i0) rng &= ~0u << (m0 & 31);
                if (i == i1 - 1 && (m1 & 31)) rng &= (1u << (m1 & 31)) - 1;
                // the target's words are the same for every lane; a missing
                // genotype sets both is0 and is2 and is no inconsistency
                const uint32_t spw = __builtin_amdgcn_readfirstlane(sp32[i] & rng);
                const uint32_t g0 = gt32[4 * (i >> 1) + (i & 1)], g2 = gt32[4 * (i >> 1) + 2 + (i & 1)];
                const uint32_t hr = __builtin_amdgcn_readfirstlane(g0 & ~g2 & ~spw & rng);
                const uint32_t ha = __builtin_amdgcn_readfirstlane(g2 & ~g0 & ~spw & rng);
                const uint32_t inc = (v & hr) | (~v & ha);
                const uint32_t pr = ~spw;
                const uint64_t sum = (uint64_t)pr + inc + cin;
                const uint32_t C = (uint32_t)sum ^ pr ^ inc;
                cin = (uint32_t)(sum >> 32);
                for (uint32_t sl = spw; sl; sl &= sl - 1) {
                    const uint32_t p = (uint32_t)__builtin_ctz(sl);
                    // the split's two rows: inconsistency (bit 0), allele (bit 1)
                    const uint32_t pair = __builtin_amdgcn_ubfe(C, p, 1u) |
                                          (__builtin_amdgcn_ubfe(v, p, 1u) << 1);
                    acc |= pair << j2;
                    j2 += 2;
                    if (j2 == 16) {
                        if (act) dst[(uint64_t)b * Kpad + k] = (uint16_t)acc;
                        acc = 0;
                        j2 = 0;
                        b++;
                    }
                }
            }
        }
        if (last) {  // row 2S: an inconsistency after the last split
            acc |= cin << j2;
            if (act) dst[(uint64_t)b * Kpad + k] = (uint16_t)acc;
        }
    }
}

// In-register transpose of two 16x16 bit blocks side by side: on entry bit c
// of a[r] is (row r, column c) for c < 16 and (row r, column 16 + c') at bit
// 16 + c'; on exit bit r of a[c] is (row r, column c) and bit 16 + r is
// (row r, column 16 + c).  Four rounds of masked block swaps (bits of one
// 16-bit half never move into the other: the masks repeat per half).
__device__ __forceinline__ void transpose16x2(uint32_t (&a)[16]) {
    static_for<0, 4>([&](auto ji) {
        constexpr int j = 8 >> decltype(ji)::value;
        constexpr uint32_t m = j == 8 ? 0x00FF00FFu : j == 4 ? 0x0F0F0F0Fu : j == 2 ? 0x33333333u
                                                                                     : 0x55555555u;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (k & j) continue;
            const uint32_t t = ((a[k] >> j) ^ a[k + j]) & m;
            a[k + j] ^= t;
            a[k] ^= t << j;
        }
    });
}

// Block b of a target taking all N haplotypes in panel order (TF_IDENTITY):
// no gather.  Each thread owns GPT groups of 64 haplotypes: it reads their
// word of the 16 staged rows, transposes the low and the high halves as two
// pairs of 16x16 bit blocks, and writes the 64 uint16 values as 128
// contiguous bytes.
template <int GPT>
__device__ __forceinline__ void cref_block_ident(const PanelDev &P, const TgtDev &T, uint32_t b,
                                                 uint64_t *srow) {
    const uint32_t tid = threadIdx.x;
    const uint32_t K = T.K, S = T.S, Kpad = T.Kpad;
    const uint32_t nw = (P.N_ad + 63) / 64;  // == Kpad / 64 (K == N_ad), at most CREF_TW
    const uint32_t sb = min(8 * b + 8, S + 1);
    cref_fill(P, T, b, 0, nw, nw, srow);
    __syncthreads();
    const uint32_t nrows = 2 * (sb - 8 * b) - (sb == S + 1 ? 1u : 0u);
    uint32_t V[GPT][32];  // two uint16 values per register: haplotypes 2i, 2i+1
#pragma unroll
    for (int e = 0; e < GPT; e++) {
        const uint32_t q = tid + CREF_NT * e;
        uint32_t lo[16], hi[16];
#pragma unroll
        for (int rr = 0; rr < 16; rr++) {
            const uint64_t wv = (q < nw && (uint32_t)rr < nrows) ? srow[(uint64_t)rr * nw + q] : 0ull;
            lo[rr] = (uint32_t)wv;
            hi[rr] = (uint32_t)(wv >> 32);
        }
        transpose16x2(lo);
        transpose16x2(hi);
        // lo[c]: haplotype c in the low half, 16 + c in the high half
#pragma unroll
        for (int i = 0; i < 8; i++) {
            V[e][i] = __builtin_amdgcn_perm(lo[2 * i + 1], lo[2 * i], 0x05040100u);
            V[e][8 + i] = __builtin_amdgcn_perm(lo[2 * i + 1], lo[2 * i], 0x07060302u);
            V[e][16 + i] = __builtin_amdgcn_perm(hi[2 * i + 1], hi[2 * i], 0x05040100u);
            V[e][24 + i] = __builtin_amdgcn_perm(hi[2 * i + 1], hi[2 * i], 0x07060302u);
        }
    }
    __syncthreads();  // the staged rows are free for the next block
    const gptr_t<uint16_t> dst = gmem(T.cref) + (uint64_t)b * Kpad;
#pragma unroll
    for (int e = 0; e < GPT; e++) {
        const uint32_t q = tid + CREF_NT * e;
        if (q < nw) {
            if (64 * q + 64 > K) {  // last group: haplotypes >= K are padding
#pragma unroll
                for (int i = 0; i < 32; i++) {
                    const uint32_t k0 = 64 * q + 2 * i;
                    if (k0 >= K) V[e][i] |= 0xFFFFu;
                    if (k0 + 1 >= K) V[e][i] |= 0xFFFF0000u;
                }
            }
            const gptr_t<u32x4> d4 = (gptr_t<u32x4>)(dst + 64 * q);
#pragma unroll
            for (int i = 0; i < 8; i++)
                d4[i] = u32x4{V[e][4 * i], V[e][4 * i + 1], V[e][4 * i + 2], V[e][4 * i + 3]};
        }
    }
}

__global__ __launch_bounds__(CREF_NT) void k_cref(PanelDev P, const TgtDev *__restrict__ tg,
                                                 const CrefItem *__restrict__ items) {
    const CrefItem it = items[blockIdx.x];
    cref_range_hap(P, tg[it.t], it.b0, it.b1);
}

template <int GPT>
__global__ __launch_bounds__(CREF_NT) void k_cref_ident(PanelDev P, const TgtDev *__restrict__ tg,
                                                       const CrefItem *__restrict__ items) {
    extern __shared__ uint64_t srow[];
    const CrefItem it = items[blockIdx.x];
    for (uint32_t b = it.b0; b < it.b1; b++) cref_block_ident<GPT>(P, tg[it.t], b, srow);
}

// ---------------------------------------------------------------------------
// a9: compressed PBWT.  One workgroup per (target, direction).
//
// The K positions of the current order are cut into chunks of 64; wave w
// holds chunks [w*NS, w*NS+NS) (NS slots, chosen per target so that all waves
// are busy), one position per lane, as a 32-bit payload: hap index << 16 |
// the 16 bits of the current cref block for that haplotype.  Positions
// K..Kpad-1 (and whole padding chunks) carry all-ones row bits: under the
// stable partition (zeros first) they never move, and they give the
// 1-padding of the last word.  A step:
//   pass 1  ballot of the row bit per slot: the slot's two data words go to
//           output lane c; the wave's zero count (scalar) to LDS
//   (barrier) zeros before the wave's first chunk and in the row (count0)
//           from the NW wave counts, in scalar registers
//   pass 2  per slot: the ones before the chunk are a scalar running sum, so
//           mbcnt of the ballot plus that base gives a one's destination and
//           a lane constant minus it a zero's; payloads scattered into LDS
//           (two contiguous runs per chunk); output lanes store their data /
//           cumulative-zero pairs
//   (barrier)
//   pass 3  read back the payloads of my positions.
// A row whose K bits are all equal keeps the order: pass 2, the second
// barrier and pass 3 are skipped (the wave counts alternate between two LDS
// halves, so the next step's counts never overwrite ones still being read).
// At each 16-row block boundary the next block's uint16 values are staged in
// LDS and gathered by hap index.
// ---------------------------------------------------------------------------
// v_writelane_b32 through the LLVM intrinsic (no clang builtin in this
// toolchain); the compiler sees it, so it inserts the wait states a VALU-written
// SGPR operand needs.
extern "C" __device__ uint32_t eimp_llvm_writelane(uint32_t val, uint32_t lane, uint32_t old)
    __asm("llvm.amdgcn.writelane");

// writes a wave-uniform value into lane L of a VGPR
template <int L>
__device__ __forceinline__ uint32_t writelane(uint32_t old, uint32_t val) {
    return eimp_llvm_writelane(val, (uint32_t)L, old);
}


// Exclusive prefix sum of v over lanes 0..N-1 (other lanes: garbage), by DPP
// row shifts and row broadcasts (no LDS).
template <int N>
__device__ __forceinline__ uint32_t lane_scan_excl(uint32_t v) {
    uint32_t x = v;
    // within rows of 16 lanes: lane i += lane i-1, i-2, i-4, i-8 (zero past the row start)
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);
    if constexpr (N > 2) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);
    if constexpr (N > 4) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);
    if constexpr (N > 8) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);
    // across rows: lane 15 of each row into rows 1 and 3, then lane 31 into rows 2 and 3
    if constexpr (N > 16) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
    if constexpr (N > 32) x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
    return x - v;
}

// LDS bytes before the PBWT payloads: 32 (NS - 1) words, the largest amount a
// slot's store address sits below its destination (pbwt_steps, pass 2).
__host__ __device__ constexpr uint32_t pbwt_pay_bias(int ns) { return 128u * (uint32_t)(ns - 1); }

// NS = slots per wave for this target.  Wave w owns chunks [w*NS, w*NS+NS);
// chunks >= Kpad/64 are padding (all-ones bits, never moving), so every slot
// is processed unconditionally.
// ABS (panel PBWT for imputation): also writes the order after every step to
// absperm (row step+1; row 0 = identity) and the step's count0 to c0out.
// WS: how each wave sums the NW wave zero counts after the first barrier --
// 0: NW LDS broadcasts read into scalars (short latency; best when the
// workgroups sharing a CU are at different points, as with targets of mixed
// K and windows), 1: one LDS read per lane and a DPP lane scan (fewer
// instructions and scalar registers; best for uniform targets).  The
// autotuner picks between the variants built with each.
//
// The kernels that call this use dynamic LDS only (the host checks that they
// have no static LDS), so LDS byte addresses are plain offsets from 0: pay[]
// starts at byte pbwt_pay_bias(NS), which keeps pass 2's store addresses
// (before their immediate offsets) non-negative.
template <int NT, int NS, bool ABS = false, int WS = 0>
__device__ __forceinline__ void pbwt_steps(const TgtDev &T, const bool rev, uint32_t *smem,
                                           int32_t *__restrict__ absperm = nullptr,
                                           int32_t *__restrict__ c0out = nullptr) {
    constexpr int NW = NT / 64;
    constexpr uint32_t NPOS = NS * NW * 64;
    static_assert(NS <= 64, "one output lane per slot");
    const uint32_t K = T.K, Mp = T.Mp, Kpad = T.Kpad, RS = T.RS;
    const uint32_t nch = Kpad / 64;
    const uint32_t npairs = (K + 31) / 32;
    const uint32_t KW = 2 * npairs;
    // NPOS payloads at LDS bytes [PBIAS, PBIAS + 4*NPOS): the partition's
    // stores address them as 4 sel + 128 c (pass 2), sel >= 0
    constexpr uint32_t PBIAS = pbwt_pay_bias(NS);
    uint32_t *pay = smem + PBIAS / 4;
    // Kpad + 2 uint16: the next cref block by haplotype (and all ones at id
    // Kpad), staged at block boundaries (when no payloads are in flight) from
    // LDS byte 0
    uint16_t *stage = (uint16_t *)smem;
    // 2 x NW: zeros of each wave's chunks, by step parity (a step that moves
    // nothing has no second barrier, so the next step writes the other half)
    uint32_t *wzero2 = pay + NPOS;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = lane_id();
    const uint32_t ch0 = wave * NS;                // first chunk of this wave
    // pass 2's per-lane constant (bytes): 4 (64 ch0 + ln + 64 (NS - 1))
    const uint32_t vbase4 = 4u * (ch0 * 64u + ln + 64u * (NS - 1));
    const gptr_t<const uint16_t> cref = gmem((const uint16_t *)T.cref);
    // wave-uniform flags as scalars: their branches stay scalar branches
    const bool dig = __builtin_amdgcn_readfirstlane(T.digest != nullptr ? 1u : 0u) != 0;

    uint32_t p[NS];
    {
        const uint32_t b0 = (rev ? Mp - 1 : 0) >> 4;
#pragma unroll
        for (int c = 0; c < NS; c++) {
            const uint32_t pos = (ch0 + (uint32_t)c) * 64 + ln;
            // the cref kernel writes all-ones for K <= k < Kpad; positions of
            // padding chunks (>= Kpad) all carry id Kpad, whose staged value
            // is all ones
            const uint32_t v = pos < Kpad ? (uint32_t)cref[(uint64_t)b0 * Kpad + pos] : 0xFFFFu;
            p[c] = ((pos < Kpad ? pos : Kpad) << 16) | v;
        }
    }
    if constexpr (ABS) {
        if (absperm) {
#pragma unroll
            for (int c = 0; c < NS; c++) {
                const uint32_t pos = (ch0 + (uint32_t)c) * 64 + ln;
                if (pos < K) absperm[pos] = (int32_t)pos;
            }
        }
    }
    const gptr_t<uint32_t> outbase = gmem(T.out) + (rev ? (uint64_t)Mp * RS : 0ull);
    const uint64_t gbase = rev ? (uint64_t)Mp * KW : 0ull;
    uint64_t dsum = 0;
    const bool raw = (__builtin_amdgcn_readfirstlane(T.flags) & TF_RAW) != 0;
    const uint32_t och = ch0 + ln;         // chunk stored by this lane
    const bool ostore = ln < (uint32_t)NS && och < nch;
    // the raw layout has one word pair per chunk, the others two
    const bool second = !raw && 2 * och + 1 < npairs;
    // digest of this lane's words: g = gbase + step*KW + 4*och (+0..3), key
    // g * kDigKey, rotation -g (alignbit uses 5 bits)
    const uint32_t g0 = (uint32_t)gbase + 4u * och;

    {
        // The padding words of every row of this direction, stored once: after
        // the logical pairs (words >= 2*npairs) zero, or in the raw layout
        // all ones after the data words (positions >= K are all ones).
        const uint32_t pw0 = raw ? 2 * nch : 2 * npairs;  // even, as RS
        const uint32_t fw = raw ? ~0u : 0u;
        for (uint32_t rr = threadIdx.x; rr < Mp; rr += NT) {
            const gptr_t<u32x2> prow = (gptr_t<u32x2>)(outbase + (uint64_t)rr * RS);
            for (uint32_t w = pw0 / 2; w < RS / 2; w++) prow[w] = u32x2{fw, fw};
        }
    }
    // this lane's byte offset of its chunk's words within a row (opaque
    // 32-bit value added to the scalar row address: stores take the
    // scalar-base form, no 64-bit vector address per step)
    uint32_t obyte;
    asm("" : "=v"(obyte) : "0"((raw ? 8u : 16u) * och));
    const uint32_t rs_bytes = RS * 4u;
    uint64_t rowaddr = (uint64_t)(size_t)outbase;
    // row bit of the current row within its 16-row block: forward rows go up
    // from bit 0, backward rows down from bit 15 (shifts as data, so the step
    // has no direction branch)
    uint32_t jbit = 1u << ((rev ? Mp - 1 : 0) & 15);
    const uint32_t jl = rev ? 0u : 1u, jr = rev ? 1u : 0u;
    // first step whose row begins a new 16-row block
    uint32_t nextb = rev ? ((Mp - 1) & 15) + 1 : 16u;
    // forward pass: count0 of the current 16-row block, lane (step & 15)
    uint32_t c0acc = 0;
    // PF: the next block boundary's cref block (Kpad uint16 = Kpad/2 words,
    // at most NPF per thread since Kpad <= NS * NT) is loaded into registers
    // one block ahead, so staging it waits for no global load.  Only the
    // 1024-thread shapes (one workgroup per CU, 128 VGPRs a wave) have the
    // registers; the others load at the boundary.
    constexpr bool PF = NT >= 1024;
    constexpr int NPF = PF ? (NS + 1) / 2 : 1;
    uint32_t pf[NPF];
    auto block_src = [&](uint32_t st) {
        const uint32_t r = rev ? Mp - 1 - st : st;
        return (gptr_t<const uint32_t>)(cref + (uint64_t)(r >> 4) * Kpad);
    };
    auto prefetch = [&](uint32_t st) {
        const gptr_t<const uint32_t> src = block_src(st);
#pragma unroll
        for (int i = 0; i < NPF; i++) {
            const uint32_t q = threadIdx.x + (uint32_t)i * NT;
            pf[i] = q < Kpad / 2 ? src[q] : 0u;
        }
    };
    if constexpr (PF) {
        if (nextb < Mp) prefetch(nextb);
    }
    for (uint32_t step = 0; step < Mp; ++step) {
        if (step == nextb) {
            nextb += 16;
            jbit = rev ? 0x8000u : 1u;
            if (!rev && wave == 0 && ln < 16u) {  // the previous block's count0
                if constexpr (ABS) gmem(c0out)[step - 16 + ln] = (int32_t)c0acc;
                else gmem(T.count0)[step - 16 + ln] = (int32_t)c0acc;
            }
            uint32_t *st32 = (uint32_t *)stage;
            __syncthreads();  // every wave has read back its payloads
            if constexpr (PF) {
#pragma unroll
                for (int i = 0; i < NPF; i++) {
                    const uint32_t q = threadIdx.x + (uint32_t)i * NT;
                    if (q < Kpad / 2) st32[q] = pf[i];
                }
                if (nextb < Mp) prefetch(nextb);
            } else {
                const gptr_t<const uint32_t> src = block_src(step);
                for (uint32_t q = threadIdx.x; q < Kpad / 2; q += NT) st32[q] = src[q];
            }
            if (threadIdx.x == 0) st32[Kpad / 2] = ~0u;  // id Kpad: the padding chunks
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NS; c++)
                p[c] = (p[c] & 0xFFFF0000u) | (uint32_t)stage[p[c] >> 16];
        }
        // pass 1: slot c's ballot words go to lane c.  Lanes >= NS of
        // olo/ohi are never read: no initial value (asm).
        uint32_t olo, ohi;
        asm volatile("" : "=v"(olo), "=v"(ohi));
        uint64_t mk[NS];  // the ballots stay in scalar registers for pass 2
        static_for<0, NS>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            const uint64_t m = __ballot((p[c] & jbit) != 0u);
            mk[c] = m;
            olo = writelane<c>(olo, (uint32_t)m);
            ohi = writelane<c>(ohi, (uint32_t)(m >> 32));
        });
        // zeros of lane c's chunk, their exclusive scan over the wave's chunks
        // (lanes >= NS: don't care) and the wave's total, written to LDS by
        // every lane (one address, one value)
        const uint32_t zl = 64u - (uint32_t)__popc(olo) - (uint32_t)__popc(ohi);
        const uint32_t zex = lane_scan_excl<NS>(zl);
        uint32_t *wzero = wzero2 + (step & 1u) * (uint32_t)NW;
        wzero[wave] = (uint32_t)__builtin_amdgcn_readlane((int)(zex + zl), NS - 1);
        __syncthreads();
        // zeros before my first chunk, and in the whole row (count0)
        uint32_t run = 0, cnt0 = 0;
        if constexpr (WS == 0) {
#pragma unroll
            for (int w2 = 0; w2 < NW; w2++) {
                const uint32_t v = __builtin_amdgcn_readfirstlane(wzero[w2]);
                cnt0 += v;
                run += (uint32_t)w2 < wave ? v : 0u;
            }
        } else {
            // lane w reads wave w's count (NW is a power of two)
            const uint32_t wv = wzero[ln & (NW - 1)];
            const uint32_t wex = lane_scan_excl<NW>(wv);
            run = (uint32_t)__builtin_amdgcn_readlane((int)wex, (int)wave);
            cnt0 = (uint32_t)__builtin_amdgcn_readlane((int)(wex + wv), NW - 1);
        }
        // count0 of the row (its zeros among the K chosen haplotypes: the
        // padding positions are all ones) into lane (step & 15); the forward
        // pass stores each block's 16 at the next block boundary
        c0acc = eimp_llvm_writelane(cnt0, step & 15u, c0acc);
        // lane c: zeros before chunk ch0 + c (lanes >= NS: don't care)
        const uint32_t zbv = run + zex;
        // A row whose K bits are all equal (about a fifth of the rows on
        // these panels) leaves the order as it is: no partition, no second
        // barrier, no read-back.  cnt0 is the same in every wave, so the whole
        // workgroup takes the same branch.
        if (cnt0 != 0u && cnt0 != K) {
            // pass 2: stable partition, zeros first.  For slot c let ob_c / run_c
            // be the destinations of the row's first one / zero at or after the
            // chunk (ob_c = cnt0 + ones before it, run_c = zeros before it, so
            // ob_c + run_c = cnt0 + 64 (ch0 + c)) and t the ones below me in the
            // chunk: a one goes to ob_c + t, a zero to run_c + ln - t.  With
            //   x = ob_c + t + 32 (NS - 1 - c)
            // dest + 32 (NS - 1 - c) is x for a one and V - x for a zero, with
            // V = cnt0 + 64 ch0 + ln + 64 (NS - 1) the same for every slot, so
            //   PBIAS + 4 dest = 4 sel + 128 c,  sel = bit ? x : V - x >= 0,
            // and 128 c is the store's immediate offset (pay[] starts at byte
            // PBIAS = 128 (NS - 1)).  In bytes: 4x = 4t + 4a_c with
            // a_c = ob_c + 32 (NS - 1 - c) (one v_lshl_add with the scalar
            // 4a_c), 4(V - x) = 4V - 4x; a_c is a scalar running sum
            // (a_{c+1} = a_c + ones_c - 32), so a slot costs
            // five vector instructions and its store.  No branch on the
            // chunk's contents: a test for all-equal chunks costs more (scalar
            // compare and taken branches) than the vector instructions it
            // would skip.
            {
                uint32_t a4 = (cnt0 + ch0 * 64u - run + 32u * (NS - 1)) * 4u;
                const uint32_t V4 = vbase4 + cnt0 * 4u;
                static_for<0, NS>([&](auto ci) {
                    constexpr int c = decltype(ci)::value;
                    const uint64_t m = mk[c];
                    const uint32_t x4 = mbcnt(m) * 4u + a4;
                    const uint32_t sel4 = __builtin_amdgcn_inverse_ballot_w64(m) ? x4 : V4 - x4;
                    *(lds_ptr_t<uint32_t>)(size_t)(sel4 + 128u * (uint32_t)c) = p[c];
                    a4 += (uint32_t)__popcll(m) * 4u - 128u;
                    // opaque scalar: keeps the running sum in bytes (left to
                    // itself the compiler factors the 4 out, one more vector op)
                    asm volatile("" : "+s"(a4));
                });
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NS; c++) p[c] = pay[(ch0 + (uint32_t)c) * 64 + ln];
        }
        if constexpr (ABS) {
            if (absperm) {
                const gptr_t<int32_t> ap = gmem(absperm) + (uint64_t)(step + 1) * K;
#pragma unroll
                for (int c = 0; c < NS; c++) {
                    const uint32_t pos = (ch0 + (uint32_t)c) * 64 + ln;
                    if (pos < K) ap[pos] = (int32_t)(p[c] >> 16);
                }
            }
        }
        // the row: chunk ch holds pairs 2ch and 2ch+1 (raw layout: words 2ch, 2ch+1)
        const uint32_t o1 = zbv + 32u - (uint32_t)__popc(olo);
        const uint32_t o3 = o1 + 32u - (uint32_t)__popc(ohi);
        // the logical words only: the padding was filled before the loop
        if (ostore) {
            const gptr_t<char> rowb = (gptr_t<char>)(size_t)rowaddr;
            *(gptr_t<u32x2>)(rowb + obyte) = u32x2{olo, raw ? ohi : o1};
            if (second) *(gptr_t<u32x2>)(rowb + obyte + 8u) = u32x2{ohi, o3};
        }
        if (dig && ostore) {
            // rotl(v ^ key, g mod 32) == rotr by -g: one v_alignbit each
            const uint32_t g = g0 + step * KW;
            const uint32_t dkey = g * kDigKey, dneg = 0u - g;
            dsum += __builtin_amdgcn_alignbit(olo ^ dkey, olo ^ dkey, dneg);
            const uint32_t x1 = o1 ^ (dkey + kDigKey);
            dsum += __builtin_amdgcn_alignbit(x1, x1, dneg - 1u);
            if (2 * och + 1 < npairs) {
                const uint32_t x2 = ohi ^ (dkey + 2u * kDigKey), x3 = o3 ^ (dkey + 3u * kDigKey);
                dsum += __builtin_amdgcn_alignbit(x2, x2, dneg - 2u);
                dsum += __builtin_amdgcn_alignbit(x3, x3, dneg - 3u);
            }
        }
        rowaddr += rs_bytes;
        jbit = (jbit << jl) >> jr;
    }
    // the forward pass's last (partial) block of count0
    if (!rev && wave == 0 && Mp > 0) {
        const uint32_t b0 = (Mp - 1) & ~15u;
        if (ln < Mp - b0) {
            if constexpr (ABS) gmem(c0out)[b0 + ln] = (int32_t)c0acc;
            else gmem(T.count0)[b0 + ln] = (int32_t)c0acc;
        }
    }
    if (dig) {
        dsum = wave_sum64(dsum);
        if (ln == 0) atomicAdd((unsigned long long *)&T.digest[0], (unsigned long long)dsum);
    }
}

// slot counts instantiated; a target uses the smallest one >= ceil(chunks / waves)
constexpr int kSlotCounts[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32};
constexpr int kNumSlotCounts = sizeof(kSlotCounts) / sizeof(kSlotCounts[0]);

__host__ __device__ inline int pbwt_slots_for(uint32_t Kpad, int nt) {
    const uint32_t need = (Kpad / 64 + nt / 64 - 1) / (nt / 64);
    for (int i = 0; i < kNumSlotCounts; i++)
        if ((uint32_t)kSlotCounts[i] >= need) return kSlotCounts[i];
    return -1;
}

template <int NT, int CPW, int WS>
__device__ __forceinline__ void pbwt_target(const TgtDev &T, const bool rev, uint32_t *smem) {
    const int ns = pbwt_slots_for(T.Kpad, NT);
    static_for<0, kNumSlotCounts>([&](auto ii) {
        constexpr int NSv = kSlotCounts[decltype(ii)::value];
        if constexpr (NSv <= CPW) {
            if (ns == NSv) pbwt_steps<NT, NSv, false, WS>(T, rev, smem);
        }
    });
}

// Waves per SIMD that one workgroup's LDS allows at the variant's capacity
// (at most 6, above which the larger variants would spill registers): the
// register budget is capped so that VGPRs do not lower the occupancy (512 x 20:
// three workgroups per CU need at most 85 VGPRs).
constexpr int pbwt_waves_hint(int NT, int CPW) {
    const int lds = (int)pbwt_pay_bias(CPW) + CPW * NT * 4 + 2 * (NT / 64) * 4;
    const int w = (163840 / lds) * (NT / 256);
    return w < 1 ? 1 : (w > 6 ? 6 : w);
}

template <int NT, int CPW, int WS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(pbwt_waves_hint(NT, CPW))))
void k_pbwt(const TgtDev *__restrict__ tg, const PbwtItem *__restrict__ items) {
    extern __shared__ uint32_t smem[];
    const PbwtItem it = items[blockIdx.x];
    pbwt_target<NT, CPW, WS>(tg[it.t], it.dir != 0, smem);
}

// ---------------------------------------------------------------------------
// Compressed PBWT for K beyond the LDS-resident kernels (src/PBWT.cpp:207-285,
// the same rows): the order lives in global memory (two arrays of Kpad
// haplotype indexes, swapped every step; L2-resident at these sizes) and only
// the current row's K bits are staged in LDS.  One workgroup of BIG_NT
// threads per item; wave w takes chunks w, w + NW, ... of 64 positions.  A
// step: stage row r; pass 1 counts the zeros of every chunk (ballots) into
// LDS; a block scan turns them into zeros-before-chunk; pass 2 recomputes
// each chunk's ballot, scatters every index to its place under the stable
// partition (zeros first) and writes the chunk's (data, cumulative zeros)
// pairs.  Positions K..Kpad-1 hold index K and read as ones (1-padding).
// ---------------------------------------------------------------------------
constexpr int BIG_NT = 1024;
constexpr uint32_t kBigMaxK = 1u << 19;  // LDS: K/8 bytes of row + K/16 of chunk counts


__device__ __forceinline__ uint32_t big_bit(const uint32_t *rowb, uint32_t idx, uint32_t K) {
    return idx < K ? (rowb[idx >> 5] >> (idx & 31)) & 1u : 1u;
}

__global__ __launch_bounds__(BIG_NT) void k_pbwt_big(const BigItem *__restrict__ items) {
    extern __shared__ uint32_t smem[];
    constexpr uint32_t NW = BIG_NT / 64;
    const BigItem I = items[blockIdx.x];
    const uint32_t K = I.K, Mp = I.Mp, RS = I.RS;
    const uint32_t Kpad = (K + 63) / 64 * 64, nch = Kpad / 64;
    const uint32_t nw32 = (K + 31) / 32, npairs = nw32, KW = 2 * npairs;
    uint32_t *rowb = smem;               // nw32 words: the current row
    uint32_t *cz = smem + (nw32 + 1) / 2 * 2;  // nch: zeros per chunk, then zeros before it
    __shared__ uint32_t s_part[NW + 1];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, ln = lane_id();
    gptr_t<uint32_t> ordA = gmem(I.ord0), ordB = gmem(I.ord1);
    for (uint32_t q = tid; q < Kpad; q += BIG_NT) {
        ordA[q] = q < K ? q : K;
        if (I.absperm && q < K) gmem(I.absperm)[q] = (int32_t)q;
    }
    {
        // the padding words of every row, stored once (as pbwt_steps)
        const uint32_t pw0 = I.raw ? 2 * nch : 2 * npairs;
        const uint32_t fw = I.raw ? ~0u : 0u;
        for (uint32_t rr = tid; rr < Mp; rr += BIG_NT) {
            gptr_t<uint32_t> prow = gmem(I.out) + (uint64_t)rr * RS;
            for (uint32_t w = pw0; w < RS; w++) prow[w] = fw;
        }
    }
    uint64_t dsum = 0;
    __syncthreads();
    for (uint32_t step = 0; step < Mp; step++) {
        const uint32_t r = I.rev ? Mp - 1 - step : step;
        const gptr_t<const uint32_t> src = (gptr_t<const uint32_t>)(I.rows + (uint64_t)r * I.rstride);
        for (uint32_t q = tid; q < nw32; q += BIG_NT) rowb[q] = src[q];
        __syncthreads();
        // pass 1: zeros of every chunk
        for (uint32_t ch = wave; ch < nch; ch += NW) {
            const uint32_t idx = ordA[ch * 64 + ln];
            const uint64_t m = __ballot(big_bit(rowb, idx, K));
            if (ln == 0) cz[ch] = 64u - (uint32_t)__popcll(m);
        }
        __syncthreads();
        // exclusive scan of cz over the chunks: thread tid owns a contiguous run
        const uint32_t per = (nch + BIG_NT - 1) / BIG_NT;
        const uint32_t c0 = min(nch, tid * per), c1 = min(nch, c0 + per);
        uint32_t mine = 0;
        for (uint32_t c = c0; c < c1; c++) mine += cz[c];
        const uint32_t incl = wave_scan_incl(mine);
        if (ln == 63) s_part[wave] = incl;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < NW; w++) {
            const uint32_t v = s_part[w];
            total += v;
            before += w < wave ? v : 0u;
        }
        uint32_t run = before + incl - mine;
        for (uint32_t c = c0; c < c1; c++) {
            const uint32_t z = cz[c];
            cz[c] = run;
            run += z;
        }
        __syncthreads();
        if (tid == 0) {
            if (I.count0) gmem(I.count0)[r] = (int32_t)total;
        }
        // pass 2: scatter the order, write the row's pairs
        const gptr_t<uint32_t> row = gmem(I.out) + (uint64_t)step * RS;
        for (uint32_t ch = wave; ch < nch; ch += NW) {
            const uint32_t pos = ch * 64 + ln;
            const uint32_t idx = ordA[pos];
            const uint32_t b = big_bit(rowb, idx, K);
            const uint64_t m = __ballot(b);
            const uint32_t zb = __builtin_amdgcn_readfirstlane(cz[ch]);
            const uint32_t t = mbcnt(m);  // ones below me in the chunk
            const uint32_t dest = b ? total + (ch * 64u - zb) + t : zb + (ln - t);
            ordB[dest] = idx;
            if (I.absperm && dest < K) gmem(I.absperm)[(uint64_t)(step + 1) * K + dest] = (int32_t)idx;
            const uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
            const uint32_t o1 = zb + 32u - (uint32_t)__popc(lo);
            const uint32_t o3 = zb + 64u - (uint32_t)__popcll(m);
            const bool second = 2 * ch + 1 < npairs;
            if (ln == 0) {
                if (I.raw) {
                    row[2 * ch] = lo;
                    row[2 * ch + 1] = hi;
                } else {
                    row[4 * ch] = lo;
                    row[4 * ch + 1] = o1;
                    if (second) {
                        row[4 * ch + 2] = hi;
                        row[4 * ch + 3] = o3;
                    }
                }
                if (I.digest) {
                    const uint64_t g = I.gbase + (uint64_t)step * KW + 4ull * ch;
                    dsum += (uint64_t)dig_word(g, lo) + dig_word(g + 1, o1);
                    if (second) dsum += (uint64_t)dig_word(g + 2, hi) + dig_word(g + 3, o3);
                }
            }
        }
        __syncthreads();  // the new order is complete (and cz free) before the next step
        const gptr_t<uint32_t> tmp = ordA;
        ordA = ordB;
        ordB = tmp;
    }
    if (I.digest && ln == 0) atomicAdd((unsigned long long *)I.digest, (unsigned long long)dsum);
}

static hipError_t raise_lds_limit(const void *fn, size_t lds);

size_t pbwt_big_lds_bytes(uint32_t K) {
    const uint32_t nw32 = (K + 31) / 32, nch = (K + 63) / 64;
    return (size_t)((nw32 + 1) / 2 * 2 + nch) * 4;
}
uint32_t pbwt_big_max_k() { return kBigMaxK; }

hipError_t launch_pbwt_big(const BigItem *items, uint32_t nitems, uint32_t maxK, hipStream_t s) {
    if (!nitems) return hipSuccess;
    const size_t lds = pbwt_big_lds_bytes(maxK);
    hipError_t e = raise_lds_limit((const void *)k_pbwt_big, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pbwt_big, dim3(nitems), dim3(BIG_NT), lds, s, items);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Panel PBWT for imputation (PBWT(refT, true), src/PBWT.cpp:53-71,131-153):
// one workgroup, forward only, K = all panel haplotypes, the absolute order
// after every site and count0 per site written out.
// ---------------------------------------------------------------------------
template <int NT, int CPW>
__global__ __launch_bounds__(NT) void k_panel_pbwt(TgtDev T, int32_t *__restrict__ absperm,
                                                  int32_t *__restrict__ c0out) {
    extern __shared__ uint32_t smem[];
    const int ns = pbwt_slots_for(T.Kpad, NT);
    static_for<0, kNumSlotCounts>([&](auto ii) {
        constexpr int NSv = kSlotCounts[decltype(ii)::value];
        if constexpr (NSv <= CPW) {
            if (ns == NSv) pbwt_steps<NT, NSv, true, 1>(T, false, smem, absperm, c0out);
        }
    });
}

// Site rows [m0, m0 + nsites) of the panel as 16-row blocks, one uint16 per
// haplotype (bit j = site m0 + 16b + j), haplotypes N..Kpad-1 all ones.
// grid (nblk, ceil(nw/256)), block 256: a thread transposes one 64-haplotype word.
__global__ __launch_bounds__(256) void k_site_blocks(PanelDev P, uint64_t m0, uint32_t nsites,
                                                    uint16_t *__restrict__ blocks, uint32_t Kpad) {
    const uint32_t b = blockIdx.x, q = blockIdx.y * 256 + threadIdx.x;
    const uint32_t nw = (P.N_ad + 63) / 64;
    if (q >= nw) return;
    uint32_t V[32];  // two uint16 values per register: haplotypes 2i, 2i+1 of the low word etc.
#pragma unroll
    for (int i = 0; i < 32; i++) V[i] = 0;
    for (uint32_t r = 0; r < 16; r++) {
        const uint32_t j = 16 * b + r;
        const uint64_t wv = j < nsites ? P.site[(m0 + j) * P.site_stride + q] : 0ull;
        const uint32_t lo = (uint32_t)wv, hi = (uint32_t)(wv >> 32);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            V[i] |= (((lo >> (2 * i)) & 1u) << r) | (((lo >> (2 * i + 1)) & 1u) << (16 + r));
            V[16 + i] |= (((hi >> (2 * i)) & 1u) << r) | (((hi >> (2 * i + 1)) & 1u) << (16 + r));
        }
    }
    if (64 * q + 64 > P.N_ad) {
#pragma unroll
        for (int i = 0; i < 32; i++) {
            const uint32_t k0 = 64 * q + 2 * i;
            if (k0 >= P.N_ad) V[i] |= 0xFFFFu;
            if (k0 + 1 >= P.N_ad) V[i] |= 0xFFFF0000u;
        }
    }
    uint4 *d4 = (uint4 *)(blocks + (uint64_t)b * Kpad + 64 * q);
#pragma unroll
    for (int i = 0; i < 8; i++) d4[i] = make_uint4(V[4 * i], V[4 * i + 1], V[4 * i + 2], V[4 * i + 3]);
}

// ---------------------------------------------------------------------------
// Device-driven batch (stream contexts): k_plan sizes every target's outputs
// from its split count, lays them out in the context's pools (targets that do
// not fit are deferred), and queues cref blocks and PBWT items (largest
// first) for the persistent kernels below.  One workgroup, n <= 1024.
// ---------------------------------------------------------------------------
constexpr int PLAN_NT = 1024;

__device__ __forceinline__ uint64_t wave_scan_incl64(uint64_t v) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t u = __shfl_up(v, o);
        if (l >= (uint32_t)o) v += u;
    }
    return v;
}

// block-wide exclusive scan of 64-bit values (PLAN_NT threads); total returned
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t *s_w, uint64_t &total) {
    const uint32_t wave = threadIdx.x >> 6, l = lane_id();
    const uint64_t inc = wave_scan_incl64(v);
    if (l == 63) s_w[wave] = inc;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < PLAN_NT / 64; w++) {
        const uint64_t x = s_w[w];
        tot += x;
        if (w < (int)wave) before += x;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
}

// block-wide maximum (PLAN_NT threads)
__device__ __forceinline__ uint64_t block_max64(uint64_t v, uint64_t *s_w) {
    const uint32_t wave = threadIdx.x >> 6, l = lane_id();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    if (l == 0) s_w[wave] = v;
    __syncthreads();
    uint64_t m = 0;
#pragma unroll
    for (int w = 0; w < PLAN_NT / 64; w++) m = s_w[w] > m ? s_w[w] : m;
    __syncthreads();
    return m;
}

__global__ __launch_bounds__(PLAN_NT) void k_plan(TgtDev *__restrict__ tg, BatchDev *__restrict__ B,
                                                 PlanArgs A) {
    __shared__ uint64_t s_w[PLAN_NT / 64];
    __shared__ uint64_t keys[2 * PLAN_NT];
    const uint32_t t = threadIdx.x;
    const uint32_t n = B->n;
    const bool valid = t < n;
    uint32_t Mp = 0, Kpad = 0, RS = 0, nblk = 0, S = 0, K = 0;
    uint64_t cb = 0, c0b = 0, ob = 0;
    if (valid) {
        const TgtDev &T = tg[t];
        S = T.stats[0];
        K = T.K;
        Mp = 2 * S + 1;
        Kpad = (K + 63) / 64 * 64;
        RS = row_stride(K, A.layout);
        nblk = (Mp + 15) / 16;
        cb = region_cref_bytes(Mp, Kpad);
        c0b = region_c0_bytes(Mp);
        ob = region_out_bytes(Mp, RS);
    }
    uint64_t tot;
    const uint64_t off = block_excl_scan64(cb + c0b + ob, s_w, tot);
    // prefix sums grow with t, so the targets that fit are a prefix of the batch
    const bool fit = valid && off + cb + c0b + ob <= A.pool_cap;
    if (valid) {
        TgtDev &T = tg[t];
        T.S = S;
        T.stats[2] = fit ? 0u : 1u;  // 1: deferred, outputs not built
        if (fit) {
            T.Mp = Mp;
            T.Kpad = Kpad;
            T.RS = RS;
            T.cref = (uint16_t *)(A.pool + off);
            T.count0 = (int32_t *)(A.pool + off + cb);
            T.out = (uint32_t *)(A.pool + off + cb + c0b);
            T.digest = A.digest ? A.dig_base + 2 * t : nullptr;
            if (A.digest) {
                A.dig_base[2 * t] = 0;
                A.dig_base[2 * t + 1] = 0;
            }
        } else {
            T.Mp = 0;
            T.digest = nullptr;
        }
    }
    // two cref queues: whole-panel targets (transpose kernel) and the others.
    // Block-major order (block b of every target before block b+1) when it
    // costs few empty items: targets with similar windows then read the same
    // panel rows at the same time, from the cache.
    const bool idk = fit && A.ident_ok && (tg[t].flags & TF_IDENTITY);
    const bool gk = fit && !idk;
    // subset targets are queued in items of CREF_G blocks
    const uint32_t ngrp = (nblk + CREF_G - 1) / CREF_G;
    uint64_t nb_tot, nb_id, ng, ni, mxg, mxi;
    const uint64_t bp = block_excl_scan64(gk ? ngrp : 0u, s_w, nb_tot);
    const uint64_t bpi = block_excl_scan64(idk ? nblk : 0u, s_w, nb_id);
    const uint64_t pg = block_excl_scan64(gk ? 1u : 0u, s_w, ng);
    const uint64_t pi = block_excl_scan64(idk ? 1u : 0u, s_w, ni);
    mxg = block_max64(gk ? ngrp : 0u, s_w);
    mxi = block_max64(idk ? nblk : 0u, s_w);
    const bool rrg = ng > 1 && ng * mxg * 4 <= nb_tot * 5;
    const bool rri = ni > 1 && ni * mxi * 4 <= nb_id * 5;
    if (valid) {
        if (rrg) {
            if (gk) A.blockpref[pg] = t;
        } else {
            A.blockpref[t] = (uint32_t)bp;
        }
        if (rri) {
            if (idk) A.blockpref_id[pi] = t;
        } else {
            A.blockpref_id[t] = (uint32_t)bpi;
        }
    }
    uint64_t nfit;
    (void)block_excl_scan64(fit ? 1u : 0u, s_w, nfit);
    if (t == 0) {
        if (!rrg) A.blockpref[n] = (uint32_t)nb_tot;
        if (!rri) A.blockpref_id[n] = (uint32_t)nb_id;
        B->ncref = (uint32_t)(rrg ? ng * mxg : nb_tot);
        B->ncref_id = (uint32_t)(rri ? ni * mxi : nb_id);
        B->rr[0] = rrg ? 1u : 0u;
        B->rr[1] = rri ? 1u : 0u;
        B->nrr[0] = (uint32_t)ng;
        B->nrr[1] = (uint32_t)ni;
        B->npbwt = 2 * (uint32_t)nfit;
        B->nfit = (uint32_t)nfit;
        B->cref_ctr = 0;
        B->cref_id_ctr = 0;
        B->pbwt_ctr = 0;
    }
    // PBWT items 2t (forward) and 2t+1 (backward), most work first:
    // key = (~(Mp * Kpad) & 40 bits) << 24 | item; unused slots sort last
    const uint64_t wk = fit ? ((~((uint64_t)Mp * Kpad)) & 0xFFFFFFFFFFull) << 24 : ~0ull;
    keys[2 * t] = fit ? (wk | (2 * t)) : ~0ull;
    keys[2 * t + 1] = fit ? (wk | (2 * t + 1)) : ~0ull;
    __syncthreads();
    // bitonic sort of 2*PLAN_NT keys, ascending
    for (uint32_t k = 2; k <= 2 * PLAN_NT; k <<= 1) {
        for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = t; i < 2 * PLAN_NT; i += PLAN_NT) {
                const uint32_t ixj = i ^ jj;
                if (ixj > i) {
                    const uint64_t a = keys[i], b = keys[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = t; i < 2 * (uint32_t)nfit; i += PLAN_NT) {
        const uint32_t item = (uint32_t)(keys[i] & 0xFFFFFFu);
        A.pitems[i] = PbwtItem{item >> 1, item & 1u};
    }
}

// next cref item from a queue: queue position i (i >= nq ends), target t,
// item b of t (items of `unit` 16-row blocks); b = ~0u for an empty item of a
// block-major queue
__device__ __forceinline__ void cref_pop(uint32_t *ctr, uint32_t nq, uint32_t n, uint32_t rr,
                                         uint32_t nrr, const TgtDev *__restrict__ tg,
                                         const uint32_t *__restrict__ pref, uint32_t unit,
                                         uint32_t &i, uint32_t &t, uint32_t &b) {
    __shared__ uint32_t s_i, s_t, s_b;
    if (threadIdx.x == 0) {
        const uint32_t ii = atomicAdd(ctr, 1u);
        uint32_t tt = 0, bb = 0;
        if (ii < nq) {
            if (rr) {  // block-major: block ii / nrr of the (ii % nrr)-th listed target
                tt = pref[ii % nrr];
                bb = ii / nrr;
                if (bb >= ((tg[tt].Mp + 15) / 16 + unit - 1) / unit) bb = ~0u;
            } else {  // last target whose first block is <= ii (it owns block ii)
                uint32_t lo = 0, hi = n;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (pref[mid] <= ii) lo = mid; else hi = mid;
                }
                tt = lo;
                bb = ii - pref[lo];
            }
        }
        s_i = ii;
        s_t = tt;
        s_b = bb;
    }
    __syncthreads();
    // the same for every lane: scalar registers, so the target's fields load
    // through the scalar cache
    i = __builtin_amdgcn_readfirstlane(s_i);
    t = __builtin_amdgcn_readfirstlane(s_t);
    b = __builtin_amdgcn_readfirstlane(s_b);
    __syncthreads();
}

// cref items (CREF_G blocks) of subset targets from their queue
__global__ __launch_bounds__(CREF_NT) void k_cref_persist(PanelDev P, const TgtDev *__restrict__ tg,
                                                         BatchDev *__restrict__ B,
                                                         const uint32_t *__restrict__ blockpref) {
    const uint32_t nq = B->ncref, n = B->n, rr = B->rr[0], nrr = B->nrr[0];
    for (;;) {
        uint32_t i, t, g;
        cref_pop(&B->cref_ctr, nq, n, rr, nrr, tg, blockpref, CREF_G, i, t, g);
        if (i >= nq) break;
        if (g != ~0u) {
            const uint32_t nblk = (tg[t].Mp + 15) / 16;
            cref_range_hap(P, tg[t], g * CREF_G, min(nblk, (g + 1) * CREF_G));
        }
    }
}

// cref blocks of whole-panel targets from their queue
template <int GPT>
__global__ __launch_bounds__(CREF_NT, 4) void k_cref_persist_ident(PanelDev P,
                                                               const TgtDev *__restrict__ tg,
                                                               BatchDev *__restrict__ B,
                                                               const uint32_t *__restrict__ blockpref) {
    extern __shared__ uint64_t srow[];
    const uint32_t nq = B->ncref_id, n = B->n, rr = B->rr[1], nrr = B->nrr[1];
    for (;;) {
        uint32_t i, t, b;
        cref_pop(&B->cref_id_ctr, nq, n, rr, nrr, tg, blockpref, 1u, i, t, b);
        if (i >= nq) break;
        if (b != ~0u) cref_block_ident<GPT>(P, tg[t], b, srow);
    }
}

// PBWT items from the queue
template <int NT, int CPW, int WS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(pbwt_waves_hint(NT, CPW))))
void k_pbwt_persist(const TgtDev *__restrict__ tg,
                                                    BatchDev *__restrict__ B,
                                                    const PbwtItem *__restrict__ items) {
    // dynamic LDS only, so the payloads start at LDS address 0; the next item's
    // index passes through smem[0], which no item uses between two items
    extern __shared__ uint32_t smem[];
    // highest wave priority for instruction issue: the PBWT chains bound the
    // launch, and other contexts' kernels may share these CUs (measured:
    // 1.3% off the launch at 200k x 500k)
    __builtin_amdgcn_s_setprio(3);
    const uint32_t np = B->npbwt;
    for (;;) {
        if (threadIdx.x == 0) smem[0] = atomicAdd(&B->pbwt_ctr, 1u);
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(smem[0]);
        __syncthreads();
        if (i >= np) break;
        const PbwtItem it = items[i];
        pbwt_target<NT, CPW, WS>(tg[it.t], it.dir != 0, smem);
        __syncthreads();
    }
}

// count0 digest (grid n, block 256)
__global__ __launch_bounds__(256) void k_digest_count0(const TgtDev *__restrict__ tg,
                                                      const uint32_t *__restrict__ d_n) {
    if (d_n && blockIdx.x >= *d_n) return;
    const TgtDev &T = tg[blockIdx.x];
    if (!T.digest || T.Mp == 0) return;
    uint64_t s = 0;
    for (uint32_t r = threadIdx.x; r < T.Mp; r += 256) s += dig_count0(r, T.count0[r]);
    s = wave_sum64(s);
    if (lane_id() == 0) atomicAdd((unsigned long long *)&T.digest[1], (unsigned long long)s);
}

// Digests recomputed from stored rows (data / cumulative-zero pairs, the
// drop-in and compact layouts) and count0: the same sums as the in-kernel
// digest, for outputs built without it.  grid (DIGR_G, n), block 256: row
// group x of target y; dig receives 2 words per target (zeroed beforehand).
constexpr uint32_t DIGR_G = 64;
__global__ __launch_bounds__(256) void k_digest_rows(const TgtDev *__restrict__ tg,
                                                    const uint32_t *__restrict__ d_n,
                                                    uint64_t *__restrict__ dig) {
    if (blockIdx.y >= *d_n) return;
    const TgtDev &T = tg[blockIdx.y];
    if (T.Mp == 0 || !T.out) return;
    const uint32_t KW = 2 * ((T.K + 31) / 32), RS = T.RS;
    const uint64_t rows = 2ull * T.Mp;
    const uint64_t per = (rows + DIGR_G - 1) / DIGR_G;
    const uint64_t r0 = per * blockIdx.x, r1 = min(rows, r0 + per);
    uint64_t sp = 0, sc = 0;
    for (uint64_t r = r0; r < r1; r++) {
        const uint32_t *row = T.out + r * RS;
        for (uint32_t w = threadIdx.x; w < KW; w += 256) sp += dig_word(r * KW + w, row[w]);
    }
    if (blockIdx.x == 0)
        for (uint32_t r = threadIdx.x; r < T.Mp; r += 256) sc += dig_count0(r, T.count0[r]);
    sp = wave_sum64(sp);
    sc = wave_sum64(sc);
    if (lane_id() == 0) {
        atomicAdd((unsigned long long *)&dig[2 * blockIdx.y], (unsigned long long)sp);
        if (blockIdx.x == 0) atomicAdd((unsigned long long *)&dig[2 * blockIdx.y + 1], (unsigned long long)sc);
    }
}

// cref blocks -> condensed reference rows (for inspection).  grid (Mp, 1), block 256
__global__ __launch_bounds__(256) void k_cref_rows(const TgtDev *__restrict__ tg, uint32_t t,
                                                  uint64_t *__restrict__ rows, uint32_t rw) {
    const TgtDev &T = tg[t];
    const uint32_t r = blockIdx.x;
    const uint16_t *blk = T.cref + (uint64_t)(r >> 4) * T.Kpad;
    for (uint32_t w = threadIdx.x; w < rw; w += 256) {
        uint64_t v = 0;
        for (uint32_t q = 0; q < 64; q++) {
            const uint32_t k = w * 64 + q;
            if (k < T.K) v |= (uint64_t)((blk[k] >> (r & 15)) & 1u) << q;
        }
        rows[(uint64_t)r * rw + w] = v;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// The dynamic-LDS limit of a kernel is process-wide: contexts of different
// sizes (and engines with different panels) only ever raise it.
static hipError_t raise_lds_limit(const void *fn, size_t lds) {
    static std::mutex mu;
    static std::map<const void *, size_t> limit;
    std::lock_guard<std::mutex> lk(mu);
    size_t &cur = limit[fn];
    if (lds <= cur) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) cur = lds;
    return e;
}

// The PBWT kernels address their payloads from LDS byte 0 (pbwt_steps):
// they must have no static LDS ahead of the dynamic buffer.
static hipError_t pbwt_lds_at_zero(const void *fn) {
    hipFuncAttributes a{};
    hipError_t e = hipFuncGetAttributes(&a, fn);
    if (e != hipSuccess) return e;
    return a.sharedSizeBytes == 0 ? hipSuccess : hipErrorInvalidKernelFile;
}

hipError_t launch_site_to_hap(const PanelDev &P, uint64_t *hap, hipStream_t s) {
    dim3 grid((unsigned)((P.hap_stride + 3) / 4), (P.N_ad + 63) / 64);
    hipLaunchKernelGGL(k_site_to_hap, grid, dim3(256), 0, s, P, hap);
    return hipGetLastError();
}
hipError_t launch_panel_mono(const PanelDev &P, uint64_t *pand, uint64_t *por, hipStream_t s) {
    hipLaunchKernelGGL(k_panel_mono, dim3((P.M + 3) / 4), dim3(256), 0, s, P, pand, por);
    return hipGetLastError();
}
hipError_t launch_distance(const PanelDev &P, const TgtDev *tg, const uint32_t *tlist,
                           uint32_t nsel, hipStream_t s) {
    if (!nsel) return hipSuccess;
    hipLaunchKernelGGL(k_distance, dim3((P.N_ad + 15) / 16, (nsel + DIST_G - 1) / DIST_G), dim3(256),
                       0, s, P, tg, tlist, nsel, (const uint32_t *)nullptr);
    return hipGetLastError();
}
hipError_t launch_select(const PanelDev &P, const TgtDev *tg, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_select, dim3(n), dim3(SEL_NT), 0, s, P, tg, (const uint32_t *)nullptr);
    return hipGetLastError();
}
hipError_t launch_monos(const PanelDev &P, const TgtDev *tg, const uint32_t *tlist, uint32_t n,
                        uint32_t max_ww, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_monos, dim3((max_ww + 63) / 64, n), dim3(MONO_NT), 0, s, P, tg, tlist,
                       (const uint32_t *)nullptr);
    return hipGetLastError();
}
hipError_t launch_splits(const PanelDev &P, const TgtDev *tg, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_splits_clear, dim3(n), dim3(64), 0, s, tg, (const uint32_t *)nullptr);
    hipLaunchKernelGGL(k_splits, dim3(n), dim3(64), 0, s, P, tg, (const uint32_t *)nullptr);
    return hipGetLastError();
}
hipError_t launch_bits_to_list(const TgtDev *tg, uint32_t n, uint32_t nbits_bh, int which,
                               hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_bits_to_list, dim3(n), dim3(SEL_NT), 0, s, tg, nbits_bh, which);
    return hipGetLastError();
}

// dynamic LDS of the cref kernels: the 16 staged rows of one tile
size_t cref_lds_bytes(const PanelDev &P, uint32_t maxK) {
    (void)maxK;
    return (size_t)16 * cref_tile_words((P.N_ad + 63) / 64) * sizeof(uint64_t);
}

hipError_t launch_cref(const PanelDev &P, const TgtDev *tg, const CrefItem *items, uint32_t nitems,
                       uint32_t maxK, bool ident, hipStream_t s) {
    if (!nitems) return hipSuccess;
    const uint32_t nw = (P.N_ad + 63) / 64;
    const size_t lds = cref_lds_bytes(P, maxK);
    if (ident) {
        const uint32_t gpt = (nw + CREF_NT - 1) / CREF_NT;
        if (gpt <= 2) {
            const void *fn = gpt <= 1 ? (const void *)k_cref_ident<1> : (const void *)k_cref_ident<2>;
            hipError_t e = raise_lds_limit(fn, lds);
            if (e != hipSuccess) return e;
            if (gpt <= 1)
                hipLaunchKernelGGL(k_cref_ident<1>, dim3(nitems), dim3(CREF_NT), lds, s, P, tg, items);
            else
                hipLaunchKernelGGL(k_cref_ident<2>, dim3(nitems), dim3(CREF_NT), lds, s, P, tg, items);
            return hipGetLastError();
        }
    }
    if (!P.hap || !P.hap_stride) return hipErrorInvalidValue;  // the haplotype-major panel is needed
    hipLaunchKernelGGL(k_cref, dim3(nitems), dim3(CREF_NT), 0, s, P, tg, items);
    return hipGetLastError();
}

// haplotype positions are 16-bit in the PBWT payload
uint32_t cref_max_k() { return 65536 - 64; }

hipError_t launch_digest_count0(const TgtDev *tg, uint32_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_digest_count0, dim3(n), dim3(256), 0, s, tg, (const uint32_t *)nullptr);
    return hipGetLastError();
}
hipError_t launch_digest_rows(const TgtDev *tg, const uint32_t *d_n, uint32_t nmax, uint64_t *dig,
                              hipStream_t s) {
    if (!nmax) return hipSuccess;
    hipLaunchKernelGGL(k_digest_rows, dim3(DIGR_G, nmax), dim3(256), 0, s, tg, d_n, dig);
    return hipGetLastError();
}
hipError_t launch_cref_rows(const TgtDev *tg, uint32_t t, uint32_t Mp, uint64_t *rows, uint32_t rw,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_cref_rows, dim3(Mp), dim3(256), 0, s, tg, t, rows, rw);
    return hipGetLastError();
}

// PBWT variants: (threads, chunks per wave), one-item-per-workgroup and queue kernels
struct PbwtVariant {
    int nt, cpw;
    void (*fn)(const TgtDev *, const PbwtItem *);
    void (*fn_persist)(const TgtDev *, BatchDev *, const PbwtItem *);
};
#define PBWT_VARIANT(NT, CPW, WS) {NT, CPW, k_pbwt<NT, CPW, WS>, k_pbwt_persist<NT, CPW, WS>}
static const PbwtVariant kPbwtVariants[] = {
    PBWT_VARIANT(256, 8, 0),  PBWT_VARIANT(256, 16, 0),  PBWT_VARIANT(512, 20, 0),
    PBWT_VARIANT(512, 32, 0), PBWT_VARIANT(1024, 24, 0),
    // the same shapes with the DPP wave-count scan (uniform targets)
    PBWT_VARIANT(512, 20, 1), PBWT_VARIANT(1024, 24, 1),
};
#undef PBWT_VARIANT
constexpr int kNumPbwtVariants = sizeof(kPbwtVariants) / sizeof(kPbwtVariants[0]);

int pbwt_variant_count() { return kNumPbwtVariants; }
int pbwt_variant_threads(int v) { return kPbwtVariants[v].nt; }
int pbwt_variant_cpw(int v) { return kPbwtVariants[v].cpw; }
uint32_t pbwt_variant_maxk(int v) {
    return (uint32_t)(kPbwtVariants[v].nt / 64) * kPbwtVariants[v].cpw * 64;
}
size_t pbwt_lds_bytes(int v, uint32_t Kpad) {
    const int nt = kPbwtVariants[v].nt;
    const int ns = pbwt_slots_for(Kpad, nt);
    if (ns < 0 || ns > kPbwtVariants[v].cpw) return ~(size_t)0;
    const size_t npos = (size_t)ns * nt;  // NS x NW x 64
    // bias, payloads (the cref stage shares them), wave counts
    return pbwt_pay_bias(ns) + npos * 4 + 2 * (nt / 64) * 4;
}
hipError_t pbwt_prepare(int v, size_t lds) {
    hipError_t e = pbwt_lds_at_zero((const void *)kPbwtVariants[v].fn);
    if (e != hipSuccess) return e;
    return raise_lds_limit((const void *)kPbwtVariants[v].fn, lds);
}
hipError_t launch_pbwt(int v, const TgtDev *tg, const PbwtItem *items, uint32_t nitems, size_t lds,
                       hipStream_t s) {
    if (!nitems) return hipSuccess;
    hipLaunchKernelGGL(kPbwtVariants[v].fn, dim3(nitems), dim3(kPbwtVariants[v].nt), lds, s, tg,
                       items);
    return hipGetLastError();
}

// ---- panel PBWT for imputation: one workgroup of the 1024-thread variant
hipError_t launch_panel_pbwt(const PanelDev &P, uint64_t m0, uint32_t nsites, uint16_t *blocks,
                             uint32_t *rows, int32_t *c0out, int32_t *absperm, hipStream_t s) {
    constexpr int NT = 1024, CPW = 24;
    TgtDev T{};
    T.K = P.N_ad;
    T.Kpad = (P.N_ad + 63) / 64 * 64;
    T.Mp = nsites;
    T.RS = row_stride(T.K, LAYOUT_COMPACT);
    T.cref = blocks;
    T.out = rows;
    const int ns = pbwt_slots_for(T.Kpad, NT);
    if (ns < 0 || ns > CPW) return hipErrorInvalidValue;
    const uint32_t nw = (P.N_ad + 63) / 64;
    hipLaunchKernelGGL(k_site_blocks, dim3((nsites + 15) / 16, (nw + 255) / 256), dim3(256), 0, s, P,
                       m0, nsites, blocks, T.Kpad);
    const size_t lds = pbwt_pay_bias(ns) + (size_t)ns * NT * 4 + 2 * (NT / 64) * 4;
    hipError_t e = pbwt_lds_at_zero((const void *)k_panel_pbwt<NT, CPW>);
    if (e != hipSuccess) return e;
    e = raise_lds_limit((const void *)k_panel_pbwt<NT, CPW>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_panel_pbwt<NT, CPW>), dim3(1), dim3(NT), lds, s, T, absperm, c0out);
    return hipGetLastError();
}
uint32_t panel_pbwt_max_k() { return 1024 / 64 * 24 * 64; }

// ---- device-driven batch (stream contexts)

// Prep phase over a batch of *d_n <= nmax targets, one kernel class per call
// (0 distances, 1 top-K, 2 monos, 3 split sites); grids cover nmax targets.
hipError_t launch_prep_dev(int cls, const PanelDev &P, const TgtDev *tg, const uint32_t *d_n,
                           uint32_t nmax, uint32_t max_ww, hipStream_t s) {
    switch (cls) {
    case 0:
        hipLaunchKernelGGL(k_distance, dim3((P.N_ad + 15) / 16, (nmax + DIST_G - 1) / DIST_G),
                           dim3(256), 0, s, P, tg, (const uint32_t *)nullptr, 0u, d_n);
        break;
    case 1:
        hipLaunchKernelGGL(k_select, dim3(nmax), dim3(SEL_NT), 0, s, P, tg, d_n);
        break;
    case 2:
        hipLaunchKernelGGL(k_monos, dim3((max_ww + 63) / 64, nmax), dim3(MONO_NT), 0, s, P, tg,
                           (const uint32_t *)nullptr, d_n);
        break;
    default:
        hipLaunchKernelGGL(k_splits_clear, dim3(nmax), dim3(64), 0, s, tg, d_n);
        hipLaunchKernelGGL(k_splits, dim3(nmax), dim3(64), 0, s, P, tg, d_n);
        break;
    }
    return hipGetLastError();
}
hipError_t launch_plan(TgtDev *tg, BatchDev *B, const PlanArgs &A, hipStream_t s) {
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(PLAN_NT), 0, s, tg, B, A);
    return hipGetLastError();
}
hipError_t launch_digest_count0_dev(const TgtDev *tg, const uint32_t *d_n, uint32_t nmax,
                                    hipStream_t s) {
    hipLaunchKernelGGL(k_digest_count0, dim3(nmax), dim3(256), 0, s, tg, d_n);
    return hipGetLastError();
}

using CrefPersistFn = void (*)(PanelDev, const TgtDev *, BatchDev *, const uint32_t *);
static CrefPersistFn cref_persist_fn(const PanelDev &P, bool ident) {
    if (!ident) return k_cref_persist;
    const uint32_t gpt = ((P.N_ad + 63) / 64 + CREF_NT - 1) / CREF_NT;
    return gpt <= 1 ? k_cref_persist_ident<1> : k_cref_persist_ident<2>;
}

// identity targets take the transpose path when every 64-haplotype group
// fits two per thread
bool cref_ident_ok(const PanelDev &P) { return (P.N_ad + 63) / 64 <= 2u * CREF_NT; }

// sets the dynamic LDS limit; occ[0], occ[1] = resident workgroups per CU of
// the subset and identity kernels
hipError_t cref_persist_prepare(const PanelDev &P, uint32_t maxK, int *occ) {
    for (int k = 0; k < 2; k++) {
        occ[k] = 0;
        if (k == 1 && !cref_ident_ok(P)) continue;
        // the subset kernel stages nothing in LDS
        const size_t lds = k == 1 ? cref_lds_bytes(P, maxK) : 0;
        const void *fn = (const void *)cref_persist_fn(P, k == 1);
        hipError_t e = raise_lds_limit(fn, lds);
        if (e != hipSuccess) return e;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[k], fn, CREF_NT, lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
// grid[0], grid[1]: workgroups of the subset and identity kernels (0: none)
hipError_t launch_cref_persist(const PanelDev &P, const TgtDev *tg, BatchDev *B,
                               const uint32_t *blockpref, const uint32_t *blockpref_id,
                               uint32_t maxK, const uint32_t *grid, hipStream_t s) {
    const size_t lds = cref_lds_bytes(P, maxK);
    if (grid[0] && (!P.hap || !P.hap_stride)) return hipErrorInvalidValue;
    if (grid[0])
        hipLaunchKernelGGL(cref_persist_fn(P, false), dim3(grid[0]), dim3(CREF_NT), 0, s, P, tg,
                           B, blockpref);
    if (grid[1] && cref_ident_ok(P))
        hipLaunchKernelGGL(cref_persist_fn(P, true), dim3(grid[1]), dim3(CREF_NT), lds, s, P, tg, B,
                           blockpref_id);
    return hipGetLastError();
}
hipError_t pbwt_persist_prepare(int v, size_t lds, int *occ) {
    const void *fn = (const void *)kPbwtVariants[v].fn_persist;
    hipError_t e = pbwt_lds_at_zero(fn);
    if (e != hipSuccess) return e;
    e = raise_lds_limit(fn, lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, fn, kPbwtVariants[v].nt, lds);
}
hipError_t launch_pbwt_persist(int v, const TgtDev *tg, BatchDev *B, const PbwtItem *items,
                               uint32_t grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(kPbwtVariants[v].fn_persist, dim3(grid), dim3(kPbwtVariants[v].nt), lds, s,
                       tg, B, items);
    return hipGetLastError();
}

}  // namespace eimp
