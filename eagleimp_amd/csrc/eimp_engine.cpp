// eimp_engine.cpp -- host side of the engine behind the C ABI in include/eimp.h.
//
// One engine = one device, one HIP stream, the resident panel and grow-only
// device pools for the per-batch data.  A batch runs in two phases:
//   A  (prep)   distances -> top-K -> monos -> split/call sites
//   B  (build)  condensed reference -> forward/backward PBWT (+ digests)
// with one small device->host read of the split counts in between, which
// sizes the variable-length outputs of phase B.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/eimp.h"
#include "eimp_internal.h"

namespace eimp {
hipError_t launch_site_to_hap(const PanelDev &P, uint64_t *hap, hipStream_t s);
hipError_t launch_panel_mono(const PanelDev &P, uint64_t *pand, uint64_t *por, hipStream_t s);
hipError_t launch_distance(const PanelDev &P, const TgtDev *tg, const uint32_t *tlist,
                           uint32_t nsel, hipStream_t s);
hipError_t launch_select(const PanelDev &P, const TgtDev *tg, uint32_t n, hipStream_t s);
hipError_t launch_monos(const PanelDev &P, const TgtDev *tg, const uint32_t *tlist, uint32_t n,
                        uint32_t max_ww, hipStream_t s);
hipError_t launch_splits(const PanelDev &P, const TgtDev *tg, uint32_t n, hipStream_t s);
hipError_t launch_bits_to_list(const TgtDev *tg, uint32_t n, uint32_t nbits_bh, int which,
                               hipStream_t s);
hipError_t launch_cref(const PanelDev &P, const TgtDev *tg, const CrefItem *items, uint32_t nitems,
                       uint32_t maxK, bool ident, hipStream_t s);
uint32_t cref_max_k();
hipError_t launch_digest_count0(const TgtDev *tg, uint32_t n, hipStream_t s);
hipError_t launch_cref_rows(const TgtDev *tg, uint32_t t, uint32_t Mp, uint64_t *rows, uint32_t rw,
                            hipStream_t s);
hipError_t launch_digest_rows(const TgtDev *tg, const uint32_t *d_n, uint32_t nmax, uint64_t *dig,
                              hipStream_t s);
size_t pbwt_big_lds_bytes(uint32_t K);
uint32_t pbwt_big_max_k();
hipError_t launch_pbwt_big(const BigItem *items, uint32_t nitems, uint32_t maxK, hipStream_t s);
int pbwt_variant_count();
int pbwt_variant_threads(int v);
int pbwt_variant_cpw(int v);
uint32_t pbwt_variant_maxk(int v);
size_t pbwt_lds_bytes(int v, uint32_t Kpad);
hipError_t pbwt_prepare(int v, size_t lds);
hipError_t launch_pbwt(int v, const TgtDev *tg, const PbwtItem *items, uint32_t nitems, size_t lds,
                       hipStream_t s);
hipError_t launch_prep_dev(int cls, const PanelDev &P, const TgtDev *tg, const uint32_t *d_n,
                           uint32_t nmax, uint32_t max_ww, hipStream_t s);
hipError_t launch_plan(TgtDev *tg, BatchDev *B, const PlanArgs &A, hipStream_t s);
hipError_t launch_digest_count0_dev(const TgtDev *tg, const uint32_t *d_n, uint32_t nmax,
                                    hipStream_t s);
size_t cref_lds_bytes(const PanelDev &P, uint32_t maxK);
hipError_t cref_persist_prepare(const PanelDev &P, uint32_t maxK, int *occ);
bool cref_ident_ok(const PanelDev &P);
hipError_t launch_panel_pbwt(const PanelDev &P, uint64_t m0, uint32_t nsites, uint16_t *blocks,
                             uint32_t *rows, int32_t *c0out, int32_t *absperm, hipStream_t s);
uint32_t panel_pbwt_max_k();
hipError_t launch_cref_persist(const PanelDev &P, const TgtDev *tg, BatchDev *B,
                               const uint32_t *blockpref, const uint32_t *blockpref_id,
                               uint32_t maxK, const uint32_t *grid, hipStream_t s);
hipError_t pbwt_persist_prepare(int v, size_t lds, int *occ);
hipError_t launch_pbwt_persist(int v, const TgtDev *tg, BatchDev *B, const PbwtItem *items,
                               uint32_t grid, size_t lds, hipStream_t s);
}  // namespace eimp

using namespace eimp;

// Scratch of eimp_engine_run_batch_noalloc: every device buffer of a batch
// except its inputs (genotype words) and outputs (rows, count0).  Each device
// that runs the library holds its own copy until the process ends.  In host
// code the name is only a stand-in: the device address of the current
// device's copy comes from hipGetSymbolAddress (scratch_for_device) and only
// that address is passed to kernels and copies.
static constexpr size_t kScratchBytes = 64ull << 20;
static constexpr size_t kNoAllocMaxTargets = 256;
static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return fail(EIMP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)

#define CHK(x)                      \
    do {                            \
        int r_ = (x);               \
        if (r_ != EIMP_OK) return r_; \
    } while (0)

inline uint64_t unit_words(uint64_t nbits) { return ((nbits + 511) / 512) * 8; }
inline uint32_t batch_layout(uint32_t flags) {
    return (flags & EIMP_BATCH_COMPACT) ? LAYOUT_COMPACT
         : (flags & EIMP_BATCH_RAW)     ? LAYOUT_RAW
                                        : LAYOUT_GPU;
}
inline uint64_t round_up(uint64_t n, uint64_t m) { return (n + m - 1) / m * m; }

// Grow-only device buffer
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    // slack: grow by an eighth more than asked, so that slowly growing batches
    // do not reallocate every call
    int ensure(size_t bytes, const char *what, bool slack = true) {
        if (bytes <= n && p) return EIMP_OK;
        release();
        size_t want = std::max<size_t>(slack ? bytes + bytes / 8 : bytes, 256);
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(EIMP_ERR_NOMEM, std::string("device allocation failed: ") + what + " (" +
                                            std::to_string(want) + " bytes)");
        }
        n = want;
        return EIMP_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename T>
    T *as() const { return reinterpret_cast<T *>(p); }
};

// Grow-only pinned host buffer
struct HostBuf {
    void *p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes) {
        if (bytes <= n && p) return EIMP_OK;
        release();
        size_t want = std::max<size_t>(bytes + bytes / 8, 256);
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(EIMP_ERR_NOMEM, "pinned host allocation failed");
        }
        n = want;
        return EIMP_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename T>
    T *as() const { return reinterpret_cast<T *>(p); }
};

struct PendingEvent {
    int cls;
    hipEvent_t a, b;
};

}  // namespace

struct eimp_engine {
    int dev = 0;
    hipStream_t stream = nullptr;
    bool initialized = false;
    // reference GPUEngine parameters
    size_t K = 0, Kwords = 0, M = 0, Ncorr = 0, Nad = 0, refTCap = 0;
    // resident panel
    DevBuf site, hap, cM, haploid, pand, por;
    bool have_cM = false, have_hap = false, have_pmono = false;
    uint64_t hap_stride = 0;
    PanelDev pd{};
    // batch pools
    DevBuf d_tg, d_gt, d_nerr, d_bh, d_bhflags, d_mono, d_splitsites, d_splits, d_cMdiffs,
        d_stats, d_cref, d_count0, d_out, d_digest, d_tlist, d_citems, d_pitems, d_tmp,
        d_crows, d_ord, d_big;  // large-K PBWT: row-major cref, orders, items
    DevBuf p_blocks, p_rows, p_c0, p_abs;  // panel PBWT
    HostBuf h_stage;
    std::vector<TgtDev> h_tg;
    std::vector<uint32_t> h_stats;
    uint32_t last_n = 0;
    TgtDev *last_tg = nullptr;    // device target table of the last batch
    uint64_t *last_dig = nullptr; // its digests
    uint64_t *scratch = nullptr;  // this device's copy of the no-allocation scratch array
    // timing
    bool timing = false;
    double ms[EIMP_NKCLASS] = {};
    uint64_t launches[EIMP_NKCLASS] = {};
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> event_pool;
    int pbwt_variant = -1;
    uint64_t panel_gen = 0;  // bumped whenever panel buffers change (stream contexts check it)
    int pbwt_prepared_variant = -1;
    size_t pbwt_prepared_lds = 0;

    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // Records a start event for a kernel class; returns the index to close.
    int tbegin(int cls) {
        if (!timing) return -1;
        PendingEvent pe{cls, get_event(), get_event()};
        if (!pe.a || !pe.b) return -1;
        (void)hipEventRecord(pe.a, stream);
        pending.push_back(pe);
        return (int)pending.size() - 1;
    }
    void tend(int idx) {
        if (idx < 0) return;
        (void)hipEventRecord(pending[idx].b, stream);
        launches[pending[idx].cls]++;
    }
    // after a stream synchronisation
    void tcollect() {
        for (auto &pe : pending) {
            float t = 0;
            if (hipEventElapsedTime(&t, pe.a, pe.b) == hipSuccess) ms[pe.cls] += t;
            event_pool.push_back(pe.a);
            event_pool.push_back(pe.b);
        }
        pending.clear();
    }
    int sync() {
        HIPCHK(hipStreamSynchronize(stream));
        tcollect();
        return EIMP_OK;
    }
    void release_pools() {
        DevBuf *bufs[] = {&d_tg, &d_gt, &d_nerr, &d_bh, &d_bhflags, &d_mono, &d_splitsites,
                          &d_splits, &d_cMdiffs, &d_stats, &d_cref, &d_count0, &d_out,
                          &d_digest, &d_tlist, &d_citems, &d_pitems, &d_tmp, &d_crows, &d_ord, &d_big,
                          &p_blocks, &p_rows, &p_c0, &p_abs};
        for (auto *b : bufs) b->release();
        h_stage.release();
    }
    void release_panel() {
        panel_gen++;
        site.release();
        hap.release();
        cM.release();
        haploid.release();
        pand.release();
        por.release();
        have_cM = have_hap = have_pmono = false;
    }
    int choose_pbwt_variant(uint32_t Kmax_pad) {
        if (pbwt_variant >= 0) {
            if (pbwt_variant >= pbwt_variant_count()) return fail(EIMP_ERR_INVALID, "bad variant");
            if (pbwt_variant_maxk(pbwt_variant) < Kmax_pad)
                return fail(EIMP_ERR_UNSUPPORTED, "PBWT variant too small for K");
            return pbwt_variant;
        }
        // fewest chunks per wave among workgroups of up to 512 threads (measured
        // best on MI355X at K = 10,000), 1024-thread workgroups only when needed
        int best = -1;
        for (int pass = 0; pass < 2 && best < 0; pass++) {
            for (int v = 0; v < pbwt_variant_count(); v++) {
                if (pbwt_variant_maxk(v) < Kmax_pad) continue;
                if (pass == 0 && pbwt_variant_threads(v) > 512) continue;
                if (best < 0 || pbwt_variant_cpw(v) < pbwt_variant_cpw(best) ||
                    (pbwt_variant_cpw(v) == pbwt_variant_cpw(best) &&
                     pbwt_variant_threads(v) < pbwt_variant_threads(best)))
                    best = v;
            }
        }
        if (best < 0) return fail(EIMP_ERR_UNSUPPORTED, "K beyond the PBWT kernels' LDS capacity");
        return best;
    }
};

namespace {

int set_device(eimp_engine *e) {
    HIPCHK(hipSetDevice(e->dev));
    return EIMP_OK;
}

int ensure_hap(eimp_engine *e) {
    if (e->have_hap) return EIMP_OK;
    e->hap_stride = unit_words(e->M);
    CHK(e->hap.ensure(e->Nad * e->hap_stride * 8, "hap-major panel"));
    HIPCHK(hipMemsetAsync(e->hap.p, 0, e->Nad * e->hap_stride * 8, e->stream));
    e->pd.hap = e->hap.as<uint64_t>();
    e->pd.hap_stride = e->hap_stride;
    HIPCHK(launch_site_to_hap(e->pd, e->hap.as<uint64_t>(), e->stream));
    e->have_hap = true;
    return e->sync();
}

int ensure_pmono(eimp_engine *e) {
    if (e->have_pmono) return EIMP_OK;
    size_t words = unit_words(e->M);
    CHK(e->pand.ensure(words * 8, "panel mono"));
    CHK(e->por.ensure(words * 8, "panel mono"));
    HIPCHK(hipMemsetAsync(e->pand.p, 0, words * 8, e->stream));
    HIPCHK(hipMemsetAsync(e->por.p, 0, words * 8, e->stream));
    HIPCHK(launch_panel_mono(e->pd, e->pand.as<uint64_t>(), e->por.as<uint64_t>(), e->stream));
    e->pd.pand = e->pand.as<uint64_t>();
    e->pd.por = e->por.as<uint64_t>();
    e->have_pmono = true;
    return e->sync();
}

// Where one batch's device buffers come from: the engine's grow-only pools
// (eimp_engine_run_batch, the drop-in call) or a bump allocator over the
// file-scope scratch array with caller-provided outputs
// (eimp_engine_run_batch_noalloc).
enum MemSlot : int {
    MS_TG, MS_NERR, MS_BH, MS_BHF, MS_MONO, MS_SS, MS_SPL, MS_CMD, MS_STATS, MS_TLIST,
    MS_CREF, MS_DIG, MS_CITEMS, MS_PITEMS, MS_CROWS, MS_ORD, MS_BIG
};
struct BatchMem {
    virtual ~BatchMem() = default;
    // device buffer of at least `bytes` for a part of the batch
    virtual int get(MemSlot s, size_t bytes, const char *what, void **p) = 0;
    // sets T.out and T.count0 of targets 0..n-1 (Mp and RS set)
    virtual int place_outputs(std::vector<TgtDev> &tg, uint32_t n) = 0;
    // false: nothing may be allocated (resident panel parts must exist already)
    virtual bool may_alloc() const { return true; }
};

struct PoolMem final : BatchMem {
    eimp_engine *e;
    explicit PoolMem(eimp_engine *e_) : e(e_) {}
    DevBuf &slot(MemSlot s) {
        switch (s) {
        case MS_TG: return e->d_tg;
        case MS_NERR: return e->d_nerr;
        case MS_BH: return e->d_bh;
        case MS_BHF: return e->d_bhflags;
        case MS_MONO: return e->d_mono;
        case MS_SS: return e->d_splitsites;
        case MS_SPL: return e->d_splits;
        case MS_CMD: return e->d_cMdiffs;
        case MS_STATS: return e->d_stats;
        case MS_TLIST: return e->d_tlist;
        case MS_CREF: return e->d_cref;
        case MS_DIG: return e->d_digest;
        case MS_CITEMS: return e->d_citems;
        case MS_CROWS: return e->d_crows;
        case MS_ORD: return e->d_ord;
        case MS_BIG: return e->d_big;
        default: return e->d_pitems;
        }
    }
    int get(MemSlot s, size_t bytes, const char *what, void **p) override {
        DevBuf &b = slot(s);
        CHK(b.ensure(bytes, what));
        *p = b.p;
        return EIMP_OK;
    }
    int place_outputs(std::vector<TgtDev> &tg, uint32_t n) override {
        uint64_t c0_bytes = 0, out_bytes = 0;
        std::vector<uint64_t> c0_off(n), out_off(n);
        for (uint32_t t = 0; t < n; t++) {
            c0_off[t] = c0_bytes;
            c0_bytes += round_up((uint64_t)tg[t].Mp * 4, 256);
            out_off[t] = out_bytes;
            out_bytes += round_up(2ull * tg[t].Mp * tg[t].RS * 4, 256);
        }
        CHK(e->d_count0.ensure(c0_bytes, "count0"));
        CHK(e->d_out.ensure(out_bytes, "cPBWT output"));
        for (uint32_t t = 0; t < n; t++) {
            tg[t].count0 = reinterpret_cast<int32_t *>(e->d_count0.as<char>() + c0_off[t]);
            tg[t].out = reinterpret_cast<uint32_t *>(e->d_out.as<char>() + out_off[t]);
        }
        return EIMP_OK;
    }
};

// Bump allocation over one device's copy of the scratch array; the outputs
// are the caller's (rows[t]: 2*Mp*RS uint32, count0[t]: Mp int32, Mp <= max_Mp).
struct ScratchMem final : BatchMem {
    char *base;
    size_t used = 0;
    uint32_t *const *rows;
    int32_t *const *c0;
    size_t max_Mp;
    ScratchMem(uint64_t *scr, uint32_t *const *rows_, int32_t *const *c0_, size_t max_Mp_)
        : base(reinterpret_cast<char *>(scr)), rows(rows_), c0(c0_), max_Mp(max_Mp_) {}
    int get(MemSlot, size_t bytes, const char *what, void **p) override {
        const size_t b = round_up(std::max<size_t>(bytes, 1), 256);
        if (used + b > kScratchBytes)
            return fail(EIMP_ERR_UNSUPPORTED, std::string("batch needs more than the ") +
                                                  std::to_string(kScratchBytes >> 20) +
                                                  " MiB no-allocation scratch (at " + what + ")");
        *p = base + used;
        used += b;
        return EIMP_OK;
    }
    int place_outputs(std::vector<TgtDev> &tg, uint32_t n) override {
        for (uint32_t t = 0; t < n; t++) {
            if (tg[t].Mp > max_Mp)
                return fail(EIMP_ERR_UNSUPPORTED, "target " + std::to_string(t) + ": Mp = " +
                                                      std::to_string(tg[t].Mp) + " exceeds max_Mp = " +
                                                      std::to_string(max_Mp));
            if (!rows[t] || !c0[t])
                return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": null output buffer");
            tg[t].out = rows[t];
            tg[t].count0 = c0[t];
        }
        return EIMP_OK;
    }
    bool may_alloc() const override { return false; }
};

// Device address of the current device's scratch copy, looked up once per
// device (the device must be current).
int scratch_for_device(int dev, uint64_t **out) {
    static std::mutex mu;
    static std::map<int, uint64_t *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        void *ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr)
            return fail(EIMP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
        it = addr.emplace(dev, static_cast<uint64_t *>(ptr)).first;
    }
    *out = it->second;
    return EIMP_OK;
}

// Phase B for targets whose phase-A fields (bh, splitsites, S) are in place.
// Sizes outputs, builds cref + PBWT in the given output layout (LAYOUT_*).
int run_build(eimp_engine *e, uint32_t n, bool digest, uint32_t layout, BatchMem &mem) {
    std::vector<TgtDev> &tg = e->h_tg;
    uint64_t cref_bytes = 0, crows_bytes = 0, ord_bytes = 0;
    std::vector<uint64_t> cref_off(n), crows_off(n, 0), ord_off(n, 0);
    uint32_t Kmax_pad = 0, Kmax = 0, Kbig = 0;
    std::vector<CrefItem> citems, citems_id;  // general / identity targets
    std::vector<PbwtItem> pitems;
    std::vector<uint32_t> big;                // targets beyond the LDS-resident PBWT kernels
    uint32_t lds_cap = 0;
    for (int v = 0; v < pbwt_variant_count(); v++) lds_cap = std::max(lds_cap, pbwt_variant_maxk(v));
    for (uint32_t t = 0; t < n; t++) {
        TgtDev &T = tg[t];
        T.Mp = 2 * T.S + 1;
        T.Kpad = (uint32_t)round_up(T.K, 64);
        T.RS = row_stride(T.K, layout);
        if (layout == LAYOUT_RAW) T.flags |= TF_RAW;
        const uint64_t nblk = (T.Mp + 15) / 16;
        cref_off[t] = cref_bytes;
        cref_bytes += round_up(nblk * T.Kpad * 2, 256);
        const bool ident = (T.flags & TF_IDENTITY) && e->Nad <= 2ull * 256 * 64;
        if (!ident) Kmax = std::max(Kmax, T.K);
        if (ident) {
            for (uint32_t b = 0; b < nblk; b++) citems_id.push_back(CrefItem{t, b, b + 1, 0});
        } else {
            for (uint32_t b = 0; b < nblk; b += kCrefGroup)
                citems.push_back(CrefItem{t, b, (uint32_t)std::min<uint64_t>(nblk, b + kCrefGroup), 0});
        }
        if (T.Kpad > lds_cap) {
            // order in global memory: rows of K bits, two orders of Kpad per direction
            if (T.K > pbwt_big_max_k())
                return fail(EIMP_ERR_UNSUPPORTED, "target " + std::to_string(t) + ": K beyond " +
                                                      std::to_string(pbwt_big_max_k()));
            big.push_back(t);
            Kbig = std::max(Kbig, T.K);
            crows_off[t] = crows_bytes;
            crows_bytes += round_up((uint64_t)T.Mp * unit_words(T.K) * 8, 256);
            ord_off[t] = ord_bytes;
            ord_bytes += 4ull * T.Kpad * 4;
        } else {
            Kmax_pad = std::max(Kmax_pad, T.Kpad);
            pitems.push_back(PbwtItem{t, 0});
            pitems.push_back(PbwtItem{t, 1});
        }
    }
    // the subset cref kernel reads the haplotype-major panel (the drop-in's
    // initialize uploads only the site-major one: transposed here, once)
    if (!citems.empty() && !e->have_hap) {
        if (!mem.may_alloc())
            return fail(EIMP_ERR_STATE, "the panel's hap-major rows are not resident");
        CHK(ensure_hap(e));
    }
    if (cref_lds_bytes(e->pd, Kmax) > 160 * 1024)
        return fail(EIMP_ERR_UNSUPPORTED, "panel too wide for the cref kernel's LDS staging");
    int v = -1;
    size_t lds = 0;
    if (!pitems.empty()) {
        v = e->choose_pbwt_variant(Kmax_pad);
        if (v < 0) return v;
        lds = pbwt_lds_bytes(v, Kmax_pad);
        if (lds > 160 * 1024)
            return fail(EIMP_ERR_UNSUPPORTED, "K beyond the PBWT kernel's slots or 160 KiB of LDS");
        if (e->pbwt_prepared_variant != v || e->pbwt_prepared_lds < lds) {
            HIPCHK(pbwt_prepare(v, lds));
            e->pbwt_prepared_variant = v;
            e->pbwt_prepared_lds = lds;
        }
    }
    if (!big.empty() && pbwt_big_lds_bytes(Kbig) > 160 * 1024)
        return fail(EIMP_ERR_UNSUPPORTED, "K beyond the large-K PBWT kernel's LDS");
    // longest targets first
    std::stable_sort(pitems.begin(), pitems.end(), [&](const PbwtItem &a, const PbwtItem &b) {
        return tg[a.t].Mp > tg[b.t].Mp;
    });
    void *p_cref = nullptr, *p_dig = nullptr, *p_ci = nullptr, *p_pi = nullptr, *p_tg = nullptr;
    void *p_crows = nullptr, *p_ord = nullptr, *p_big = nullptr;
    CHK(mem.get(MS_CREF, cref_bytes, "cref", &p_cref));
    CHK(mem.get(MS_DIG, (size_t)n * 16, "digest", &p_dig));
    const size_t nci = citems.size();
    citems.insert(citems.end(), citems_id.begin(), citems_id.end());
    CHK(mem.get(MS_CITEMS, std::max<size_t>(citems.size(), 1) * sizeof(CrefItem), "cref items", &p_ci));
    CHK(mem.get(MS_PITEMS, std::max<size_t>(pitems.size(), 1) * sizeof(PbwtItem), "pbwt items", &p_pi));
    if (!big.empty()) {
        CHK(mem.get(MS_CROWS, crows_bytes, "large-K cref rows", &p_crows));
        CHK(mem.get(MS_ORD, ord_bytes, "large-K orders", &p_ord));
        CHK(mem.get(MS_BIG, 2 * big.size() * sizeof(BigItem), "large-K items", &p_big));
    }
    CHK(mem.get(MS_TG, n * sizeof(TgtDev), "target table", &p_tg));
    CHK(mem.place_outputs(tg, n));
    uint64_t *d_dig = static_cast<uint64_t *>(p_dig);
    TgtDev *d_tg = static_cast<TgtDev *>(p_tg);
    CrefItem *d_ci = static_cast<CrefItem *>(p_ci);
    PbwtItem *d_pi = static_cast<PbwtItem *>(p_pi);
    for (uint32_t t = 0; t < n; t++) {
        TgtDev &T = tg[t];
        T.cref = reinterpret_cast<uint16_t *>(static_cast<char *>(p_cref) + cref_off[t]);
        T.digest = digest ? d_dig + 2 * t : nullptr;
    }
    std::vector<BigItem> bitems;
    for (uint32_t t : big) {
        const TgtDev &T = tg[t];
        const uint64_t *rows = reinterpret_cast<const uint64_t *>(static_cast<char *>(p_crows) + crows_off[t]);
        uint32_t *ord = reinterpret_cast<uint32_t *>(static_cast<char *>(p_ord) + ord_off[t]);
        const uint32_t KW = 2 * ((T.K + 31) / 32);
        for (uint32_t dir = 0; dir < 2; dir++) {
            BigItem B{};
            B.rows = rows;
            B.rstride = unit_words(T.K);
            B.out = T.out + (dir ? (uint64_t)T.Mp * T.RS : 0ull);
            B.count0 = dir ? nullptr : T.count0;
            B.absperm = nullptr;
            B.ord0 = ord + (uint64_t)dir * 2 * T.Kpad;
            B.ord1 = B.ord0 + T.Kpad;
            B.digest = T.digest;
            B.gbase = dir ? (uint64_t)T.Mp * KW : 0ull;
            B.K = T.K;
            B.Mp = T.Mp;
            B.RS = T.RS;
            B.rev = dir;
            B.raw = layout == LAYOUT_RAW ? 1u : 0u;
            bitems.push_back(B);
        }
    }
    if (digest) HIPCHK(hipMemsetAsync(d_dig, 0, (size_t)n * 16, e->stream));
    HIPCHK(hipMemcpyAsync(d_tg, tg.data(), n * sizeof(TgtDev), hipMemcpyHostToDevice, e->stream));
    if (!citems.empty())
        HIPCHK(hipMemcpyAsync(d_ci, citems.data(), citems.size() * sizeof(CrefItem),
                              hipMemcpyHostToDevice, e->stream));
    if (!pitems.empty())
        HIPCHK(hipMemcpyAsync(d_pi, pitems.data(), pitems.size() * sizeof(PbwtItem),
                              hipMemcpyHostToDevice, e->stream));
    if (!bitems.empty())
        HIPCHK(hipMemcpyAsync(p_big, bitems.data(), bitems.size() * sizeof(BigItem),
                              hipMemcpyHostToDevice, e->stream));
    int ti = e->tbegin(EIMP_KCLASS_CREF);
    HIPCHK(launch_cref(e->pd, d_tg, d_ci, (uint32_t)nci, Kmax, false, e->stream));
    HIPCHK(launch_cref(e->pd, d_tg, d_ci + nci, (uint32_t)(citems.size() - nci), 0, true, e->stream));
    // large-K targets read their condensed reference as rows of K bits
    for (uint32_t t : big) {
        const TgtDev &T = tg[t];
        uint64_t *rows = reinterpret_cast<uint64_t *>(static_cast<char *>(p_crows) + crows_off[t]);
        HIPCHK(launch_cref_rows(d_tg, t, T.Mp, rows, (uint32_t)unit_words(T.K), e->stream));
    }
    e->tend(ti);
    ti = e->tbegin(EIMP_KCLASS_PBWT);
    if (!pitems.empty())
        HIPCHK(launch_pbwt(v, d_tg, d_pi, (uint32_t)pitems.size(), lds, e->stream));
    if (!bitems.empty())
        HIPCHK(launch_pbwt_big(static_cast<const BigItem *>(p_big), (uint32_t)bitems.size(), Kbig,
                               e->stream));
    e->tend(ti);
    if (digest) HIPCHK(launch_digest_count0(d_tg, n, e->stream));
    e->last_tg = d_tg;
    e->last_dig = d_dig;
    return EIMP_OK;
}

// Validates target t and fills its input fields and flags.  Ncorr is the
// engine's (the caller accounts for the skipped sample,
// src/PBWTPhaser.cpp:629); localK = min(Ncorr, K) (:683).
int target_inputs(const eimp_engine *e, const eimp_target &x, uint32_t t, float cMmaxSplit,
                  TgtDev &T) {
    if (!x.gt) return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": null genotypes");
    if (x.W < 1 || x.s0 >= e->M || x.W > e->M - x.s0)
        return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": window outside the panel");
    if (x.K < 1) return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": K < 1");
    if (x.ignoreSample >= (int64_t)(e->Nad / 2))
        return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": ignoreSample out of range");
    const uint64_t ncorr = e->Ncorr;
    T.K = (uint32_t)std::min<uint64_t>(x.K, ncorr);
    T.W = (uint32_t)x.W;
    T.s0 = x.s0;
    T.ignore = (int32_t)std::max<int64_t>(x.ignoreSample, -1);
    T.cMmaxSplit = cMmaxSplit;
    const bool select = T.K < ncorr;
    const bool ident = !select && e->Nad == e->Ncorr && x.ignoreSample < 0 && T.K == e->Nad;
    T.flags = (select ? TF_SELECT : 0u) | (ident ? TF_IDENTITY : 0u);
    return EIMP_OK;
}

// Copies the outputs of one target (host record T, device table d_tg) to
// the caller's buffers; any pointer may be null.
int copy_target_out(const TgtDev &T, const TgtDev *d_tg, uint32_t t, uint64_t Nad, hipStream_t s,
                    DevBuf &tmp, uint32_t *besthaps, uint64_t *bhflags, uint32_t *splitsites,
                    uint64_t *splits, float *cMdiffs, int32_t *count0, uint32_t *pbwt,
                    uint64_t *cref_rows) {
    if ((count0 || pbwt || cref_rows) && T.Mp == 0)
        return fail(EIMP_ERR_STATE, "target " + std::to_string(t) + " was deferred: outputs not built");
    if (besthaps && T.bh) HIPCHK(hipMemcpy(besthaps, T.bh, (size_t)T.K * 4, hipMemcpyDeviceToHost));
    if (bhflags && T.bhflags)
        HIPCHK(hipMemcpy(bhflags, T.bhflags, unit_words(Nad) * 8, hipMemcpyDeviceToHost));
    if (splitsites && T.S) HIPCHK(hipMemcpy(splitsites, T.splitsites, (size_t)T.S * 4, hipMemcpyDeviceToHost));
    if (splits && T.splits) HIPCHK(hipMemcpy(splits, T.splits, unit_words(T.W) * 8, hipMemcpyDeviceToHost));
    if (cMdiffs && T.S && T.cMdiffs)
        HIPCHK(hipMemcpy(cMdiffs, T.cMdiffs, ((size_t)T.S + 1) * 4, hipMemcpyDeviceToHost));
    if (count0) HIPCHK(hipMemcpy(count0, T.count0, (size_t)T.Mp * 4, hipMemcpyDeviceToHost));
    if (pbwt) HIPCHK(hipMemcpy(pbwt, T.out, 2ull * T.Mp * T.RS * 4, hipMemcpyDeviceToHost));
    if (cref_rows) {
        const uint32_t rw = (uint32_t)unit_words(T.K);
        CHK(tmp.ensure((size_t)T.Mp * rw * 8, "cref rows"));
        HIPCHK(launch_cref_rows(d_tg, t, T.Mp, tmp.as<uint64_t>(), rw, s));
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipMemcpy(cref_rows, tmp.p, (size_t)T.Mp * rw * 8, hipMemcpyDeviceToHost));
    }
    return EIMP_OK;
}

// Mono bits of a target (Target::findMonos, src/Target.cpp:878-897: site m
// of the window is mono when all chosen haplotypes agree there) from the
// device's AND / OR words: mono = and | ~or.  Subset targets keep window-
// relative words (mand_w / mor_w), whole-panel targets read the panel-level
// words at bit offset s0 (src_words readable words either way).  monos
// receives roundToMultiple(W,512)/64 words, bits past W zero.
int copy_monos_out(const TgtDev &T, uint64_t src_words, uint64_t *monos) {
    if (!T.mand || !T.mor) return fail(EIMP_ERR_STATE, "no mono words for this target");
    const uint64_t W = T.W, nw = (W + 63) / 64, uw = unit_words(W);
    const uint64_t q0 = T.mono_off >> 6, sh = T.mono_off & 63;
    const uint64_t ncopy = std::min<uint64_t>(nw + 1, src_words - q0);
    std::vector<uint64_t> a(nw + 1, 0), o(nw + 1, 0);
    HIPCHK(hipMemcpy(a.data(), T.mand + q0, ncopy * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(o.data(), T.mor + q0, ncopy * 8, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < uw; i++) {
        uint64_t v = 0;
        if (i < nw) {
            const uint64_t av = sh ? (a[i] >> sh) | (a[i + 1] << (64 - sh)) : a[i];
            const uint64_t ov = sh ? (o[i] >> sh) | (o[i + 1] << (64 - sh)) : o[i];
            v = av | ~ov;
            if (i == nw - 1 && (W & 63)) v &= (1ull << (W & 63)) - 1;
        }
        monos[i] = v;
    }
    return EIMP_OK;
}

// The whole pre-stage for n targets (eimp_engine_run_batch and its
// no-allocation variant): phase A (distances, top-K, monos, split sites), one
// read of the split counts, phase B (run_build).  dev_gt: targets[].gt are
// device pointers (used in place), else host words staged through pinned
// memory.
int run_batch_impl(eimp_engine *e, size_t n, const eimp_target *targets, float cMmaxSplit,
                   uint32_t flags, eimp_target_result *results, BatchMem &mem, bool dev_gt) {
    if (!targets || !results) return fail(EIMP_ERR_INVALID, "null targets/results");
    if (!e->have_cM) return fail(EIMP_ERR_STATE, "genetic positions not set (eimp_engine_set_panel_extras)");
    if (n > (1u << 24)) return fail(EIMP_ERR_UNSUPPORTED, "too many targets in one batch");
    if ((flags & EIMP_BATCH_COMPACT) && (flags & EIMP_BATCH_RAW))
        return fail(EIMP_ERR_INVALID, "EIMP_BATCH_COMPACT and EIMP_BATCH_RAW exclude each other");
    const uint32_t nt = (uint32_t)n;
    const uint64_t Nw = unit_words(e->Nad);
    // ---- sizes and validation
    std::vector<uint64_t> gt_off(nt), ww_off(nt), ss_off(nt), nerr_off(nt), bh_off(nt), bhf_off(nt),
        mono_off(nt);
    uint64_t gt_words = 0, w_words = 0, ss_elems = 0, nerr_elems = 0, bh_elems = 0, mono_words = 0;
    bool need_hap = false, need_pmono = false;
    std::vector<uint32_t> sel_list, mono_list;
    uint32_t max_ww = 0;
    e->h_tg.assign(nt, TgtDev{});
    for (uint32_t t = 0; t < nt; t++) {
        const eimp_target &x = targets[t];
        TgtDev &T = e->h_tg[t];
        CHK(target_inputs(e, x, t, cMmaxSplit, T));
        const bool select = (T.flags & TF_SELECT) != 0;
        const bool ident = (T.flags & TF_IDENTITY) != 0;
        gt_off[t] = gt_words;
        gt_words += 2 * unit_words(x.W);
        ww_off[t] = w_words;
        w_words += unit_words(x.W);
        ss_off[t] = ss_elems;
        ss_elems += round_up(x.W + 1, 64);
        nerr_off[t] = nerr_elems;
        if (select) {
            nerr_elems += round_up(e->Nad, 64);
            sel_list.push_back(t);
            need_hap = true;
        }
        bh_off[t] = bh_elems;
        bh_elems += round_up(T.K, 64);
        bhf_off[t] = (uint64_t)t * Nw;
        mono_off[t] = mono_words;
        if (ident) {
            need_pmono = true;
        } else {
            mono_words += unit_words(x.W);
            mono_list.push_back(t);
            need_hap = true;
            max_ww = std::max<uint32_t>(max_ww, (uint32_t)((x.W + 63) / 64));
        }
    }
    if (!mem.may_alloc() && ((need_hap && !e->have_hap) || (need_pmono && !e->have_pmono)))
        return fail(EIMP_ERR_STATE, "the panel's hap-major rows / mono bits are not resident: pass "
                                    "hapData to eimp_engine_set_panel_extras (or run one "
                                    "eimp_engine_run_batch with such targets) first");
    if (need_hap) CHK(ensure_hap(e));
    if (need_pmono) CHK(ensure_pmono(e));
    // ---- device buffers
    void *p_tg, *p_nerr, *p_bh, *p_bhf, *p_mono, *p_ss, *p_spl, *p_cmd, *p_stats, *p_tl;
    CHK(mem.get(MS_TG, nt * sizeof(TgtDev), "target table", &p_tg));
    CHK(mem.get(MS_NERR, std::max<uint64_t>(nerr_elems, 1) * 4, "distances", &p_nerr));
    CHK(mem.get(MS_BH, bh_elems * 4, "best haps", &p_bh));
    CHK(mem.get(MS_BHF, (uint64_t)nt * Nw * 8, "best-hap flags", &p_bhf));
    CHK(mem.get(MS_MONO, std::max<uint64_t>(mono_words, 1) * 16, "monos", &p_mono));
    CHK(mem.get(MS_SS, ss_elems * 4, "split sites", &p_ss));
    CHK(mem.get(MS_SPL, w_words * 8, "split bitmap", &p_spl));
    CHK(mem.get(MS_CMD, ss_elems * 4, "cM diffs", &p_cmd));
    CHK(mem.get(MS_STATS, (uint64_t)nt * 16, "stats", &p_stats));
    CHK(mem.get(MS_TLIST, (uint64_t)(sel_list.size() + mono_list.size() + 2) * 4, "target lists", &p_tl));
    TgtDev *d_tg = static_cast<TgtDev *>(p_tg);
    // ---- genotypes: in place (device) or one copy from pinned staging
    const uint64_t *d_gt = nullptr;
    if (!dev_gt) {
        CHK(e->d_gt.ensure(gt_words * 8, "genotypes"));
        CHK(e->h_stage.ensure(gt_words * 8));
        for (uint32_t t = 0; t < nt; t++) {
            const uint64_t w = 2 * unit_words(targets[t].W);
            std::memcpy(e->h_stage.as<uint64_t>() + gt_off[t], targets[t].gt, w * 8);
        }
        HIPCHK(hipMemcpyAsync(e->d_gt.p, e->h_stage.p, gt_words * 8, hipMemcpyHostToDevice, e->stream));
        d_gt = e->d_gt.as<uint64_t>();
    }
    uint64_t *mono_base = static_cast<uint64_t *>(p_mono);
    const uint64_t mono_total = std::max<uint64_t>(mono_words, 1);
    for (uint32_t t = 0; t < nt; t++) {
        TgtDev &T = e->h_tg[t];
        T.gt = dev_gt ? targets[t].gt : d_gt + gt_off[t];
        T.nerr = static_cast<uint32_t *>(p_nerr) + nerr_off[t];
        T.bh = static_cast<uint32_t *>(p_bh) + bh_off[t];
        T.bhflags = static_cast<uint64_t *>(p_bhf) + bhf_off[t];
        if (T.flags & TF_IDENTITY) {
            T.mand = e->pand.as<uint64_t>();
            T.mor = e->por.as<uint64_t>();
            T.mand_w = nullptr;
            T.mor_w = nullptr;
            T.mono_off = T.s0;
        } else {
            T.mand_w = mono_base + mono_off[t];
            T.mor_w = mono_base + mono_total + mono_off[t];
            T.mand = T.mand_w;
            T.mor = T.mor_w;
            T.mono_off = 0;
        }
        T.splitsites = static_cast<uint32_t *>(p_ss) + ss_off[t];
        T.splits = static_cast<uint64_t *>(p_spl) + ww_off[t];
        T.cMdiffs = static_cast<float *>(p_cmd) + ss_off[t];
        T.stats = static_cast<uint32_t *>(p_stats) + 4 * t;
    }
    std::vector<uint32_t> lists(sel_list);
    lists.insert(lists.end(), mono_list.begin(), mono_list.end());
    if (!lists.empty())
        HIPCHK(hipMemcpyAsync(p_tl, lists.data(), lists.size() * 4, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(d_tg, e->h_tg.data(), nt * sizeof(TgtDev), hipMemcpyHostToDevice, e->stream));
    const uint32_t *d_sel = static_cast<const uint32_t *>(p_tl);
    const uint32_t *d_mono_list = d_sel + sel_list.size();
    // ---- phase A
    int ti = e->tbegin(EIMP_KCLASS_DISTANCE);
    HIPCHK(launch_distance(e->pd, d_tg, d_sel, (uint32_t)sel_list.size(), e->stream));
    e->tend(ti);
    ti = e->tbegin(EIMP_KCLASS_SELECT);
    HIPCHK(launch_select(e->pd, d_tg, nt, e->stream));
    e->tend(ti);
    ti = e->tbegin(EIMP_KCLASS_MONOS);
    HIPCHK(launch_monos(e->pd, d_tg, d_mono_list, (uint32_t)mono_list.size(), max_ww, e->stream));
    e->tend(ti);
    ti = e->tbegin(EIMP_KCLASS_SPLITS);
    HIPCHK(launch_splits(e->pd, d_tg, nt, e->stream));
    e->tend(ti);
    e->h_stats.resize(4 * (size_t)nt);
    HIPCHK(hipMemcpyAsync(e->h_stats.data(), p_stats, (size_t)nt * 16, hipMemcpyDeviceToHost, e->stream));
    CHK(e->sync());
    for (uint32_t t = 0; t < nt; t++) e->h_tg[t].S = e->h_stats[4 * t];
    // ---- phase B
    const bool digest = (flags & EIMP_BATCH_DIGEST) != 0;
    CHK(run_build(e, nt, digest, batch_layout(flags), mem));
    std::vector<uint64_t> dig(digest ? 2 * (size_t)nt : 0);
    if (digest)
        HIPCHK(hipMemcpyAsync(dig.data(), e->last_dig, (size_t)nt * 16, hipMemcpyDeviceToHost, e->stream));
    CHK(e->sync());
    for (uint32_t t = 0; t < nt; t++) {
        const TgtDev &T = e->h_tg[t];
        results[t].S = T.S;
        results[t].Mp = T.Mp;
        results[t].ncalls = e->h_stats[4 * t + 1];
        results[t].K = T.K;
        results[t].digest_pbwt = digest ? dig[2 * t] : 0;
        results[t].digest_count0 = digest ? dig[2 * t + 1] : 0;
    }
    e->last_n = nt;
    return EIMP_OK;
}

}  // namespace


// ---------------------------------------------------------------------------
// Stream context: one HIP stream, fixed-capacity buffers, and the batch's
// launch sequence captured once as two hipGraphs: the preparation (distances,
// top-K, monos, call locations; the context's own buffers) and the build
// (planner, condensed reference, PBWT; the output pool).  A submission copies
// the batch header, the target table and the genotypes (one pinned staging
// buffer), replays the graphs, and copies back the split counts, the table
// (output placement) and the digests.  Nothing in the sequence depends on
// host-side sizes: kernels read the batch size from the header and the
// planner sizes the outputs on the device.
//
// Contexts may share one output pool (eimp_stream_create_shared): each
// build then waits for the pool's previous build, whichever context ran it,
// while the preparation of a submission runs beside the other context's
// build.  A context's pool outputs stay valid until a context sharing the
// pool submits again.
// ---------------------------------------------------------------------------
// Contexts sharing a pool may be driven from different host threads: `mu`
// guards seq and last_build, and is held from the wait on the previous build
// to the record of the new one, and across every synchronous read of pool
// rows, so a build is never enqueued while another context reads them.
struct StreamPool {
    DevBuf buf;
    uint64_t cap = 0;
    std::mutex mu;
    uint64_t seq = 0;                 // builds submitted into the pool
    hipEvent_t last_build = nullptr;  // the last submitted build's event (its context owns it)
};

struct eimp_stream {
    eimp_engine *e = nullptr;
    hipStream_t stream = nullptr;
    uint32_t nmax = 0, Wmax = 0, Kmax = 0, flags = 0, max_ww = 0;
    uint64_t gen = 0;
    bool timed = false;
    hipEvent_t kev[EIMP_NKCLASS + 1] = {};  // timed contexts: events between kernel classes
    hipEvent_t kev_prep_end = nullptr;      // timed contexts: end of the preparation
    double kms[EIMP_NKCLASS] = {};
    // device buffers; d_in = header (256 B) | TgtDev[nmax] | genotypes
    DevBuf d_in, d_nerr, d_bh, d_bhflags, d_mono, d_splitsites, d_splits, d_cMdiffs, d_stats,
        d_digest, d_blockpref, d_pitems, d_tmp;
    std::shared_ptr<StreamPool> pool;  // the output pool, maybe shared with other contexts
    uint64_t tab_off = 0, gt_off = 0, gt_cap_words = 0;
    HostBuf h_in, h_out;  // h_out = stats | digests | TgtDev table
    // launch configuration
    int pbwt_v = -1;
    size_t pbwt_lds = 0;
    uint32_t cref_grid[2] = {0, 0}, pbwt_grid = 0;
    hipGraph_t graph_prep = nullptr, graph_build = nullptr;
    hipGraphExec_t exec_prep = nullptr, exec_build = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t build_done = nullptr;
    // submission state
    bool pending = false, collected = false;
    uint32_t n_last = 0;
    uint64_t build_seq = 0;  // pool->seq of this context's last build

    // the pool still holds this context's last outputs (pool->mu held)
    bool outputs_current() const { return pool && build_seq == pool->seq; }

    BatchDev *d_batch() const { return d_in.as<BatchDev>(); }
    TgtDev *d_tg() const { return reinterpret_cast<TgtDev *>(d_in.as<char>() + tab_off); }
    uint32_t *h_stats() const { return h_out.as<uint32_t>(); }
    uint64_t *h_digest() const { return reinterpret_cast<uint64_t *>(h_out.as<char>() + 16ull * nmax); }
    TgtDev *h_tg() const { return reinterpret_cast<TgtDev *>(h_out.as<char>() + 32ull * nmax); }
};

namespace {

const char *const kPoolReused =
    "the output pool holds a later submission of a context sharing it";

// the preparation: a3..a6 in the context's own buffers
int stream_launch_prep(eimp_stream *s) {
    eimp_engine *e = s->e;
    const uint32_t *d_n = &s->d_batch()->n;
    for (int cls = EIMP_KCLASS_DISTANCE; cls <= EIMP_KCLASS_SPLITS; cls++) {
        if (s->timed) HIPCHK(hipEventRecord(s->kev[cls], s->stream));
        HIPCHK(launch_prep_dev(cls, e->pd, s->d_tg(), d_n, s->nmax, s->max_ww, s->stream));
    }
    if (s->timed) HIPCHK(hipEventRecord(s->kev_prep_end, s->stream));
    return EIMP_OK;
}

// the build: planner, a7, a9 (and the count0 digest) into the output pool
int stream_launch_build(eimp_stream *s) {
    eimp_engine *e = s->e;
    const uint32_t *d_n = &s->d_batch()->n;
    PlanArgs A{};
    A.pool = s->pool->buf.as<char>();
    A.pool_cap = s->pool->cap;
    A.dig_base = s->d_digest.as<uint64_t>();
    A.blockpref = s->d_blockpref.as<uint32_t>();
    A.blockpref_id = s->d_blockpref.as<uint32_t>() + s->nmax + 1;
    A.ident_ok = cref_ident_ok(e->pd) ? 1u : 0u;
    A.pitems = s->d_pitems.as<PbwtItem>();
    A.layout = batch_layout(s->flags);
    A.digest = (s->flags & EIMP_BATCH_DIGEST) ? 1u : 0u;
    // timed contexts record an event before each kernel class (the planner
    // counts with the cref class, the count0 digest with the PBWT class)
    auto mark = [&](int cls) -> int {
        if (s->timed) HIPCHK(hipEventRecord(s->kev[cls], s->stream));
        return EIMP_OK;
    };
    CHK(mark(EIMP_KCLASS_CREF));
    HIPCHK(launch_plan(s->d_tg(), s->d_batch(), A, s->stream));
    HIPCHK(launch_cref_persist(e->pd, s->d_tg(), s->d_batch(), A.blockpref, A.blockpref_id, s->Kmax,
                               s->cref_grid, s->stream));
    CHK(mark(EIMP_KCLASS_PBWT));
    HIPCHK(launch_pbwt_persist(s->pbwt_v, s->d_tg(), s->d_batch(), s->d_pitems.as<PbwtItem>(),
                               s->pbwt_grid, s->pbwt_lds, s->stream));
    if (A.digest) HIPCHK(launch_digest_count0_dev(s->d_tg(), d_n, s->nmax, s->stream));
    CHK(mark(EIMP_KCLASS_COPY));
    return EIMP_OK;
}

void stream_release(eimp_stream *s) {
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->pool) {
        {
            // the build event dies with this context: later builds need not wait for it
            std::lock_guard<std::mutex> lk(s->pool->mu);
            if (s->build_done && s->pool->last_build == s->build_done) s->pool->last_build = nullptr;
        }
        if (s->pool.use_count() == 1) s->pool->buf.release();
        s->pool.reset();
    }
    for (hipGraphExec_t x : {s->exec_prep, s->exec_build})
        if (x) (void)hipGraphExecDestroy(x);
    for (hipGraph_t g : {s->graph_prep, s->graph_build})
        if (g) (void)hipGraphDestroy(g);
    for (auto &ev : s->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &ev : s->kev)
        if (ev) (void)hipEventDestroy(ev);
    if (s->kev_prep_end) (void)hipEventDestroy(s->kev_prep_end);
    if (s->build_done) (void)hipEventDestroy(s->build_done);
    DevBuf *bufs[] = {&s->d_in, &s->d_nerr, &s->d_bh, &s->d_bhflags, &s->d_mono, &s->d_splitsites,
                      &s->d_splits, &s->d_cMdiffs, &s->d_stats, &s->d_digest, &s->d_blockpref,
                      &s->d_pitems, &s->d_tmp};
    for (auto *b : bufs) b->release();
    s->h_in.release();
    s->h_out.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// capture one part of the launch sequence as a graph
int stream_capture(eimp_stream *s, int (*part)(eimp_stream *), hipGraph_t *graph,
                   hipGraphExec_t *exec) {
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int rc = part(s);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s->stream, &g);
    if (rc != EIMP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(EIMP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    *graph = g;
    HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return EIMP_OK;
}

// pool: an existing pool to share, or null to allocate one of pool_bytes
int stream_setup(eimp_stream *s, size_t pool_bytes, const std::shared_ptr<StreamPool> &pool) {
    eimp_engine *e = s->e;
    HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (auto &ev : s->ev) HIPCHK(hipEventCreate(&ev));
    HIPCHK(hipEventCreateWithFlags(&s->build_done, hipEventDisableTiming));
    if (s->timed) {
        for (auto &ev : s->kev) HIPCHK(hipEventCreate(&ev));
        HIPCHK(hipEventCreate(&s->kev_prep_end));
    }
    CHK(ensure_hap(e));
    if (s->Kmax == e->Nad && e->Ncorr == e->Nad) CHK(ensure_pmono(e));
    s->gen = e->panel_gen;
    const uint64_t nmax = s->nmax, Nw = unit_words(e->Nad), Ww = unit_words(s->Wmax);
    s->max_ww = (uint32_t)((s->Wmax + 63) / 64);
    s->tab_off = 256;
    s->gt_off = round_up(s->tab_off + nmax * sizeof(TgtDev), 256);
    s->gt_cap_words = nmax * 2 * Ww;
    const uint64_t in_bytes = s->gt_off + s->gt_cap_words * 8;
    CHK(s->d_in.ensure(in_bytes, "stream inputs"));
    CHK(s->h_in.ensure(in_bytes));
    std::memset(s->h_in.p, 0, s->gt_off);
    // any target with K < Ncorr selects (K varies per target up to Kmax)
    CHK(s->d_nerr.ensure(nmax * round_up(e->Nad, 64) * 4, "distances"));
    CHK(s->d_bh.ensure(nmax * round_up(s->Kmax, 64) * 4, "best haps"));
    CHK(s->d_bhflags.ensure(nmax * Nw * 8, "best-hap flags"));
    CHK(s->d_mono.ensure(nmax * 2 * Ww * 8, "monos"));
    CHK(s->d_splitsites.ensure(nmax * round_up(s->Wmax + 1, 64) * 4, "split sites"));
    CHK(s->d_splits.ensure(nmax * Ww * 8, "split bitmap"));
    CHK(s->d_cMdiffs.ensure(nmax * round_up(s->Wmax + 1, 64) * 4, "cM diffs"));
    CHK(s->d_stats.ensure(nmax * 16, "stats"));
    CHK(s->d_digest.ensure(nmax * 16, "digests"));
    CHK(s->d_blockpref.ensure(2 * (nmax + 1) * 4, "block prefix"));
    CHK(s->d_pitems.ensure(2 * nmax * sizeof(PbwtItem), "pbwt items"));
    if (pool) {
        s->pool = pool;
    } else {
        s->pool = std::make_shared<StreamPool>();
        CHK(s->pool->buf.ensure(pool_bytes, "output pool", false));
        s->pool->cap = pool_bytes;
    }
    CHK(s->h_out.ensure(32 * nmax + nmax * sizeof(TgtDev)));
    // kernels and grids
    const uint32_t Kpad = (uint32_t)round_up(s->Kmax, 64);
    if (s->Kmax > cref_max_k()) return fail(EIMP_ERR_UNSUPPORTED, "K beyond the cref kernel limit");
    if (cref_lds_bytes(e->pd, s->Kmax) > 160 * 1024)
        return fail(EIMP_ERR_UNSUPPORTED, "panel too wide for the cref kernel's LDS staging");
    int v = e->choose_pbwt_variant(Kpad);
    if (v < 0) return v;
    const size_t lds = pbwt_lds_bytes(v, Kpad);
    if (lds > 160 * 1024)
        return fail(EIMP_ERR_UNSUPPORTED, "K beyond the PBWT kernel's slots or 160 KiB of LDS");
    s->pbwt_v = v;
    s->pbwt_lds = lds;
    int ncu = 0, occ_c[2] = {0, 0}, occ_p = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, e->dev));
    HIPCHK(cref_persist_prepare(e->pd, s->Kmax, occ_c));
    HIPCHK(pbwt_persist_prepare(v, lds, &occ_p));
    s->cref_grid[0] = (uint32_t)(std::max(ncu, 1) * std::max(occ_c[0], 1));
    // identity targets need K == N_ad
    const bool ident_possible = cref_ident_ok(e->pd) && s->Kmax == e->Nad && e->Ncorr == e->Nad;
    s->cref_grid[1] = ident_possible ? (uint32_t)(std::max(ncu, 1) * std::max(occ_c[1], 1)) : 0u;
    s->pbwt_grid = (uint32_t)std::min<uint64_t>(2 * nmax, (uint64_t)std::max(ncu, 1) * std::max(occ_p, 1));
    // fixed per-slot pointers of the target table
    TgtDev *tab = reinterpret_cast<TgtDev *>(s->h_in.as<char>() + s->tab_off);
    for (uint64_t t = 0; t < nmax; t++) {
        TgtDev &T = tab[t];
        T.nerr = s->d_nerr.as<uint32_t>() + t * round_up(e->Nad, 64);
        T.bh = s->d_bh.as<uint32_t>() + t * round_up(s->Kmax, 64);
        T.bhflags = s->d_bhflags.as<uint64_t>() + t * Nw;
        T.mand_w = s->d_mono.as<uint64_t>() + t * 2 * Ww;
        T.mor_w = T.mand_w + Ww;
        T.splitsites = s->d_splitsites.as<uint32_t>() + t * round_up(s->Wmax + 1, 64);
        T.splits = s->d_splits.as<uint64_t>() + t * Ww;
        T.cMdiffs = s->d_cMdiffs.as<float>() + t * round_up(s->Wmax + 1, 64);
        T.stats = s->d_stats.as<uint32_t>() + 4 * t;
    }
    if (!(s->flags & (EIMP_STREAM_NO_GRAPH | EIMP_STREAM_TIMED))) {
        CHK(stream_capture(s, stream_launch_prep, &s->graph_prep, &s->exec_prep));
        CHK(stream_capture(s, stream_launch_build, &s->graph_build, &s->exec_build));
    }
    return EIMP_OK;
}

int stream_create(eimp_engine *e, size_t max_targets, size_t max_window, size_t max_K,
                  size_t pool_bytes, uint32_t flags, const std::shared_ptr<StreamPool> &pool,
                  eimp_stream **out) {
    if (!out) return fail(EIMP_ERR_INVALID, "null stream pointer");
    *out = nullptr;
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    if (!e->have_cM) return fail(EIMP_ERR_STATE, "genetic positions not set (eimp_engine_set_panel_extras)");
    if (max_targets < 1 || max_targets > 1024)
        return fail(EIMP_ERR_INVALID, "max_targets must be in [1, 1024]");
    if (max_window < 1 || max_window > e->M) return fail(EIMP_ERR_INVALID, "max_window must be in [1, M]");
    if (max_K < 1) return fail(EIMP_ERR_INVALID, "max_K < 1");
    if (!pool && pool_bytes < 256) return fail(EIMP_ERR_INVALID, "pool_bytes too small");
    if (flags & ~(EIMP_BATCH_DIGEST | EIMP_BATCH_COMPACT | EIMP_BATCH_RAW | EIMP_STREAM_NO_GRAPH |
                  EIMP_STREAM_TIMED))
        return fail(EIMP_ERR_INVALID, "unknown stream flags");
    if ((flags & EIMP_BATCH_COMPACT) && (flags & EIMP_BATCH_RAW))
        return fail(EIMP_ERR_INVALID, "EIMP_BATCH_COMPACT and EIMP_BATCH_RAW exclude each other");
    CHK(set_device(e));
    eimp_stream *s = new eimp_stream();
    s->e = e;
    s->nmax = (uint32_t)max_targets;
    s->Wmax = (uint32_t)max_window;
    s->Kmax = (uint32_t)std::min<uint64_t>(max_K, e->Ncorr);
    s->flags = flags;
    s->timed = (flags & EIMP_STREAM_TIMED) != 0;
    const int rc = stream_setup(s, pool_bytes, pool);
    if (rc != EIMP_OK) {
        const std::string msg = g_err;
        stream_release(s);
        delete s;
        (void)hipGetLastError();
        return fail(rc, msg);
    }
    *out = s;
    return EIMP_OK;
}

}  // namespace

namespace eimp {
// error message for the C ABI entry points of other translation units
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace eimp

extern "C" {

const char *eimp_last_error(void) { return g_err.c_str(); }

int eimp_device_count(int *count) {
    if (!count) return fail(EIMP_ERR_INVALID, "null count");
    int c = 0;
    hipError_t err = hipGetDeviceCount(&c);
    if (err != hipSuccess) {
        *count = 0;
        (void)hipGetLastError();
        return fail(EIMP_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(err));
    }
    *count = c;
    return EIMP_OK;
}

int eimp_engine_create(int device_index, eimp_engine **engine) {
    if (!engine) return fail(EIMP_ERR_INVALID, "null engine pointer");
    *engine = nullptr;
    int c = 0;
    CHK(eimp_device_count(&c));
    if (device_index < 0 || device_index >= c)
        return fail(EIMP_ERR_INVALID, "device index " + std::to_string(device_index) +
                                          " out of range (" + std::to_string(c) + " devices)");
    eimp_engine *e = new eimp_engine();
    e->dev = device_index;
    if (set_device(e) != EIMP_OK || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(EIMP_ERR_HIP, "stream creation failed");
    }
    // this device's copy of the no-allocation scratch (device current)
    const int rc = scratch_for_device(device_index, &e->scratch);
    if (rc != EIMP_OK) {
        const std::string msg = g_err;
        (void)hipStreamDestroy(e->stream);
        delete e;
        return fail(rc, msg);
    }
    *engine = e;
    return EIMP_OK;
}

int eimp_engine_free(eimp_engine *e) {
    if (!e) return fail(EIMP_ERR_INVALID, "null engine");
    CHK(set_device(e));
    (void)hipStreamSynchronize(e->stream);
    e->tcollect();
    e->release_pools();
    e->release_panel();
    e->initialized = false;
    return EIMP_OK;
}

int eimp_engine_destroy(eimp_engine *e) {
    if (!e) return EIMP_OK;
    eimp_engine_free(e);
    for (auto ev : e->event_pool) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return EIMP_OK;
}

int eimp_engine_initialize(eimp_engine *e, size_t K, size_t Kwords, size_t M, size_t Nrefhapscorr,
                           size_t NrefhapsAD, size_t refTCapacity, const uint64_t *refTData) {
    if (!e) return fail(EIMP_ERR_INVALID, "null engine");
    if (!refTData || M == 0 || NrefhapsAD == 0) return fail(EIMP_ERR_INVALID, "empty panel");
    if (Nrefhapscorr > NrefhapsAD) return fail(EIMP_ERR_INVALID, "Nrefhapscorr > NrefhapsAD");
    if (K < 1 || K > Nrefhapscorr) return fail(EIMP_ERR_INVALID, "K must be in [1, Nrefhapscorr]");
    if (Kwords != unit_words(K))
        return fail(EIMP_ERR_INVALID, "Kwords must be roundToMultiple(K,512)/64");
    if (refTCapacity % 64 || refTCapacity * 8 < NrefhapsAD)
        return fail(EIMP_ERR_INVALID, "refTCapacity must be a multiple of 64 bytes holding NrefhapsAD bits");
    if (NrefhapsAD >= (1u << 31) || M >= (1u << 31)) return fail(EIMP_ERR_UNSUPPORTED, "panel too large");
    CHK(set_device(e));
    if (e->initialized) {
        (void)hipStreamSynchronize(e->stream);
        e->release_panel();
    }
    e->panel_gen++;
    e->K = K;
    e->Kwords = Kwords;
    e->M = M;
    e->Ncorr = Nrefhapscorr;
    e->Nad = NrefhapsAD;
    e->refTCap = refTCapacity;
    const size_t bytes = M * refTCapacity;
    CHK(e->site.ensure(bytes, "site-major panel"));
    HIPCHK(hipMemcpy(e->site.p, refTData, bytes, hipMemcpyHostToDevice));
    e->pd = PanelDev{};
    e->pd.site = e->site.as<uint64_t>();
    e->pd.site_stride = refTCapacity / 8;
    e->pd.N_ad = (uint32_t)NrefhapsAD;
    e->pd.N_corr = (uint32_t)Nrefhapscorr;
    e->pd.M = (uint32_t)M;
    e->initialized = true;
    return EIMP_OK;
}

int eimp_engine_block_size(eimp_engine *e, size_t *bs) {
    if (!e || !bs) return fail(EIMP_ERR_INVALID, "null argument");
    // Targets per call: the reference uses 256 (src/GPUEngine.cu:66); the
    // engine takes any number, 256 keeps the drop-in's host buffers bounded.
    *bs = 256;
    return EIMP_OK;
}

int eimp_engine_device_index(eimp_engine *e, int *index) {
    if (!e || !index) return fail(EIMP_ERR_INVALID, "null argument");
    *index = e->dev;
    return EIMP_OK;
}

int eimp_engine_set_panel_extras(eimp_engine *e, const float *cM, const uint64_t *hapData,
                                 size_t hapStride, const uint8_t *haploidRef) {
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    CHK(set_device(e));
    e->panel_gen++;
    if (cM) {
        CHK(e->cM.ensure(e->M * sizeof(float), "cM"));
        HIPCHK(hipMemcpy(e->cM.p, cM, e->M * sizeof(float), hipMemcpyHostToDevice));
        e->pd.cM = e->cM.as<float>();
        e->have_cM = true;
    }
    if (hapData) {
        if (hapStride < unit_words(e->M) || hapStride % 8)
            return fail(EIMP_ERR_INVALID, "hapStride must be >= roundToMultiple(M,512)/64 and a multiple of 8");
        CHK(e->hap.ensure(e->Nad * hapStride * 8, "hap-major panel"));
        HIPCHK(hipMemcpy(e->hap.p, hapData, e->Nad * hapStride * 8, hipMemcpyHostToDevice));
        e->hap_stride = hapStride;
        e->pd.hap = e->hap.as<uint64_t>();
        e->pd.hap_stride = hapStride;
        e->have_hap = true;
    }
    if (haploidRef) {
        size_t ns = e->Nad / 2;
        CHK(e->haploid.ensure(ns ? ns : 1, "haploid flags"));
        HIPCHK(hipMemcpy(e->haploid.p, haploidRef, ns, hipMemcpyHostToDevice));
        e->pd.haploid = e->haploid.as<uint8_t>();
    }
    // per-site all-ones / any-one bits of the whole panel (monos of
    // whole-panel targets), built here so that later batches need no allocation
    CHK(ensure_pmono(e));
    return EIMP_OK;
}

int eimp_engine_run_batch_noalloc(eimp_engine *e, size_t n, const eimp_target *targets,
                                  float cMmaxSplit, uint32_t flags, size_t max_Mp,
                                  uint32_t *const *rows, int32_t *const *count0,
                                  eimp_target_result *results) {
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    if (n == 0) return EIMP_OK;
    if (n > kNoAllocMaxTargets)
        return fail(EIMP_ERR_UNSUPPORTED, "at most " + std::to_string(kNoAllocMaxTargets) +
                                              " targets per no-allocation batch");
    if (!rows || !count0) return fail(EIMP_ERR_INVALID, "null output arrays");
    if (!e->scratch) return fail(EIMP_ERR_STATE, "no scratch address for this device");
    CHK(set_device(e));
    // event timing would create events: off for this call
    const bool timing = e->timing;
    e->timing = false;
    ScratchMem mem(e->scratch, rows, count0, max_Mp);
    const int rc = run_batch_impl(e, n, targets, cMmaxSplit, flags, results, mem, true);
    e->timing = timing;
    return rc;
}

int eimp_engine_run_batch(eimp_engine *e, size_t n, const eimp_target *targets, float cMmaxSplit,
                          uint32_t flags, eimp_target_result *results) {
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    if (n == 0) return EIMP_OK;
    CHK(set_device(e));
    PoolMem mem(e);
    return run_batch_impl(e, n, targets, cMmaxSplit, flags, results, mem, false);
}

int eimp_engine_run_cref_pbwt(eimp_engine *e, size_t n, const uint64_t *const *tgtData,
                              const uint64_t *const *splitdata, const uint64_t *const *besthapdata,
                              const uint32_t *crefTOffsets, uint32_t *const *cPBWT,
                              int32_t *const *count0s) {
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    if (n == 0) return EIMP_OK;
    if (!tgtData || !splitdata || !besthapdata || !crefTOffsets || !cPBWT || !count0s)
        return fail(EIMP_ERR_INVALID, "null argument");
    CHK(set_device(e));
    const uint32_t nt = (uint32_t)n;
    const uint64_t M = e->M;
    const uint64_t gtw = 2 * unit_words(M);                 // our per-target GT stride (words)
    const uint64_t gt_read = round_up(2 * M, 512) / 64;     // words the reference copies (tgtGTSize)
    const uint64_t spw = unit_words(M);                     // split bitmap words
    const uint64_t bhw = e->refTCap / 8;                    // best-hap bitmap words (refTCapacity)
    for (uint32_t t = 0; t < nt; t++) {
        if (crefTOffsets[t + 1] < crefTOffsets[t] || ((crefTOffsets[t + 1] - crefTOffsets[t]) % 2) == 0)
            return fail(EIMP_ERR_INVALID, "crefTOffsets must grow by odd Mp = 2*S+1");
    }
    e->h_tg.assign(nt, TgtDev{});
    const uint64_t stage_words = (uint64_t)nt * (gtw + spw + bhw);
    CHK(e->h_stage.ensure(stage_words * 8));
    uint64_t *st = e->h_stage.as<uint64_t>();
    std::memset(st, 0, stage_words * 8);
    for (uint32_t t = 0; t < nt; t++) {
        std::memcpy(st + (uint64_t)t * gtw, tgtData[t], std::min(gt_read, gtw) * 8);
        std::memcpy(st + (uint64_t)nt * gtw + (uint64_t)t * spw, splitdata[t], spw * 8);
        std::memcpy(st + (uint64_t)nt * (gtw + spw) + (uint64_t)t * bhw, besthapdata[t], bhw * 8);
    }
    CHK(e->d_tg.ensure(nt * sizeof(TgtDev), "target table"));
    CHK(e->d_gt.ensure(stage_words * 8, "drop-in inputs"));
    CHK(e->d_bh.ensure((uint64_t)nt * round_up(e->K, 64) * 4, "best haps"));
    CHK(e->d_splitsites.ensure((uint64_t)nt * round_up(M + 1, 64) * 4, "split sites"));
    CHK(e->d_stats.ensure((uint64_t)nt * 16, "stats"));
    HIPCHK(hipMemcpyAsync(e->d_gt.p, st, stage_words * 8, hipMemcpyHostToDevice, e->stream));
    const bool ident = e->K == e->Nad;
    for (uint32_t t = 0; t < nt; t++) {
        TgtDev &T = e->h_tg[t];
        T.gt = e->d_gt.as<uint64_t>() + (uint64_t)t * gtw;
        T.s0 = 0;
        T.W = (uint32_t)M;
        T.K = (uint32_t)e->K;
        T.ignore = -1;
        T.flags = ident ? TF_IDENTITY : 0u;
        T.splits = e->d_gt.as<uint64_t>() + (uint64_t)nt * gtw + (uint64_t)t * spw;
        T.bhflags = e->d_gt.as<uint64_t>() + (uint64_t)nt * (gtw + spw) + (uint64_t)t * bhw;
        T.bh = e->d_bh.as<uint32_t>() + (uint64_t)t * round_up(e->K, 64);
        T.splitsites = e->d_splitsites.as<uint32_t>() + (uint64_t)t * round_up(M + 1, 64);
        T.stats = e->d_stats.as<uint32_t>() + 4 * t;
    }
    HIPCHK(hipMemcpyAsync(e->d_tg.p, e->h_tg.data(), nt * sizeof(TgtDev), hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(launch_bits_to_list(e->d_tg.as<TgtDev>(), nt, (uint32_t)e->Nad, 0, e->stream));
    HIPCHK(launch_bits_to_list(e->d_tg.as<TgtDev>(), nt, (uint32_t)e->Nad, 1, e->stream));
    e->h_stats.resize(4 * (size_t)nt);
    HIPCHK(hipMemcpyAsync(e->h_stats.data(), e->d_stats.p, (size_t)nt * 16, hipMemcpyDeviceToHost,
                          e->stream));
    CHK(e->sync());
    for (uint32_t t = 0; t < nt; t++) {
        const uint32_t S = e->h_stats[4 * t], nbh = e->h_stats[4 * t + 3];
        if (nbh != e->K)
            return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": best-hap bitmap has " +
                                              std::to_string(nbh) + " bits set, K = " + std::to_string(e->K));
        if (2ull * S + 1 != crefTOffsets[t + 1] - crefTOffsets[t])
            return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) +
                                              ": crefTOffsets disagree with the split bitmap");
        e->h_tg[t].S = S;
    }
    PoolMem mem(e);
    CHK(run_build(e, nt, false, LAYOUT_GPU, mem));
    // results to the caller's host buffers, as src/GPUEngine.cu:182-211
    for (uint32_t t = 0; t < nt; t++) {
        const TgtDev &T = e->h_tg[t];
        const int ti = e->tbegin(EIMP_KCLASS_COPY);
        HIPCHK(hipMemcpyAsync(cPBWT[t], T.out, 2ull * T.Mp * T.RS * 4, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipMemcpyAsync(count0s[t], T.count0, (size_t)T.Mp * 4, hipMemcpyDeviceToHost, e->stream));
        e->tend(ti);
    }
    CHK(e->sync());
    e->last_n = nt;
    return EIMP_OK;
}

int eimp_engine_copy_target(eimp_engine *e, size_t t, uint32_t *besthaps, uint64_t *bhflags,
                            uint32_t *splitsites, uint64_t *splits, float *cMdiffs, int32_t *count0,
                            uint32_t *pbwt, uint64_t *cref_rows) {
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    if (t >= e->last_n || t >= e->h_tg.size()) return fail(EIMP_ERR_INVALID, "target index out of range");
    CHK(set_device(e));
    return copy_target_out(e->h_tg[t], e->last_tg, (uint32_t)t, e->Nad, e->stream, e->d_tmp,
                           besthaps, bhflags, splitsites, splits, cMdiffs, count0, pbwt, cref_rows);
}

int eimp_engine_copy_monos(eimp_engine *e, size_t t, uint64_t *monos) {
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    if (t >= e->last_n || t >= e->h_tg.size()) return fail(EIMP_ERR_INVALID, "target index out of range");
    if (!monos) return fail(EIMP_ERR_INVALID, "null monos");
    CHK(set_device(e));
    const TgtDev &T = e->h_tg[t];
    return copy_monos_out(T, T.mand == T.mand_w ? unit_words(T.W) : unit_words(e->M), monos);
}

int eimp_engine_ref_pbwt(eimp_engine *e, size_t m0, size_t nsites, uint32_t *rows, int32_t *count0,
                         int32_t *absperm, double *device_ms) {
    if (!e || !e->initialized) return fail(EIMP_ERR_STATE, "engine not initialized");
    if (!rows || !count0) return fail(EIMP_ERR_INVALID, "null rows/count0");
    if (nsites < 1 || m0 >= e->M || nsites > e->M - m0)
        return fail(EIMP_ERR_INVALID, "site range outside the panel");
    if (nsites >= (1u << 31)) return fail(EIMP_ERR_UNSUPPORTED, "too many sites");
    const uint64_t N = e->Nad, Kpad = round_up(N, 64);
    // panels up to one workgroup's LDS run the LDS-resident kernel; wider ones
    // the large-K kernel (order in global memory)
    const bool big = Kpad > panel_pbwt_max_k();
    if (big && (N > pbwt_big_max_k() || pbwt_big_lds_bytes((uint32_t)N) > 160 * 1024))
        return fail(EIMP_ERR_UNSUPPORTED, "panel wider than " + std::to_string(pbwt_big_max_k()) +
                                              " haplotypes");
    CHK(set_device(e));
    const uint64_t nblk = (nsites + 15) / 16, RS = 2 * ((N + 31) / 32);
    if (!big) CHK(e->p_blocks.ensure(nblk * Kpad * 2, "panel PBWT blocks"));
    CHK(e->p_rows.ensure(nsites * RS * 4, "panel PBWT rows"));
    CHK(e->p_c0.ensure(nsites * 4, "panel PBWT count0"));
    if (absperm) CHK(e->p_abs.ensure((nsites + 1) * N * 4, "panel PBWT permutations"));
    if (big) {
        CHK(e->d_ord.ensure(2 * Kpad * 4, "panel PBWT orders"));
        CHK(e->d_big.ensure(sizeof(BigItem), "panel PBWT item"));
        BigItem B{};
        B.rows = e->site.as<uint64_t>() + m0 * (e->refTCap / 8);
        B.rstride = e->refTCap / 8;
        B.out = e->p_rows.as<uint32_t>();
        B.count0 = e->p_c0.as<int32_t>();
        B.absperm = absperm ? e->p_abs.as<int32_t>() : nullptr;
        B.ord0 = e->d_ord.as<uint32_t>();
        B.ord1 = B.ord0 + Kpad;
        B.K = (uint32_t)N;
        B.Mp = (uint32_t)nsites;
        B.RS = (uint32_t)RS;
        HIPCHK(hipMemcpyAsync(e->d_big.p, &B, sizeof(BigItem), hipMemcpyHostToDevice, e->stream));
    }
    hipEvent_t a = e->get_event(), b = e->get_event();
    if (a) (void)hipEventRecord(a, e->stream);
    if (big)
        HIPCHK(launch_pbwt_big(e->d_big.as<BigItem>(), 1, (uint32_t)N, e->stream));
    else
        HIPCHK(launch_panel_pbwt(e->pd, m0, (uint32_t)nsites, e->p_blocks.as<uint16_t>(), e->p_rows.as<uint32_t>(),
                                 e->p_c0.as<int32_t>(), absperm ? e->p_abs.as<int32_t>() : nullptr, e->stream));
    if (b) (void)hipEventRecord(b, e->stream);
    HIPCHK(hipStreamSynchronize(e->stream));
    if (device_ms) {
        float ms = 0;
        if (a && b) (void)hipEventElapsedTime(&ms, a, b);
        *device_ms = ms;
    }
    if (a) e->event_pool.push_back(a);
    if (b) e->event_pool.push_back(b);
    HIPCHK(hipMemcpy(rows, e->p_rows.p, nsites * RS * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(count0, e->p_c0.p, nsites * 4, hipMemcpyDeviceToHost));
    if (absperm) HIPCHK(hipMemcpy(absperm, e->p_abs.p, (nsites + 1) * N * 4, hipMemcpyDeviceToHost));
    return EIMP_OK;
}

int eimp_engine_set_timing(eimp_engine *e, int enable) {
    if (!e) return fail(EIMP_ERR_INVALID, "null engine");
    e->timing = enable != 0;
    return EIMP_OK;
}

int eimp_engine_get_timing(eimp_engine *e, double *ms, uint64_t *launches) {
    if (!e) return fail(EIMP_ERR_INVALID, "null engine");
    for (int i = 0; i < EIMP_NKCLASS; i++) {
        if (ms) ms[i] = e->ms[i];
        if (launches) launches[i] = e->launches[i];
    }
    return EIMP_OK;
}

int eimp_engine_reset_timing(eimp_engine *e) {
    if (!e) return fail(EIMP_ERR_INVALID, "null engine");
    for (int i = 0; i < EIMP_NKCLASS; i++) {
        e->ms[i] = 0;
        e->launches[i] = 0;
    }
    return EIMP_OK;
}

int eimp_engine_set_pbwt_variant(eimp_engine *e, int variant) {
    if (!e) return fail(EIMP_ERR_INVALID, "null engine");
    if (variant >= pbwt_variant_count()) return fail(EIMP_ERR_INVALID, "variant out of range");
    e->pbwt_variant = variant < 0 ? -1 : variant;
    return EIMP_OK;
}

int eimp_pbwt_variant_count(void) { return pbwt_variant_count(); }

int eimp_pbwt_variant_info(int v, int *threads, int *cpw, int *max_k) {
    if (v < 0 || v >= pbwt_variant_count()) return fail(EIMP_ERR_INVALID, "variant out of range");
    if (threads) *threads = pbwt_variant_threads(v);
    if (cpw) *cpw = pbwt_variant_cpw(v);
    if (max_k) *max_k = (int)pbwt_variant_maxk(v);
    return EIMP_OK;
}


int eimp_engine_mem_info(eimp_engine *e, size_t *free_bytes, size_t *total_bytes) {
    if (!e) return fail(EIMP_ERR_INVALID, "null engine");
    CHK(set_device(e));
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return EIMP_OK;
}

int eimp_stream_create(eimp_engine *e, size_t max_targets, size_t max_window, size_t max_K,
                       size_t pool_bytes, uint32_t flags, eimp_stream **out) {
    return stream_create(e, max_targets, max_window, max_K, pool_bytes, flags, nullptr, out);
}

int eimp_stream_create_shared(eimp_stream *other, size_t max_targets, size_t max_window,
                              size_t max_K, uint32_t flags, eimp_stream **out) {
    if (out) *out = nullptr;
    if (!other) return fail(EIMP_ERR_INVALID, "null stream to share the pool of");
    return stream_create(other->e, max_targets, max_window, max_K, 0, flags, other->pool, out);
}

int eimp_stream_submit(eimp_stream *s, size_t n, const eimp_target *targets, float cMmaxSplit) {
    if (!s) return fail(EIMP_ERR_INVALID, "null stream");
    if (s->pending) return fail(EIMP_ERR_STATE, "previous submission not collected (eimp_stream_wait)");
    eimp_engine *e = s->e;
    if (s->gen != e->panel_gen || !e->initialized)
        return fail(EIMP_ERR_STATE, "panel changed since the context was created");
    if (n > s->nmax) return fail(EIMP_ERR_INVALID, "more targets than the context's max_targets");
    if (n && !targets) return fail(EIMP_ERR_INVALID, "null targets");
    CHK(set_device(e));
    const uint32_t nt = (uint32_t)n;
    BatchDev *hb = s->h_in.as<BatchDev>();
    *hb = BatchDev{};
    hb->n = nt;
    TgtDev *tab = reinterpret_cast<TgtDev *>(s->h_in.as<char>() + s->tab_off);
    uint64_t *gtst = reinterpret_cast<uint64_t *>(s->h_in.as<char>() + s->gt_off);
    uint64_t *gtdev = reinterpret_cast<uint64_t *>(s->d_in.as<char>() + s->gt_off);
    uint64_t gw = 0;
    for (uint32_t t = 0; t < nt; t++) {
        const eimp_target &x = targets[t];
        TgtDev &T = tab[t];
        CHK(target_inputs(e, x, t, cMmaxSplit, T));
        if (batch_layout(s->flags) == LAYOUT_RAW) T.flags |= TF_RAW;
        if (x.W > s->Wmax) return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": window beyond max_window");
        if (T.K > s->Kmax) return fail(EIMP_ERR_INVALID, "target " + std::to_string(t) + ": K beyond max_K");
        const uint64_t w = 2 * unit_words(x.W);
        std::memcpy(gtst + gw, x.gt, w * 8);
        T.gt = gtdev + gw;
        gw += w;
        if (T.flags & TF_IDENTITY) {
            T.mand = e->pand.as<uint64_t>();
            T.mor = e->por.as<uint64_t>();
            T.mono_off = T.s0;
        } else {
            T.mand = T.mand_w;
            T.mor = T.mor_w;
            T.mono_off = 0;
        }
        T.cref = nullptr;
        T.count0 = nullptr;
        T.out = nullptr;
        T.digest = nullptr;
        T.Mp = T.Kpad = T.RS = T.S = 0;
    }
    HIPCHK(hipEventRecord(s->ev[0], s->stream));
    HIPCHK(hipMemcpyAsync(s->d_in.p, s->h_in.p, s->tab_off + (size_t)nt * sizeof(TgtDev),
                          hipMemcpyHostToDevice, s->stream));
    if (gw) HIPCHK(hipMemcpyAsync(gtdev, gtst, gw * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipEventRecord(s->ev[1], s->stream));
    if (nt) {
        if (s->exec_prep) HIPCHK(hipGraphLaunch(s->exec_prep, s->stream));
        else CHK(stream_launch_prep(s));
        // the pool's previous build (this context's or another's) ends first
        std::lock_guard<std::mutex> lk(s->pool->mu);
        if (s->pool->last_build) HIPCHK(hipStreamWaitEvent(s->stream, s->pool->last_build, 0));
        if (s->exec_build) HIPCHK(hipGraphLaunch(s->exec_build, s->stream));
        else CHK(stream_launch_build(s));
        HIPCHK(hipEventRecord(s->build_done, s->stream));
        s->pool->last_build = s->build_done;
        s->build_seq = ++s->pool->seq;
    }
    HIPCHK(hipEventRecord(s->ev[2], s->stream));
    if (nt) {
        HIPCHK(hipMemcpyAsync(s->h_stats(), s->d_stats.p, (size_t)nt * 16, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipMemcpyAsync(s->h_tg(), s->d_tg(), (size_t)nt * sizeof(TgtDev), hipMemcpyDeviceToHost,
                              s->stream));
        if (s->flags & EIMP_BATCH_DIGEST)
            HIPCHK(hipMemcpyAsync(s->h_digest(), s->d_digest.p, (size_t)nt * 16, hipMemcpyDeviceToHost,
                                  s->stream));
    }
    HIPCHK(hipEventRecord(s->ev[3], s->stream));
    s->pending = true;
    s->collected = false;
    s->n_last = nt;
    return EIMP_OK;
}

int eimp_stream_query(eimp_stream *s) {
    if (!s) return fail(EIMP_ERR_INVALID, "null stream");
    if (!s->pending) return EIMP_OK;
    CHK(set_device(s->e));
    const hipError_t q = hipEventQuery(s->ev[3]);
    if (q == hipSuccess) return EIMP_OK;
    if (q == hipErrorNotReady) return EIMP_BUSY;
    return fail(EIMP_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
}

int eimp_stream_wait(eimp_stream *s, eimp_target_result *results, size_t *ndeferred, double *times) {
    if (!s) return fail(EIMP_ERR_INVALID, "null stream");
    if (!s->pending) return fail(EIMP_ERR_STATE, "no submission to wait for");
    CHK(set_device(s->e));
    HIPCHK(hipEventSynchronize(s->ev[3]));
    s->pending = false;
    s->collected = true;
    const uint32_t nt = s->n_last;
    size_t nd = 0;
    const bool digest = (s->flags & EIMP_BATCH_DIGEST) != 0;
    for (uint32_t t = 0; t < nt; t++) {
        TgtDev &T = s->h_tg()[t];
        const uint32_t *st = s->h_stats() + 4 * t;
        const bool deferred = st[2] != 0;
        if (deferred) {
            T.Mp = 0;
            nd++;
        }
        if (results) {
            results[t].S = st[0];
            results[t].Mp = deferred ? 0 : T.Mp;
            results[t].ncalls = st[1];
            results[t].K = T.K;
            results[t].digest_pbwt = digest && !deferred ? s->h_digest()[2 * t] : 0;
            results[t].digest_count0 = digest && !deferred ? s->h_digest()[2 * t + 1] : 0;
        }
    }
    if (ndeferred) *ndeferred = nd;
    if (times) {
        float a = 0, b = 0;
        HIPCHK(hipEventElapsedTime(&a, s->ev[1], s->ev[2]));
        HIPCHK(hipEventElapsedTime(&b, s->ev[0], s->ev[3]));
        times[0] = a;
        times[1] = b;
    }
    for (int c = 0; c < EIMP_NKCLASS; c++) s->kms[c] = 0;
    if (s->timed && nt) {
        for (int c = EIMP_KCLASS_DISTANCE; c < EIMP_KCLASS_COPY; c++) {
            // the call locations end with the preparation: the wait for the
            // pool's previous build, if any, belongs to no class
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, s->kev[c],
                                       c == EIMP_KCLASS_SPLITS ? s->kev_prep_end : s->kev[c + 1]));
            s->kms[c] = ms;
        }
        float ms = 0;  // copies: header/table/genotypes in, results out
        HIPCHK(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
        s->kms[EIMP_KCLASS_COPY] = ms;
        HIPCHK(hipEventElapsedTime(&ms, s->ev[2], s->ev[3]));
        s->kms[EIMP_KCLASS_COPY] += ms;
    }
    return EIMP_OK;
}

int eimp_stream_kernel_ms(eimp_stream *s, double *ms) {
    if (!s || !ms) return fail(EIMP_ERR_INVALID, "null argument");
    if (!s->timed) return fail(EIMP_ERR_STATE, "context not created with EIMP_STREAM_TIMED");
    for (int c = 0; c < EIMP_NKCLASS; c++) ms[c] = s->kms[c];
    return EIMP_OK;
}

int eimp_stream_copy_target(eimp_stream *s, size_t t, uint32_t *besthaps, uint64_t *bhflags,
                            uint32_t *splitsites, uint64_t *splits, float *cMdiffs, int32_t *count0,
                            uint32_t *pbwt, uint64_t *cref_rows) {
    if (!s) return fail(EIMP_ERR_INVALID, "null stream");
    if (s->pending || !s->collected) return fail(EIMP_ERR_STATE, "no collected submission");
    if (t >= s->n_last) return fail(EIMP_ERR_INVALID, "target index out of range");
    // the copies are synchronous: holding the lock keeps other builds out of the pool meanwhile
    std::unique_lock<std::mutex> lk;
    if (count0 || pbwt || cref_rows) {
        lk = std::unique_lock<std::mutex>(s->pool->mu);
        if (!s->outputs_current()) return fail(EIMP_ERR_STATE, kPoolReused);
    }
    CHK(set_device(s->e));
    return copy_target_out(s->h_tg()[t], s->d_tg(), (uint32_t)t, s->e->Nad, s->stream, s->d_tmp,
                           besthaps, bhflags, splitsites, splits, cMdiffs, count0, pbwt, cref_rows);
}

int eimp_stream_copy_monos(eimp_stream *s, size_t t, uint64_t *monos) {
    if (!s) return fail(EIMP_ERR_INVALID, "null stream");
    if (s->pending || !s->collected) return fail(EIMP_ERR_STATE, "no collected submission");
    if (t >= s->n_last) return fail(EIMP_ERR_INVALID, "target index out of range");
    if (!monos) return fail(EIMP_ERR_INVALID, "null monos");
    CHK(set_device(s->e));
    const TgtDev &T = s->h_tg()[t];
    return copy_monos_out(T, T.mand == T.mand_w ? unit_words(T.W) : unit_words(s->e->M), monos);
}

int eimp_stream_digest_outputs(eimp_stream *s, uint64_t *digests) {
    if (!s || !digests) return fail(EIMP_ERR_INVALID, "null argument");
    if (s->pending || !s->collected) return fail(EIMP_ERR_STATE, "no collected submission");
    if (batch_layout(s->flags) == LAYOUT_RAW)
        return fail(EIMP_ERR_UNSUPPORTED, "raw rows carry no zero counts to digest");
    std::lock_guard<std::mutex> lk(s->pool->mu);  // held until the digests are read back
    if (!s->outputs_current()) return fail(EIMP_ERR_STATE, kPoolReused);
    CHK(set_device(s->e));
    const uint32_t nt = s->n_last;
    if (!nt) return EIMP_OK;
    HIPCHK(hipMemsetAsync(s->d_digest.p, 0, (size_t)nt * 16, s->stream));
    HIPCHK(launch_digest_rows(s->d_tg(), &s->d_batch()->n, nt, s->d_digest.as<uint64_t>(), s->stream));
    HIPCHK(hipMemcpyAsync(digests, s->d_digest.p, (size_t)nt * 16, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EIMP_OK;
}

int eimp_stream_destroy(eimp_stream *s) {
    if (!s) return EIMP_OK;
    (void)set_device(s->e);
    stream_release(s);
    delete s;
    return EIMP_OK;
}

}  // extern "C"
