// eimp_qref.cpp -- reader of the Qref reference panel format (host code):
// the panel loader in front of the resident-panel engine.
//
// Layout (qref_spec.md of the reference; written by src/VCFData.cpp:2006-2104,
// read by src/VCFData.cpp:2107-2376): an 8-byte header "QREF", major and
// minor version, 0, chromosome; Nref, NrefHap, Mref, MrefMA as uint64; one
// haploid flag byte per sample when some but not all samples are haploid or
// the chromosome is X (23); Mref int64 positions; Mref float allele
// frequencies; 2*Mref allele strings and Mref IDs, each zero-terminated; the
// multi-allelic flags as a bit vector rounded to 512 bits; then per variant a
// uint64 encoded size and the haplotype bits of all 2*Nref haplotypes (bit h
// = allele of haplotype h, vectors rounded to 512 bits) -- raw when the size
// is 0, else run-length coded (src/utils.h:133-190): runs of alternating bits
// starting with 0, one byte for runs below 128, two bytes (low 7 bits with
// the top bit set, then the high 8 bits) up to 32767, longer runs split by
// zero-length runs of the other bit.
//
// The file is memory-mapped; variant offsets are indexed once at open; site
// rows decode in parallel straight into the engine's site-major layout.
//
// Sample masking (participant withdrawals, src/VCFData.cpp:2139-2208 and
// 2418-2530; sample lists src/VCFData.cpp:3698-3746): the masked samples'
// two haplotype bits leave every row, the later haplotypes move down, the
// haploid flags of the kept samples stay, and each variant's allele
// frequency is corrected by the variant alleles the masked samples carried.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/eimp.h"

namespace eimp {
void set_last_error(const std::string &msg);  // eimp_engine.cpp
}

struct eimp_qref {
    int fd = -1;
    const uint8_t *base = nullptr;
    size_t size = 0;
    eimp_qref_info info{};
    size_t hap_words = 0;              // uint64 words per variant: roundToMultiple(2*Nref,512)/64
    const uint8_t *haploid = nullptr;  // Nref flag bytes, or null (flags not stored)
    size_t off_pos = 0, off_af = 0, off_ma = 0;
    std::vector<const char *> alleles;  // 2 per variant, then
    std::vector<const char *> ids;      // 1 per variant (zero-terminated, inside the mapping)
    std::vector<uint64_t> var_off;      // offset of each variant's encoded size
    // as stored in the file, before any mask
    uint64_t file_Nref = 0, file_NrefHap = 0;
    // masked samples (ascending) and whether each is haploid
    std::vector<uint64_t> mask;
    std::vector<uint8_t> mask_haploid;
    ~eimp_qref() {
        if (base) munmap(const_cast<uint8_t *>(base), size);
        if (fd >= 0) close(fd);
    }
};

namespace {

constexpr int kMajor = 1, kMinor = 1;  // src/VCFData.h:413-414
constexpr int kChrX = 23;              // src/VCFData.h:418

int fail(int code, const std::string &msg) {
    eimp::set_last_error(msg);
    return code;
}

uint64_t u64_at(const uint8_t *p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

// Decodes one run-length coded haplotype vector into `words` uint64 words
// (pre-zeroed).  Returns false if the code ends inside a word.
bool rle_decode(const uint8_t *enc, size_t n, uint64_t *out, size_t words) {
    const uint64_t total = (uint64_t)words * 64;
    uint64_t pos = 0;
    bool bit = false;  // the first run is of zeros
    size_t i = 0;
    while (i < n) {
        uint64_t run = enc[i] & 0x7fu;
        if (enc[i] & 0x80u) {
            if (++i >= n) return false;  // two-byte run cut off
            run |= (uint64_t)enc[i] << 7;
        }
        i++;
        if (bit && run && pos < total) {  // set bits [pos, min(pos+run, total))
            uint64_t a = pos, b = std::min(pos + run, total);
            while (a < b) {
                const uint64_t w = a >> 6, lo = a & 63;
                const uint64_t take = std::min<uint64_t>(64 - lo, b - a);
                const uint64_t mask = (take == 64 ? ~0ull : ((1ull << take) - 1)) << lo;
                out[w] |= mask;
                a += take;
            }
        }
        pos += run;
        bit = !bit;
    }
    // the coder covers the whole rounded vector; a code that stops at a word
    // boundary leaves the rest zero, as the reference's decoder does
    // (src/utils.h:227-233)
    return pos >= total || pos % 64 == 0;
}

bool file_haploid(const eimp_qref *q, uint64_t s) {
    if (q->haploid) return q->haploid[s] != 0;
    return q->file_NrefHap == q->file_Nref;  // all haploid or all diploid (src/VCFData.cpp:2196-2208)
}

// Bit `pos` onwards of a vector of `words` words, 64 bits (zero past its end).
inline uint64_t bits_at(const uint64_t *v, size_t words, uint64_t pos) {
    const uint64_t w = pos >> 6, sh = pos & 63;
    if (w >= words) return 0;
    uint64_t x = v[w] >> sh;
    if (sh && w + 1 < words) x |= v[w + 1] << (64 - sh);
    return x;
}

// OR bits [s, s+len) of src into dst starting at bit d (dst pre-zeroed, dst
// bits past dst_words*64 dropped).
void copy_bits(const uint64_t *src, size_t src_words, uint64_t s, uint64_t *dst, size_t dst_words,
               uint64_t d, uint64_t len) {
    const uint64_t dend = std::min<uint64_t>(d + len, (uint64_t)dst_words * 64);
    while (d < dend) {
        const uint64_t w = d >> 6, sh = d & 63;
        const uint64_t take = std::min<uint64_t>(64 - sh, dend - d);
        uint64_t x = bits_at(src, src_words, s);
        if (take < 64) x &= (1ull << take) - 1;
        dst[w] |= x << sh;
        d += take;
        s += take;
    }
}

// One variant's record (size word at p) decoded into full (q->hap_words
// words, all 2*file_Nref haplotypes).  Returns false for a truncated code.
bool decode_record(const eimp_qref *q, const uint8_t *p, uint64_t *full, bool *raw) {
    const uint64_t enc = u64_at(p);
    *raw = enc == 0;
    if (enc == 0) {
        std::memcpy(full, p + 8, q->hap_words * 8);
        return true;
    }
    std::fill(full, full + q->hap_words, 0ull);
    return rle_decode(p + 8, enc, full, q->hap_words);
}

// Variant alleles of the masked samples in one decoded record.  Run-length
// coded records count a haploid sample's allele once, from its second bit
// after a set first bit (src/utils.h:245-262: only (1,1) counts); raw
// records count one for any set bit and a second for a diploid sample with
// both set (src/VCFData.cpp:2468-2475: a haploid (1,0) or (0,1) counts one).
uint64_t masked_alleles(const eimp_qref *q, const uint64_t *full, bool raw) {
    uint64_t c = 0;
    for (size_t i = 0; i < q->mask.size(); i++) {
        const uint64_t h = 2 * q->mask[i];
        const uint64_t a = (full[h >> 6] >> (h & 63)) & 1u, b = (full[h >> 6] >> ((h & 63) + 1)) & 1u;
        if (q->mask_haploid[i])
            c += raw ? (a | b) : (a & b);
        else
            c += a + b;
    }
    return c;
}

// The kept haplotypes of a decoded record, in order, into dst (pre-zeroed).
void compact_masked(const eimp_qref *q, const uint64_t *full, uint64_t *dst, size_t dst_words) {
    uint64_t s = 0, d = 0;
    for (uint64_t ms : q->mask) {
        const uint64_t h = 2 * ms;
        copy_bits(full, q->hap_words, s, dst, dst_words, d, h - s);
        d += h - s;
        s = h + 2;
    }
    copy_bits(full, q->hap_words, s, dst, dst_words, d, 2 * q->file_Nref - s);
}

// The allele frequency the reference holds after masking
// (src/VCFData.cpp:2426-2432, 2524-2529): new_ac = old_an * old_af -
// masked alleles, af = new_ac / new_an, in single precision.  The reference
// is built with -march=native at -O3 (src/CMakeLists.txt:165, README), where
// GCC contracts the product and the subtraction into one fused
// multiply-subtract on every host with FMA; fmaf gives that single rounding.
float corrected_af(const eimp_qref *q, float old_af, uint64_t masked) {
    const uint64_t old_an = 2 * q->file_Nref - q->file_NrefHap;
    const uint64_t new_an = 2 * q->info.Nref - q->info.NrefHap;
    const float new_ac = std::fma((float)old_an, old_af, -(float)masked);
    return new_ac / (float)new_an;
}

}  // namespace

extern "C" {

int eimp_qref_open(const char *path, eimp_qref **out, eimp_qref_info *info) {
    if (!path || !out) return fail(EIMP_ERR_INVALID, "qref: null argument");
    *out = nullptr;
    auto *q = new eimp_qref();
    auto bad = [&](const std::string &m) {
        delete q;
        return fail(EIMP_ERR_INVALID, std::string("qref: ") + path + ": " + m);
    };
    q->fd = open(path, O_RDONLY);
    if (q->fd < 0) return bad("cannot open");
    struct stat st;
    if (fstat(q->fd, &st) != 0) return bad("cannot stat");
    q->size = (size_t)st.st_size;
    if (q->size < 40) return bad("too short for a header");
    void *m = mmap(nullptr, q->size, PROT_READ, MAP_PRIVATE, q->fd, 0);
    if (m == MAP_FAILED) return bad("cannot map");
    q->base = static_cast<const uint8_t *>(m);
    const uint8_t *b = q->base;
    // A) header (src/VCFData.cpp:2117-2127)
    if (std::memcmp(b, "QREF", 4) != 0 || b[6] != 0) return bad("not a Qref file");
    q->info.v_major = (int8_t)b[4];
    q->info.v_minor = (int8_t)b[5];
    if (q->info.v_major != kMajor || q->info.v_minor > kMinor)
        return bad("unsupported version " + std::to_string(q->info.v_major) + "." +
                   std::to_string(q->info.v_minor) + " (this reader: 1.<=1)");
    q->info.chrom = (int8_t)b[7];
    // B) constants
    q->info.Nref = u64_at(b + 8);
    q->info.NrefHap = u64_at(b + 16);
    q->info.Mref = u64_at(b + 24);
    q->info.MrefMA = u64_at(b + 32);
    const uint64_t N = q->info.Nref, M = q->info.Mref;
    q->file_Nref = N;
    q->file_NrefHap = q->info.NrefHap;
    if (q->info.NrefHap > N) return bad("more haploid samples than samples");
    if (N > (1ull << 40) || M > (1ull << 40)) return bad("implausible sizes");
    q->hap_words = (2 * N + 511) / 512 * 8;
    size_t off = 40;
    auto need = [&](uint64_t n) { return n <= q->size && off <= q->size - n; };
    if (M == 0) {  // no variants: nothing else is stored (src/VCFData.cpp:2158-2159)
        if (info) *info = q->info;
        *out = q;
        return EIMP_OK;
    }
    // C) haploid flags (src/VCFData.cpp:2174)
    q->info.has_haploid_flags =
        ((q->info.NrefHap > 0 && q->info.NrefHap != N) || q->info.chrom == kChrX) ? 1 : 0;
    if (q->info.has_haploid_flags) {
        if (!need(N)) return bad("truncated haploid flags");
        q->haploid = b + off;
        off += N;
    }
    // D) positions, E) allele frequencies
    if (!need(8 * M)) return bad("truncated positions");
    q->off_pos = off;
    off += 8 * M;
    if (!need(4 * M)) return bad("truncated allele frequencies");
    q->off_af = off;
    off += 4 * M;
    // F) alleles, G) IDs: zero-terminated strings
    auto strings = [&](size_t n, std::vector<const char *> &v) {
        v.resize(n);
        for (size_t i = 0; i < n; i++) {
            if (off >= q->size) return false;
            const void *z = std::memchr(b + off, 0, q->size - off);
            if (!z) return false;
            v[i] = reinterpret_cast<const char *>(b + off);
            off = (size_t)(static_cast<const uint8_t *>(z) - b) + 1;
        }
        return true;
    };
    if (!strings(2 * M, q->alleles)) return bad("truncated allele strings");
    if (!strings(M, q->ids)) return bad("truncated variant IDs");
    // H) multi-allelic flags, rounded to 512 bits
    const uint64_t ma_bytes = (M + 511) / 512 * 64;
    if (!need(ma_bytes)) return bad("truncated multi-allelic flags");
    q->off_ma = off;
    off += ma_bytes;
    // I) haplotype data: index every variant
    q->var_off.resize(M);
    const uint64_t raw_bytes = q->hap_words * 8;
    for (uint64_t i = 0; i < M; i++) {
        if (!need(8)) return bad("truncated haplotypes at variant " + std::to_string(i));
        q->var_off[i] = off;
        const uint64_t enc = u64_at(b + off);
        off += 8;
        const uint64_t len = enc ? enc : raw_bytes;
        if (!need(len)) return bad("truncated haplotypes at variant " + std::to_string(i));
        off += len;
    }
    q->info.hap_words = q->hap_words;
    if (info) *info = q->info;
    *out = q;
    return EIMP_OK;
}

int eimp_qref_close(eimp_qref *q) {
    delete q;
    return EIMP_OK;
}

int eimp_qref_metadata(eimp_qref *q, int64_t *positions, float *af, uint8_t *haploid,
                       uint64_t *ma_flags) {
    if (!q) return fail(EIMP_ERR_INVALID, "qref: null handle");
    const uint64_t N = q->info.Nref, M = q->info.Mref;
    if (positions && M) std::memcpy(positions, q->base + q->off_pos, 8 * M);
    if (af && M) std::memcpy(af, q->base + q->off_af, 4 * M);
    if (haploid) {  // the kept samples' flags, in order (src/VCFData.cpp:2170-2192)
        size_t k = 0, mi = 0;
        for (uint64_t s = 0; s < q->file_Nref; s++) {
            if (mi < q->mask.size() && q->mask[mi] == s) {
                mi++;
                continue;
            }
            haploid[k++] = file_haploid(q, s) ? 1 : 0;
        }
        (void)N;
    }
    if (ma_flags && M) std::memcpy(ma_flags, q->base + q->off_ma, (M + 63) / 64 * 8);
    return EIMP_OK;
}

int eimp_qref_variant(eimp_qref *q, size_t m, const char **ref, const char **alt,
                      const char **id) {
    if (!q) return fail(EIMP_ERR_INVALID, "qref: null handle");
    if (m >= q->info.Mref) return fail(EIMP_ERR_INVALID, "qref: variant index out of range");
    if (ref) *ref = q->alleles[2 * m];
    if (alt) *alt = q->alleles[2 * m + 1];
    if (id) *id = q->ids[m];
    return EIMP_OK;
}

int eimp_qref_read_sites_af(eimp_qref *q, size_t m0, size_t n, uint64_t *rows, size_t row_words,
                            float *af, uint64_t *masked, int threads) {
    if (!q) return fail(EIMP_ERR_INVALID, "qref: null handle");
    if (n == 0) return EIMP_OK;
    if (m0 > q->info.Mref || n > q->info.Mref - m0)
        return fail(EIMP_ERR_INVALID, "qref: sites past the last variant");
    const size_t need_words = (2 * q->info.Nref + 63) / 64;
    if (rows && row_words < need_words)
        return fail(EIMP_ERR_INVALID, "qref: row_words below ceil(2*Nref/64)");
    const bool masking = !q->mask.empty();
    if (!masking && !rows) {  // nothing to decode
        if (af) std::memcpy(af, q->base + q->off_af + 4 * m0, 4 * n);
        if (masked) std::fill(masked, masked + n, 0ull);
        return EIMP_OK;
    }
    const size_t HW = q->hap_words;
    size_t nt = threads > 0 ? (size_t)threads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, n);
    std::atomic<size_t> next{0};
    std::atomic<int64_t> bad{-1};
    auto work = [&]() {
        std::vector<uint64_t> tmp(HW);
        for (size_t i; (i = next.fetch_add(1)) < n;) {
            const uint64_t m = m0 + i;
            bool raw;
            if (!decode_record(q, q->base + q->var_off[m], tmp.data(), &raw)) {
                int64_t expect = -1;
                bad.compare_exchange_strong(expect, (int64_t)m);
                continue;
            }
            float f;
            std::memcpy(&f, q->base + q->off_af + 4 * m, 4);
            uint64_t cnt = 0;
            if (masking) {
                cnt = masked_alleles(q, tmp.data(), raw);
                f = corrected_af(q, f, cnt);
            }
            if (af) af[i] = f;
            if (masked) masked[i] = cnt;
            if (!rows) continue;
            uint64_t *dst = rows + i * row_words;
            if (masking) {
                std::fill(dst, dst + row_words, 0ull);
                compact_masked(q, tmp.data(), dst, std::min(row_words, need_words));
            } else {
                const size_t c = std::min(row_words, HW);
                std::memcpy(dst, tmp.data(), c * 8);
            }
            // bits past the 2*Nref haplotypes are padding: zero
            const uint64_t nb = 2 * q->info.Nref;
            if (nb % 64) dst[nb / 64] &= (1ull << (nb % 64)) - 1;
            for (size_t w = need_words; w < row_words; w++) dst[w] = 0;
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (bad >= 0)
        return fail(EIMP_ERR_INVALID,
                    "qref: run-length code of variant " + std::to_string(bad.load()) + " is truncated");
    return EIMP_OK;
}

int eimp_qref_read_sites(eimp_qref *q, size_t m0, size_t n, uint64_t *rows, size_t row_words,
                         int threads) {
    if (!q) return fail(EIMP_ERR_INVALID, "qref: null handle");
    if (n && !rows && row_words) return fail(EIMP_ERR_INVALID, "qref: null rows");
    return eimp_qref_read_sites_af(q, m0, n, rows, row_words, nullptr, nullptr, threads);
}

int eimp_qref_set_mask(eimp_qref *q, const uint64_t *samples, size_t n, eimp_qref_info *info) {
    if (!q) return fail(EIMP_ERR_INVALID, "qref: null handle");
    if (n && !samples) return fail(EIMP_ERR_INVALID, "qref: null sample list");
    for (size_t i = 0; i < n; i++) {
        if (samples[i] >= q->file_Nref)
            return fail(EIMP_ERR_INVALID, "qref: masked sample " + std::to_string(samples[i]) +
                                              " past the panel's " + std::to_string(q->file_Nref) +
                                              " samples");
        if (i && samples[i] <= samples[i - 1])
            return fail(EIMP_ERR_INVALID, "qref: masked samples not strictly ascending");
    }
    q->mask.assign(samples, samples + n);
    q->mask_haploid.resize(n);
    uint64_t hap = 0;
    for (size_t i = 0; i < n; i++) hap += q->mask_haploid[i] = file_haploid(q, samples[i]) ? 1 : 0;
    q->info.Nref = q->file_Nref - n;
    q->info.NrefHap = q->file_NrefHap - hap;
    q->info.hap_words = q->info.Mref ? (2 * q->info.Nref + 511) / 512 * 8 : 0;
    if (info) *info = q->info;
    return EIMP_OK;
}

int eimp_qref_parse_sample_list(const char *path, uint64_t *masked, size_t cap, size_t *nmasked,
                                size_t *nsamples) {
    if (!path || !nmasked || !nsamples) return fail(EIMP_ERR_INVALID, "sample list: null argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(EIMP_ERR_INVALID, std::string("sample list: cannot open ") + path);
    const size_t L = std::strlen(path);
    const bool bgen = L >= 7 && std::strcmp(path + L - 7, ".sample") == 0;
    int64_t idx = bgen ? -2 : 0;  // a bgen .sample file has two header lines
    size_t k = 0;
    std::string line;
    bool ok = true;
    for (int c = 0; c != EOF;) {
        line.clear();
        while ((c = std::fgetc(f)) != EOF && c != '\n') line.push_back((char)c);
        if (c == EOF && line.empty()) break;
        if (idx >= 0) {
            // first whitespace-separated token; the second one in a .sample file
            auto token = [&](size_t &at) {
                while (at < line.size() && std::isspace((unsigned char)line[at])) at++;
                const size_t b = at;
                while (at < line.size() && !std::isspace((unsigned char)line[at])) at++;
                return line.substr(b, at - b);
            };
            size_t at = 0;
            std::string sid = token(at);
            if (!sid.empty() && bgen) {
                sid = token(at);
                if (sid.empty()) {
                    ok = false;
                    break;
                }
            }
            // a withdrawal: 'W' or '-' followed by at least one digit and only digits
            if (sid.size() >= 2 && (sid[0] == 'W' || sid[0] == '-') &&
                std::all_of(sid.begin() + 1, sid.end(),
                            [](char ch) { return std::isdigit((unsigned char)ch) != 0; })) {
                if (masked && k < cap) masked[k] = (uint64_t)idx;
                k++;
            }
        }
        idx++;
    }
    std::fclose(f);
    if (!ok) return fail(EIMP_ERR_INVALID, "sample list: not a valid bgen .sample file");
    *nmasked = k;
    *nsamples = idx > 0 ? (size_t)idx : 0;
    if (masked && k > cap) return fail(EIMP_ERR_INVALID, "sample list: more withdrawals than cap");
    return EIMP_OK;
}

int eimp_qref_mask_sample_list(eimp_qref *q, const char *path, eimp_qref_info *info) {
    if (!q) return fail(EIMP_ERR_INVALID, "qref: null handle");
    size_t nm = 0, ns = 0;
    int rc = eimp_qref_parse_sample_list(path, nullptr, 0, &nm, &ns);
    if (rc != EIMP_OK) return rc;
    std::vector<uint64_t> ms(nm);
    rc = eimp_qref_parse_sample_list(path, ms.data(), nm, &nm, &ns);
    if (rc != EIMP_OK) return rc;
    if (nm && ns != q->file_Nref)  // src/VCFData.cpp:2139-2142
        return fail(EIMP_ERR_INVALID, "qref: sample list holds " + std::to_string(ns) +
                                          " samples, the panel " + std::to_string(q->file_Nref));
    return eimp_qref_set_mask(q, ms.data(), nm, info);
}

}  // extern "C"
