// ref_harness.cpp -- C entry points around the reference's own PBWT and
// GenotypeVector code (src/PBWT.cpp, src/Datatypes.cpp, src/RingBuffer.cpp of
// ikmb/eagleimp), compiled in place from /root/reference by oracle/Makefile
// into oracle/_ref/libeimp_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used to write golden vectors (tests/golden) and
// to cross-check the oracle where the reference sources are present.
//
// The reference keeps the compressed PBWT in a private member; it is located
// through MyMalloc's public allocation map (src/MyMalloc.h:57), where
// PBWT::init registers it as "comprPBWT_t<ntarget>" (src/PBWT.cpp:134).
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "Datatypes.h"
#include "MyMalloc.h"
#include "PBWT.h"
#include "RingBuffer.h"
#include "StatusFile.h"

// static members normally defined by the reference's main.cpp (src/main.cpp:48-50)
MyMalloc MyMalloc::mm;
StatusFile StatusFile::instance;

namespace {

size_t roundUp(size_t n, size_t m) { return (n % m) ? n + m - (n % m) : n; }

const uint32_t *findAlloc(const std::string &id) {
    for (const auto &e : MyMalloc::getMap())
        if (e.second.second == id) return reinterpret_cast<const uint32_t *>(e.first);
    return nullptr;
}

size_t g_next_id = 1000000;

}  // namespace

extern "C" {

// Builds the reference PBWT (forward + backward) from condensed rows.
// cref: Mp rows of cref_stride uint64 words (bits >= K zero); count0: Mp ints.
// out: 2*Mp*2*ceil(K/32) uint32 in the reference's own layout.
// Optional interval queries: nq queries of (reverse, m, l, r) -> out4 per query.
int ref_build_pbwt(const uint64_t *cref, uint64_t cref_stride, uint64_t K, uint64_t Mp,
                   const int32_t *count0, uint32_t *out, uint64_t nq, const int32_t *queries,
                   int32_t *qout) {
    size_t capacity = roundUp(K, UNITWORDS * 64) / 8;  // bytes per row as src/Target.cpp:254
    std::vector<uint64_t> buf(Mp * capacity / 8, 0ull);
    for (size_t m = 0; m < Mp; m++)
        std::memcpy(&buf[m * capacity / 8], cref + m * cref_stride, capacity);
    RingBuffer<BooleanVector> refincT(Mp, Mp);
    uint64_t *cur = buf.data();
    for (auto &row : refincT) {
        row.setData(cur, capacity, K);
        cur += capacity / 8;
    }
    auto *c0 = new std::vector<int>(count0, count0 + Mp);
    size_t id = g_next_id++;
    int rc = 0;
    {
        PBWT pbwt(&refincT, c0, id);  // takes ownership of c0
        const uint32_t *data = findAlloc(std::string("comprPBWT_t") + std::to_string(id));
        if (!data) return -1;
        size_t words = 2 * Mp * 2 * ((K + 31) / 32);
        std::memcpy(out, data, words * sizeof(uint32_t));
        // forward queries first; switchToReverse (src/PBWT.cpp:196-200) is one-way
        for (int pass = 0; pass < 2; pass++) {
            if (pass == 1) pbwt.switchToReverse();
            for (uint64_t q = 0; q < nq; q++) {
                const int32_t *qq = queries + 4 * q;
                if ((qq[0] != 0) != (pass == 1)) continue;
                PBWTInterval from(qq[2], qq[3]), to0, to1;
                pbwt.mapInterval((size_t)qq[1], from, to0, to1);
                qout[4 * q + 0] = to0.getLeft();
                qout[4 * q + 1] = to0.getRight();
                qout[4 * q + 2] = to1.getLeft();
                qout[4 * q + 3] = to1.getRight();
            }
        }
    }
    return rc;
}

// Encodes genotypes (0 HomRef, 1 Het, 2 HomAlt, 3 Miss) with the reference
// GenotypeVector::push_back and returns its words; counts errors against a hap
// row with GenotypeVector::compare_withPreInit.
int ref_genotype_compare(const uint8_t *g, uint64_t W, const uint64_t *hap, uint64_t hap_len,
                         uint64_t *gt_words_out, uint64_t gt_words_cap, uint64_t *nerr_out) {
    GenotypeVector gv;
    for (uint64_t m = 0; m < W; m++) gv.push_back(static_cast<Genotype>(g[m]));
    const auto &d = gv.getData();
    if (d.size() > gt_words_cap) return -1;
    std::memcpy(gt_words_out, d.data(), d.size() * sizeof(uint64_t));
    size_t capacity = roundUp(hap_len, UNITWORDS * 64) / 8;
    std::vector<uint64_t> hb(capacity / 8, 0ull);
    std::memcpy(hb.data(), hap, ((hap_len + 63) / 64) * sizeof(uint64_t));
    BooleanVector hv(hb.data(), capacity, hap_len);
    *nerr_out = gv.compare_withPreInit(hv);
    return (int)d.size();
}

// Ingests a raw cPBWT buffer (2*M rows of kwords_raw data words, the FPGA
// handler's layout, src/FPGAHandler.cpp:896-900) with the reference's
// PBWT(uint32_t*, K, Kwords, M, ntarget) (src/PBWT.cpp:72-110), which inserts
// the cumulative zero counts; returns its rows (2*M x 2*kwords_raw uint32)
// and count0 (M).
int ref_ingest_raw(const uint32_t *raw, uint64_t K, uint64_t kwords_raw, uint64_t M, uint32_t *out,
                   int32_t *count0) {
    std::vector<uint32_t> buf(raw, raw + 2 * M * kwords_raw);  // the constructor copies it
    size_t id = g_next_id++;
    PBWT pbwt(buf.data(), K, kwords_raw, M, id);
    const uint32_t *data = findAlloc(std::string("comprPBWT_t") + std::to_string(id));
    if (!data) return -1;
    std::memcpy(out, data, 2 * M * 2 * kwords_raw * sizeof(uint32_t));
    for (size_t m = 0; m < M; m++) count0[m] = pbwt.getCount0(m);
    return 0;
}

// The reference panel PBWT (refPBWT: forward only, absolute permutations kept,
// src/PBWT.cpp:53-71,131-153) over M site rows of K haplotypes (rows of
// row_stride uint64 words).  out: M rows of 2*ceil(K/32) uint32; count0: M;
// absperm: (M+1) x K int32 (row m = order before site m), or NULL.
int ref_build_ref_pbwt(const uint64_t *rows, uint64_t row_stride, uint64_t K, uint64_t M,
                       uint32_t *out, int32_t *count0, int32_t *absperm) {
    size_t capacity = roundUp(K, UNITWORDS * 64) / 8;
    std::vector<uint64_t> buf(M * capacity / 8, 0ull);
    for (size_t m = 0; m < M; m++)
        std::memcpy(&buf[m * capacity / 8], rows + m * row_stride, ((K + 63) / 64) * 8);
    RingBuffer<BooleanVector> refT(M, M);
    uint64_t *cur = buf.data();
    for (auto &row : refT) {
        row.setData(cur, capacity, K);
        cur += capacity / 8;
    }
    size_t id = g_next_id++;
    PBWT pbwt(&refT, true, id);
    const uint32_t *data = findAlloc(std::string("comprPBWT_t") + std::to_string(id));
    if (!data) return -1;
    std::memcpy(out, data, M * 2 * ((K + 31) / 32) * sizeof(uint32_t));
    for (size_t m = 0; m < M; m++) count0[m] = pbwt.getCount0(m);
    if (absperm)
        for (size_t m = 0; m <= M; m++) {
            const auto &o = pbwt.getSortOrder(m);
            std::memcpy(absperm + m * K, o.data(), K * sizeof(int32_t));
        }
    return 0;
}

}  // extern "C"
