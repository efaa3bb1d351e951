// ref_harness.cpp -- C entry points around the reference's own PBWT,
// GenotypeVector and beam search code (src/PBWT.cpp, src/Datatypes.cpp,
// src/RingBuffer.cpp, src/Beam.cpp, src/BeamPath.cpp, src/HaplotypePath.cpp,
// src/BooleanHistory.cpp of ikmb/eagleimp), compiled in place from
// /root/reference by oracle/Makefile into oracle/_ref/libeimp_ref.so.
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

#include <cmath>
#include <iostream>  // the reference headers expect it included first

#include "Beam.h"
#include "Datatypes.h"
#include "MinMaxHeap.h"
#include "MyMalloc.h"
#include "PBWT.h"
#include "RingBuffer.h"
#include "StatusFile.h"
#include "eimp.h"
#include "utils.h"

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

// The Qref run-length coder (src/utils.h:133-190): codes nwords uint64 of
// haplotype bits; returns the code length, or -1 when the coder gives up
// (the code would not be shorter), -2 when cap is too small.
int64_t ref_rle_encode(const uint64_t *words, uint64_t nwords, uint8_t *out, uint64_t cap) {
    std::vector<char> enc;
    if (!runlengthEncode(words, nwords, enc)) return -1;
    if (enc.size() > cap) return -2;
    std::memcpy(out, enc.data(), enc.size());
    return (int64_t)enc.size();
}

// Its decoder (src/utils.h:193-315) without sample masking: nwords uint64
// (zeroed here first).
int ref_rle_decode(const uint8_t *code, uint64_t n, uint64_t *words, uint64_t nwords) {
    std::vector<char> enc(code, code + n);
    std::memset(words, 0, nwords * sizeof(uint64_t));
    std::vector<size_t> masked;
    std::vector<bool> masked_haploid;
    size_t masked_alleles = 0;
    runlengthDecode(enc, words, nwords, masked, masked_haploid, masked_alleles);
    return 0;
}

// Its decoder with sample masking (src/utils.h:193-315): the haplotypes of
// the masked samples (ascending sample indices) are dropped, their variant
// alleles counted into *masked_alleles (a haploid masked sample counts once).
int ref_rle_decode_masked(const uint8_t *code, uint64_t n, const uint64_t *masked, uint64_t nmasked,
                          const uint8_t *masked_haploid, uint64_t *words, uint64_t nwords,
                          uint64_t *masked_alleles) {
    std::vector<char> enc(code, code + n);
    std::memset(words, 0, nwords * sizeof(uint64_t));
    std::vector<size_t> ms(masked, masked + nmasked);
    std::vector<bool> mh(nmasked);
    for (uint64_t i = 0; i < nmasked; i++) mh[i] = masked_haploid[i] != 0;
    size_t cnt = 0;
    runlengthDecode(enc, words, nwords, ms, mh, cnt);
    *masked_alleles = cnt;
    return 0;
}

// Target::phaseExtPBWT (src/Target.cpp:291-777) restated around the
// reference's own Beam over PBWT(cpbwt, count0, K, row_stride, Mp)
// (src/GPUHandler.cpp:259-260).  Target.cpp itself needs VCFData (HTSlib),
// so its driver loop is written out here, statement for statement; monos:
// the target's mono bits over the window (src/Target.cpp:866-897).
int ref_phase_target(const eimp_phase_params *P, const eimp_phase_input *in,
                     const uint64_t *monos, eimp_phase_output *out) {
    struct Conf {  // src/Target.h:181-186
        size_t constraint_idx;
        fp_type p;
        Conf(size_t i, fp_type p_) : constraint_idx(i), p(p_) {}
        bool operator<(const Conf &o) const { return p < o.p; }
    };
    const size_t MFull = in->W, S = (in->Mp - 1) / 2, K = in->K;
    auto geno = [&](size_t m) {
        const uint64_t b = 1ull << (m & 63);
        const bool is0 = in->gt[2 * (m >> 6)] & b, is2 = in->gt[2 * (m >> 6) + 1] & b;
        return is0 ? (is2 ? Genotype::Miss : Genotype::HomRef)
                   : (is2 ? Genotype::HomAlt : Genotype::Het);
    };
    auto mono = [&](size_t m) { return (monos[m >> 6] >> (m & 63)) & 1ull; };
    std::vector<size_t> csplits(in->splitsites, in->splitsites + S);
    std::vector<fp_type> cMsplitDiffs(in->cMdiffs, in->cMdiffs + (S ? S + 1 : 0));
    std::vector<size_t> callLocs;
    GenotypeVector target;
    for (size_t i = 0; i < S; i++) {
        target.push_back(geno(csplits[i]));
        if (geno(csplits[i]) == Genotype::Het) callLocs.push_back(i);
    }
    // the PBWT object owns (and frees) its buffers
    size_t bytes = 2 * in->Mp * (size_t)in->row_stride * sizeof(uint32_t);
    uint32_t *cp = (uint32_t *)MyMalloc::malloc(bytes, "phase_cpbwt");
    std::memcpy(cp, in->pbwt, bytes);
    auto *c0 = new std::vector<int>(in->count0, in->count0 + in->Mp);
    PBWT pbwt(cp, c0, K, in->row_stride, in->Mp, g_next_id++);

    std::vector<float> dmat(MFull, 0.0f), dpat(MFull, 0.0f);
    std::vector<bool> pmat, ppat;
    fp_type totalconf = 0;
    size_t ncallsites = 0;
    int totalmissings = 0, totalmonhets = 0;
    const bool doPrePhasing = P->pre_phasing, doRevPhasing = P->rev_phasing;
    const bool lastIter = P->last_iter, impMissing = P->imp_missing;
    const fp_type hist = P->hist, pErr = P->p_err, pLimit = P->p_limit;
    const fp_type expectIBDcM = P->expect_ibd_cM;
    const fp_type minFixSize = P->min_fix_size, maxFixSize = P->max_fix_size;
    const fp_type minFixThresh = P->min_fix_thresh, maxFixThresh = P->max_fix_thresh;
    const size_t HistorySize = P->history, HistorySizeFast = P->history_fast;

    std::vector<Constraint> constraints(csplits.size(), Constraint(Constraint::Type::NoConstr, 0));
    Haplotype currMatPhase = in->start_phase ? Haplotype::Alt : Haplotype::Ref;
    Haplotype lastMatPhase = currMatPhase;

    if (doPrePhasing) {
        size_t history = static_cast<size_t>((HistorySizeFast * hist) + 0.5);
        Beam beam(history, P->beam_width_fast, P->delta_fast, pErr, pLimit, expectIBDcM,
                  cMsplitDiffs, callLocs, target, pbwt, constraints);
        if (csplits.size() > 0) constraints[0] = Constraint(Constraint::Type::HetKeep, 0);
        MinMaxHeap<Conf> callheap, homheap;
        size_t next_call_idx = 1;
        for (size_t si = 1; si < csplits.size(); si++) {
            if (si == callLocs[next_call_idx]) {
                fp_type p = beam.callKeepPhaseProb(next_call_idx) - 0.5;
                Haplotype ref = currMatPhase;
                Constraint c(Constraint::Type::NoConstr,
                             callLocs[next_call_idx] - callLocs[next_call_idx - 1] - 1);
                if (next_call_idx > 1) {
                    fp_type p2 = beam.callKeepPhaseProb(next_call_idx, true) - 0.5;
                    if (std::abs(p2) > std::abs(p)) {
                        p = p2;
                        ref = lastMatPhase;
                        c.refidx = callLocs[next_call_idx] - callLocs[next_call_idx - 2] - 1;
                    }
                }
                lastMatPhase = currMatPhase;
                if (p < 0) {
                    currMatPhase = !ref;
                    c.type = Constraint::Type::HetSwitch;
                } else {
                    currMatPhase = ref;
                    c.type = Constraint::Type::HetKeep;
                }
                constraints[si] = c;
                p = std::abs(p);
                callheap.emplace(si, p);
                next_call_idx++;
            } else {
                fp_type p = beam.callDosage(si);
                if (target[si] == Genotype::HomRef) {
                    constraints[si] = Constraint(Constraint::Type::HomRef, 0);
                    p = 2.0 - p;
                } else
                    constraints[si] = Constraint(Constraint::Type::HomAlt, 0);
                homheap.emplace(si, p);
            }
        }
        size_t sizedivdmin = callheap.getSize() * minFixSize;
        size_t sizedivdmax = callheap.getSize() * (1.0 - maxFixSize);
        size_t si;
        for (si = 0; si < sizedivdmin; si++) {
            Conf cc = callheap.popMin();
            if (cc.p >= maxFixThresh - 0.5 || (cc.p >= minFixThresh - 0.5 && si >= sizedivdmax))
                break;
            constraints[cc.constraint_idx].type = Constraint::Type::NoConstr;
        }
        if (si == sizedivdmin) {
            for (; si < callheap.getSize(); si++) {
                Conf cc = callheap.popMin();
                if (cc.p > 0.0) break;
                constraints[cc.constraint_idx].type = Constraint::Type::NoConstr;
            }
        }
        sizedivdmin = homheap.getSize() * minFixSize;
        sizedivdmax = homheap.getSize() * (1.0 - maxFixSize);
        for (si = 0; si < sizedivdmin; si++) {
            Conf cc = homheap.popMin();
            if (cc.p >= maxFixThresh * 2.0 || (cc.p >= minFixThresh * 2.0 && si >= sizedivdmax))
                break;
            constraints[cc.constraint_idx].type = Constraint::Type::NoConstr;
        }
    }

    currMatPhase = in->start_phase ? Haplotype::Alt : Haplotype::Ref;
    std::vector<fp_type> phaseProbs;
    phaseProbs.push_back(0);
    auto plain_site = [&](size_t m, Haplotype &mat, Haplotype &pat) {
        switch (geno(m)) {
        case Genotype::HomAlt:
            mat = Haplotype::Alt;
            pat = Haplotype::Alt;
            dmat[m] = 1.0;
            dpat[m] = 1.0;
            break;
        case Genotype::Miss:
            if (impMissing) {
                float af = in->af[m];
                dmat[m] = af;
                dpat[m] = af;
                mat = af > 0.5 ? Haplotype::Alt : Haplotype::Ref;
                pat = mat;
            }
            if (lastIter) totalmissings++;
            break;
        default:
            break;
        }
    };
    {
        size_t history = static_cast<size_t>((HistorySize * hist) + 0.5);
        Beam beam(history, P->beam_width, P->delta, pErr, pLimit, expectIBDcM, cMsplitDiffs,
                  callLocs, target, pbwt, constraints);
        size_t m = 0;
        while (csplits.size() > 0 ? m < csplits[callLocs[0]] : m < MFull) {
            Haplotype mat = Haplotype::Ref, pat = Haplotype::Ref;
            if (geno(m) == Genotype::Het) {
                mat = Haplotype::Ref;
                pat = Haplotype::Alt;
                dmat[m] = 0.5;
                dpat[m] = 0.5;
            } else
                plain_site(m, mat, pat);
            pmat.push_back(toBool(mat));
            ppat.push_back(toBool(pat));
            m++;
        }
        if (m < MFull) {
            pmat.push_back(toBool(currMatPhase));
            ppat.push_back(!toBool(currMatPhase));
            m++;
        }
        size_t next_call_idx = 1;
        for (; m < MFull; m++) {
            Haplotype mat = Haplotype::Ref, pat = Haplotype::Ref;
            if (geno(m) == Genotype::Het) {
                if (!mono(m)) {
                    fp_type p = beam.callKeepPhaseProb(next_call_idx) - 0.5;
                    Haplotype ref = currMatPhase;
                    phaseProbs.push_back(p);
                    if (p < 0)
                        currMatPhase = !ref;
                    else
                        currMatPhase = ref;
                    next_call_idx++;
                } else {
                    dmat[m] = 0.5;
                    dpat[m] = 0.5;
                    if (lastIter) totalmonhets++;
                }
                mat = currMatPhase;
                pat = !currMatPhase;
            } else
                plain_site(m, mat, pat);
            pmat.push_back(toBool(mat));
            ppat.push_back(toBool(pat));
        }
    }

    std::vector<fp_type> callConfs;
    if (lastIter && csplits.size() > 0) {
        callConfs.resize(callLocs.size(), 0.0);
        if (doRevPhasing) {
            callConfs[callLocs.size() - 1] = std::abs(phaseProbs.back()) + 0.5;
        } else {
            auto cc = callConfs.begin();
            for (const auto &p : phaseProbs) {
                *cc = std::abs(p) + 0.5;
                cc++;
            }
        }
    }
    if (lastIter && doRevPhasing && csplits.size() > 0) {
        pbwt.switchToReverse();
        size_t history = static_cast<size_t>((HistorySize * hist) + 0.5);
        std::vector<Constraint> revconstraints(csplits.size(),
                                               Constraint(Constraint::Type::NoConstr, 0));
        for (size_t cidx = 1; cidx < csplits.size(); cidx++) {
            Constraint &c = constraints[cidx];
            switch (c.type) {
            case Constraint::Type::HetKeep:
            case Constraint::Type::HetSwitch: {
                Constraint &rc = revconstraints[cidx - c.refidx - 1];
                if (rc.type != Constraint::Type::NoConstr) {
                    size_t ridx = cidx - c.refidx + rc.refidx;
                    revconstraints[ridx].refidx = c.refidx - rc.refidx - 1;
                    if (c.type != rc.type)
                        revconstraints[ridx].type = Constraint::Type::HetSwitch;
                    else
                        revconstraints[ridx].type = Constraint::Type::HetKeep;
                } else {
                    rc.type = c.type;
                    rc.refidx = c.refidx;
                }
            } break;
            case Constraint::Type::HomRef:
            case Constraint::Type::HomAlt:
                revconstraints[cidx].type = c.type;
                break;
            default:
                break;
            }
        }
        revconstraints[csplits.size() - 1] = Constraint(Constraint::Type::HetKeep, 0);
        float dref = phaseProbs.back() < 0.0 ? 0.5 + phaseProbs.back() : 0.5 - phaseProbs.back();
        float dalt = phaseProbs.back() < 0.0 ? 0.5 - phaseProbs.back() : 0.5 + phaseProbs.back();
        dmat[csplits[callLocs.back()]] = currMatPhase == Haplotype::Ref ? dref : dalt;
        dpat[csplits[callLocs.back()]] = currMatPhase == Haplotype::Ref ? dalt : dref;
        Beam beam(history, P->beam_width, P->delta, pErr, pLimit, expectIBDcM, cMsplitDiffs,
                  callLocs, target, pbwt, revconstraints, true);
        for (int rci = callLocs.size() - 2; rci >= 0; rci--) {
            fp_type p = beam.callKeepPhaseProb(rci) - 0.5;
            Haplotype ref = currMatPhase;
            if (std::abs(phaseProbs[rci + 1]) > std::abs(p)) p = phaseProbs[rci + 1];
            if (p < 0)
                currMatPhase = !ref;
            else
                currMatPhase = ref;
            pmat[csplits[callLocs[rci]]] = toBool(currMatPhase);
            ppat[csplits[callLocs[rci]]] = !toBool(currMatPhase);
            dref = p < 0.0 ? 0.5 + p : 0.5 - p;
            dalt = p < 0.0 ? 0.5 - p : 0.5 + p;
            dmat[csplits[callLocs[rci]]] = currMatPhase == Haplotype::Ref ? dref : dalt;
            dpat[csplits[callLocs[rci]]] = currMatPhase == Haplotype::Ref ? dalt : dref;
            callConfs[rci] = std::abs(p) + 0.5;
        }
    }
    if (lastIter && csplits.size() > 0) {
        for (size_t cci = 0; cci < callConfs.size(); cci++) totalconf += callConfs[cci];
        ncallsites += callConfs.size();
    }

    const size_t words = (MFull + 63) / 64;
    std::memset(out->mat, 0, words * 8);
    std::memset(out->pat, 0, words * 8);
    for (size_t m = 0; m < MFull; m++) {
        if (pmat[m]) out->mat[m >> 6] |= 1ull << (m & 63);
        if (ppat[m]) out->pat[m >> 6] |= 1ull << (m & 63);
    }
    std::memcpy(out->dosage_mat, dmat.data(), MFull * sizeof(float));
    std::memcpy(out->dosage_pat, dpat.data(), MFull * sizeof(float));
    if (out->call_conf && !callConfs.empty())
        std::memcpy(out->call_conf, callConfs.data(), callConfs.size() * sizeof(float));
    out->total_conf = totalconf;
    out->ncalls = (uint32_t)callLocs.size();
    out->ncallsites = ncallsites;
    out->nmissing = totalmissings;
    out->nmonhets = totalmonhets;
    return 0;
}

}  // extern "C"
