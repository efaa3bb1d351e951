// eimp_phase.cpp -- beam-search phasing of one target over its compressed
// PBWT (host code): the consumer of the engine's output.
//
// The reference wraps each target's cPBWT as PBWT(cpbwt, count0, K,
// 4*Kwords, Mp) and runs Target::phaseExtPBWT on it (src/GPUHandler.cpp:
// 259-266, src/Target.cpp:291-777): an optional fast pre-phasing beam that
// fixes confident calls as constraints, a forward beam that calls the phase
// at every call site, and a reverse beam whose calls replace the forward
// ones where they are more confident.  A beam holds pairs of haplotype paths
// (src/Beam.cpp, src/BeamPath.cpp); a haplotype path's probability of being
// extended by 0 or 1 sums, over its history, the sizes of the PBWT intervals
// that its recent haplotypes map to, weighted by recombination
// probabilities (src/HaplotypePath.cpp:260-479, src/PBWT.cpp:288-382).
//
// Results are identical to the reference's: every float is computed with the
// same operand types and in the same order (some intermediates are double,
// as the reference's literals and pow() calls make them), the beam keeps the
// reference's order of paths (pruning and merging depend on it where
// probabilities tie), and the selection heap makes the same comparisons.
// Compiled with -ffp-contract=off.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/eimp.h"

namespace eimp {
void set_last_error(const std::string &msg);  // eimp_engine.cpp
}

namespace {

using fp = float;  // the reference computes in float (src/Datatypes.h:68)

enum Gt : uint8_t { GT_HOMREF = 0, GT_HET = 1, GT_HOMALT = 2, GT_MISS = 3 };

// Constraint kinds, in the order of src/Datatypes.h:88-99
enum Kind : uint8_t { C_HOMREF, C_KEEP, C_SWITCH, C_HOMALT, C_NONE };
struct Constr {
    Kind kind = C_NONE;
    int ref = 0;  // history index (before extension) of the reference call
};

struct Failure {
    int code;
    std::string msg;
};
[[noreturn]] void raise(int code, const std::string &msg) { throw Failure{code, msg}; }

// GenotypeVector words: (is0, is2) word pairs (src/Datatypes.cpp:79-90)
inline Gt gt_at(const uint64_t *gt, size_t m) {
    const uint64_t b = 1ull << (m & 63);
    const bool is0 = gt[2 * (m >> 6)] & b, is2 = gt[2 * (m >> 6) + 1] & b;
    return is0 ? (is2 ? GT_MISS : GT_HOMREF) : (is2 ? GT_HOMALT : GT_HET);
}

// ---------------------------------------------------------------------------
// Interval of PBWT positions, borders inclusive; empty when l > r.
struct Iv {
    int l = 0, r = -1;
    bool empty() const { return l > r; }
    int size() const { return r - l + 1; }
    bool operator==(const Iv &o) const { return l == o.l && r == o.r; }
};

// Compressed PBWT: rows of (data, cumulative zeros) uint32 pairs.  Forward
// row m at m, backward row m at 2*Mp-1-m.
struct Cpbwt {
    const uint32_t *rows;
    size_t stride, Mp;
    const int32_t *c0;
    int K;
    fp kinv;
    bool rev = false;

    int count0(size_t m) const { return c0[m]; }

    // position at site m of the sequence at position x before it
    static int next_pos(const uint32_t *row, int gc0, int x) {
        const uint32_t data = row[2 * (x >> 5)], zeros_through = row[2 * (x >> 5) + 1];
        const uint32_t b = 1u << (x & 31);
        const int zeros_before =
            (int)(zeros_through - (uint32_t)__builtin_popcount(~(data | (b - 1))));
        return (data & b) ? gc0 + (x - zeros_before) : zeros_before;
    }

    // sequences in `from` extended by 0 land in to0, by 1 in to1
    void map(size_t m, const Iv &from, Iv &to0, Iv &to1) const {
        const uint32_t *row = rows + (rev ? (2 * Mp - 1 - m) : m) * stride;
        const int g = c0[m];
        to0 = Iv();
        to1 = Iv();
        if (from.empty()) return;
        if (from.l == from.r) {
            const int i = next_pos(row, g, from.l);
            (i < g ? to0 : to1) = Iv{i, i};
            return;
        }
        // each border maps directly into the side of its own bit; the other
        // side's border follows from the count of the other bit before it
        int a = next_pos(row, g, from.l), a_other = from.l - a + g;
        if (a >= g) std::swap(a, a_other);
        int b = next_pos(row, g, from.r), b_other = from.r - b + g - 1;
        if (b >= g) std::swap(b, b_other);
        to0 = Iv{a, b};
        to1 = Iv{a_other, b_other};
    }
};

// ---------------------------------------------------------------------------
// Haplotype history: the most recent haplotype at bit 0 of two words, unused
// bits held at one (src/BooleanHistory.cpp:24-77).  Reads past the 128 bits
// return what the reference reads there: the words after its bit array hold
// the history's fill count and its last index (src/BooleanHistory.h:82-85).
struct HapHist {
    uint64_t w[2];
    uint64_t fill[2];
    int last, histlast;

    void init(size_t H) {
        histlast = (int)H - 1;
        const size_t lw = (H - 1) / 64;
        const uint64_t lm = (H % 64) ? ~((1ull << (H % 64)) - 1ull) : 0ull;
        for (size_t i = 0; i < 2; i++) fill[i] = i < lw ? 0ull : i == lw ? lm : ~0ull;
        last = -1;
        w[0] = fill[0];
        w[1] = fill[1];
    }
    void push(bool b) {
        w[1] = (w[1] << 1) | (w[0] >> 63);
        w[0] = (w[0] << 1) | (b ? 1ull : 0ull);
        if (last != histlast) last++;
        w[0] |= fill[0];
        w[1] |= fill[1];
    }
    bool at(size_t pos) const {
        uint64_t word;
        if (pos < 128)
            word = w[pos >> 6];
        else if (pos < 192)
            word = (uint64_t)(uint32_t)last | ((uint64_t)(uint32_t)histlast << 32);
        else
            raise(EIMP_ERR_UNSUPPORTED, "phasing: a constraint or call looks " +
                                            std::to_string(pos) +
                                            " sites back into a haplotype history of at most 128");
        return (word >> (pos & 63)) & 1ull;
    }
    bool same(const HapHist &o) const {
        return w[0] == o.w[0] && w[1] == o.w[1] && last == o.last && histlast == o.histlast;
    }
    // number of equal recent sites (128 if all are equal)
    size_t common(const HapHist &o) const {
        uint64_t t = w[0] ^ o.w[0];
        if (t) return (size_t)__builtin_ctzll(t);
        t = w[1] ^ o.w[1];
        return t ? 64 + (size_t)__builtin_ctzll(t) : 128;
    }
};

// Probabilities along a path, newest at 0, in a power-of-two ring
// (src/History.h).
struct ProbRing {
    std::vector<fp> d;
    size_t mask = 0, start = 0;
    void init(size_t H) {
        size_t n = 1;
        while (n < H) n <<= 1;
        d.assign(n, 0.0f);
        mask = n - 1;
        start = 0;
    }
    void push(fp v) {
        start = (start - 1) & mask;
        d[start] = v;
    }
    fp &at(size_t i) { return d[(i + start) & mask]; }
    fp at(size_t i) const { return d[(i + start) & mask]; }
};

// ---------------------------------------------------------------------------
// One haplotype path (src/HaplotypePath.cpp)
struct HapPath {
    size_t H = 0;
    HapHist hist;
    fp p = 1.0f;        // path probability ...
    int32_t sexp = 0;   // ... times 2^sexp
    ProbRing ph;
    std::vector<Iv> iv;  // intervals of the path's recent sub-sequences, newest first
    fp ext_p[2] = {0.0f, 0.0f};
    int32_t ext_sexp = 0;
    std::vector<Iv> ext_iv[2];
    bool active[2] = {false, false};
    bool want[2] = {false, false};
    int refs[2] = {0, 0};

    void start(size_t H_, bool hap) {
        H = H_;
        hist.init(H);
        hist.push(hap);
        p = 1.0f;
        sexp = 0;
        ph.init(H);
        ph.push(1.0f);
        iv.clear();
        clear_ext();
        ext_iv[0].clear();
        ext_iv[1].clear();
        reserve();
    }
    // interval lists hold at most H entries: reserved once per pooled object
    void reserve() {
        iv.reserve(H + 1);
        ext_iv[0].reserve(H + 1);
        ext_iv[1].reserve(H + 1);
    }
    void clear_ext() {
        ext_p[0] = ext_p[1] = 0.0f;
        ext_sexp = 0;
        active[0] = active[1] = false;
        want[0] = want[1] = false;
        refs[0] = refs[1] = 0;
    }
    // this path continued by extension e of `o` (o keeps its extensions)
    void copy_extended(const HapPath &o, int e) {
        H = o.H;
        hist = o.hist;
        p = o.ext_p[e];
        sexp = o.ext_sexp;
        ph = o.ph;
        iv.assign(o.ext_iv[e].begin(), o.ext_iv[e].end());
        clear_ext();
        ext_iv[0].clear();
        ext_iv[1].clear();
        reserve();
        hist.push(e == 1);
        ph.push(o.ext_p[e]);
    }
    void apply(int e) {
        hist.push(e == 1);
        p = ext_p[e];
        sexp = ext_sexp;
        ph.push(ext_p[e]);
        iv.swap(ext_iv[e]);
        clear_ext();
    }
    void drop_ref(int e) {
        if (--refs[e] == 0) active[e] = false;
    }
    bool history_bit(int idx) const { return hist.at((size_t)idx); }

    void mark_required(const Constr &c) {
        // only ever sets flags: another beam path may already have set one
        if (c.kind == C_NONE) {
            want[0] = want[1] = true;
            return;
        }
        if ((c.kind == C_KEEP && !history_bit(c.ref)) || (c.kind == C_SWITCH && history_bit(c.ref)) ||
            c.kind == C_HOMREF)
            want[0] = true;
        else if ((c.kind == C_KEEP && history_bit(c.ref)) ||
                 (c.kind == C_SWITCH && !history_bit(c.ref)) || c.kind == C_HOMALT)
            want[1] = true;
    }

    // rescales the probability history when the path probability's binary
    // exponent reaches +-20 (src/HaplotypePath.cpp:226-258)
    void rescale() {
        const int e = std::ilogb(p);
        int32_t ns = sexp;
        if (e != INT_MIN && (e >= 20 || e <= -20)) {
            ns += e;
            const double f = std::ldexp(1.0, -e);
            for (size_t i = 0; i <= iv.size(); i++) ph.at(i) = (fp)((double)ph.at(i) * f);
        }
        ext_sexp = ns;
    }

    // probabilities of the 0- and 1-extensions at split site `cursor`; the
    // previously extended path `prev` shares its newest common sites' intervals
    int32_t extend(const std::vector<fp> &prec, size_t cursor, const HapPath *prev,
                   const Cpbwt &pb) {
        rescale();
        size_t ncopy = 0;
        if (prev) ncopy = std::min(H, 1 + hist.common(prev->hist));
        std::vector<Iv> &o0 = ext_iv[0], &o1 = ext_iv[1];
        o0.clear();
        o1.clear();
        const bool calc0 = want[0], calc1 = want[1];
        fp p0 = 0.0f, p1 = 0.0f;
        const size_t depth = std::min(H, iv.size() + 1);
        // weight of history depth i: its path probability times the
        // recombination probability (computed only for the depths visited)
        auto weight = [&](size_t i) -> fp { return ph.at(i) * prec[i]; };

        // depth 0: every conditioning haplotype, split by its allele here
        const int c0 = pb.count0(2 * cursor + 1);
        const fp w0 = weight(0);
        if (c0 > 0) {
            if (calc0) p0 = (fp)c0 * w0;
            o0.push_back(Iv{0, c0 - 1});
        }
        if ((size_t)c0 < (size_t)pb.K) {
            if (calc1) p1 = (fp)((size_t)pb.K - (size_t)c0) * w0;
            o1.push_back(Iv{c0, pb.K - 1});
        }
        bool stop = false;
        size_t i = 1;
        for (const size_t n = std::min(ncopy, iv.size() + 1); !stop && i < n; i++) {
            stop = true;
            const bool h0 = prev->ext_iv[0].size() > i, h1 = prev->ext_iv[1].size() > i;
            const fp wi = (h0 && calc0) || (h1 && calc1) ? weight(i) : 0.0f;
            if (h0) {
                const Iv &t = prev->ext_iv[0][i];
                if (calc0) p0 += (fp)t.size() * wi;
                o0.push_back(t);
                stop = false;
            }
            if (h1) {
                const Iv &t = prev->ext_iv[1][i];
                if (calc1) p1 += (fp)t.size() * wi;
                o1.push_back(t);
                stop = false;
            }
        }
        const size_t seg = pb.rev ? 2 * (cursor + 1) : 2 * cursor;  // inconsistency row
        for (; !stop && i < depth; i++) {
            const Iv &from = iv[i - 1];
            Iv t0, t1;
            if (i > 1 && iv[i - 2] == from) {  // same interval one site back: same result
                if (o0.size() == i) t0 = o0.back();
                if (o1.size() == i) t1 = o1.back();
            } else {
                Iv con, incon;
                pb.map(seg, from, con, incon);
                pb.map(2 * cursor + 1, con, t0, t1);
            }
            stop = true;
            const bool h0 = !t0.empty(), h1 = !t1.empty();
            const fp wi = (h0 && calc0) || (h1 && calc1) ? weight(i) : 0.0f;
            if (h0) {
                if (calc0) p0 += (fp)t0.size() * wi;
                o0.push_back(t0);
                stop = false;
            }
            if (h1) {
                if (calc1) p1 += (fp)t1.size() * wi;
                o1.push_back(t1);
                stop = false;
            }
        }
        ext_p[0] = p0;
        ext_p[1] = p1;
        active[0] = want[0];
        active[1] = want[1];
        return ext_sexp;
    }
};

// ---------------------------------------------------------------------------
// A pair of haplotype paths (maternal, paternal) in the beam (src/BeamPath.cpp)
struct PairPath {
    int mat = 0, pat = 0;  // indices into the beam's haplotype paths
    fp prob = 0.0f, weight = 0.0f;
    fp perr_base = 0.0f;   // per-error factor
    int nerr = 0;
    fp perr = 0.0f, perr1 = 0.0f;  // perr_base^nerr, ^(nerr+1)
    fp ext[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // extensions 00, 01, 10, 11 (mat, pat)
    bool act[4] = {false, false, false, false};
    int top = -1;  // highest active extension, -1 none
    Gt g = GT_MISS;  // target genotype at the extension site

    static bool inconsistent(int e, Gt g) {
        switch (g) {
        case GT_HET: return e == 0 || e == 3;
        case GT_HOMREF: return e > 0;
        case GT_HOMALT: return e < 3;
        default: return true;
        }
    }

    void clear_ext() {
        ext[0] = ext[1] = ext[2] = ext[3] = 0.0f;
        act[0] = act[1] = act[2] = act[3] = false;
        top = -1;
        g = GT_MISS;
    }

    void count_errors(int e, bool forget_one) {
        const int add = (inconsistent(e, g) ? 1 : 0) - (forget_one ? 1 : 0);
        nerr += add;
        if (add < 0) {
            perr1 = perr;
            perr = perr / perr_base;
        } else {
            for (int a = 0; a < add; a++) {
                perr = perr1;
                perr1 = perr1 * perr_base;
            }
        }
    }

    void scale_weight(int wse) {
        if (wse != 0) weight = (fp)((double)weight * std::ldexp(1.0, wse));
    }

    // continues this pair with extension e, its paths now at (m, p)
    void apply(int m, int p, int e, int wse, bool forget_one) {
        mat = m;
        pat = p;
        prob = ext[e];
        scale_weight(wse);
        count_errors(e, forget_one);
        clear_ext();
    }

    // a new pair: `o` continued by extension e (o unchanged)
    static PairPath branch(const PairPath &o, int m, int p, int e, int wse, bool forget_one) {
        PairPath n;
        n.mat = m;
        n.pat = p;
        n.prob = o.ext[e];
        n.weight = o.weight;
        n.perr_base = o.perr_base;
        n.nerr = o.nerr;
        n.perr = o.perr;
        n.perr1 = o.perr1;
        n.g = o.g;
        n.scale_weight(wse);
        n.count_errors(e, forget_one);
        n.clear_ext();
        return n;
    }

    int next_active_below(int x) const {
        for (int i = x - 1; i >= 0; i--)
            if (act[i]) return i;
        return -1;
    }
    void refresh_top() {
        top = -1;
        for (int i = 3; i >= 0; i--)
            if (act[i]) {
                top = i;
                break;
            }
    }
};

// ---------------------------------------------------------------------------
// Min-max heap (Atkinson, Sack, Santoro, Strothotte 1986): min at the root,
// max at one of its children.  The pruning keeps the beam's best paths with
// it; which of several equal-probability paths is dropped depends on the
// heap's exact comparisons, so these are the reference's: strict "<" on min
// levels, its negation on max levels (src/MinMaxHeap.h).
template <class T>
class MinMaxHeap {
    std::vector<T> h;

    static bool lt(const T &a, const T &b) { return a.p < b.p; }
    static bool min_level(size_t i) { return ((63 - __builtin_clzll(i + 1)) & 1) == 0; }

    void up_levels(size_t i, bool max_lvl) {
        while (i != 0) {
            const size_t par = (i - 1) / 2;
            if (par == 0) return;
            const size_t gp = (par - 1) / 2;
            if (lt(h[i], h[gp]) == max_lvl) return;
            std::swap(h[gp], h[i]);
            i = gp;
        }
    }
    void up(size_t i) {
        if (i == 0) return;
        const size_t par = (i - 1) / 2;
        if (min_level(i)) {
            if (lt(h[par], h[i])) {
                std::swap(h[par], h[i]);
                up_levels(par, true);
            } else {
                up_levels(i, false);
            }
        } else {
            if (lt(h[i], h[par])) {
                std::swap(h[par], h[i]);
                up_levels(par, false);
            } else {
                up_levels(i, true);
            }
        }
    }
    void down(size_t i) {
        const bool mx = !min_level(i);
        for (;;) {
            size_t x = i;
            const size_t c = 2 * i + 1, n = h.size();
            if (c < n && lt(h[c], h[x]) != mx) x = c;
            if (c + 1 < n && lt(h[c + 1], h[x]) != mx) x = c + 1;
            const size_t gc = 2 * c + 1;
            for (size_t k = 0; k < 4 && gc + k < n; k++)
                if (lt(h[gc + k], h[x]) != mx) x = gc + k;
            if (x == i) return;
            std::swap(h[i], h[x]);
            if (x - c <= 1) return;  // a child: done
            const size_t px = (x - 1) / 2;
            if (lt(h[px], h[x]) != mx) std::swap(h[px], h[x]);
            i = x;
        }
    }
    size_t max_index() const {
        if (h.size() <= 2) return h.size() - 1;
        return lt(h[1], h[2]) ? 2 : 1;
    }
    void erase_at(size_t i) {
        if (i + 1 == h.size()) {
            h.pop_back();
            return;
        }
        std::swap(h[i], h.back());
        h.pop_back();
        down(i);
    }

public:
    void clear() { h.clear(); }
    size_t size() const { return h.size(); }
    bool empty() const { return h.empty(); }
    void push(const T &v) {
        h.push_back(v);
        up(h.size() - 1);
    }
    const T &min() const { return h[0]; }
    const T &max() const { return h[max_index()]; }
    T pop_min() {
        T v = h[0];
        erase_at(0);
        return v;
    }
};

struct Cand {  // a pair path's extension competing for a beam slot
    size_t path;
    fp p;
    int e;
};
struct Conf {  // a constraint and its confidence
    size_t idx;
    fp p;
};

// ---------------------------------------------------------------------------
// Beam over split sites (src/Beam.cpp)
class Beam {
public:
    Beam(size_t H, size_t P, size_t delta, fp perr, fp plimit, fp ibd, const std::vector<fp> &cmd,
         const std::vector<size_t> &calls, const std::vector<Gt> &target, Cpbwt &pb,
         const std::vector<Constr> &constraints, bool reverse)
        : H_(H), P_(P), delta_(delta), perr_(perr), plimit_(plimit), ibd_(ibd), cmd_(cmd),
          calls_(calls), target_(target), cons_(constraints), pb_(pb), rev_(reverse) {
        prec_.assign(H, 0.0f);
        cursor_ = rev_ ? (int)cons_.size() : -1;
        HapPath *m = take();
        m->start(H_, false);
        HapPath *p = take();
        p->start(H_, true);
        haps_ = {m, p};
        PairPath bp;
        bp.mat = 0;
        bp.pat = 1;
        bp.weight = 1.0f;
        bp.perr_base = perr_;
        bp.nerr = 0;
        bp.perr = 1.0f;
        bp.perr1 = 1.0f * perr_;
        // 2^(0+0) * 1 * 1 * 1, evaluated in double as the reference does
        bp.prob = (fp)(std::ldexp(1.0, m->sexp + p->sexp) * (double)m->p * (double)p->p * 1.0);
        bp.clear_ext();
        beam_.push_back(bp);
    }

    // probability of keeping the phase from the previous call site (or the
    // one before it) to call site ci, at the beam front delta sites ahead
    fp keep_prob(size_t ci, bool two_back) {
        if (ci >= calls_.size()) raise(EIMP_ERR_INVALID, "phasing: call index out of range");
        const int front = rev_ ? std::max((int)calls_[ci] - (int)delta_, 0)
                               : (int)std::min(calls_[ci] + delta_, cons_.size() - 1);
        const int cur = rev_ ? (int)calls_[ci] - front : front - (int)calls_[ci];
        int ref = (int)(rev_ ? (two_back ? calls_[ci + 2] : calls_[ci + 1])
                             : (two_back ? calls_[ci - 2] : calls_[ci - 1]));
        ref = rev_ ? ref - front : front - ref;
        advance_to(front);
        fp keep = 0.0f, sw = 0.0f;
        for (const PairPath &bp : beam_) {
            const HapHist &hm = haps_[bp.mat]->hist, &hp = haps_[bp.pat]->hist;
            const bool cm = hm.at((size_t)cur), cp = hp.at((size_t)cur);
            const bool rm = hm.at((size_t)ref), rp = hp.at((size_t)ref);
            const fp p = bp.weight * bp.prob;
            if (cm != cp && rm != rp) {
                if (cm == rm)
                    keep += p;
                else
                    sw += p;
            } else {  // homozygous at one of the two: half and half
                keep += p / 2;
                sw += p / 2;
            }
        }
        fp r = 0.5f;
        if (keep + sw > 0) r = keep / (keep + sw);
        return r;
    }

    // diploid alt dosage at split site si
    fp dosage(size_t si) {
        if (si >= cons_.size()) raise(EIMP_ERR_INVALID, "phasing: split index out of range");
        const int front = rev_ ? std::max((int)si - (int)delta_, 0)
                               : (int)std::min(si + delta_, cons_.size() - 1);
        const int cur = rev_ ? (int)si - front : front - (int)si;
        advance_to(front);
        fp dos = 0.0f, tot = 0.0f;
        for (const PairPath &bp : beam_) {
            const bool cm = haps_[bp.mat]->hist.at((size_t)cur);
            const bool cp = haps_[bp.pat]->hist.at((size_t)cur);
            const fp p = bp.weight * bp.prob;
            tot += p;
            if (cm == cp) {
                if (cm) dos += 2 * p;
            } else {
                dos += p;
            }
        }
        fp r = 1.0f;
        if (tot > 0.0f) r = dos / tot;
        return r;
    }

    ~Beam() {
        for (HapPath *h : haps_) delete h;
        for (HapPath *h : spare_) delete h;
    }

private:
    HapPath *take() {
        if (spare_.empty()) return new HapPath();
        HapPath *h = spare_.back();
        spare_.pop_back();
        return h;
    }
    void give(HapPath *h) { spare_.push_back(h); }

    void advance_to(int front) {
        while ((rev_ && cursor_ > front) || (!rev_ && cursor_ < front)) advance();
    }

    // recombination probabilities for the history depths at the new cursor
    // (src/Beam.cpp:541-572)
    void recombination() {
        int c = rev_ ? cursor_ + 1 : cursor_;
        const fp cursordiff = rev_ ? cmd_[cursor_] : cmd_[cursor_ + 1];
        fp d = 0.0f;
        for (int i = 0; i < (int)H_; i++) {
            const fp t1 = (fp)(1.0 + (double)(d / ibd_));
            fp pr = (fp)(1.0 / (double)(t1 * t1));
            const fp t2 = (fp)(1.0 + (double)((d + cursordiff) / ibd_));
            pr = (fp)((double)pr - 1.0 / (double)(t2 * t2));
            pr = std::max(std::min(pr, (fp)1.0), (fp)0.000001);
            prec_[i] = pr * pb_.kinv;
            d += cmd_[c];
            if (rev_) {
                c++;
                if (c >= (int)cmd_.size()) break;
            } else {
                if (c == 0) break;
                c--;
            }
        }
    }

    void advance() {
        cursor_ += rev_ ? -1 : 1;
        recombination();
        Constr c = cons_[cursor_];
        for (;;) {
            for (PairPath &bp : beam_) {
                haps_[bp.mat]->mark_required(c);
                haps_[bp.pat]->mark_required(c);
            }
            const HapPath *prev = nullptr;
            int32_t minse = 0, maxse = INT32_MIN;
            for (HapPath *hp : haps_) {
                if (!(hp->want[0] || hp->want[1])) continue;
                const int32_t se = hp->extend(prec_, (size_t)cursor_, prev, pb_);
                prev = hp;
                minse = std::min(minse, se);
                maxse = std::max(maxse, se);
            }
            // normalises the product of the most and the least probable paths
            const int32_t bpse = (int32_t)(0u - ((uint32_t)maxse + (uint32_t)minse));
            fp maxP = 0.0f;
            for (PairPath &bp : beam_) {
                const fp m = extend_pair(bp, target_[cursor_], c, bpse);
                if (maxP < m) maxP = m;
            }
            if (prune(maxP, c)) break;
            if (c.kind == C_NONE) raise(EIMP_ERR_STATE, "phasing: the beam is empty");
            c.kind = C_NONE;  // drop the constraint and try again
        }
        merge();
    }

    fp extend_pair(PairPath &bp, Gt g, const Constr &c, int32_t bpse) {
        bp.g = g;
        HapPath *pm = haps_[bp.mat], *pp = haps_[bp.pat];
        const fp m0 = pm->ext_p[0], m1 = pm->ext_p[1], q0 = pp->ext_p[0], q1 = pp->ext_p[1];
        const bool am0 = pm->active[0], am1 = pm->active[1];
        const bool ap0 = pp->active[0], ap1 = pp->active[1];
        const fp scale = (fp)std::ldexp(1.0, pm->ext_sexp + pp->ext_sexp + bpse);
        fp maxP = 0.0f;
        auto single = [&](int id, fp a, fp b) {
            bp.ext[id] = a * b * scale;
            if (PairPath::inconsistent(id, g))
                bp.ext[id] *= bp.perr1;
            else
                bp.ext[id] *= bp.perr;
            if (maxP < bp.ext[id]) maxP = bp.ext[id];
            bp.act[id] = true;
            bp.top = id;
        };
        const bool none = c.kind == C_NONE;
        if (none || c.kind == C_HOMREF) {
            if (am0 && ap0) {
                single(0, m0, q0);
                pm->refs[0]++;
                pp->refs[0]++;
            }
        }
        if (none || (c.kind == C_KEEP && !pm->history_bit(c.ref)) ||
            (c.kind == C_SWITCH && pm->history_bit(c.ref))) {
            if (am0 && ap1) {
                single(1, m0, q1);
                pm->refs[0]++;
                pp->refs[1]++;
            }
        }
        if (none || (c.kind == C_KEEP && pm->history_bit(c.ref)) ||
            (c.kind == C_SWITCH && !pm->history_bit(c.ref))) {
            if (am1 && ap0) {
                single(2, m1, q0);
                pm->refs[1]++;
                pp->refs[0]++;
            }
        }
        if (none || c.kind == C_HOMALT) {
            if (am1 && ap1) {
                single(3, m1, q1);
                pm->refs[1]++;
                pp->refs[1]++;
            }
        }
        return maxP;
    }

    void drop_ext(PairPath &bp, int e) {
        bp.act[e] = false;
        haps_[bp.mat]->drop_ref(e / 2);
        haps_[bp.pat]->drop_ref(e % 2);
        bp.refresh_top();
    }

    // keeps the P most probable extensions above pLimit * max, then applies
    // them to the haplotype paths and the pairs (src/Beam.cpp:576-775)
    bool prune(fp maxP, const Constr &c) {
        const fp limit = plimit_ * maxP;
        heap_.clear();
        for (size_t i = 0; i < beam_.size(); i++) {
            PairPath &bp = beam_[i];
            int e = bp.top;
            while (e >= 0) {
                const fp p = bp.ext[e];
                if (p >= limit) {
                    if (heap_.size() < P_ || (heap_.size() == P_ && p > heap_.min().p)) {
                        if (heap_.size() == P_) {
                            const Cand out = heap_.pop_min();
                            drop_ext(beam_[out.path], out.e);
                        }
                        heap_.push(Cand{i, p, e});
                    } else {
                        drop_ext(bp, e);
                    }
                } else {
                    drop_ext(bp, e);
                }
                e = c.kind == C_NONE ? e - 1 : -1;
            }
        }
        if (heap_.empty()) return false;

        const int wse = -std::ilogb(beam_[heap_.min().path].weight);
        const bool forget = beam_[heap_.max().path].nerr > 0;

        // haplotype paths: 0-extensions first, in order, then 1-extensions
        old_ = haps_;
        size_t n0 = 0;
        for (const HapPath *hp : haps_)
            if (hp->active[0]) n0++;
        map_.clear();
        haps_.resize(n0);
        size_t one = n0, zero = 0;
        for (HapPath *hp : old_) {
            map_.emplace_back((int)zero, (int)one);
            if (!hp->active[0] && !hp->active[1]) {
                give(hp);
                continue;
            }
            if (hp->active[0] && hp->active[1]) {
                HapPath *cp = take();
                cp->copy_extended(*hp, 1);
                haps_.push_back(cp);
                one++;
                hp->apply(0);
                haps_[zero++] = hp;
                continue;
            }
            if (hp->active[0]) {
                hp->apply(0);
                haps_[zero++] = hp;
            } else {
                hp->apply(1);
                haps_.push_back(hp);
                one++;
            }
        }

        // pairs: survivors keep their order, branches follow in creation order
        next_.clear();
        branches_.clear();
        for (PairPath &bp : beam_) {
            int e = bp.top;
            if (e < 0) continue;
            const int m[2] = {map_[bp.mat].first, map_[bp.mat].second};
            const int p[2] = {map_[bp.pat].first, map_[bp.pat].second};
            int nx = bp.next_active_below(e);
            while (nx != -1) {
                branches_.push_back(PairPath::branch(bp, m[e / 2], p[e % 2], e, wse, forget));
                e = nx;
                nx = bp.next_active_below(nx);
            }
            bp.apply(m[e / 2], p[e % 2], e, wse, forget);
            next_.push_back(bp);
        }
        next_.insert(next_.end(), branches_.begin(), branches_.end());
        beam_.swap(next_);
        return true;
    }

    // pairs whose haplotype paths have equal histories are merged into the
    // more probable one (src/Beam.cpp:778-836)
    void merge() {
        keys_.clear();
        alive_.assign(beam_.size(), 1);
        for (size_t i = 0; i < beam_.size(); i++) {
            size_t a = (size_t)beam_[i].mat;
            while (a > 0 && haps_[a]->hist.same(haps_[a - 1]->hist)) a--;
            size_t b = (size_t)beam_[i].pat;
            while (b > 0 && haps_[b]->hist.same(haps_[b - 1]->hist)) b--;
            if (a > b) std::swap(a, b);
            const size_t key = (a << 32) | b;
            auto it = keys_.find(key);
            if (it == keys_.end()) {
                keys_.emplace(key, i);
                continue;
            }
            PairPath &o = beam_[it->second];
            PairPath &n = beam_[i];
            if (o.prob > n.prob) {
                o.weight += n.weight * (n.prob / o.prob);
                alive_[i] = 0;
            } else {
                n.weight += o.weight * (o.prob / n.prob);
                alive_[it->second] = 0;
                it->second = i;
            }
        }
        size_t k = 0;
        for (size_t i = 0; i < beam_.size(); i++)
            if (alive_[i]) beam_[k++] = beam_[i];
        beam_.resize(k);
    }

    size_t H_, P_, delta_;
    fp perr_, plimit_, ibd_;
    const std::vector<fp> &cmd_;
    const std::vector<size_t> &calls_;
    const std::vector<Gt> &target_;
    const std::vector<Constr> &cons_;  // read live: the phasing loop fills it while this beam runs
    Cpbwt &pb_;
    bool rev_;
    int cursor_;
    std::vector<PairPath> beam_, next_, branches_;
    std::vector<HapPath *> haps_, old_, spare_;
    std::vector<std::pair<int, int>> map_;
    std::vector<fp> prec_;
    MinMaxHeap<Cand> heap_;
    std::unordered_map<size_t, size_t> keys_;
    std::vector<uint8_t> alive_;
};

// ---------------------------------------------------------------------------
// Target::phaseExtPBWT (src/Target.cpp:291-777)
struct Phaser {
    Phaser(const eimp_phase_params &P_, const eimp_phase_input &in_, eimp_phase_output &out_)
        : P(P_), in(in_), out(out_) {}

    const eimp_phase_params &P;
    const eimp_phase_input &in;
    eimp_phase_output &out;

    size_t W = 0, S = 0;
    std::vector<size_t> splits, calls;
    std::vector<Gt> target;
    std::vector<fp> cmd;
    std::vector<uint8_t> is_call;  // per window site

    void set_bit(uint64_t *v, size_t m, bool b) {
        if (b)
            v[m >> 6] |= 1ull << (m & 63);
        else
            v[m >> 6] &= ~(1ull << (m & 63));
    }

    size_t history(uint32_t base) const {
        const size_t H = (size_t)((double)((fp)base * P.hist) + 0.5);
        if (H == 0 || H > 128)
            raise(EIMP_ERR_INVALID, "phasing: history length " + std::to_string(H) +
                                        " outside 1..128");
        return H;
    }

    // allele and dosages at a site without a phase call
    void fixed_site(size_t m, Gt g, bool &mat, bool &pat) {
        mat = pat = false;
        switch (g) {
        case GT_HOMALT:
            mat = pat = true;
            out.dosage_mat[m] = 1.0f;
            out.dosage_pat[m] = 1.0f;
            break;
        case GT_MISS:
            if (P.imp_missing) {
                const float af = in.af[m];
                out.dosage_mat[m] = af;
                out.dosage_pat[m] = af;
                mat = pat = af > 0.5;
            }
            if (P.last_iter) out.nmissing++;
            break;
        default:
            break;
        }
    }

    void run() {
        W = in.W;
        S = (in.Mp - 1) / 2;
        splits.assign(in.splitsites, in.splitsites + S);
        cmd.assign(in.cMdiffs, in.cMdiffs + (S ? S + 1 : 0));
        target.resize(S);
        is_call.assign(W, 0);
        for (size_t i = 0; i < S; i++) {
            target[i] = gt_at(in.gt, splits[i]);
            if (target[i] == GT_HET) {
                calls.push_back(i);
                is_call[splits[i]] = 1;
            }
        }
        Cpbwt pb{in.pbwt, in.row_stride, in.Mp, in.count0, (int)in.K,
                 (fp)(1.0 / (double)(fp)in.K)};
        const size_t words = (W + 63) / 64;
        std::memset(out.mat, 0, words * 8);
        std::memset(out.pat, 0, words * 8);
        std::fill(out.dosage_mat, out.dosage_mat + W, 0.0f);
        std::fill(out.dosage_pat, out.dosage_pat + W, 0.0f);
        out.total_conf = 0.0f;
        out.ncalls = (uint32_t)calls.size();
        out.ncallsites = 0;
        out.nmissing = 0;
        out.nmonhets = 0;

        std::vector<Constr> cons(S);
        bool cur = in.start_phase != 0, lastp = cur;

        if (P.pre_phasing) {
            Beam beam(history(P.history_fast), P.beam_width_fast, P.delta_fast, P.p_err,
                      P.p_limit, P.expect_ibd_cM, cmd, calls, target, pb, cons, false);
            if (S > 0) cons[0] = Constr{C_KEEP, 0};
            MinMaxHeap<Conf> callheap, homheap;
            size_t nci = 1;
            for (size_t si = 1; si < S; si++) {
                if (nci < calls.size() && si == calls[nci]) {
                    fp p = (fp)((double)beam.keep_prob(nci, false) - 0.5);
                    bool ref = cur;
                    Constr c{C_NONE, (int)(calls[nci] - calls[nci - 1] - 1)};
                    if (nci > 1) {
                        const fp p2 = (fp)((double)beam.keep_prob(nci, true) - 0.5);
                        if (std::abs(p2) > std::abs(p)) {
                            p = p2;
                            ref = lastp;
                            c.ref = (int)(calls[nci] - calls[nci - 2] - 1);
                        }
                    }
                    lastp = cur;
                    if (p < 0) {
                        cur = !ref;
                        c.kind = C_SWITCH;
                    } else {
                        cur = ref;
                        c.kind = C_KEEP;
                    }
                    cons[si] = c;
                    callheap.push(Conf{si, std::abs(p)});
                    nci++;
                } else {
                    fp p = beam.dosage(si);
                    if (target[si] == GT_HOMREF) {
                        cons[si] = Constr{C_HOMREF, 0};
                        p = (fp)(2.0 - (double)p);
                    } else {
                        cons[si] = Constr{C_HOMALT, 0};
                    }
                    homheap.push(Conf{si, p});
                }
            }
            // keep at least min_fix_size of the calls fixed, more of them if
            // confident (src/Target.cpp:401-435)
            size_t lo = (size_t)((fp)callheap.size() * P.min_fix_size);
            size_t hi = (size_t)((double)callheap.size() * (1.0 - (double)P.max_fix_size));
            size_t si;
            for (si = 0; si < lo; si++) {
                const Conf cc = callheap.pop_min();
                if ((double)cc.p >= (double)P.max_fix_thresh - 0.5 ||
                    ((double)cc.p >= (double)P.min_fix_thresh - 0.5 && si >= hi))
                    break;
                cons[cc.idx].kind = C_NONE;
            }
            if (si == lo) {
                for (; si < callheap.size(); si++) {
                    const Conf cc = callheap.pop_min();
                    if (cc.p > 0.0f) break;
                    cons[cc.idx].kind = C_NONE;
                }
            }
            lo = (size_t)((fp)homheap.size() * P.min_fix_size);
            hi = (size_t)((double)homheap.size() * (1.0 - (double)P.max_fix_size));
            for (si = 0; si < lo; si++) {
                const Conf cc = homheap.pop_min();
                if ((double)cc.p >= (double)P.max_fix_thresh * 2.0 ||
                    ((double)cc.p >= (double)P.min_fix_thresh * 2.0 && si >= hi))
                    break;
                cons[cc.idx].kind = C_NONE;
            }
        }

        // forward phasing and phase calls
        cur = in.start_phase != 0;
        std::vector<fp> probs;
        probs.reserve(calls.size());
        probs.push_back(0.0f);
        const size_t H = history(P.history);
        {
            Beam beam(H, P.beam_width, P.delta, P.p_err, P.p_limit, P.expect_ibd_cM, cmd, calls,
                      target, pb, cons, false);
            size_t m = 0;
            const size_t first = S > 0 ? splits[calls[0]] : W;
            for (; m < first; m++) {  // up to the first call site nothing is phased
                const Gt g = gt_at(in.gt, m);
                bool a, b;
                if (g == GT_HET) {  // a mono het
                    a = false;
                    b = true;
                    out.dosage_mat[m] = 0.5f;
                    out.dosage_pat[m] = 0.5f;
                } else {
                    fixed_site(m, g, a, b);
                }
                set_bit(out.mat, m, a);
                set_bit(out.pat, m, b);
            }
            if (m < W) {  // the first call site: the start phase
                set_bit(out.mat, m, cur);
                set_bit(out.pat, m, !cur);
                m++;
            }
            size_t nci = 1;
            for (; m < W; m++) {
                const Gt g = gt_at(in.gt, m);
                bool a, b;
                if (g == GT_HET) {
                    if (is_call[m]) {
                        const fp p = (fp)((double)beam.keep_prob(nci, false) - 0.5);
                        const bool ref = cur;
                        probs.push_back(p);
                        cur = p < 0 ? !ref : ref;
                        nci++;
                    } else {  // mono het
                        out.dosage_mat[m] = 0.5f;
                        out.dosage_pat[m] = 0.5f;
                        if (P.last_iter) out.nmonhets++;
                    }
                    a = cur;
                    b = !cur;
                } else {
                    fixed_site(m, g, a, b);
                }
                set_bit(out.mat, m, a);
                set_bit(out.pat, m, b);
            }
        }

        std::vector<fp> conf;
        const size_t nc = calls.size();
        if (P.last_iter && S > 0) {
            conf.assign(nc, 0.0f);
            if (P.rev_phasing) {
                conf[nc - 1] = (fp)((double)std::abs(probs.back()) + 0.5);
            } else {
                for (size_t i = 0; i < probs.size(); i++)
                    conf[i] = (fp)((double)std::abs(probs[i]) + 0.5);
            }
        }

        // reverse phasing: a call is replaced where the reverse beam is more
        // confident (src/Target.cpp:622-741)
        if (P.last_iter && P.rev_phasing && S > 0) {
            pb.rev = true;
            std::vector<Constr> rc(S);
            for (size_t ci = 1; ci < S; ci++) {
                const Constr &c = cons[ci];
                switch (c.kind) {
                case C_KEEP:
                case C_SWITCH: {
                    Constr &r = rc[ci - c.ref - 1];  // the site the constraint refers to
                    if (r.kind != C_NONE) {  // chain through that site's own constraint
                        const size_t ridx = ci - c.ref + r.ref;
                        rc[ridx].ref = c.ref - r.ref - 1;
                        rc[ridx].kind = c.kind != r.kind ? C_SWITCH : C_KEEP;
                    } else {
                        r.kind = c.kind;
                        r.ref = c.ref;
                    }
                    break;
                }
                case C_HOMREF:
                case C_HOMALT:
                    rc[ci].kind = c.kind;
                    break;
                default:
                    break;
                }
            }
            rc[S - 1] = Constr{C_KEEP, 0};

            const fp pb_last = probs.back();
            fp dref = (fp)(pb_last < 0.0f ? 0.5 + (double)pb_last : 0.5 - (double)pb_last);
            fp dalt = (fp)(pb_last < 0.0f ? 0.5 - (double)pb_last : 0.5 + (double)pb_last);
            const size_t lastsite = splits[calls.back()];
            out.dosage_mat[lastsite] = !cur ? dref : dalt;
            out.dosage_pat[lastsite] = !cur ? dalt : dref;

            Beam beam(H, P.beam_width, P.delta, P.p_err, P.p_limit, P.expect_ibd_cM, cmd, calls,
                      target, pb, rc, true);
            for (int ci = (int)nc - 2; ci >= 0; ci--) {
                fp p = (fp)((double)beam.keep_prob((size_t)ci, false) - 0.5);
                const bool ref = cur;
                if (std::abs(probs[ci + 1]) > std::abs(p)) p = probs[ci + 1];
                cur = p < 0 ? !ref : ref;
                const size_t site = splits[calls[ci]];
                set_bit(out.mat, site, cur);
                set_bit(out.pat, site, !cur);
                dref = (fp)(p < 0.0f ? 0.5 + (double)p : 0.5 - (double)p);
                dalt = (fp)(p < 0.0f ? 0.5 - (double)p : 0.5 + (double)p);
                out.dosage_mat[site] = !cur ? dref : dalt;
                out.dosage_pat[site] = !cur ? dalt : dref;
                conf[ci] = (fp)((double)std::abs(p) + 0.5);
            }
        }

        if (P.last_iter && S > 0) {
            for (size_t i = 0; i < conf.size(); i++) out.total_conf += conf[i];
            out.ncallsites += conf.size();
            if (out.call_conf) std::memcpy(out.call_conf, conf.data(), conf.size() * sizeof(fp));
        }
    }
};

// The FPGA handler's raw rows (data words only, bits past K one) turned into
// (data, cumulative zero count) pairs as PBWT(cpbwt, K, Kwords, M) does, with
// count0 taken from each forward row's count after its last logical word
// (src/PBWT.cpp:72-110).
struct PairRows {
    std::vector<uint32_t> rows;
    std::vector<int32_t> c0;
};
void raw_to_pairs(const eimp_phase_input &in, PairRows &out) {
    const size_t nw = (in.K + 31) / 32, Mp = in.Mp;
    out.rows.resize(2 * Mp * 2 * nw);
    out.c0.resize(Mp);
    for (size_t r = 0; r < 2 * Mp; r++) {
        const uint32_t *src = in.pbwt + r * (size_t)in.row_stride;
        uint32_t *dst = out.rows.data() + r * 2 * nw;
        uint32_t zeros = 0;
        for (size_t k = 0; k < nw; k++) {
            zeros += 32u - (uint32_t)__builtin_popcount(src[k]);
            dst[2 * k] = src[k];
            dst[2 * k + 1] = zeros;
        }
        if (r < Mp) out.c0[r] = (int32_t)zeros;
    }
}

void validate(const eimp_phase_params *P, const eimp_phase_input *in,
              const eimp_phase_output *out) {
    if (!P || !in || !out) raise(EIMP_ERR_INVALID, "phasing: null argument");
    if (!in->pbwt || (!in->count0 && !in->raw_rows) || !in->gt || !out->mat || !out->pat ||
        !out->dosage_mat || !out->dosage_pat)
        raise(EIMP_ERR_INVALID, "phasing: null buffer");
    if (in->K == 0 || in->Mp == 0 || in->Mp % 2 == 0)
        raise(EIMP_ERR_INVALID, "phasing: K > 0 and an odd Mp = 2*S+1 are required");
    if (in->raw_rows ? (uint64_t)in->row_stride < (in->K + 31) / 32
                     : (uint64_t)in->row_stride < 2ull * ((in->K + 31) / 32))
        raise(EIMP_ERR_INVALID, in->raw_rows ? "phasing: raw row stride below ceil(K/32)"
                                             : "phasing: row stride below 2*ceil(K/32)");
    const size_t S = (in->Mp - 1) / 2;
    if (S > 0 && (!in->splitsites || !in->cMdiffs))
        raise(EIMP_ERR_INVALID, "phasing: split sites and cM differences required");
    for (size_t i = 0; i < S; i++)
        if (in->splitsites[i] >= in->W || (i > 0 && in->splitsites[i] <= in->splitsites[i - 1]))
            raise(EIMP_ERR_INVALID, "phasing: split sites must ascend inside the window");
    if (!in->raw_rows)
        for (size_t m = 0; m < in->Mp; m++)
            if (in->count0[m] < 0 || (uint32_t)in->count0[m] > in->K)
                raise(EIMP_ERR_INVALID, "phasing: count0 outside 0..K");
    if (P->imp_missing && !in->af) raise(EIMP_ERR_INVALID, "phasing: imp_missing needs af");
    if (P->beam_width == 0 || (P->pre_phasing && P->beam_width_fast == 0))
        raise(EIMP_ERR_INVALID, "phasing: beam width 0");
    if (S > 0) {
        bool any = false;
        for (size_t i = 0; i < S && !any; i++) any = gt_at(in->gt, in->splitsites[i]) == GT_HET;
        if (!any) raise(EIMP_ERR_INVALID, "phasing: split sites without a call site");
        if (gt_at(in->gt, in->splitsites[S - 1]) != GT_HET)
            raise(EIMP_ERR_INVALID, "phasing: the last split site must be a call site");
    }
}

int phase_one(const eimp_phase_params *P, const eimp_phase_input *in, eimp_phase_output *out,
              std::string &err) {
    try {
        validate(P, in, out);
        if (in->raw_rows) {
            PairRows pr;
            raw_to_pairs(*in, pr);
            eimp_phase_input conv = *in;
            conv.pbwt = pr.rows.data();
            conv.count0 = pr.c0.data();
            conv.row_stride = (uint32_t)(2 * ((in->K + 31) / 32));
            conv.raw_rows = 0;
            Phaser ph(*P, conv, *out);
            ph.run();
        } else {
            Phaser ph(*P, *in, *out);
            ph.run();
        }
        return EIMP_OK;
    } catch (const Failure &f) {
        err = f.msg;
        return f.code;
    } catch (const std::bad_alloc &) {
        err = "phasing: out of host memory";
        return EIMP_ERR_NOMEM;
    }
}

}  // namespace

extern "C" {

void eimp_phase_default_params(eimp_phase_params *p) {
    if (!p) return;
    p->history = 100;
    p->history_fast = 30;
    p->beam_width = 50;
    p->beam_width_fast = 30;
    p->delta = 20;
    p->delta_fast = 10;
    p->hist = 1.0f;
    p->p_err = 0.003f;
    p->p_limit = 0.003f;
    p->expect_ibd_cM = 2.0f;
    p->min_fix_size = 0.5f;
    p->max_fix_size = 0.9f;
    p->min_fix_thresh = 0.99f;
    p->max_fix_thresh = 1.0f;
    p->pre_phasing = 0;
    p->rev_phasing = 1;
    p->last_iter = 1;
    p->imp_missing = 0;
}

int eimp_phase_target(const eimp_phase_params *params, const eimp_phase_input *in,
                      eimp_phase_output *out) {
    std::string err;
    const int rc = phase_one(params, in, out, err);
    if (rc != EIMP_OK) eimp::set_last_error(err);
    return rc;
}

int eimp_phase_batch(const eimp_phase_params *params, size_t n, const eimp_phase_input *in,
                     eimp_phase_output *out, int threads) {
    if (n == 0) return EIMP_OK;
    if (!in || !out) {
        eimp::set_last_error("phasing: null argument");
        return EIMP_ERR_INVALID;
    }
    size_t nt = threads > 0 ? (size_t)threads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, n);
    std::atomic<size_t> next{0};
    std::atomic<int> first_rc{EIMP_OK};
    std::string first_err;
    std::atomic<bool> err_set{false};
    auto work = [&]() {
        std::string err;
        for (size_t i; (i = next.fetch_add(1)) < n;) {
            const int rc = phase_one(params, &in[i], &out[i], err);
            if (rc != EIMP_OK) {
                bool expect = false;
                if (err_set.compare_exchange_strong(expect, true)) {
                    first_err = "target " + std::to_string(i) + ": " + err;
                    first_rc = rc;
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (first_rc != EIMP_OK) eimp::set_last_error(first_err);
    return first_rc;
}

}  // extern "C"
