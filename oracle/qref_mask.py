"""Sample masking of Qref panels restated in numpy (the checker of
eimp_qref_set_mask / eimp_qref_read_sites_af).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ -- never by the product package.

Reference behaviour restated here:
  * rows: the masked samples' bit pairs (haplotypes 2s, 2s+1) leave the
    record and the later haplotypes move down (src/utils.h:245-268 for
    run-length coded records, src/VCFData.cpp:2456-2520 for raw ones);
  * masked variant alleles: coded records count a haploid sample only when
    both of its bits are set (src/utils.h:250-256), raw records count one for
    any set bit and a second one for a diploid sample with both set
    (src/VCFData.cpp:2468-2475);
  * haploid flags of the kept samples and the reduced haploid count
    (src/VCFData.cpp:2170-2208);
  * the corrected allele frequency (src/VCFData.cpp:2426-2432, 2524-2529),
    with the product and subtraction fused into one rounding as GCC compiles
    them under the reference's -O3 -march=native build on an FMA host.

Pinning: the coded-record rows and counts against the reference's own
decoder (oracle/ref_harness.cpp, live and through tests/golden/
ref_rle_masked.npz) and against brute force; the fused rounding against the
same expression compiled by g++ with the reference's flags; the frequency
against the closed form (ac - masked) / new_an where old_an is a power of
two.  The raw-record count of a haploid sample stored heterozygous ((1,0) or
(0,1), which the reference's writer never produces) is parity unpinned: no
reference code for it runs here.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def _bits(words: np.ndarray) -> np.ndarray:
    return np.unpackbits(np.ascontiguousarray(words, np.uint64).view(np.uint8), bitorder="little")


def masked_record(full_words, n_samples, mask, mask_haploid, raw, out_words):
    """(rows, masked variant alleles) of one decoded record of n_samples
    samples with the ascending sample indices `mask` removed."""
    bits = _bits(full_words)[:2 * n_samples]
    keep = np.ones(2 * n_samples, bool)
    cnt = 0
    for s, h in zip(mask, mask_haploid):
        s = int(s)
        a, b = int(bits[2 * s]), int(bits[2 * s + 1])
        keep[2 * s] = keep[2 * s + 1] = False
        if h:
            cnt += (a | b) if raw else (a & b)
        else:
            cnt += a + b
    kept = bits[keep]
    out = np.zeros(out_words * 64, np.uint8)
    n = min(kept.size, out.size)
    out[:n] = kept[:n]
    return np.packbits(out, bitorder="little").view(np.uint64), cnt


def kept_haploid(haploid, mask):
    """Haploid flags of the kept samples."""
    keep = np.ones(len(haploid), bool)
    keep[np.asarray(mask, np.int64)] = False
    return np.asarray(haploid)[keep]


def round_f32(x: Fraction) -> np.float32:
    """x rounded to the nearest float32, ties to even (normal range)."""
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    elif Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)
    scaled = a / Fraction(2) ** (e - 23)  # in [2^23, 2^24) for normal values
    n = scaled.numerator // scaled.denominator
    rem = scaled - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(sign * float(Fraction(n) * Fraction(2) ** (e - 23)))


def corrected_af(old_af, old_an: int, new_an: int, masked: int) -> np.float32:
    """The reference's masked allele frequency: round(old_an*old_af -
    masked) (one rounding), then divided by new_an, all in float32."""
    f_an = Fraction(float(round_f32(Fraction(old_an))))
    f_af = Fraction(float(np.float32(old_af)))
    f_cnt = Fraction(float(round_f32(Fraction(masked))))
    ac = round_f32(f_an * f_af - f_cnt)
    return round_f32(Fraction(float(ac)) / Fraction(float(round_f32(Fraction(new_an)))))
