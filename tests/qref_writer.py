"""Writes Qref panel files for tests, as the reference writes them
(src/VCFData.cpp:2006-2104; run-length coder src/utils.h:133-190): the file
layout of qref_spec.md, haplotype vectors run-length coded unless the code
would not be shorter than the raw vector."""
from __future__ import annotations

import struct

import numpy as np

MAXRUN = 32767


def _add_run(run: int, out: bytearray):
    while run > MAXRUN:  # longest run, then an empty run of the other bit
        out += b"\xff\xff\x00"
        run -= MAXRUN
    if run >= 128:
        out += bytes(((run & 0x7F) | 0x80, run >> 7))
    else:
        out.append(run)


def rle_encode(words: np.ndarray):
    """Run-length code of a haplotype vector (uint64 words, LSB first):
    alternating runs starting with zeros.  None when the code reaches the raw
    size (the reference's coder then gives up, src/utils.h:160-186)."""
    words = np.ascontiguousarray(words, np.uint64)
    limit = words.size * 8
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    change = np.flatnonzero(np.diff(bits)) + 1
    bounds = np.concatenate(([0], change, [bits.size]))
    runs = np.diff(bounds).tolist()
    if bits.size and bits[0] == 1:
        runs = [0] + runs  # the code starts with a run of zeros
    out = bytearray()
    for r in runs[:-1]:
        _add_run(r, out)
        if len(out) >= limit:
            return None
    if runs and runs[-1]:
        _add_run(runs[-1], out)
    if len(out) > limit:
        return None
    return bytes(out)


def rle_decode(code: bytes, nwords: int) -> np.ndarray:
    bits = np.zeros(nwords * 64, np.uint8)
    pos, bit, i = 0, 0, 0
    while i < len(code):
        run = code[i] & 0x7F
        if code[i] & 0x80:
            i += 1
            run |= code[i] << 7
        i += 1
        if bit:
            bits[pos:pos + run] = 1
        pos += run
        bit ^= 1
    return np.packbits(bits, bitorder="little").view(np.uint64)


def write_qref(path, rows, Nref, positions, af, alleles, ids, *, chrom=1, haploid=None,
               ma_flags=None, force_raw=False, version=(1, 1)):
    """rows: Mref site rows of 2*Nref haplotype bits (uint64 words, any width
    >= ceil(2*Nref/64)); haploid: None or Nref flags; alleles: Mref (ref, alt)."""
    M = len(positions)
    hw = (2 * Nref + 511) // 512 * 8
    hap = np.zeros((M, hw), np.uint64)
    rows = np.asarray(rows, np.uint64)
    w = min(rows.shape[1], hw) if M else 0
    if M:
        hap[:, :w] = rows[:, :w]
    nh = 0 if haploid is None else int(np.count_nonzero(haploid))
    with open(path, "wb") as f:
        f.write(b"QREF" + bytes((version[0], version[1], 0, chrom)))
        ma = np.zeros(M, bool) if ma_flags is None else np.asarray(ma_flags, bool)
        f.write(struct.pack("<4Q", Nref, nh, M, int(ma.sum())))
        if (nh > 0 and nh != Nref) or chrom == 23:
            flags = np.zeros(Nref, np.uint8) if haploid is None else \
                np.where(np.asarray(haploid) != 0, 0xFF, 0).astype(np.uint8)
            f.write(flags.tobytes())
        f.write(np.asarray(positions, "<i8").tobytes())
        f.write(np.asarray(af, "<f4").tobytes())
        for r, a in alleles:
            f.write(r.encode() + b"\0" + a.encode() + b"\0")
        for s in ids:
            f.write(s.encode() + b"\0")
        mab = np.zeros((M + 511) // 512 * 512, np.uint8)
        mab[:M] = ma
        f.write(np.packbits(mab, bitorder="little").tobytes())
        for m in range(M):
            code = None if force_raw else rle_encode(hap[m])
            if code is None:
                f.write(struct.pack("<Q", 0))
                f.write(hap[m].tobytes())
            else:
                f.write(struct.pack("<Q", len(code)))
                f.write(code)
