"""Beam-search phasing over compressed PBWTs (eimp_phase_target, host code):
bit-identical to the reference's own beam search (src/Beam.cpp,
src/BeamPath.cpp, src/HaplotypePath.cpp driven as Target::phaseExtPBWT,
src/Target.cpp:291-777) on hard synthetic targets, against golden outputs
written by the reference, and consistent with the genotypes it phases."""
import os

import numpy as np
import pytest

import oracle
import phase_cases as pc
from eagleimp_amd import phase
from eagleimp_amd.engine import EimpError

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_phase.npz"))

needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="reference harness (oracle/_ref) not built")


def run(c, **kw):
    r = c["r"]
    return phase.phase_target(pc.params_for(c), r["pbwt"], r["count0"], r["K"], c["gt"], c["W"],
                              r["splitsites"], r["cMdiffs"], af=c["af"], start_phase=c["start"],
                              **kw)


@needs_ref
@pytest.mark.parametrize("name", list(pc.CASES))
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matches_reference_beam(name, seed):
    c = pc.make_case(name, seed)
    got = run(c)
    want = oracle.ref_phase_target(pc.params_for(c), pc.job_for(c), c["r"]["monos"])
    assert got.same(want)


def test_golden_outputs():
    for i in range(int(GOLD["ncases"])):
        name, seed = str(GOLD[f"c{i}_name"]), int(GOLD[f"c{i}_seed"])
        c = pc.make_case(name, seed)
        assert pc.input_hash(c) == str(GOLD[f"c{i}_inhash"]), f"{name}: inputs drifted"
        got = run(c)
        W = c["W"]
        assert np.array_equal(got.mat, np.unpackbits(GOLD[f"c{i}_mat"], bitorder="little")[:W])
        assert np.array_equal(got.pat, np.unpackbits(GOLD[f"c{i}_pat"], bitorder="little")[:W])
        for a, k in ((got.dosage_mat, "dm"), (got.dosage_pat, "dp"), (got.call_conf, "cc")):
            assert np.array_equal(a.view(np.uint32), GOLD[f"c{i}_{k}"].view(np.uint32)), (name, k)
        assert np.float32(got.total_conf) == GOLD[f"c{i}_tot"]
        assert [got.ncalls, got.ncallsites, got.nmissing, got.nmonhets] == \
            list(GOLD[f"c{i}_counts"])


@pytest.mark.parametrize("name", ["default", "missing_imputed_no_reverse", "roh_hom_splits_prephase"])
def test_consistent_with_genotypes(name):
    c = pc.make_case(name, 0)
    res = run(c)
    codes, r = c["codes"], c["r"]
    imp = bool(c["params"].get("imp_missing", 0))
    het, hr, ha, mi = codes == 1, codes == 0, codes == 2, codes == 3
    assert np.all(res.mat[het] != res.pat[het])
    assert not res.mat[hr].any() and not res.pat[hr].any()
    assert res.mat[ha].all() and res.pat[ha].all()
    assert np.all(res.dosage_mat[ha] == 1.0) and np.all(res.dosage_pat[ha] == 1.0)
    if imp:
        assert np.array_equal(res.dosage_mat[mi], c["af"][mi])
        assert np.array_equal(res.mat[mi], c["af"][mi] > 0.5)
    else:
        assert not res.mat[mi].any() and np.all(res.dosage_mat[mi] == 0.0)
    assert res.nmissing == int(mi.sum())
    # the call sites are the heterozygous split sites; the other hets are mono
    calls = r["splitsites"][r["callLocs"]]
    assert res.ncalls == len(calls) == res.ncallsites
    # mono hets are counted from the first call site on
    after = np.arange(len(codes)) > (calls[0] if len(calls) else len(codes))
    is_call = np.zeros(len(codes), bool)
    is_call[calls] = True
    assert res.nmonhets == int(np.sum(het & after & ~is_call))
    assert np.all((res.call_conf >= 0.5) & (res.call_conf <= 1.0))
    acc = np.float32(0)
    for v in res.call_conf:
        acc = np.float32(acc + v)
    assert np.float32(res.total_conf) == acc
    if c["params"].get("rev_phasing", 1):  # call-site dosages: the phase probabilities
        dm, dp = res.dosage_mat[calls], res.dosage_pat[calls]
        assert np.allclose(dm + dp, 1.0, atol=1e-6)
        assert np.allclose(np.maximum(dm, dp), res.call_conf, atol=1e-6)


def test_drop_in_layout_and_batch():
    """The consumer reads the engine's drop-in rows (4*Kwords uint32, tail
    pairs zero) as it reads the compact CPU rows; a threaded batch equals
    single calls."""
    cases = [pc.make_case(n, 1) for n in ("default", "roh_hom_splits_prephase", "short_history")]
    P = phase.default_params(pre_phasing=1)
    single = []
    jobs = []
    for c in cases:
        r = c["r"]
        K = int(r["K"])
        wide = np.zeros((r["pbwt"].shape[0], 4 * oracle.unit_words(K)), np.uint32)
        wide[:, :r["pbwt"].shape[1]] = r["pbwt"]
        a = phase.phase_target(P, r["pbwt"], r["count0"], K, c["gt"], c["W"], r["splitsites"],
                               r["cMdiffs"], af=c["af"], start_phase=c["start"])
        b = phase.phase_target(P, wide, r["count0"], K, c["gt"], c["W"], r["splitsites"],
                               r["cMdiffs"], af=c["af"], start_phase=c["start"])
        assert a.same(b)
        single.append(a)
        jobs.append(pc.job_for(c, pbwt=wide))
    for got, want in zip(phase.phase_batch(P, jobs, threads=3), single):
        assert got.same(want)


@pytest.mark.parametrize("name", ["default", "roh_hom_splits_prephase", "missing_imputed_no_reverse"])
def test_raw_rows_as_the_fpga_handler_passes_them(name):
    """Rows of data words only (bits past K one), turned into pairs the way
    PBWT(cpbwt, K, Kwords, M) does, phase exactly as the pair rows."""
    c = pc.make_case(name, 2)
    r = c["r"]
    K = int(r["K"])
    raw = oracle.raw_rows(r["pbwt"], K)
    a = run(c)
    b = phase.phase_target(pc.params_for(c), raw, None, K, c["gt"], c["W"], r["splitsites"],
                           r["cMdiffs"], af=c["af"], start_phase=c["start"], raw=True)
    assert a.same(b)
    with pytest.raises(EimpError):  # a raw row shorter than ceil(K/32) words
        phase.phase_target(pc.params_for(c), raw[:, : (K + 31) // 32 - 1], None, K, c["gt"],
                           c["W"], r["splitsites"], r["cMdiffs"], af=c["af"], raw=True)


def test_rejects_bad_inputs():
    c = pc.make_case("default", 0)
    r = c["r"]
    K, W = int(r["K"]), c["W"]
    P = phase.default_params()

    def call(**kw):
        a = dict(pbwt=r["pbwt"], count0=r["count0"], K=K, gt=c["gt"], W=W,
                 splitsites=r["splitsites"], cMdiffs=r["cMdiffs"])
        a.update(kw)
        return phase.phase_target(kw.pop("P", P), a["pbwt"], a["count0"], a["K"], a["gt"], a["W"],
                                  a["splitsites"], a["cMdiffs"])

    with pytest.raises(EimpError):  # row stride below 2*ceil(K/32)
        call(K=K + 64)
    bad = r["count0"].copy()
    bad[3] = K + 1
    with pytest.raises(EimpError):
        call(count0=bad)
    ss = r["splitsites"].copy()
    ss[[1, 2]] = ss[[2, 1]]
    with pytest.raises(EimpError):  # not ascending
        call(splitsites=ss)
    with pytest.raises(EimpError):  # history above 128
        call(P=phase.default_params(hist=2.0))
    with pytest.raises(EimpError):  # imp_missing without allele frequencies
        call(P=phase.default_params(imp_missing=1))
    with pytest.raises(ValueError):  # rows are forward + backward: an even count
        call(pbwt=r["pbwt"][:-1])
    with pytest.raises(AttributeError):
        phase.default_params(beamwidth=3)


def test_too_deep_history_is_rejected():
    """A call whose reference lies more than 192 split sites back reads past
    the reference's history storage: rejected with EIMP_ERR_UNSUPPORTED."""
    c = pc.make_case("too_deep", 0, pc.TOO_DEEP)
    with pytest.raises(EimpError) as e:
        run(c)
    assert e.value.code == -5
