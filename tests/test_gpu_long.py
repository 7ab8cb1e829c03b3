"""GPU parity for inputs beyond the staged sweep: reads of more than 511 bases and haplotypes of
more than 2048 bases, which the reference accepts like any other (crates/pairhmm/src/lib.rs:55-67
validates only emptiness and matching lengths).  Their regions run every pair in the strip
kernel; results are held to the same bounds as everywhere else: FP64 within 1e-12 relative of
the oracle, FP32 with FP64 rescue within 1e-4 absolute in log10.
"""
from __future__ import annotations

import numpy as np
import pytest

import fgkl_amd as F
from fgkl_amd.synthetic import PdRegion, Region, Rng, batch_from_regions
from _oracle import batch_reference, pd_reference_all, reference_all

pytestmark = pytest.mark.gpu

FLOAT_TOL = 1e-4
DOUBLE_REL = 1e-12


def assert_close(actual, expected, tol, what=""):
    d = np.abs(np.asarray(actual) - np.asarray(expected))
    assert np.all(np.isfinite(actual)), f"{what}: non-finite results"
    i = int(np.argmax(d))
    assert d[i] <= tol, f"{what}: pair {i}: got {actual[i]!r}, expected {expected[i]!r}"


def assert_rel(actual, expected, what=""):
    actual, expected = np.asarray(actual), np.asarray(expected)
    assert np.all(np.isfinite(actual)), f"{what}: non-finite results"
    bound = DOUBLE_REL * np.maximum(1.0, np.abs(expected))
    d = np.abs(actual - expected)
    i = int(np.argmax(d / bound))
    assert d[i] <= bound[i], f"{what}: pair {i}: got {actual[i]!r}, expected {expected[i]!r}"


def _check_region(pd, reads, haps, expected, what):
    for precision in (F.Precision.DOUBLE, F.Precision.FLOAT):
        eng = F.PdPairHmm if pd else F.PairHmm
        out = eng(F.Config(precision=precision)).compute_log10_likelihoods(reads, haps)
        if precision is F.Precision.DOUBLE:
            assert_rel(out, expected, f"{what} per call double")
        else:
            assert_close(out, expected, FLOAT_TOL, f"{what} per call float")
    data = batch_from_regions([(reads, haps)], pd=pd)
    b = F.Batch(data)
    b.run(F.MODE_FP64, share=True)
    assert_rel(b.results(), expected, f"{what} batch double")
    b.run(F.MODE_FP32_RESCUE, share=True)
    assert_close(b.results(), expected, FLOAT_TOL, f"{what} batch float")
    b.close()


@pytest.mark.parametrize("n", [512, 600, 1023, 2000])
def test_long_reads_match_the_oracle(gpu, n):
    """Strips of 64 lanes x 8 rows: two (the last of one lane, of 12 lanes, of all 64 with a
    single identity row) and four."""
    data = Region.generate(40 + n, 5, n, 3, n + 300)
    reads, haps = data.region_reads_list(), data.region_haps_list()
    _check_region(False, reads, haps, reference_all(reads, haps), f"n={n}")


@pytest.mark.parametrize("n", [600, 1000, 2000])
def test_long_pd_reads_match_the_pd_oracle(gpu, n):
    data = PdRegion.generate(70 + n, 4, n, 3, n + 250)
    reads, haps = data.region_reads_list(), data.region_pd_haps_list()
    assert any(any(h.flags) for h in haps)
    _check_region(True, reads, haps, pd_reference_all(reads, haps), f"pd n={n}")


def test_long_haplotypes_match_the_oracle(gpu):
    """Haplotypes of 2049 to 5000 bases with reads of 150 and 700 bases."""
    rng = Rng(9)
    base = bytes(rng.base() for _ in range(5000))
    haps = [base[:2049], base[:3500], base[300:5000], base[:2048]]
    reads = []  # from the span every haplotype holds, so every likelihood is finite in FP64
    for start, n in ((400, 150), (1700, 150), (1300, 700), (500, 700)):
        b = base[start:start + n]
        reads.append(F.ReadRef(b, [int(20 + rng.below(21)) for _ in range(n)], [40] * n,
                               [42] * n, [10] * n))
    _check_region(False, reads, haps, reference_all(reads, haps), "long haplotypes")
    pd_haps = []
    for h in haps:
        flags = [0] * len(h)
        for j in range(5, len(h) - 10, 97):
            flags[j] = F.SNP | F.ALT_T
            flags[j + 3] = F.DEL_START
            flags[j + 5] = F.DEL_END
        pd_haps.append(F.PdHaplotype(h, flags))
    _check_region(True, reads, pd_haps, pd_reference_all(reads, pd_haps), "pd long haplotypes")


def test_long_regions_do_not_change_the_bits_of_the_other_regions(gpu):
    """lib.rs:971-1017: a batch with long regions between ordinary ones gives the ordinary
    regions the bits they get without them (FP32 with rescue and FP64, eager and captured),
    and the FP64 per-region baseline covers the long regions too."""
    short = [Region.generate(200 + i, 12 + i, 100 + 20 * i, 4 + i % 3, 260 + 30 * i)
             for i in range(6)]
    long = [Region.generate(300, 4, 900, 3, 1100), Region.generate(301, 6, 150, 3, 2600)]
    as_pairs = lambda rs: [(r.region_reads_list(), r.region_haps_list()) for r in rs]
    mixed = as_pairs(short[:3]) + as_pairs(long[:1]) + as_pairs(short[3:]) + as_pairs(long[1:])
    data = batch_from_regions(mixed)
    alone = batch_from_regions(as_pairs(short))
    expected = batch_reference(data)
    b, a = F.Batch(data), F.Batch(alone)
    base, abase = data.out_base(), alone.out_base()
    short_slots = [0, 1, 2, 4, 5, 6]
    for mode in (F.MODE_FP32_RESCUE, F.MODE_FP64):
        a.run(mode, share=True)
        ra = a.results()
        for graph in (False, True, True):
            b.run(mode, share=True, graph=graph)
            rb = b.results()
            if mode == F.MODE_FP64:
                assert_rel(rb, expected, "mixed double")
            else:
                assert_close(rb, expected, FLOAT_TOL, "mixed float")
            for i, g in enumerate(short_slots):
                assert np.array_equal(rb[base[g]:base[g + 1]], ra[abase[i]:abase[i + 1]]), (mode, g)
    b.run(F.MODE_FP64_PER_REGION)
    assert_rel(b.results(), expected, "per-region double")
    b.run_pairs(F.MODE_FP32_RESCUE)
    assert_close(b.results(), expected, FLOAT_TOL, "pairs float")
    b.close()
    a.close()


def test_underflowing_long_reads_are_rescued_in_double(gpu):
    """kernel.rs:897-914 and lib.rs:467-549 on long reads: reads with a mismatch every ten
    bases fall below the FP32 threshold (raw < 1e-28, log10 L < -64) but not below FP64's
    range, so they are recomputed in FP64 (the fallback count is the number of pairs) and the
    result is the FP64 value."""
    rng = Rng(17)
    hap = bytes(rng.base() for _ in range(1400))
    foreign = []
    for start, n in ((50, 600), (200, 1100)):
        b = bytearray(hap[start:start + n])
        for j in range(0, n, 10):
            b[j] = ord("A") if b[j] != ord("A") else ord("C")
        foreign.append(F.ReadRef(bytes(b), [35] * n, [45] * n, [45] * n, [10] * n))
    haps = [hap, hap[:1300]]
    expected = reference_all(foreign, haps)
    assert np.all(np.isfinite(expected)) and np.all(expected < -70)
    hmm = F.PairHmm(F.Config(precision=F.Precision.FLOAT))
    out = hmm.compute_log10_likelihoods(foreign, haps)
    assert_rel(out, expected, "rescued")
    assert hmm.fallback_pairs() == len(expected)
    nofb = F.PairHmm(F.Config(precision=F.Precision.FLOAT, double_fallback=False))
    assert np.all(np.isnan(nofb.compute_log10_likelihoods(foreign, haps)))


def test_a_region_beyond_the_snapshot_slots_shares_nothing_and_matches(gpu):
    """A region of 200 haplotypes, each the same sequence with one base changed at its own
    position, needs 199 snapshot columns to share prefixes, beyond the 128 a region may use:
    it runs unshared (and without suffix sharing), matches the oracle, and leaves the other
    region of the batch its own bits."""
    rng = Rng(23)
    ref = bytes(rng.base() for _ in range(260))
    haps = []
    for i in range(200):
        h = bytearray(ref)
        j = 20 + i
        h[j] = ord("A") if h[j] != ord("A") else ord("G")
        haps.append(bytes(h))
    reads = [F.ReadRef(ref[s:s + 120], [30] * 120, [45] * 120, [45] * 120, [10] * 120)
             for s in (0, 70, 140)]
    other = Region.generate(5, 20, 110, 5, 240)
    regions = [(reads, haps), (other.region_reads_list(), other.region_haps_list())]
    data = batch_from_regions(regions)
    expected = batch_reference(data)
    b = F.Batch(data)
    b.run(F.MODE_FP64, share=True)
    assert_rel(b.results(), expected, "many haplotypes double")
    b.run(F.MODE_FP32_RESCUE, share=True)
    got = b.results()
    assert_close(got, expected, FLOAT_TOL, "many haplotypes float")
    b.close()
    o = F.Batch(batch_from_regions(regions[1:]))
    o.run(F.MODE_FP32_RESCUE, share=True)
    base = data.out_base()
    assert np.array_equal(got[base[1]:], o.results())
    o.close()
