"""GPU parity: the HIP engine against the CPU oracle on the reference's own test cases.

Restates crates/pairhmm/src/lib.rs:583-1097 and pd.rs:112-465 through the engine's C-ABI.
Tolerances: FP64 mode within 1e-12 relative of the oracle (|d| <= 1e-12 * max(1, |ref|));
FP32 mode (with FP64 rescue) within 1e-4 absolute in log10 on every pair; rescued pairs are
FP64 results and are held to the FP64 bound where the reference holds them to 1e-9.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

import fgkl_amd as F
from fgkl_amd.synthetic import PdRegion, Region, Rng, batch_from_regions, workload
from _oracle import batch_reference, pd_reference_all, reference_all

pytestmark = pytest.mark.gpu

FLOAT_TOL = 1e-4
DOUBLE_REL = 1e-12


def assert_close(actual, expected, tol, what=""):
    actual, expected = np.asarray(actual), np.asarray(expected)
    assert actual.shape == expected.shape
    assert np.all(np.isfinite(actual)), f"{what}: non-finite results"
    d = np.abs(actual - expected)
    i = int(np.argmax(d))
    assert d[i] <= tol, f"{what}: pair {i}: got {actual[i]!r}, expected {expected[i]!r} (diff {d[i]})"


def assert_rel(actual, expected, rel=DOUBLE_REL, what=""):
    actual, expected = np.asarray(actual), np.asarray(expected)
    assert actual.shape == expected.shape
    assert np.all(np.isfinite(actual)), f"{what}: non-finite results"
    bound = rel * np.maximum(1.0, np.abs(expected))
    d = np.abs(actual - expected)
    i = int(np.argmax(d / bound))
    assert d[i] <= bound[i], f"{what}: pair {i}: got {actual[i]!r}, expected {expected[i]!r} (diff {d[i]})"


def _shared_batch_agrees(pd, precision, reads, haps, per_call):
    """The per-call entry runs each pair in its own wave; the same region as a one-region batch
    runs the prefix-shared sweeps.  Both must give the same likelihoods within the precision's
    bound, so every per-call case here also exercises prefix sharing."""
    if not reads or not haps:
        return
    from fgkl_amd.synthetic import batch_from_regions

    b = F.Batch(batch_from_regions([(reads, haps)], pd=pd))
    double = precision is F.Precision.DOUBLE
    b.run(F.MODE_FP64 if double else F.MODE_FP32_RESCUE, share=True)
    shared = b.results()
    b.close()
    ok = np.isfinite(per_call)
    assert np.array_equal(ok, np.isfinite(shared))
    d = np.abs(shared[ok] - per_call[ok])
    bound = (DOUBLE_REL * np.maximum(1.0, np.abs(per_call[ok]))) if double else FLOAT_TOL
    assert np.all(d <= bound), float(d.max())


def compute(precision, reads, haps, fallback=True):
    hmm = F.PairHmm(F.Config(precision=precision, double_fallback=fallback))
    out = hmm.compute_log10_likelihoods(reads, haps)
    if fallback:
        _shared_batch_agrees(False, precision, reads, haps, out)
    return out, hmm.fallback_pairs()


def pd_compute(precision, reads, haps):
    hmm = F.PdPairHmm(F.Config(precision=precision))
    out = hmm.compute_log10_likelihoods(reads, haps)
    _shared_batch_agrees(True, precision, reads, haps, out)
    return out, hmm.fallback_pairs()


def read(bases: bytes, q=30, i=45, d=45, g=10):
    n = len(bases)
    return F.ReadRef(bases, [q] * n, [i] * n, [d] * n, [g] * n)


def foreign_reads(seed, n, length):
    """lib.rs:879-893"""
    rng = Rng(seed)
    out = []
    for _ in range(n):
        bases = bytes(rng.base() for _ in range(length))
        quals = [20 + rng.below(21) for _ in range(length)]
        out.append(F.ReadRef(bases, quals, [45] * length, [45] * length, [10] * length))
    return out


def region():
    r = Region.generate(11, 70, 120, 24, 300)
    return r.region_reads_list(), r.region_haps_list()


# ---- the engine's own lane layout --------------------------------------------------------------
def test_dpp_wave_shift_moves_each_lane_to_the_next(gpu):
    import ctypes as C
    out = (C.c_int32 * 64)()
    assert F._lib.lib().fgkl_selftest_dpp(out) == 0
    assert out[0] == -1
    assert [out[i] for i in range(1, 64)] == [1000 + i - 1 for i in range(1, 64)]


# ---- lib.rs:620-645 ----------------------------------------------------------------------------
def test_double_precision_matches_reference(gpu):
    reads, haps = region()
    out, _ = compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, reference_all(reads, haps), what="double")


def test_single_precision_matches_reference_within_float_tolerance(gpu):
    reads, haps = region()
    out, _ = compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, reference_all(reads, haps), FLOAT_TOL, "float")


def test_haplotype_order_does_not_change_results(gpu):  # lib.rs:647-663
    reads, haps = region()
    fwd, _ = compute(F.Precision.DOUBLE, reads, haps)
    rev, _ = compute(F.Precision.DOUBLE, reads, haps[::-1])
    n = len(haps)
    a = fwd.reshape(-1, n)
    b = rev.reshape(-1, n)[:, ::-1]
    assert np.max(np.abs(a - b)) < 1e-11


def test_prefix_sharing_handles_nested_and_repeated_prefixes(gpu):  # lib.rs:665-681
    haps = [b"ABBC", b"AAAC", b"AC", b"AAAB", b"ABBB", b"AAAC", b"ABBCC"]
    reads = Region.generate(3, 20, 6, 1, 8).region_reads_list()
    expected = reference_all(reads, haps)
    for p in (F.Precision.DOUBLE, F.Precision.FLOAT):
        out, _ = compute(p, reads, haps)
        if p is F.Precision.DOUBLE:
            assert_rel(out, expected, what="nested double")
        else:
            assert_close(out, expected, FLOAT_TOL, "nested float")


def test_haplotypes_of_different_lengths_share_prefixes_correctly(gpu):  # lib.rs:683-691
    haps = [b"ACGTACGTAC", b"ACGTACGT", b"ACGTACGTACGTACGT", b"ACGT", b"A"]
    reads = Region.generate(9, 40, 12, 1, 16).region_reads_list()
    out, _ = compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, reference_all(reads, haps), what="lengths")


def test_non_acgtn_bases_match_only_themselves(gpu):  # lib.rs:693-716
    reads = [read(b"ACRTN")]
    haps = [b"ACRTG", b"ACGTG", b"ACNTG", b"RRRRR"]
    expected = reference_all(reads, haps)
    out, _ = compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, expected, what="non-acgtn")
    assert expected[0] > expected[1] and expected[2] >= expected[0]


def test_float_underflow_falls_back_to_double(gpu):  # lib.rs:718-739
    reads = [read(b"A" * 80, q=40)]
    haps = [b"C" * 80]
    expected = reference_all(reads, haps)
    assert np.isfinite(expected[0]) and expected[0] < -50
    out, fb = compute(F.Precision.FLOAT, reads, haps)
    assert_rel(out, expected, what="underflow")
    assert fb == 1


def test_without_fallback_underflowing_pairs_are_nan(gpu):  # lib.rs:221-226
    reads = [read(b"A" * 80, q=40)]
    out, fb = compute(F.Precision.FLOAT, reads, [b"C" * 80], fallback=False)
    assert np.isnan(out[0]) and fb == 1


def test_reads_unrelated_to_every_haplotype_are_recomputed_in_double(gpu):  # lib.rs:895-914
    _, haps = region()
    reads = foreign_reads(5, 40, 120)
    expected = reference_all(reads, haps)
    assert np.all(expected < -64.0)
    out, fb = compute(F.Precision.FLOAT, reads, haps)
    assert_rel(out, expected, what="foreign")
    assert fb == len(expected)


def test_matching_and_mismatching_reads_of_varied_lengths_share_a_batch(gpu):  # lib.rs:916-938
    r = Region.generate(11, 70, 120, 24, 300)
    haps = r.region_haps_list()
    owned = r.region_reads_list()[:20] + foreign_reads(6, 20, 120)
    reads = []
    for i, rd in enumerate(owned):
        n = 30 + (i * 7) % 90
        reads.append(F.ReadRef(rd.bases[:n], rd.quals[:n], rd.ins_gop[:n], rd.del_gop[:n], rd.gcp[:n]))
    out, fb = compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, reference_all(reads, haps), FLOAT_TOL, "mixed")
    assert fb > 0


def test_a_read_gets_the_same_bits_whatever_shares_its_batch(gpu):  # lib.rs:940-955
    r = Region.generate(11, 70, 120, 24, 300)
    haps = r.region_haps_list()
    alone = r.region_reads_list()[:3]
    batched = alone + foreign_reads(7, 29, 120)
    for p in (F.Precision.FLOAT, F.Precision.DOUBLE):
        solo, _ = compute(p, alone, haps)
        with_others, _ = compute(p, batched, haps)
        assert np.array_equal(solo, with_others[: len(solo)]), p


def test_a_read_gets_the_same_bits_alone_and_in_any_batch_size(gpu):  # lib.rs:957-985
    r = Region.generate(5, 40, 60, 4, 100)
    reads, haps = r.region_reads_list(), r.region_haps_list()
    for p in (F.Precision.FLOAT, F.Precision.DOUBLE):
        solo = [compute(p, [rd], haps)[0] for rd in reads]
        for n in (8, 13, 16, 17, 31, 33, 40):
            allr, _ = compute(p, reads[:n], haps)
            for i in range(n):
                assert np.array_equal(allr[i * len(haps):(i + 1) * len(haps)], solo[i]), (p, n, i)


@pytest.mark.parametrize("share", [True, False])
def test_a_pd_read_gets_the_same_bits_whatever_shares_its_batch(gpu, share):  # pd.rs:400-426
    """Three PD reads alone, and with 29 reads of random bases (which underflow in FP32 and are
    rescued) in one region, as a batch with and without prefix sharing: the three reads' results
    are the same bits in both precisions.  The random reads change the pairing of reads into
    waves and the rows per lane the three run at."""
    region = PdRegion.generate(31, 70, 120, 24, 300)
    haps = region.region_pd_haps_list()
    alone = region.region_reads_list()[:3]
    rng = Rng(77)
    others = [F.ReadRef(bytes(rng.base() for _ in range(120)), [30] * 120, [45] * 120,
                        [45] * 120, [10] * 120) for _ in range(29)]
    for mode in (F.MODE_FP32_RESCUE, F.MODE_FP64):
        outs = []
        for reads in (alone, alone + others):
            b = F.Batch(batch_from_regions([(reads, haps)], pd=True))
            st = b.run(mode, share=share)
            outs.append(b.results())
            b.close()
        n = len(alone) * len(haps)
        assert np.array_equal(outs[0], outs[1][:n]), (mode, share)
        if mode == F.MODE_FP32_RESCUE:
            assert st.rescued_pairs > 0, "the random reads must underflow"


def test_a_read_matching_only_the_previous_haplotype_is_recomputed_in_double(gpu):  # lib.rs:1019-1056
    rng = Rng(8)
    reference = bytes(rng.base() for _ in range(300))
    tail = bytes(rng.base() for _ in range(150))
    h1 = reference
    h2 = reference[:150] + tail
    h3 = bytearray(reference)
    h3[200] = ord("C") if h3[200] == ord("A") else ord("A")
    haps = [h1, h2, bytes(h3)]
    reads = []
    for k in range(12):
        src = reference if k % 2 == 0 else h2
        start = 100 + (k * 13) % 80
        reads.append(read(src[start:start + 100], q=35))
    expected = reference_all(reads, haps)
    assert np.any(expected < -64.0) and np.any(expected > -10.0)
    out, fb = compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, expected, FLOAT_TOL, "previous haplotype")
    assert 0 < fb <= 24


def test_more_reads_than_lanes_and_single_read_both_work(gpu):  # lib.rs:1058-1068
    for seed, n in ((21, 1), (22, 37)):
        r = Region.generate(seed, n, 50, 3, 60)
        reads, haps = r.region_reads_list(), r.region_haps_list()
        out, _ = compute(F.Precision.FLOAT, reads, haps)
        assert_close(out, reference_all(reads, haps), FLOAT_TOL, f"{n} reads")


def test_long_reads_far_below_single_precision(gpu):  # lib.rs:827-877 shapes
    rng = Rng(99)
    reference = bytes(rng.base() for _ in range(700))
    haps = [reference]
    for pos in (20, 40, 60):
        h = bytearray(reference)
        h[pos] = ord("C") if h[pos] == ord("A") else ord("A")
        haps.append(bytes(h))
    reads = [read(bytes(rng.base() for _ in range(n)), q=40) for n in (290, 360, 500)]
    expected = reference_all(reads, haps)
    assert np.any(expected < -400.0)
    for p in (F.Precision.DOUBLE, F.Precision.FLOAT):
        out, _ = compute(p, reads, haps)
        assert_rel(out, expected, what=str(p))


def test_every_lane_layout_matches_reference(gpu):
    # Read lengths around every rows-per-lane boundary (R = ceil((n+1)/64)) and identity-row
    # count, including n a multiple of R (identity rows fill lane 0) and the 511-base maximum.
    rng = Rng(77)
    lengths = sorted({1, 2, 3, 5, 8, 31, 32, 62, 63, 64, 65, 100, 126, 127, 128, 129, 150, 189,
                      190, 191, 192, 250, 252, 254, 255, 256, 300, 320, 383, 384, 448, 500, 510,
                      511})
    ref = bytes(rng.base() for _ in range(700))
    haps = [ref[:620], ref[5:640], ref[:600] + b"ACGTACGT"]
    reads = []
    for n in lengths:
        start = rng.below(100)
        bases = bytearray(ref[start:start + n])
        for i in range(0, n, 17):
            bases[i] = rng.base()
        reads.append(F.ReadRef(bytes(bases), [20 + rng.below(21) for _ in range(n)],
                               [30 + rng.below(16) for _ in range(n)],
                               [30 + rng.below(16) for _ in range(n)], [10] * n))
    expected = reference_all(reads, haps)
    out, _ = compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, expected, what="lane layouts double")
    out, _ = compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, expected, FLOAT_TOL, "lane layouts float")


def test_random_small_haplotype_sets_match_reference(gpu):  # lib.rs:761-798
    from fgkl_amd.synthetic import MASK64  # noqa: F401
    rng = Rng(31)
    letter = lambda: ord("A") if rng.chance(0.5) else ord("C")
    for _ in range(60):
        haps = [bytes(letter() for _ in range(6 + rng.below(30)))]
        for _ in range(rng.below(12)):
            src = haps[rng.below(len(haps))]
            haps.append(_edited(rng, b"AC", src))
        reads = []
        for _ in range(1 + rng.below(20)):
            n = 1 + rng.below(25)
            bases = bytes(letter() for _ in range(n))
            quals = [10 + rng.below(31) for _ in range(n)]
            ins = [20 + rng.below(26) for _ in range(n)]
            dele = [20 + rng.below(26) for _ in range(n)]
            gcp = [1 + rng.below(10) for _ in range(n)]
            reads.append(F.ReadRef(bases, quals, ins, dele, gcp))
        expected = reference_all(reads, haps)
        out, _ = compute(F.Precision.DOUBLE, reads, haps)
        assert_rel(out, expected, what="random small")


def _edited(rng, alphabet, base):  # synthetic.rs:84-98
    hap = bytearray(base)
    for _ in range(1 + rng.below(3)):
        pos = rng.below(len(hap))
        kind = rng.below(3)
        if kind == 0:
            hap[pos] = alphabet[rng.below(len(alphabet))]
        elif kind == 1:
            hap.insert(pos, alphabet[rng.below(len(alphabet))])
        elif len(hap) > 1:
            del hap[pos]
    return bytes(hap)


# ---- pd.rs:161-436 ----------------------------------------------------------------------------
def pd_region():
    r = PdRegion.generate(31, 70, 120, 24, 300)
    return r.region_reads_list(), r.region_pd_haps_list()


def test_pd_double_precision_matches_the_pd_reference(gpu):
    reads, haps = pd_region()
    out, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, pd_reference_all(reads, haps), what="pd double")


def test_pd_single_precision_matches_within_float_tolerance(gpu):
    reads, haps = pd_region()
    out, _ = pd_compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, pd_reference_all(reads, haps), FLOAT_TOL, "pd float")


def test_pd_flagged_alternates_and_deletions_change_the_result(gpu):
    reads, haps = pd_region()
    unflagged = [F.PdHaplotype(h.bases, [0] * len(h.bases)) for h in haps]
    w, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    wo, _ = pd_compute(F.Precision.DOUBLE, reads, unflagged)
    assert np.sum(np.abs(w - wo) > 1.0) >= 20
    assert np.all(w >= wo)


def test_pd_haplotypes_without_flags_agree_with_the_plain_kernel(gpu):
    r = Region.generate(5, 40, 100, 8, 200)
    reads, plain = r.region_reads_list(), r.region_haps_list()
    pd = [F.PdHaplotype(h, [0] * len(h)) for h in plain]
    for p, tol in ((F.Precision.DOUBLE, 1e-12), (F.Precision.FLOAT, FLOAT_TOL)):
        a, _ = pd_compute(p, reads, pd)
        b, _ = compute(p, reads, plain)
        assert_close(a, b, tol, str(p))


def test_pd_flags_on_the_last_columns_carry_their_state_into_the_next_row(gpu):
    hap = b"ACGTTGCAAGGCTTAGGCTTACG"
    n = len(hap)
    ends_del = [0] * n; ends_del[n - 3] = F.DEL_START; ends_del[n - 1] = F.DEL_END
    starts_del = [0] * n; starts_del[n - 1] = F.DEL_START
    single_del = [0] * n; single_del[n - 1] = F.DEL_START | F.DEL_END
    first_del = [0] * n; first_del[0] = F.DEL_START | F.DEL_END; first_del[1] = F.SNP | F.ALT_G
    snp = [0] * n; snp[n - 1] = F.SNP | F.ALT_C | F.ALT_T
    haps = [F.PdHaplotype(hap, f) for f in (ends_del, starts_del, single_del, first_del, snp)]
    reads = [read(b"GTTGCAAGGCTTAG"), read(b"GGCTTAGGCTTACG"), read(b"GGCTTAGGCTTC"),
             read(b"GGCTTAGGCTTA"), read(b"CGTTGCAAGG"), read(hap)]
    expected = pd_reference_all(reads, haps)
    out, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, expected, what="pd last columns double")
    out, _ = pd_compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, expected, FLOAT_TOL, "pd last columns float")


def _pd_class_region(rng, others: bytes):
    """A region whose reads and haplotypes mix upper- and lowercase bases, 'N' and the bytes in
    `others`, against SNP alternates (some lowercase-accepting) and deletion runs."""
    n = 160
    hap = bytearray(b"ACGT"[rng.below(4)] for _ in range(n))
    for j in range(0, n, 11):
        hap[j] = hap[j] | 0x20  # lowercase
    for j in (7, 50, 93):
        hap[j] = ord("N")
    for j, b in zip(range(20, n, 23), others):
        hap[j] = b
    flags = [0] * n
    for j in range(5, n, 13):
        flags[j] = F.SNP | (F.ALT_A, F.ALT_C, F.ALT_G, F.ALT_T)[j % 4] | (F.ALT_C if j % 3 else 0)
    for j in (30, 70, 120):
        flags[j] |= F.DEL_START
        flags[j + 2] |= F.DEL_END
    haps = [F.PdHaplotype(bytes(hap), flags),
            F.PdHaplotype(bytes(hap).upper(), [f & ~F.SNP for f in flags])]
    reads = []
    for k in range(10):
        lo = rng.below(n - 60)
        rb = bytearray(hap[lo:lo + 40 + rng.below(20)])
        for i in range(len(rb)):
            u = rng.below(20)
            if u == 0:
                rb[i] = ord("N")
            elif u == 1:
                rb[i] = rb[i] ^ 0x20 if chr(rb[i]).isalpha() else rb[i]
            elif u == 2 and others:
                rb[i] = others[rng.below(len(others))]
        reads.append(read(bytes(rb)))
    return reads, haps


@pytest.mark.parametrize("others", [b"", b"acgtR*", b"acgtnRYKM\x07"])
def test_pd_read_base_classes_match_the_reference_for_any_byte(gpu, others):
    """A staged PD column carries the classes of read bases it accepts (A, C, G, T and up to
    eight other bytes of the region's reads; MatchSet::accepts, pd_kernel.rs:196-214): lowercase
    read bases against SNP alternates and the haplotype's own lowercase bytes, 'N' on either side,
    other bytes that match only themselves and 'N'.  Nine other bytes send the region to the
    long kernel, which compares bytes.  Both precisions against the oracle, per call and batched
    with prefix sharing."""
    reads, haps = _pd_class_region(Rng(77 + len(others)), others)
    expected = pd_reference_all(reads, haps)
    out, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, expected, what=f"pd classes double {others!r}")
    out, _ = pd_compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, expected, FLOAT_TOL, f"pd classes float {others!r}")


def test_pd_carried_deletion_state_keeps_later_rows_from_skipping(gpu):
    """The oracle's hand-worked carry case (haplotype GCTT, DEL_END on column 3, DEL_START on
    column 4, read GCT): rows 2-3 start inside a deletion, so the skip of column 3 does not count
    and the likelihood is that of the one perfect alignment; without the trailing DEL_START it
    counts twice.  Reads of 3..9 rows and lane layouts with identity rows at every offset."""
    import math

    haps = [F.PdHaplotype(b"GCTT", [0, 0, F.DEL_END, F.DEL_START]),
            F.PdHaplotype(b"GCTT", [0, 0, F.DEL_END, 0]),
            F.PdHaplotype(b"GCTTAGCTT", [0, 0, F.DEL_END, F.DEL_START, 0, 0, F.DEL_END, 0,
                                         F.DEL_START])]
    reads = [read(b"GCT"), read(b"GCTA"), read(b"GCTAGCT"), read(b"CTTAGCTTA")]
    expected = pd_reference_all(reads, haps)
    assert abs(expected[1] - expected[0] - math.log10(2.0)) < 1e-4
    out, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, expected, what="pd carry double")
    out, _ = pd_compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, expected, FLOAT_TOL, "pd carry float")


@pytest.mark.parametrize("n", [70, 130, 200, 333, 480])
def test_pd_rows_sharing_maxima_match_the_reference(gpu, n):
    """A PD haplotype whose deletion state carried out of a row is NORMAL (read row 1 reads the
    same state bits as the rows below it, pd_reference.rs:46) takes the sweep in which each row
    reuses, as its left maxima, the diagonal maxima the row below formed from the same values;
    a haplotype ending inside or just after a deletion takes the full sweep.  Reads of 2 to 8 rows
    per lane over both kinds of haplotype, batched (prefix sharing, two reads to a wave) and per
    call, in both precisions, against the PD oracle."""
    rng = Rng(300 + n)
    m = n + 120
    base = bytes(rng.base() for _ in range(m))

    def flags(kind):
        f = [0] * m
        for j in range(7, m - 14, 23):  # alternates and deletion runs inside the haplotype
            f[j] = F.SNP | F.ALT_G
            f[j + 4] = F.DEL_START
            f[j + 5 + rng.below(5)] |= F.DEL_END
        if kind == 1:
            f[m - 3] = F.DEL_START
            f[m - 1] = F.DEL_END    # AFTER_DEL carried into the next row
        elif kind == 2:
            f[m - 2] = F.DEL_START  # INSIDE_DEL carried into the next row
        return f

    haps = [F.PdHaplotype(base, flags(k)) for k in (0, 1, 2, 0)]
    reads = []
    for length in (n, n - 3, max(1, n // 2), max(1, n // 3)):
        start = rng.below(m - length + 1)
        b = bytearray(base[start:start + length])
        for _ in range(max(1, length // 40)):
            b[rng.below(length)] = b"ACGT"[rng.below(4)]
        reads.append(F.ReadRef(bytes(b), [20 + rng.below(21) for _ in range(length)],
                               [40] * length, [40] * length, [10] * length))
    expected = pd_reference_all(reads, haps)
    out, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, expected, what=f"pd shared maxima per call double n={n}")
    out, _ = pd_compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, expected, FLOAT_TOL, f"pd shared maxima per call float n={n}")
    b = F.Batch(batch_from_regions([(reads, haps)], pd=True))
    b.run(F.MODE_FP64, share=True)
    assert_rel(b.results(), expected, what=f"pd shared maxima batch double n={n}")
    b.run(F.MODE_FP32_RESCUE, share=True)
    assert_close(b.results(), expected, FLOAT_TOL, f"pd shared maxima batch float n={n}")
    b.close()


def test_pd_adjacent_and_single_base_deletions_match_the_reference(gpu):
    hap = b"ACGTTGCAAGGCTTAGGCTTACGATCGATTACA"
    n = len(hap)
    b2b = [0] * n
    b2b[5] = F.DEL_START; b2b[7] = F.DEL_END; b2b[8] = F.DEL_START | F.DEL_END
    b2b[9] = F.DEL_START; b2b[12] = F.DEL_END
    singles = [0] * n
    for j in range(3, n, 4):
        singles[j] = F.DEL_START | F.DEL_END
    snp_in_del = [0] * n
    snp_in_del[10] = F.DEL_START; snp_in_del[11] = F.SNP | F.ALT_T
    snp_in_del[13] = F.DEL_END | F.SNP | F.ALT_G
    haps = [F.PdHaplotype(hap, f) for f in (b2b, singles, snp_in_del)]
    reads = PdRegion.generate(9, 30, 25, 1, 33).region_reads_list()
    out, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, pd_reference_all(reads, haps), what="pd adjacent")


def test_pd_prefix_sharing_handles_nested_prefixes_lengths_and_flag_splits(gpu):
    base = b"ACGTTGCAAGGCTTAGGCTTACGATCGATTACAGGT"
    n = len(base)
    plain = [0] * n
    snp_early = plain.copy(); snp_early[5] = F.SNP | F.ALT_T
    snp_late = plain.copy(); snp_late[30] = F.SNP | F.ALT_C
    del_mid = plain.copy(); del_mid[12] = F.DEL_START; del_mid[15] = F.DEL_END
    del_mid_snp = del_mid.copy(); del_mid_snp[25] = F.SNP | F.ALT_G
    del_last = plain.copy(); del_last[n - 2] = F.DEL_START; del_last[n - 1] = F.DEL_END
    short = base[:20]
    short_del = [0] * 20; short_del[12] = F.DEL_START; short_del[15] = F.DEL_END
    longer = base + b"ACGTAC"
    longer_flags = [0] * len(longer); longer_flags[12] = F.DEL_START; longer_flags[15] = F.DEL_END
    haps = [F.PdHaplotype(base, del_last), F.PdHaplotype(base, snp_late),
            F.PdHaplotype(short, short_del), F.PdHaplotype(base, plain),
            F.PdHaplotype(base, del_mid_snp), F.PdHaplotype(longer, longer_flags),
            F.PdHaplotype(base, del_mid), F.PdHaplotype(base, snp_early),
            F.PdHaplotype(base, plain)]
    reads = PdRegion.generate(13, 40, 30, 1, 36).region_reads_list()
    expected = pd_reference_all(reads, haps)
    out, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    assert_rel(out, expected, what="pd sharing double")
    out, _ = pd_compute(F.Precision.FLOAT, reads, haps)
    assert_close(out, expected, FLOAT_TOL, "pd sharing float")


def test_pd_haplotype_order_does_not_change_results(gpu):
    reads, haps = pd_region()
    fwd, _ = pd_compute(F.Precision.DOUBLE, reads, haps)
    rev, _ = pd_compute(F.Precision.DOUBLE, reads, haps[::-1])
    n = len(haps)
    assert np.max(np.abs(fwd.reshape(-1, n) - rev.reshape(-1, n)[:, ::-1])) < 1e-11


def test_pd_float_underflow_falls_back_to_double(gpu):
    flags = [0] * 80
    flags[10] = F.DEL_START
    flags[20] = F.DEL_END
    reads = [read(b"A" * 80, q=40)]
    haps = [F.PdHaplotype(b"C" * 80, flags)]
    expected = pd_reference_all(reads, haps)
    assert expected[0] < -50
    out, fb = pd_compute(F.Precision.FLOAT, reads, haps)
    assert_rel(out, expected, what="pd underflow")
    assert fb == 1


def test_pd_more_reads_than_lanes_and_single_read_both_work(gpu):
    for seed, n in ((41, 1), (42, 37), (43, 65)):
        r = PdRegion.generate(seed, n, 50, 3, 60)
        reads, haps = r.region_reads_list(), r.region_pd_haps_list()
        out, _ = pd_compute(F.Precision.FLOAT, reads, haps)
        assert_close(out, pd_reference_all(reads, haps), FLOAT_TOL, f"pd {n}")


# ---- the JNI-shaped entry (FgklPairHmmTest.java:130-211) ---------------------------------------
def _native_arrays(data):
    ro = data.read_offsets.astype(np.int32)
    ho = data.hap_offsets.astype(np.int32)
    return ro, ho


def test_native_entry_matches_reference_in_both_precisions(gpu):
    data = Region.generate(1, 60, 120, 20, 300)
    reads, haps = data.region_reads_list(), data.region_haps_list()
    expected = reference_all(reads, haps)
    ro, ho = _native_arrays(data)
    for dp, tol in ((True, None), (False, FLOAT_TOL)):
        out = np.zeros(len(expected))
        F.compute_native(data.bases, data.quals, data.ins, data.dels, data.gcp, ro,
                         data.hap_bases, ho, dp, out)
        if tol is None:
            assert_rel(out, expected, what="native double")
        else:
            assert_close(out, expected, tol, "native float")


def test_a_native_call_gives_each_read_its_lone_bits_whatever_it_is_paired_with(gpu):
    """lib.rs:971-1017 for the per-call path: a call runs each (read, haplotype) pair in a wave
    of its own or, when two reads of the region fit one wave at the launch's rows per lane, both
    reads on the haplotype in one wave.  Short reads beside a long one pair up here; every read's
    results must be the bits of calling it alone, in both precisions."""
    rng = np.random.default_rng(2024)
    lens = [300, 60, 70, 55, 250, 80, 64, 127, 30, 41, 9, 1]
    def seq(alphabet, n, p=None):
        return rng.choice(np.frombuffer(alphabet, np.uint8), size=n, p=p).astype(np.uint8).tobytes()

    haps = [seq(b"ACGT", int(rng.integers(320, 380))) for _ in range(5)]
    reads = []
    for n in lens:
        reads.append((seq(b"ACGTN", n, [0.24, 0.24, 0.24, 0.24, 0.04]),
                      rng.integers(2, 42, size=n).astype(np.uint8),
                      rng.integers(10, 46, size=n).astype(np.uint8),
                      rng.integers(10, 46, size=n).astype(np.uint8),
                      np.full(n, 10, np.uint8)))

    def call(rs, dp):
        ro = np.concatenate([[0], np.cumsum([len(r[0]) for r in rs])]).astype(np.int32)
        ho = np.concatenate([[0], np.cumsum([len(h) for h in haps])]).astype(np.int32)
        cat = lambda i: np.concatenate([np.frombuffer(r[0], np.uint8) if i == 0 else r[i]
                                        for r in rs]).astype(np.uint8)
        out = np.zeros(len(rs) * len(haps))
        F.compute_native(cat(0), cat(1), cat(2), cat(3), cat(4), ro,
                         np.frombuffer(b"".join(haps), np.uint8), ho, dp, out)
        return out.reshape(len(rs), len(haps))

    for dp in (False, True):
        together = call(reads, dp)
        for i, r in enumerate(reads):
            alone = call([r], dp)[0]
            assert np.array_equal(together[i], alone), (dp, lens[i])
        if not dp:
            refs = [F.ReadRef(r[0], list(r[1]), list(r[2]), list(r[3]), list(r[4])) for r in reads]
            assert_close(together.ravel(), reference_all(refs, haps), FLOAT_TOL, "paired per-call")


@pytest.mark.parametrize("aggregate", [0, 8])
def test_concurrent_native_calls_give_identical_results(gpu, aggregate):  # FgklPairHmmTest.java:193-211
    """Eight concurrent calls of one region give bitwise the results of a lone call, with each
    call on its own and with the calls gathered into shared batches."""
    data = Region.generate(8, 200, 150, 30, 400)
    ro, ho = _native_arrays(data)
    n = data.n_pairs()

    def run():
        out = np.zeros(n)
        F.compute_native(data.bases, data.quals, data.ins, data.dels, data.gcp, ro,
                         data.hap_bases, ho, False, out)
        return out

    expected = run()
    results = [None] * 8
    errors = []
    barrier = threading.Barrier(8)

    def worker(i):
        try:
            barrier.wait()
            results[i] = run()
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    b0, r0 = F.aggregate_stats()
    F.aggregate(aggregate, 20000)
    try:
        threads = [threading.Thread(target=worker, args=(i,)) for i in range(8)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    finally:
        F.aggregate(0)
    assert not errors, errors
    for r in results:
        assert np.array_equal(r, expected)
    b1, r1 = F.aggregate_stats()
    assert r1 - r0 == (8 if aggregate else 0)


def _region_native(data, g):
    """Region g of a batch as the JNI-shaped entry takes it."""
    r0, r1 = int(data.region_reads[g]), int(data.region_reads[g + 1])
    h0, h1 = int(data.region_haps[g]), int(data.region_haps[g + 1])
    lo, hi = int(data.read_offsets[r0]), int(data.read_offsets[r1])
    hlo, hhi = int(data.hap_offsets[h0]), int(data.hap_offsets[h1])
    ro = (data.read_offsets[r0:r1 + 1] - lo).astype(np.int32)
    ho = (data.hap_offsets[h0:h1 + 1] - hlo).astype(np.int32)
    arrs = [a[lo:hi] for a in (data.bases, data.quals, data.ins, data.dels, data.gcp)]
    return arrs, ro, data.hap_bases[hlo:hhi], data.hap_flags[hlo:hhi], ho, (r1 - r0) * (h1 - h0)


@pytest.mark.parametrize("pd", [False, True])
def test_aggregated_regions_get_their_lone_call_bits_in_large_gathered_batches(gpu, pd):
    """lib.rs:971-1017, pd.rs:400-426 for the aggregator.  The largest regions of a draw of the
    batched uneven workload (PD: with flags), 1.6e10 cells together and reads of mixed lengths,
    run as one gathered batch the way the aggregator runs one (every pair in a wave of its own,
    each rows-per-lane bucket at its own rows), and arrive together from one thread each with
    aggregation on; each region gets exactly the bits of a lone call (one launch at the region's
    largest rows per lane), in both precisions."""
    draw = workload(3 if pd else 1, seed=23, n_regions=600)
    rr, hh = draw.read_offsets[draw.region_reads], draw.hap_offsets[draw.region_haps]
    cells = (rr[1:] - rr[:-1]) * (hh[1:] - hh[:-1])
    picked, total = [], 0
    for g in np.argsort(-cells, kind="stable"):
        picked.append(int(g))
        total += int(cells[g])
        if total > 1.6e10:
            break
    regions = [(draw.region_reads_list(g),
                draw.region_pd_haps_list(g) if pd else draw.region_haps_list(g)) for g in picked]
    data = batch_from_regions(regions, pd=pd)
    assert data.cells() > 1.5e10 and data.n_pairs() <= 1 << 20
    native = [_region_native(data, g) for g in range(data.n_regions)]

    def call(g, dp):
        arrs, ro, hb, hf, ho, n = native[g]
        out = np.zeros(n)
        if pd:
            F.pd_compute_native(*arrs, ro, hb, hf, ho, dp, out)
        else:
            F.compute_native(*arrs, ro, hb, ho, dp, out)
        return out

    base = data.out_base()
    b = F.Batch(data)
    for dp in (False, True):
        alone = [call(g, dp) for g in range(len(regions))]
        b.run_pairs(F.MODE_FP64 if dp else F.MODE_FP32_RESCUE)
        whole = b.results()
        for g in range(len(regions)):
            assert np.array_equal(whole[base[g]:base[g + 1]], alone[g]), f"batch: region {g} double={dp}"
            assert np.all(np.isfinite(alone[g]))

        got, errors = {}, []
        barrier = threading.Barrier(len(regions))

        def worker(g):
            try:
                barrier.wait()
                got[g] = call(g, dp)
            except Exception as e:  # pragma: no cover - reported below
                errors.append(e)

        b0, r0 = F.aggregate_stats()
        F.aggregate(len(regions), 200000)
        try:
            threads = [threading.Thread(target=worker, args=(g,)) for g in range(len(regions))]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        finally:
            F.aggregate(0)
        assert not errors, errors
        b1, r1 = F.aggregate_stats()
        assert r1 - r0 == len(regions)
        assert b1 - b0 < len(regions), "no two calls shared a batch"
        for g in range(len(regions)):
            assert np.array_equal(got[g], alone[g]), f"aggregated: region {g} double={dp}"
    b.close()


def test_per_call_batches_are_pooled_across_threads_that_exit(gpu):
    """Per-call batches are checked out of a process-wide pool: four threads that call and exit
    leave their batches idle in the pool (nothing is freed at thread exit), later calls reuse
    them with the same results, and release_idle frees them."""
    data = Region.generate(5, 40, 120, 6, 260)
    ro, ho = _native_arrays(data)

    def run():
        out = np.zeros(data.n_pairs())
        F.compute_native(data.bases, data.quals, data.ins, data.dels, data.gcp, ro,
                         data.hap_bases, ho, False, out)
        return out

    expected = run()
    F.release_idle()
    barrier = threading.Barrier(4)
    results = []

    def worker():
        barrier.wait()
        results.append(run())

    threads = [threading.Thread(target=worker) for _ in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    idle, out, reused0 = F.pool_stats(F.POOL_PAIRHMM)
    assert out == 0 and 1 <= idle <= 4
    assert all(np.array_equal(r, expected) for r in results)
    assert np.array_equal(run(), expected)
    _, _, reused1 = F.pool_stats(F.POOL_PAIRHMM)
    assert reused1 == reused0 + 1
    assert F.release_idle() >= idle
    assert F.pool_stats(F.POOL_PAIRHMM)[0] == 0
    assert np.array_equal(run(), expected)


def test_pd_native_entry_matches_reference(gpu):
    data = PdRegion.generate(31, 70, 120, 24, 300)
    expected = pd_reference_all(data.region_reads_list(), data.region_pd_haps_list())
    ro, ho = _native_arrays(data)
    out = np.zeros(len(expected))
    F.pd_compute_native(data.bases, data.quals, data.ins, data.dels, data.gcp, ro,
                        data.hap_bases, data.hap_flags, ho, True, out)
    assert_rel(out, expected, what="pd native")


def test_aggregated_native_calls_match_the_oracle(gpu):
    """Calls from several threads gathered into shared batches give each caller its own region's
    likelihoods, in both precisions and for PD, within the precision's bound of the oracle."""
    plain = [Region.generate(40 + i, 30 + 7 * i, 80 + 11 * i, 6 + i, 200 + 9 * i) for i in range(6)]
    pdr = [PdRegion.generate(60 + i, 20 + 5 * i, 70 + 9 * i, 5 + i, 180) for i in range(4)]
    cases = [(False, dp, r) for r in plain for dp in (True, False)]
    cases += [(True, dp, r) for r in pdr for dp in (True, False)]
    expected = {}
    for i, (pd, dp, r) in enumerate(cases):
        if pd:
            expected[i] = pd_reference_all(r.region_reads_list(), r.region_pd_haps_list())
        else:
            expected[i] = reference_all(r.region_reads_list(), r.region_haps_list())
    results, errors = {}, []

    def call(i):
        pd, dp, r = cases[i]
        ro, ho = _native_arrays(r)
        out = np.zeros(r.n_pairs())
        if pd:
            F.pd_compute_native(r.bases, r.quals, r.ins, r.dels, r.gcp, ro, r.hap_bases,
                                r.hap_flags, ho, dp, out)
        else:
            F.compute_native(r.bases, r.quals, r.ins, r.dels, r.gcp, ro, r.hap_bases, ho, dp, out)
        results[i] = out

    def worker(idx):
        try:
            for i in idx:
                call(i)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    b0, r0 = F.aggregate_stats()
    F.aggregate(8, 20000)
    try:
        threads = [threading.Thread(target=worker, args=(list(range(t, len(cases), 6)),))
                   for t in range(6)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        # an invalid region fails only its own call, with the entry point's own message
        bad = Region.generate(3, 4, 30, 2, 60)
        ro, ho = _native_arrays(bad)
        with pytest.raises(F.InvalidArgument, match="elements but reads x haplotypes"):
            F.compute_native(bad.bases, bad.quals, bad.ins, bad.dels, bad.gcp, ro,
                             bad.hap_bases, ho, True, np.zeros(bad.n_pairs() + 1))
    finally:
        F.aggregate(0)
    assert not errors, errors
    b1, r1 = F.aggregate_stats()
    assert r1 - r0 == len(cases)
    for i, (pd, dp, r) in enumerate(cases):
        what = f"case {i} pd={pd} double={dp}"
        if dp:
            assert_rel(results[i], expected[i], what=what)
        else:
            assert_close(results[i], expected[i], FLOAT_TOL, what)
    # off again: calls run on their own and the counters stay put
    pd, dp, r = cases[0]
    call(0)
    assert F.aggregate_stats() == (b1, r1)


def test_aggregated_calls_released_together_share_batches(gpu):
    """Eight threads released together, four times, on one kind: the first call of a round runs at
    once on the idle GPU (at least 0.1 ms) while the other seven queue behind it, so the rounds
    need fewer batches than calls; every caller still gets its own region's likelihoods."""
    regions = [Region.generate(90 + i, 40, 120, 8, 260) for i in range(8)]
    expected = [reference_all(r.region_reads_list(), r.region_haps_list()) for r in regions]
    barrier = threading.Barrier(8)
    errors, results = [], {}

    def worker(i):
        r = regions[i]
        ro, ho = _native_arrays(r)
        try:
            for k in range(4):
                out = np.zeros(r.n_pairs())
                barrier.wait()
                F.compute_native(r.bases, r.quals, r.ins, r.dels, r.gcp, ro, r.hap_bases, ho,
                                 False, out)
                results[(i, k)] = out
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    b0, r0 = F.aggregate_stats()
    F.aggregate(8, 5000)
    try:
        threads = [threading.Thread(target=worker, args=(i,)) for i in range(8)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    finally:
        F.aggregate(0)
    assert not errors, errors
    b1, r1 = F.aggregate_stats()
    assert r1 - r0 == 32
    assert b1 - b0 < 32, "no two calls shared a batch"
    for (i, k), out in results.items():
        assert_close(out, expected[i], FLOAT_TOL, f"thread {i} round {k}")


# ---- the batched engine ------------------------------------------------------------------------
def test_a_batch_planned_on_several_host_threads_matches_the_oracle_and_its_parts(gpu):
    """1200 small regions are planned in ranges on several host threads and joined in region
    order; the results match the oracle, and every pair has the bits it gets when its regions
    are planned in batches of 200 (one planning thread each)."""
    regions = []
    for i in range(1200):
        r = Region.generate(500 + i, 3 + i % 4, 40 + i % 50, 2 + i % 3, 90 + i % 40)
        regions.append((r.region_reads_list(), r.region_haps_list()))
    data = batch_from_regions(regions)
    b = F.Batch(data)
    b.run(F.MODE_FP64, share=True)
    assert_rel(b.results(), batch_reference(data), what="fp64")
    b.run(F.MODE_FP32_RESCUE, share=True)
    whole = b.results()
    b.close()
    parts = []
    for k in range(0, 1200, 200):
        pb = F.Batch(batch_from_regions(regions[k:k + 200]))
        pb.run(F.MODE_FP32_RESCUE, share=True)
        parts.append(pb.results())
        pb.close()
    assert np.array_equal(whole, np.concatenate(parts))

@pytest.mark.parametrize("wl,regions", [(0, 1), (1, 6), (2, 3), (3, 4), (4, 6)])
def test_batch_workloads_match_the_oracle(gpu, wl, regions):
    data = workload(wl, seed=1234 + wl, n_regions=regions)
    expected = batch_reference(data)
    b = F.Batch(data)
    st = b.run(F.MODE_FP64, share=True)
    assert_rel(b.results(), expected, what=f"workload {wl} fp64")
    st = b.run(F.MODE_FP32_RESCUE, share=True)
    assert_close(b.results(), expected, FLOAT_TOL, f"workload {wl} fp32")
    if wl == 2:
        assert st.rescued_pairs >= 0.25 * data.n_pairs()
    b.run(F.MODE_FP64_PER_REGION)
    assert_rel(b.results(), expected, what=f"workload {wl} per-region")
    b.close()


def _wave_pairs():
    """(R, A, B): read lengths that fill one wave at exactly R rows per lane, for R = 1..8, with
    1, about R/2 and R identity rows in read B's first lane (and a varying count in read A's)."""
    out = []
    for R in range(1, 9):
        for pad_b in sorted({1, (R + 1) // 2, R}):
            lanes_b = 6 + 3 * R
            pad_a = 1 + (R + pad_b) % R
            out.append((R, R * (64 - lanes_b) - pad_a, R * lanes_b - pad_b))
    return out


@pytest.mark.parametrize("pd", [False, True])
def test_two_reads_sharing_a_wave_match_the_oracle(gpu, pd):
    # One region per length pair, so the planner can only pair its two reads or not; at these
    # lengths pairing always lowers the modelled cost, so each region's cells land in bucket R.
    rng = Rng(404 + pd)
    regions, want = [], np.zeros(8, np.int64)
    for R, na, nb in _wave_pairs():
        ref = bytes(rng.base() for _ in range(na + 120))
        reads = []
        for n in (na, nb):
            start = rng.below(60)
            bases = bytearray(ref[start:start + n])
            for i in range(rng.below(11), n, 13):
                bases[i] = rng.base()
            reads.append(F.ReadRef(bytes(bases), [15 + rng.below(26) for _ in range(n)],
                                   [30 + rng.below(16) for _ in range(n)],
                                   [30 + rng.below(16) for _ in range(n)],
                                   [5 + rng.below(8) for _ in range(n)]))
        haps = [ref, ref[:na // 2] + ref[na // 2 + 3:], ref[:40] + b"G" + ref[41:]]
        if pd:
            flagged = []
            for h in haps:
                f = [0] * len(h)
                for c in range(5 + rng.below(20), len(h), 37):
                    f[c] = F.SNP | (F.ALT_A, F.ALT_C, F.ALT_G, F.ALT_T)[rng.below(4)]
                d0 = 10 + rng.below(len(h) - 20)
                f[d0] |= F.DEL_START
                f[d0 + 1 + rng.below(4)] |= F.DEL_END
                flagged.append(F.PdHaplotype(h, f))
            haps = flagged
        regions.append((reads, haps))
        want[R - 1] += (na + nb) * sum(len(h.bases if pd else h) for h in haps)
    data = batch_from_regions(regions, pd=pd)
    expected = batch_reference(data)
    b = F.Batch(data)
    for share in (True, False):
        st = b.run(F.MODE_FP64, share=share)
        assert_rel(b.results(), expected, what=f"paired reads fp64 share={share}")
        st = b.run(F.MODE_FP32_RESCUE, share=share)
        assert_close(b.results(), expected, FLOAT_TOL, f"paired reads fp32 share={share}")
    assert list(st.bucket_cells) == list(want)
    assert sum(st.bucket_cells) == data.cells()
    b.close()


def _launch_space(R):
    rows = [R] + ([R + 1] if R < 8 else [])
    return [(r, b, o) for r in rows for b in (256, 128, 64) for o in (False, True)]


@pytest.mark.parametrize("wl", [1, 3])
def test_every_launch_configuration_gives_the_same_bits(gpu, wl):
    """Every rows per lane (R and R + 1), block size and build of every FP32 bucket gives the
    bits of the default launch, with prefix (and, plain, suffix) sharing; PD included."""
    data = workload(wl, seed=808 + wl, n_regions=5)
    b = F.Batch(data)
    for R in range(1, 9):
        b.set_launch(R)  # the default, whatever an earlier tuning chose
    b.run(F.MODE_FP32_RESCUE, share=True)
    ref = b.results()
    assert_close(ref, batch_reference(data), FLOAT_TOL, f"workload {wl}")
    st = b.run(F.MODE_FP32_RESCUE, share=True)
    used = [R for R in range(1, 9) if st.bucket_cells[R - 1] > 0]
    assert len(used) >= 3
    for R in used:
        for rows, block, boost in _launch_space(R):
            b.set_launch(R, rows, block, boost)
            assert b.get_launch(R) == (rows, block, boost)
            b.run(F.MODE_FP32_RESCUE, share=True)
            got = b.results()
            assert np.array_equal(got, ref), f"bucket {R}: rows {rows}, block {block}, boost {boost}"
        b.set_launch(R)
    b.close()


def test_tuning_keeps_results_and_is_inherited_by_later_batches(gpu):
    data = workload(1, seed=4242, n_regions=6)
    b = F.Batch(data)
    b.run(F.MODE_FP32_RESCUE, share=True, graph=True)
    before = b.results()
    rep = b.tune(share=True)
    st = b.run(F.MODE_FP32_RESCUE, share=True)
    assert rep.candidates > 0 and rep.seconds > 0
    for R in range(1, 9):
        if st.bucket_cells[R - 1] == 0:
            continue
        assert rep.rows[R - 1] in (R, min(R + 1, 8))
        assert rep.block[R - 1] in (64, 256)
        assert 0 < rep.tuned_ms[R - 1] <= rep.default_ms[R - 1]
        assert b.get_launch(R) == (rep.rows[R - 1], rep.block[R - 1], bool(rep.boost[R - 1]))
    assert np.array_equal(b.results(), before)
    for _ in range(2):  # graphs captured after tuning use the tuned launches
        b.run(F.MODE_FP32_RESCUE, share=True, graph=True)
        assert np.array_equal(b.results(), before)
        b.run(F.MODE_FP64, share=True)
    later = F.Batch(workload(1, seed=4243, n_regions=6))
    for R in range(1, 9):
        if st.bucket_cells[R - 1]:
            assert later.get_launch(R) == b.get_launch(R)
    with pytest.raises(F.InvalidArgument):
        b.set_launch(3, rows=5)
    with pytest.raises(F.InvalidArgument):
        b.set_launch(2, block=512)
    pd = F.Batch(workload(3, seed=4244, n_regions=2))
    pd.set_launch(1, rows=2)  # PD buckets take R + 1 like plain ones
    assert pd.get_launch(1)[0] == 2
    with pytest.raises(F.InvalidArgument):
        pd.set_launch(1, rows=3)
    pd.close()
    later.close()
    b.close()


# ---- suffix sharing (plan.rs, kernel.rs:579-602, 744-891, 1029-1126) ---------------------------
@pytest.mark.parametrize("wl,regions", [(1, 6), (4, 8), (0, 1)])
def test_suffix_sharing_matches_the_oracle(gpu, wl, regions):
    data = workload(wl, seed=515 + wl, n_regions=regions)
    expected = batch_reference(data)
    b = F.Batch(data)
    work = b.work()
    assert work["backward"] > 0 and work["forward"] < work["prefix"] < work["nominal"]
    st = b.run(F.MODE_FP32_RESCUE, share=True)
    assert_close(b.results(), expected, FLOAT_TOL, f"suffix sharing, workload {wl}")
    shared = b.results()
    b.run(F.MODE_FP32_RESCUE, share=False)
    assert_close(b.results(), shared, FLOAT_TOL, "prefix-only vs suffix sharing")
    b.run(F.MODE_FP64, share=True)  # FP64 shares prefixes and suffixes
    fp64_shared = b.results()
    assert_rel(fp64_shared, expected, what="fp64 with suffix sharing")
    b.run(F.MODE_FP64, share=False)
    assert_rel(fp64_shared, b.results(), what="fp64 suffix sharing vs unshared")
    b.run(F.MODE_FP64, share=True, graph=True)
    assert np.array_equal(b.results(), fp64_shared)
    # The scratch now holds FP64 records: the FP32 pass must read none it did not write.
    b.run(F.MODE_FP32_RESCUE, share=True)
    assert np.array_equal(b.results(), shared)
    for _ in range(2):  # graph replays give the eager bits
        b.run(F.MODE_FP32_RESCUE, share=True, graph=True)
        assert np.array_equal(b.results(), shared)
    b.close()


def _edited_sets(seed, cases):
    rng = Rng(seed)
    regions = []
    for case in range(cases):
        alphabet = b"AC" if case % 2 == 0 else b"ACGT"
        base = bytes(alphabet[rng.below(len(alphabet))] for _ in range(30 + rng.below(150)))
        haps = [base]
        for _ in range(1 + rng.below(12)):
            src = haps[rng.below(len(haps))]
            if rng.chance(0.15):
                haps.append(src)
                continue
            h = bytearray(src)
            at = rng.below(len(h))
            h[at:at + rng.below(4)] = bytes(alphabet[rng.below(len(alphabet))]
                                             for _ in range(rng.below(4)))
            haps.append(bytes(h) if h else src)
        reads = []
        for _ in range(1 + rng.below(12)):
            src = haps[rng.below(len(haps))]
            n = 1 + rng.below(len(src))
            s0 = rng.below(len(src) - n + 1)
            reads.append(F.ReadRef(src[s0:s0 + n], [10 + rng.below(31) for _ in range(n)],
                                   [20 + rng.below(26) for _ in range(n)],
                                   [20 + rng.below(26) for _ in range(n)],
                                   [1 + rng.below(10) for _ in range(n)]))
        regions.append((reads, haps))
    return regions


def test_suffix_sharing_on_edited_haplotype_sets(gpu):
    # Random edits of a base sequence over small alphabets, with copies (as the reference's plan
    # tests build them, plan.rs:320-340): cuts at every depth, resumed backward sweeps, duplicate
    # haplotypes, reads of every length from 1 base, alone and two to a wave.
    data = batch_from_regions(_edited_sets(2718, 60))
    expected = batch_reference(data)
    b = F.Batch(data)
    assert b.work()["backward"] > 0
    b.run(F.MODE_FP32_RESCUE, share=True)
    got = b.results()
    assert_close(got, expected, FLOAT_TOL, "edited sets, suffix sharing")
    b.run(F.MODE_FP32_NO_RESCUE, share=True)
    nores = b.results()
    ok = np.isfinite(nores)
    assert np.array_equal(nores[ok], got[ok])
    b.run(F.MODE_FP64, share=True)
    assert_rel(b.results(), expected, what="edited sets, fp64 suffix sharing")
    b.close()


@pytest.mark.parametrize("mode", [F.MODE_FP32_RESCUE, F.MODE_FP64])
def test_suffix_shared_bits_do_not_depend_on_the_other_regions(gpu, mode):
    regions = _edited_sets(3141, 12)
    whole = F.Batch(batch_from_regions(regions))
    whole.run(mode, share=True)
    all_bits = whole.results()
    whole.close()
    for g in (0, 5, 11):
        alone = F.Batch(batch_from_regions([regions[g]]))
        alone.run(mode, share=True)
        bits = alone.results()
        alone.close()
        start = sum(len(r) * len(h) for r, h in regions[:g])
        assert np.array_equal(bits, all_bits[start:start + len(bits)])


def test_fp64_joins_below_1e24_are_rerun_without_suffix_sharing(gpu):
    """simd.rs:136 and lib.rs:467-487: FP64 backward values are unscaled, so a joined FP64 result
    below 1e24 (a likelihood below about 1e-283 once the 2^1020 scale is divided out) may have
    lost precision and is recomputed without suffix sharing.  Reads with a substitution every
    third base at Phred 40 and gap continuation Phred 20 (so that one long insertion costs about
    as much: log10 L near -360 and -390) against haplotypes that share a suffix are rerun
    (the run counts them) and match the oracle at the FP64 bound, as do the region's ordinary
    reads, whose joins are accepted."""
    rng = Rng(61)
    base = bytes(b"ACGT"[rng.below(4)] for _ in range(320))
    haps = [base]
    for j in (8, 16, 24, 32):  # distinct prefixes, one suffix from column 33 on
        h = bytearray(base)
        h[j] = ord("A") if h[j] != ord("A") else ord("C")
        haps.append(bytes(h))
    reads = [F.ReadRef(base[s:s + 150], [30] * 150, [45] * 150, [45] * 150, [10] * 150)
             for s in (0, 120, 160)]
    for s, n in ((20, 240), (40, 260)):
        b = bytearray(base[s:s + n])
        for j in range(0, n, 3):
            b[j] = ord("A") if b[j] != ord("A") else ord("G")
        reads.append(F.ReadRef(bytes(b), [40] * n, [45] * n, [45] * n, [20] * n))
    data = batch_from_regions([(reads, haps)])
    expected = batch_reference(data)
    H = len(haps)
    assert np.all(expected[3 * H:] < -300) and np.all(expected[:3 * H] > -20)
    bt = F.Batch(data)
    assert bt.work()["backward"] > 0
    st = bt.run(F.MODE_FP64, share=True)
    got = bt.results()
    assert_rel(got, expected, what="fp64 joins with reruns")
    assert st.rescued_pairs > 0
    bt.run(F.MODE_FP64, share=False)
    unshared = bt.results()
    assert np.array_equal(got[3 * H:], unshared[3 * H:])  # the reruns are the unshared sweeps
    bt.close()


def test_batch_graph_replays_match_eager_runs(gpu):
    data = workload(1, seed=99, n_regions=5)
    b = F.Batch(data)
    b.run(F.MODE_FP32_RESCUE, share=True, graph=False)
    eager = b.results()
    for _ in range(3):
        b.run(F.MODE_FP32_RESCUE, share=True, graph=True)
        assert np.array_equal(b.results(), eager)
        b.run(F.MODE_FP64, share=True, graph=False)  # eager work between replays
    b.close()


def test_batch_prefix_sharing_only_moves_results_within_tolerance(gpu):
    data = workload(1, seed=5, n_regions=4)
    b = F.Batch(data)
    b.run(F.MODE_FP64, share=True)
    shared = b.results()
    b.run(F.MODE_FP64, share=False)
    unshared = b.results()
    assert_rel(shared, unshared, what="share vs no-share")
    b.close()


@pytest.mark.parametrize("wl,regions", [(1, 6), (3, 4), (2, 2)])
def test_snapshot_records_in_lds_give_the_bits_of_hbm_records(gpu, wl, regions):
    """Prefix-sharing launches keep one snapshot record per wave in LDS where the plan marks one
    and the record costs no resident workgroups.  The restored state, with the row above each
    lane's first row moved down by DPP rather than stored, gives the bits of the HBM records in
    both precisions (workload 2 reruns its dense reads in FP64 with prefix sharing)."""
    data = workload(wl, seed=606 + wl, n_regions=regions)
    b = F.Batch(data)
    runs = {}
    for lds in (True, False):
        b.set_lds_records(lds)
        before = b.snapshots()["lds_launches"]
        for mode in (F.MODE_FP32_RESCUE, F.MODE_FP64):
            b.run(mode, share=True)
            runs[(lds, mode)] = b.results()
        used = b.snapshots()["lds_launches"] - before
        assert (used > 0) == lds, f"lds={lds}: {used} launches kept an LDS record"
    assert b.snapshots()["lds"] > 0
    for mode in (F.MODE_FP32_RESCUE, F.MODE_FP64):
        assert np.array_equal(runs[(True, mode)], runs[(False, mode)]), f"mode {mode}"
    expected = batch_reference(data)
    assert_close(runs[(True, F.MODE_FP32_RESCUE)], expected, FLOAT_TOL, f"workload {wl} fp32")
    assert_rel(runs[(True, F.MODE_FP64)], expected, what=f"workload {wl} fp64")
    b.close()


def _perturbed_gaps(data, every: int, seed: int):
    """A copy of `data` in which one base of every `every`-th read (of at least two bases) has
    its insertion quality raised by one: those reads no longer have uniform gap qualities."""
    import dataclasses

    ins = data.ins.copy()
    rng = np.random.default_rng(seed)
    for r in range(0, len(data.read_offsets) - 1, every):
        o0, o1 = int(data.read_offsets[r]), int(data.read_offsets[r + 1])
        if o1 - o0 >= 2:
            i = int(rng.integers(o0 + 1, o1))
            ins[i] = min(int(ins[i]) + 1, 60)
    return dataclasses.replace(data, ins=ins)


@pytest.mark.parametrize("case", ["uniform", "mixed", "rescue"])
def test_uniform_gap_build_gives_the_bits_of_the_general_build(gpu, case):
    """Items whose reads have the same insertion, deletion and gap-continuation quality at every
    base run with one copy of the gap terms per lane (the uniform-gap build); the others, and
    items pairing such a read with one that varies, with the general build after them in the same
    bucket.  Every run gives the general build's bits: FP32 rescued or not and FP64 (prefix and
    suffix sharing), with and without LDS snapshot records, from a graph, and within the bounds of
    the oracle."""
    if case == "rescue":  # underflowing reads: lost pairs, dense reads rerun in FP64
        data = workload(2, seed=913, n_regions=2)
    else:  # edited haplotype sets: prefix and suffix sharing
        data = workload(1, seed=911, n_regions=6)
        if case == "mixed":
            data = _perturbed_gaps(data, 3, 5)
    b = F.Batch(data)
    runs = {}
    for ug in (True, False):
        b.set_uniform_gaps(ug)
        for lds in (True, False):
            b.set_lds_records(lds)
            for mode in (F.MODE_FP32_NO_RESCUE, F.MODE_FP32_RESCUE, F.MODE_FP64):
                b.run(mode, share=True)
                runs[(ug, lds, mode)] = b.results()
    for lds in (True, False):
        for mode in (F.MODE_FP32_NO_RESCUE, F.MODE_FP32_RESCUE, F.MODE_FP64):
            assert np.array_equal(runs[(True, lds, mode)], runs[(False, lds, mode)],
                                  equal_nan=True), (case, lds, mode)
    b.set_uniform_gaps(True)
    b.set_lds_records(True)
    for _ in range(2):
        b.run(F.MODE_FP32_RESCUE, share=True, graph=True)
        assert np.array_equal(b.results(), runs[(True, True, F.MODE_FP32_RESCUE)])
    expected = batch_reference(data)
    assert_close(runs[(True, True, F.MODE_FP32_RESCUE)], expected, FLOAT_TOL, case)
    assert_rel(runs[(True, True, F.MODE_FP64)], expected, what=f"{case} fp64")
    b.close()


def _regapped(data, seed: int):
    """A copy of `data` whose reads take uniform gap qualities from wide ranges (insertion and
    deletion 5..90, gap continuation 1..40, so some reads fall below the gap-continuation quality
    of 4 that scaled insertion values need), every fourth read with the gap-continuation quality of
    one base changed, so its gaps vary."""
    import dataclasses

    ins, dels, gcp = data.ins.copy(), data.dels.copy(), data.gcp.copy()
    rng = np.random.default_rng(seed)
    for r in range(len(data.read_offsets) - 1):
        o0, o1 = int(data.read_offsets[r]), int(data.read_offsets[r + 1])
        ins[o0:o1] = rng.choice([5, 20, 45, 70, 90])
        dels[o0:o1] = rng.choice([5, 20, 45, 70, 90])
        gcp[o0:o1] = rng.choice([1, 2, 3, 4, 5, 10, 40])
        if r % 4 == 3 and o1 - o0 >= 2:
            i = int(rng.integers(o0 + 1, o1))
            gcp[i] = gcp[i] + 1
    return dataclasses.replace(data, ins=ins, dels=dels, gcp=gcp)


@pytest.mark.parametrize("pd", [False, True])
def test_scaled_gap_values_give_one_set_of_bits_on_every_path(gpu, pd):
    """Reads with uniform gaps and a gap-continuation quality of at least 4 keep their insertion
    and deletion values divided by the match-to-insertion and match-to-deletion probabilities in
    both sweep builds (PD: the branch values alike).  Over wide gap qualities (scaled, unscaled and
    varying reads, paired in waves), the uniform-gap and the general build give the same bits in
    FP32 (rescued) and FP64 with prefix sharing (plain: and suffix sharing) and in the
    one-pair-per-wave runs; a region alone in a batch, all of whose reads may take the uniform-gap
    build in every launch, gets its bits in the whole batch's pair runs; all lie within the
    oracle's bounds."""
    data = _regapped(workload(3 if pd else 1, seed=1717, n_regions=6), 3)
    expected = batch_reference(data)
    b = F.Batch(data)
    for mode in (F.MODE_FP32_RESCUE, F.MODE_FP64):
        runs, pair_runs = {}, {}
        for ug in (True, False):
            b.set_uniform_gaps(ug)
            b.run(mode, share=True)
            runs[ug] = b.results()
            b.run_pairs(mode)
            pair_runs[ug] = b.results()
        assert np.array_equal(runs[True], runs[False], equal_nan=True), mode
        assert np.array_equal(pair_runs[True], pair_runs[False], equal_nan=True), mode
        pairs = pair_runs[True]
        if mode == F.MODE_FP64:
            assert_rel(runs[True], expected, what="fp64 shared")
            assert_rel(pairs, expected, what="fp64 pairs")
        else:
            assert_close(runs[True], expected, FLOAT_TOL, "fp32 shared")
            assert_close(pairs, expected, FLOAT_TOL, "fp32 pairs")
    b.set_uniform_gaps(True)
    b.run_pairs(F.MODE_FP32_RESCUE)
    whole = b.results()
    base = data.out_base()
    for g in range(data.n_regions):  # a region alone in a batch: the pair runs' bits
        haps = data.region_pd_haps_list(g) if pd else data.region_haps_list(g)
        one = F.Batch(batch_from_regions([(data.region_reads_list(g), haps)], pd=pd))
        one.run_pairs(F.MODE_FP32_RESCUE)
        assert np.array_equal(one.results(), whole[base[g]:base[g + 1]]), g
        one.close()
    b.close()
