"""Smith-Waterman on the device against the oracle (oracle/sw.c, GATK's aligner restated), bit
for bit: CIGAR and offset of every pair.  The cases follow the reference's own tests
(crates/smithwaterman/src/lib.rs:854-1060; FgklSmithWatermanTest.java): its known answers,
related and unrelated random pairs under the three weight sets and four strategies, odd match
values, a reused aligner as pair sizes shrink and grow, plus the layouts the device kernel has
that the CPU has not: references longer than one strip of lanes, every rows-per-lane bucket,
alternates shorter than the wave and single-base sequences."""
from __future__ import annotations

import numpy as np
import pytest

import fgkl_amd.sw as S
from fgkl_amd.synthetic import Rng
from _oracle import sw_align
from _sw_cases import (DANGLING, HAP_TO_REF, ORIGINAL, READ_TO_HAP, STRATEGIES, random_seq,
                       related_pair)

pytestmark = pytest.mark.gpu


def _check_batch(pairs, params, strategies=STRATEGIES, label=""):
    batch = S.SwBatch(pairs)
    for strategy in strategies:
        code = S.OverhangStrategy[strategy]
        got = batch.align(params, code)
        for k, ((ref, alt), a) in enumerate(zip(pairs, got)):
            want = sw_align(ref, alt, params, strategy)
            assert (a.cigar_string(), a.offset) == want, (
                f"{label} pair {k} {strategy} {params} n={len(ref)} m={len(alt)}\n"
                f"ref={ref.decode()}\nalt={alt.decode()}")
    return batch


def test_known_answers(gpu):  # lib.rs:927-968
    al = S.Aligner()
    a = al.align(b"AAACGTACGTAAA", b"ACGTACGT", HAP_TO_REF, S.OverhangStrategy.SOFTCLIP)
    assert (a.cigar_string(), a.offset) == ("8M", 2)
    assert al.align(b"ACGTACGTACGT", b"ACGT", HAP_TO_REF, S.OverhangStrategy.IGNORE).offset == 8
    ref = b"NNNNNNNNNNACGTTTGCAAGGCTTAGGCTNNNNNNNNNN"
    a = al.align(ref, b"NNNNNNNNNNACGTTTGCGGCTTAGGCTNNNNNNNNNN", HAP_TO_REF, S.OverhangStrategy.INDEL)
    assert (a.cigar_string(), a.offset) == ("18M2D20M", 0)
    a = al.align(ref, b"NNNNNNNNNNACGTTTGCAAGGCTTTTAGGCTNNNNNNNNNN", HAP_TO_REF,
                 S.OverhangStrategy.INDEL)
    assert (a.cigar_string(), a.offset) == ("23M2I17M", 0)
    read = b"GGGGACGTACGTAGGCCTTAGCAGGGG"
    a = al.align(b"ACGTACGTAGGCCTTAGCA", read, READ_TO_HAP, S.OverhangStrategy.SOFTCLIP)
    assert (a.cigar_string(), a.offset) == ("4S19M4S", 0)
    a = al.align(b"TTTTTTTTTTACGTACGTAGGCCTTAGCATTTTTTTTTT", read, READ_TO_HAP,
                 S.OverhangStrategy.SOFTCLIP)
    assert a.cigar_string() == "4I19M4I"
    # the hand-worked ties of the oracle's tests
    assert S.align_native(b"A", b"C", (1, -100, -3, -1), S.OverhangStrategy.INDEL) == ("1D1I", 0)
    assert S.align_native(b"AA", b"CC", ORIGINAL, S.OverhangStrategy.SOFTCLIP) == ("1S1M", 0)


def test_empty_sequences_are_rejected(gpu):  # lib.rs:970-981
    with pytest.raises(S.InvalidArgument):
        S.align_native(b"", b"A", HAP_TO_REF, 0)
    with pytest.raises(S.InvalidArgument):
        S.SwBatch([(b"ACGT", b"")])


@pytest.mark.parametrize("params", [HAP_TO_REF, READ_TO_HAP, ORIGINAL], ids=["h2r", "r2h", "orig"])
def test_random_related_pairs(gpu, params):  # lib.rs:897-912
    rng = Rng(7)
    pairs = []
    for case in range(400):
        if case % 2 == 0:
            pairs.append(related_pair(rng, 40 + rng.below(300), 20 + rng.below(200)))
        else:
            pairs.append(related_pair(rng, 5 + rng.below(30), 3 + rng.below(30)))
    _check_batch(pairs, params, label="related")


def test_unrelated_pairs(gpu):  # lib.rs:914-925
    rng = Rng(11)
    pairs = [(random_seq(rng, 5 + rng.below(80)), random_seq(rng, 5 + rng.below(80)))
             for _ in range(100)]
    _check_batch(pairs, HAP_TO_REF, label="unrelated")
    _check_batch(pairs, READ_TO_HAP, label="unrelated")


def test_odd_match_values(gpu):  # lib.rs:883-894
    rng = Rng(31)
    pairs = [related_pair(rng, 40 + rng.below(200), 20 + rng.below(150)) for _ in range(50)]
    _check_batch(pairs, DANGLING, label="dangling")


def test_long_references_span_several_strips(gpu):
    """References beyond 512 bases run in several strips of 64 lanes x 8 rows; the cases put the
    end cell, the last row and gaps on either side of strip boundaries."""
    rng = Rng(41)
    pairs = []
    for ref_len in (513, 600, 1024, 1025, 1300, 2100):
        for alt_len in (30, 200, 700):
            pairs.append(related_pair(rng, ref_len, alt_len))
    pairs.append((random_seq(rng, 700), random_seq(rng, 700)))   # lib.rs:854-866
    r = random_seq(rng, 600)
    a = bytearray(r[380:])
    a[100] = ord("C") if a[100] == ord("A") else ord("A")
    pairs.append((r, bytes(a)))                                  # lib.rs:868-881
    for params in (HAP_TO_REF, READ_TO_HAP):
        _check_batch(pairs, params, label="strips")


def test_every_bucket_and_short_alternates(gpu):
    """Reference lengths across the rows-per-lane buckets (1..8 rows per lane), alternates from one
    base to beyond the wave, single-base references."""
    rng = Rng(43)
    pairs = []
    for ref_len in (1, 2, 7, 63, 64, 65, 128, 129, 200, 256, 257, 320, 383, 448, 511, 512):
        for alt_len in (1, 2, 5, 63, 64, 65, 150):
            if rng.chance(0.5):
                pairs.append(related_pair(rng, ref_len, alt_len))
            else:
                pairs.append((random_seq(rng, ref_len), random_seq(rng, alt_len)))
    for params in (HAP_TO_REF, ORIGINAL):
        _check_batch(pairs, params, label="buckets")


def test_a_reused_aligner_as_sizes_shrink_and_grow(gpu):  # lib.rs:1022-1060
    rng = Rng(11)
    sizes = [(600, 520), (500, 150), (40, 30), (320, 310), (8, 5), (590, 600), (60, 200)]
    pairs = [(random_seq(rng, r), random_seq(rng, a)) if k == 0 else related_pair(rng, r, a)
             for k, (r, a) in enumerate(sizes)]
    al = S.Aligner()
    for ref, alt in pairs:
        for params in (HAP_TO_REF, READ_TO_HAP):
            for strategy in STRATEGIES:
                a = al.align(ref, alt, params, S.OverhangStrategy[strategy])
                assert (a.cigar_string(), a.offset) == sw_align(ref, alt, params, strategy)


def test_one_batch_runs_again_with_other_weights(gpu):
    rng = Rng(47)
    pairs = [related_pair(rng, 100 + rng.below(400), 50 + rng.below(300)) for _ in range(60)]
    batch = S.SwBatch(pairs)
    for params, strategy in ((HAP_TO_REF, "INDEL"), (READ_TO_HAP, "SOFTCLIP"),
                             (HAP_TO_REF, "INDEL"), (ORIGINAL, "IGNORE")):
        got = batch.align(params, S.OverhangStrategy[strategy])
        for (ref, alt), a in zip(pairs, got):
            assert (a.cigar_string(), a.offset) == sw_align(ref, alt, params, strategy)


def test_batch_stats_count_buckets_and_shortcuts(gpu):
    rng = Rng(53)
    pairs = [related_pair(rng, 300, 100) for _ in range(20)]
    pairs += [(b"AAACGTACGTAAA", b"ACGTACGT")] * 3
    batch = S.SwBatch(pairs)
    st = batch.run(HAP_TO_REF, S.OverhangStrategy.SOFTCLIP)
    assert sum(st.bucket_pairs) == len(pairs)
    assert sum(st.bucket_cells) == sum(len(r) * len(a) for r, a in pairs)
    assert st.substring_pairs >= 3
    assert st.total_ms > 0
    st = batch.run(HAP_TO_REF, S.OverhangStrategy.INDEL)
    assert st.substring_pairs == 0
