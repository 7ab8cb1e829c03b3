"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Checks, in order: the model.rs:108-155 known answers; reference.rs:66-115 known answers and
invariants; an independent exact path-sum enumerator over every alignment path of tiny pairs
(which fixes the recurrence: free start C/m in row 0, the three transitions, the free end on
the last row); closed forms for a one-base read; the Jacobian table's quantisation bound over
the whole match-to-match table; the PD known answers of pd_reference.rs:120-184 and pd.rs; and
the gradient oracle against central finite differences.
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

import _oracle as O
import fgkl_amd as F
from fgkl_amd.synthetic import PdRegion, Region, Rng

o = O.oracle()


def read(bases: bytes, q=30, i=45, d=45, g=10):
    n = len(bases)
    return F.ReadRef(bases, [q] * n, [i] * n, [d] * n, [g] * n)


# ---- model.rs:108-155 --------------------------------------------------------------------------
def test_error_prob_matches_phred_definition():
    assert abs(o.fgklo_error_prob(0) - 1.0) < 1e-15
    assert abs(o.fgklo_error_prob(10) - 0.1) < 1e-15
    assert abs(o.fgklo_error_prob(30) - 0.001) < 1e-15


def test_quals_above_max_are_clamped():
    assert o.fgklo_error_prob(255) == o.fgklo_error_prob(254)
    assert o.fgklo_match_to_match(255, 10) == o.fgklo_match_to_match(254, 10)


def test_match_to_match_is_symmetric_and_close_to_one_minus_gap_opens():
    assert o.fgklo_match_to_match(45, 30) == o.fgklo_match_to_match(30, 45)
    expected = 1.0 - (o.fgklo_error_prob(45) + o.fgklo_error_prob(30))
    assert abs(o.fgklo_match_to_match(45, 30) - expected) < 1e-6


def test_match_to_match_at_zero_quals_is_zero():
    assert o.fgklo_match_to_match(0, 0) == 0.0


def test_transitions_use_gap_continuation_for_extension():
    mm, im, mi, ii, md, dd = O.transitions(45, 40, 10)
    assert abs(ii - 0.1) < 1e-15 and abs(dd - 0.1) < 1e-15 and abs(im - 0.9) < 1e-15
    assert abs(mi - o.fgklo_error_prob(45)) < 1e-15 and abs(md - o.fgklo_error_prob(40)) < 1e-15
    assert mm == o.fgklo_match_to_match(45, 40)


def test_approximate_log10_sum_matches_exact_for_moderate_differences():
    a, b = -3.0, -2.5
    exact = math.log10(10 ** a + 10 ** b)
    assert abs(o.fgklo_approx_log10_sum_log10(a, b) - exact) < 1e-4
    assert o.fgklo_approx_log10_sum_log10(a, b) == o.fgklo_approx_log10_sum_log10(b, a)
    assert o.fgklo_approx_log10_sum_log10(-20.0, -2.0) == -2.0
    assert o.fgklo_approx_log10_sum_log10(-math.inf, -2.0) == -2.0


def test_whole_match_to_match_table_is_within_the_quantisation_bound():
    # mm(i, j) = 1 - 10^approx, where approx is log10(e_i + e_j) with the Jacobian term looked
    # up at a difference rounded to 1e-4 (and dropped beyond 8).  Rounding moves the term by
    # at most 0.5e-4 * d/dx log10(1 + 10^-x) <= 0.5e-4 * 0.5, so 10^approx is within a relative
    # 1e-4 * ln 10 / 4 of e_i + e_j, plus 10^-8 relative for the dropped tail.  A wrong index,
    # a swapped argument or a missing term breaks this everywhere.
    for i in range(0, 255, 3):
        for j in range(0, i + 1, 3):
            s = 10 ** (-i / 10) + 10 ** (-j / 10)
            exact = 1.0 - min(s, 1.0)
            got = o.fgklo_match_to_match(i, j)
            tol = min(s, 1.0) * (1e-4 * math.log(10) / 4 + 1.1e-8) + 1e-15
            assert abs(got - exact) <= tol, (i, j, got, exact)


def test_priors_are_match_and_a_third_of_the_error():
    pm, px = O.priors(20)
    assert pm == 1.0 - o.fgklo_error_prob(20) and px == o.fgklo_error_prob(20) / 3.0


# ---- reference.rs:66-115 -----------------------------------------------------------------------
def test_initial_condition_is_two_to_the_1020():
    assert o.fgklo_initial_condition() == 2.0 ** 1020


def test_identical_single_base_has_likelihood_near_log10_of_match_probability():
    lk = O.log10_likelihood(b"A", read(b"A"))
    assert abs(lk - math.log10((1 - 1e-3) * 0.9)) < 1e-12


def test_mismatch_is_less_likely_than_match():
    matched = O.log10_likelihood(b"ACGT", read(b"ACGT"))
    mismatched = O.log10_likelihood(b"ACGT", read(b"ACTT"))
    assert matched > mismatched and matched <= 0.0


def test_n_matches_every_base_wherever_it_is_aligned():
    exact = O.log10_likelihood(b"ACGT", read(b"ACGT"))
    n_read = O.log10_likelihood(b"ACGT", read(b"NNNN"))
    assert n_read == O.log10_likelihood(b"TTTT", read(b"NNNN"))
    n_hap = O.log10_likelihood(b"NNNN", read(b"ACGT"))
    assert n_hap == O.log10_likelihood(b"NNNN", read(b"GGGG"))
    assert n_read >= exact and n_hap >= exact


def test_one_base_read_closed_form():
    # A one-base read against m columns: every path starts in row 0 at column j-1 (C/m), moves
    # diagonally into M[1][j] with prior * (1 - g), or is an insertion from M[0][j] = 0.  So
    # L = sum_j prior_j * (1 - g) / m.
    hap = b"ACGTTA"
    for q, g in ((30, 10), (12, 25), (40, 3)):
        r = read(b"T", q=q, g=g)
        pm, px = O.priors(q)
        im = 1.0 - o.fgklo_error_prob(g)
        expected = sum((pm if h == ord("T") else px) for h in hap) * im / len(hap)
        assert abs(O.log10_likelihood(hap, r) - math.log10(expected)) < 1e-13


# ---- exact path-sum enumerator (independent of the DP ordering) ---------------------------------
def path_sum(hap: bytes, rd) -> float:
    """Sum over every M/I/D alignment path of the pair HMM, by explicit enumeration.

    States live on a grid of rows 0..n (read positions) and columns 0..m (haplotype positions).
    A path starts in the deletion state of row 0 at any column with weight 1/m, and moves:
    diagonally into a match cell (emitting the read base against the haplotype base, with
    match-to-match or indel-to-match), down into an insertion cell (match-to-insertion or
    insertion-to-insertion), or right into a deletion cell in rows >= 1 (match-to-deletion or
    deletion-to-deletion); row-0 deletions only start paths.  The likelihood is the weight of
    every path whose final cell is a match or insertion in the last row.
    """
    n, m = len(rd.bases), len(hap)
    T = [O.transitions(rd.ins_gop[i], rd.del_gop[i], rd.gcp[i]) for i in range(n)]
    P = [O.priors(rd.quals[i]) for i in range(n)]

    def prior(i, j):  # 1-based row i, column j
        rb, hb = rd.bases[i - 1], hap[j - 1]
        match = rb == hb or rb == ord("N") or hb == ord("N")
        return P[i - 1][0] if match else P[i - 1][1]

    total = 0.0
    # stack of (state, i, j, weight); state in "MXY"
    stack = [("Y", 0, j, 1.0 / m) for j in range(m + 1)]
    while stack:
        s, i, j, w = stack.pop()
        if i == n and s in "MX" and j >= 1:
            total += w
            continue
        if i < n and j < m:  # diagonal into M[i+1][j+1]
            mm, im = T[i][0], T[i][1]
            stack.append(("M", i + 1, j + 1, w * prior(i + 1, j + 1) * (mm if s == "M" else im)))
        if i < n and i >= 1 and s in "MX" and j >= 1:  # down into X[i+1][j]
            stack.append(("X", i + 1, j, w * (T[i][2] if s == "M" else T[i][3])))
        if i >= 1 and s in "MY" and j < m:  # right into Y[i][j+1]
            stack.append(("Y", i, j + 1, w * (T[i - 1][4] if s == "M" else T[i - 1][5])))
    return total


def test_oracle_equals_the_exact_sum_over_alignment_paths():
    rng = Rng(2024)
    for _ in range(40):
        n, m = 1 + rng.below(4), 1 + rng.below(4)
        hap = bytes(b"ACGN"[rng.below(4)] for _ in range(m))
        rd = F.ReadRef(bytes(b"ACGTN"[rng.below(5)] for _ in range(n)),
                       [2 + rng.below(39) for _ in range(n)], [10 + rng.below(36) for _ in range(n)],
                       [10 + rng.below(36) for _ in range(n)], [1 + rng.below(30) for _ in range(n)])
        got = O.log10_likelihood(hap, rd)
        assert abs(got - math.log10(path_sum(hap, rd))) < 1e-12, (hap, rd)


def test_a_planted_index_mistake_is_caught_by_the_enumerator():
    # The enumerator distinguishes the GATK recurrence from one that takes the insertion
    # transitions of the wrong row: the check above is not vacuous.
    rd = F.ReadRef(b"ACG", [30, 20, 10], [20, 40, 30], [45, 15, 30], [10, 20, 5])
    hap = b"AGCG"
    wrong = read(b"ACG")  # other penalties, same bases
    assert abs(math.log10(path_sum(hap, rd)) - O.log10_likelihood(hap, wrong)) > 1e-3


# ---- reference test regions regenerate and pass their own invariants -------------------------------
def test_reference_regions_are_finite_and_below_zero():
    d = Region.generate(11, 70, 120, 24, 300)
    e = O.batch_reference(d)
    assert np.all(np.isfinite(e)) and np.all(e <= 0.0)


# ---- pd_reference.rs:120-184, pd.rs:184-202 -----------------------------------------------------
def test_pd_without_flags_equals_the_plain_pairhmm():
    hap = b"ACGTTGCAAGGCTTAGGCTTACG"
    r = read(b"GTTGCAAGGCTTAG")
    assert abs(O.pd_log10_likelihood(hap, [0] * len(hap), r) - O.log10_likelihood(hap, r)) < 1e-12


def test_pd_snp_alternate_makes_the_alternate_base_match():
    hap = b"ACGTTGCAAGGCTTAGGCTTACG"
    flags = [0] * len(hap)
    flags[8] = F.SNP | F.ALT_T
    alt = read(b"GTTGCATGGCTTAG")
    w = O.pd_log10_likelihood(hap, flags, alt)
    wo = O.pd_log10_likelihood(hap, [0] * len(hap), alt)
    exact = O.pd_log10_likelihood(hap, flags, read(b"GTTGCAAGGCTTAG"))
    assert w > wo + 1.0 and abs(w - exact) < 0.2


def test_pd_deletion_span_lets_the_read_skip_bases():
    hap = b"ACGTTGCAAGGCTTAGGCTTACG"
    flags = [0] * len(hap)
    flags[7] = F.DEL_START
    flags[10] = F.DEL_END
    r = read(b"GTTGCACTTAGGC")
    assert O.pd_log10_likelihood(hap, flags, r) > O.pd_log10_likelihood(hap, [0] * len(hap), r)


def test_pd_base_matching_honours_only_flagged_alternates():
    bm = o.fgklo_base_matches_pd
    assert bm(ord("A"), F.SNP | F.ALT_A)
    assert bm(ord("a"), F.SNP | F.ALT_A | F.ALT_G)
    assert not bm(ord("C"), F.SNP | F.ALT_A)
    assert not bm(ord("A"), F.ALT_A)
    assert not bm(ord("N"), F.SNP | F.ALT_A | F.ALT_C | F.ALT_G | F.ALT_T)


def test_pd_flags_never_lower_a_likelihood_on_a_reference_region():
    d = PdRegion.generate(31, 20, 120, 8, 300)
    reads, haps = d.region_reads_list(), d.region_pd_haps_list()
    w = O.pd_reference_all(reads, haps)
    wo = O.pd_reference_all(reads, [F.PdHaplotype(h.bases, [0] * len(h.bases)) for h in haps])
    assert np.all(w >= wo) and np.sum(np.abs(w - wo) > 1.0) >= 5


def test_pd_deletion_equals_skipping_the_span():
    # A read matching the haplotype with a flagged span removed sees the deletion branch carry
    # the mass from before the span across it: the PD likelihood equals the plain likelihood
    # against the physically shortened haplotype, with the row-0 start density C/m of the full
    # length (a factor (m - s) / m), up to the small mass of paths that do not skip.  Ignoring
    # the branch values (a wrong diagonal or left index) lands 4-10 log10 units lower.
    rng = Rng(3)
    for _ in range(12):
        m = 80
        hap = bytes(rng.base() for _ in range(m))
        a, s = 20 + rng.below(30), 1 + rng.below(8)
        flags = [0] * m
        flags[a] |= F.DEL_START
        flags[a + s - 1] |= F.DEL_END
        shortened = hap[:a] + hap[a + s:]
        r = read(shortened[a - 15:a + 15])
        pd = O.pd_log10_likelihood(hap, flags, r)
        skip = O.log10_likelihood(shortened, r) + math.log10((m - s) / m)
        assert 0.0 <= pd - skip < 1e-4, (a, s, pd, skip)


def test_pd_single_base_deletion_on_one_column_is_max_of_skip_and_keep():
    # A one-base read against a two-column haplotype whose second column is a one-base deletion:
    # flags only change branch bookkeeping; with a one-row read the result equals the plain one.
    r = read(b"A")
    hap = b"CA"
    assert abs(O.pd_log10_likelihood(hap, [0, F.DEL_START | F.DEL_END], r) -
               O.log10_likelihood(hap, r)) < 1e-12


# pd_reference.rs:46: the deletion state is carried from the end of one row into the next.
# Hand-worked case: haplotype GCTT with column 3 flagged DEL_END and column 4 DEL_START, read GCT
# (Q30, ins/del 45, GCP 10).  Row 1 starts NORMAL; column 3 (DEL_END) leaves it AFTER_DEL and
# column 4 (DEL_START) INSIDE_DEL, so rows 2 and 3 start inside a deletion whose branch values
# hold column 0's zeros until column 3 ends it.  The read can therefore not skip column 3: the
# skip at column 4 takes max(branch, M) on the diagonal, and row 2's branch at column 3 is zero.
# The only perfect alignment is G C T on columns 1-3:
#     L = (C/4) * im * pm * mm * pm * mm * pm   (im = 1 - 10^-1, pm = 1 - 10^-3, mm = mm(45, 45))
# plus paths with a mismatch or a gap, each below 3.4e-4 of it (e/3 = 3.3e-4, 10^-4.5 gaps).
# Resetting the state at each row would start rows 2 and 3 NORMAL, their branch at column 3 would
# hold M of column 2, and the skip G C _ T would add a second perfect path: L twice as large.
# Two-row reads cannot show the difference (a branch value of row 2 is read only by row 3).
PD_CARRY_KAT = (b"GCTT", [0, 0, F.DEL_END, F.DEL_START], b"GCT")


def _pd_carry_main_path():
    pm = 1.0 - o.fgklo_error_prob(30)
    mm = o.fgklo_match_to_match(45, 45)
    im = 1.0 - o.fgklo_error_prob(10)
    return math.log10(pm ** 3 * mm ** 2 * im / 4.0)


def test_pd_deletion_state_carries_into_the_next_row():
    hap, flags, rb = PD_CARRY_KAT
    got = O.pd_log10_likelihood(hap, flags, read(rb))
    assert 0.0 <= got - _pd_carry_main_path() < 1e-4, got
    # the same flags without the trailing DEL_START leave rows 2-3 NORMAL: the skip counts
    no_start = O.pd_log10_likelihood(hap, [0, 0, F.DEL_END, 0], read(rb))
    assert abs(no_start - (_pd_carry_main_path() + math.log10(2.0))) < 1e-4, no_start


def test_a_planted_per_row_state_reset_fails_the_carry_kat(tmp_path):
    """The oracle rebuilt with the deletion state reset at every row (a planted mistake) misses
    the known answer by log10 2: the check above is not vacuous."""
    import ctypes as C
    import subprocess

    so = tmp_path / "planted.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11",
                    "-DFGKLO_PLANT_PD_ROW_RESET", "-o", str(so), str(O.ORACLE_DIR / "oracle.c"),
                    "-lm"], check=True)
    planted = C.CDLL(str(so))
    planted.fgklo_pd_log10_likelihood.restype = C.c_double
    planted.fgklo_pd_log10_likelihood.argtypes = o.fgklo_pd_log10_likelihood.argtypes
    hap, flags, rb = PD_CARRY_KAT
    h, f = O._u8(hap), O._u8(flags)
    r = read(rb)
    b, q, i, d, g = (O._u8(x) for x in (r.bases, r.quals, r.ins_gop, r.del_gop, r.gcp))
    bad = planted.fgklo_pd_log10_likelihood(O._p(h), O._p(f), len(h), O._p(b), O._p(q), O._p(i),
                                            O._p(d), O._p(g), len(b))
    assert abs(bad - (_pd_carry_main_path() + math.log10(2.0))) < 1e-4, bad
    assert bad - O.pd_log10_likelihood(hap, flags, r) > 0.3


# ---- gradient oracle ---------------------------------------------------------------------------------
def _cont_inputs(rd):
    e = np.array([o.fgklo_error_prob(q) for q in rd.quals])
    di = np.array([o.fgklo_error_prob(q) for q in rd.ins_gop])
    dd = np.array([o.fgklo_error_prob(q) for q in rd.del_gop])
    return e, di, dd


def _p(a):
    return a.ctypes.data_as(O.f64p)


def test_continuous_forward_is_within_table_error_of_the_quantised_one():
    d = Region.generate(11, 6, 60, 3, 100)
    for rd in d.region_reads_list():
        for hap in d.region_haps_list():
            e, di, dd = _cont_inputs(rd)
            h = np.frombuffer(hap, np.uint8).copy()
            b = np.frombuffer(rd.bases, np.uint8).copy()
            cont = o.fgklo_log10_likelihood_cont(O._p(h), len(h), O._p(b), len(b), _p(e), _p(di),
                                                 _p(dd), 0.1)
            assert abs(cont - O.log10_likelihood(hap, rd)) < 1e-4


def test_gradient_oracle_matches_central_differences():
    d = Region.generate(17, 4, 40, 2, 60)
    rng = Rng(5)
    for rd in d.region_reads_list():
        hap = d.region_haps_list()[rng.below(2)]
        e, di, dd = _cont_inputs(rd)
        h = np.frombuffer(hap, np.uint8).copy()
        b = np.frombuffer(rd.bases, np.uint8).copy()
        n = len(b)
        g = 0.1
        ge, gi, gd = np.zeros(n), np.zeros(n), np.zeros(n)
        gg = np.zeros(1)
        o.fgklo_log10_likelihood_grad(O._p(h), len(h), O._p(b), n, _p(e), _p(di), _p(dd), g,
                                      _p(ge), _p(gi), _p(gd), _p(gg))

        def f(e_, di_, dd_, g_):
            return o.fgklo_log10_likelihood_cont(O._p(h), len(h), O._p(b), n, _p(e_), _p(di_),
                                                 _p(dd_), g_)

        for idx in {0, n // 2, n - 1}:
            for arr, grad in ((e, ge), (di, gi), (dd, gd)):
                # Relative step: the probabilities span 1e-5..1e-1, and a central difference's
                # rounding noise is ~1e-16 |f| / step.
                step = 1e-3 * arr[idx]
                up, dn = arr.copy(), arr.copy()
                up[idx] += step
                dn[idx] -= step
                args_up = [up if a is arr else a for a in (e, di, dd)]
                args_dn = [dn if a is arr else a for a in (e, di, dd)]
                fd = (f(*args_up, g) - f(*args_dn, g)) / (2 * step)
                assert abs(fd - grad[idx]) <= 1e-5 * max(1.0, abs(fd)), (idx, fd, grad[idx])
        fd = (f(e, di, dd, g + 1e-6) - f(e, di, dd, g - 1e-6)) / 2e-6
        assert abs(fd - gg[0]) <= 1e-5 * max(1.0, abs(fd))
