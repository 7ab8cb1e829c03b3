"""Host planning of prefix-sharing snapshots (no device).

A haplotype after the first sorted one resumes at the column it shares with its predecessor,
from the state the last earlier sweep to pass that column left there (kernel.rs:236-259,
556-562, 641).  The plan records only those states, keeps as many of them in one LDS record per
wave as an earliest-last-reader schedule fits, and writes the rest to HBM slots.  These tests
restate that rule independently in Python and compare the counts fgkl_plan_snapshots reports.
"""
from __future__ import annotations

import pytest

import fgkl_amd as F
from fgkl_amd.synthetic import batch_from_regions, workload


def _lcp(a: bytes, b: bytes) -> int:
    n = min(len(a), len(b))
    i = 0
    while i < n and a[i] == b[i]:
        i += 1
    return i


def expected_snapshots(data) -> dict:
    """Per FP32 pass, summed over reads: records written to HBM, records kept in LDS, and
    sweeps that resume from a record.  Plain regions sharing prefixes (reads of at most 511
    bases, haplotypes of at most 2,048, at most 128 distinct shared columns)."""
    out = {"hbm": 0, "lds": 0, "restarts": 0}
    for g in range(data.n_regions):
        h0, h1 = int(data.region_haps[g]), int(data.region_haps[g + 1])
        nreads = int(data.region_reads[g + 1] - data.region_reads[g])
        haps = sorted(data.hap(h) for h in range(h0, h1))  # bytes order: memcmp, then length
        if len(haps) < 2:
            continue
        lcp = [0] + [_lcp(haps[k - 1], haps[k]) for k in range(1, len(haps))]
        dup = [False] + [len(haps[k]) == len(haps[k - 1]) == lcp[k] for k in range(1, len(haps))]
        sweeps = [k for k in range(len(haps)) if not dup[k]]
        assert len({lcp[k] for k in sweeps[1:] if lcp[k] > 0}) <= 128
        last = {}  # (writer, column) -> last reader
        for i, r in enumerate(sweeps[1:], 1):
            c = lcp[r]
            if c == 0:
                continue
            # the last earlier sweep that starts before c and reaches it
            w = [j for j in sweeps[:i] if (lcp[j] if j else 0) < c <= len(haps[j])][-1]
            last[(w, c)] = r
            out["restarts"] += nreads
        free_from = -1
        for (w, c), r in sorted(last.items(), key=lambda kv: (kv[1], kv[0][0])):
            if w >= free_from:
                out["lds"] += nreads
                free_from = r
            else:
                out["hbm"] += nreads
    return out


def test_a_stale_record_is_not_written():
    """AAAA, ACAA, ACAC: sweep 0 passes column 3, which ACAC resumes from, but ACAA passes it
    later with the prefix ACAC shares; only ACAA records it.  Both records live one sweep each
    and fit the one LDS record in turn."""
    rd = F.ReadRef(b"ACGT", [30] * 4, [45] * 4, [45] * 4, [10] * 4)
    d = batch_from_regions([([rd, rd, rd], [b"ACAC", b"AAAA", b"ACAA"])])
    assert F.plan_snapshots(d) == {"hbm": 0, "lds": 6, "restarts": 6}


def test_records_live_in_lds_by_earliest_last_reader():
    """AAAA, AAAC, AACC, ACCC resume at columns 3, 2 and 1, all from sweep 0: three records live
    at once, the one read first takes the LDS record."""
    rd = F.ReadRef(b"ACGT", [30] * 4, [45] * 4, [45] * 4, [10] * 4)
    d = batch_from_regions([([rd], [b"AACC", b"ACCC", b"AAAA", b"AAAC"])])
    assert F.plan_snapshots(d) == {"hbm": 2, "lds": 1, "restarts": 3}


def test_duplicates_and_a_lone_haplotype_record_nothing():
    rd = F.ReadRef(b"ACGT", [30] * 4, [45] * 4, [45] * 4, [10] * 4)
    d = batch_from_regions([([rd], [b"ACGTAC", b"ACGTAC"]), ([rd], [b"ACGT"])])
    assert F.plan_snapshots(d) == {"hbm": 0, "lds": 0, "restarts": 0}


@pytest.mark.parametrize("wl,seed,regions", [(1, 7, 48), (1, 31, 24), (4, 5, 32), (0, 7, 1)])
def test_snapshot_plan_matches_the_last_writer_rule(wl, seed, regions):
    data = workload(wl, seed=seed, n_regions=regions)
    got = F.plan_snapshots(data)
    assert got == expected_snapshots(data)
    assert got["restarts"] > 0 and got["lds"] > 0
