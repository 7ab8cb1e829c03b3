"""The allocation-free entry point (fgkl_pairhmm_compute_native_noalloc) against the pooled one
it varies (fgkl_pairhmm_compute_native): the same likelihoods, bit for bit, on consecutive calls
of decreasing size (the shared static scratch keeps the previous call's contents), through the
FP64 rescue, in both precisions; its limits are clear errors."""
from __future__ import annotations

import numpy as np
import pytest

import fgkl_amd as F
from fgkl_amd.synthetic import Region, Rng

pytestmark = pytest.mark.gpu


def _arrays(data):
    return (data.bases, data.quals, data.ins, data.dels, data.gcp,
            data.read_offsets.astype(np.int32), data.hap_bases, data.hap_offsets.astype(np.int32))


def _both(data, double):
    b, q, i, d, g, ro, hb, ho = _arrays(data)
    want = np.full(data.n_pairs(), -1.0)
    got = np.full(data.n_pairs(), -2.0)
    F.compute_native(b, q, i, d, g, ro, hb, ho, double, want)
    F.compute_native_noalloc(b, q, i, d, g, ro, hb, ho, double, got)
    return want, got


def _foreign_region(seed, n_reads, n, m):
    """Reads of unrelated bases: every pair falls below the FP32 threshold and is rescued."""
    rng = Rng(seed)
    from fgkl_amd.synthetic import batch_from_regions

    hap = bytes(rng.base() for _ in range(m))
    reads = [F.ReadRef(bytes(rng.base() for _ in range(n)), [40] * n, [45] * n, [45] * n,
                       [10] * n) for _ in range(n_reads)]
    return batch_from_regions([(reads, [hap, hap[: m - 7]])])


def test_noalloc_entry_matches_the_pooled_entry_on_consecutive_calls(gpu):
    larger = Region.generate(61, 300, 250, 12, 600)
    smaller = Region.generate(62, 40, 90, 3, 170)
    rescued = _foreign_region(63, 20, 80, 300)
    for double in (False, True):
        for data in (larger, smaller, rescued, larger, smaller):
            want, got = _both(data, double)
            assert np.all(np.isfinite(want))
            assert np.array_equal(want, got), (double, data.n_pairs())


def test_noalloc_entry_limits_are_clear_errors(gpu):
    n = 600  # beyond the staged sweep
    b = np.full(n, ord("A"), np.uint8)
    out = np.zeros(1)
    with pytest.raises(F.InvalidArgument, match="at most 511 bases"):
        F.compute_native_noalloc(b, b, b, b, b, np.array([0, n], np.int32), b,
                                 np.array([0, n], np.int32), False, out)
    # more bytes than the scratch holds: 14000 reads of 500 bases against one haplotype
    reads = 14000
    rb = np.full(reads * 500, ord("C"), np.uint8)
    ro = (np.arange(reads + 1) * 500).astype(np.int32)
    h = np.full(700, ord("C"), np.uint8)
    with pytest.raises(F.InvalidArgument, match="32 MiB"):
        F.compute_native_noalloc(rb, rb, rb, rb, rb, ro, h, np.array([0, 700], np.int32), False,
                                 np.zeros(reads))
    # the same validation messages as the pooled entry
    with pytest.raises(F.InvalidArgument, match="output has 3 elements but reads x haplotypes is 1"):
        F.compute_native_noalloc(b, b, b, b, b, np.array([0, n], np.int32), b,
                                 np.array([0, n], np.int32), False, np.zeros(3))
    # and a good call afterwards still matches
    want, got = _both(Region.generate(64, 30, 120, 4, 260), False)
    assert np.array_equal(want, got)
