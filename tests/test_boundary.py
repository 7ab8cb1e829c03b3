"""CPU tests of the drop-in boundary: the C-ABI library loads and exports every symbol
include/fgkl.h declares, argument validation reports the reference's error classes and
messages (crates/jni/src/lib.rs:105-363, crates/pairhmm/src/lib.rs:305-344, pd.rs:69-95), empty
inputs are no-ops, and the product fails loudly (never computes on the CPU) without a GPU.
"""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import threading

import numpy as np
import pytest

import fgkl_amd as F
from fgkl_amd import _lib
from fgkl_amd.synthetic import PdRegion, Region, Rng, splitmix, workload

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> set[str]:
    names = set()
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names |= set(re.findall(r"\b((?:fgkl_|Java_)[A-Za-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert "fgkl_pairhmm_compute_native" in names and "fgkl_batch_run" in names
    assert "Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_computeNative" in names
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = names - exported
    assert not missing, missing
    assert {n for n in names if n.startswith("fgkl_")} <= set(_lib.SIGNATURES)


def test_backend_name():
    assert F.backend_name() == "hip-gfx950"


# ---- JNI-shaped entry: ReadBatch::fetch / split / check_output ----------------------------------
def _good():
    d = Region.generate(4, 5, 50, 4, 80)
    return d, d.read_offsets.astype(np.int32), d.hap_offsets.astype(np.int32)


def _call(d, ro, ho, out=None, **over):
    args = dict(bases=d.bases, quals=d.quals, ins=d.ins, dels=d.dels, gcp=d.gcp, read_offsets=ro,
                hap_bases=d.hap_bases, hap_offsets=ho)
    args.update(over)
    if out is None:
        out = np.zeros(d.n_pairs())
    F.compute_native(args["bases"], args["quals"], args["ins"], args["dels"], args["gcp"],
                     args["read_offsets"], args["hap_bases"], args["hap_offsets"], False, out)


def test_read_arrays_of_different_total_length_are_rejected():
    d, ro, ho = _good()
    with pytest.raises(F.InvalidArgument, match="read bases, qualities and penalties differ in total length"):
        _call(d, ro, ho, quals=d.quals[:-1])


@pytest.mark.parametrize("offs,msg", [
    ([1, 4], "offset table must start at zero"),
    ([], "offset table must start at zero"),
    ([0, 3], "offset table must end at the array length"),
    ([0, 5, 2, 6], "offset table must be non-decreasing"),
])
def test_offset_tables_are_validated(offs, msg):  # crates/jni/src/lib.rs:373-394
    d, ro, ho = _good()
    total = 6 if msg.endswith("non-decreasing") else 4
    bases = np.zeros(total, np.uint8) + ord("A")
    with pytest.raises(F.InvalidArgument, match=msg):
        F.compute_native(bases, bases, bases, bases, bases, np.array(offs, np.int32), d.hap_bases,
                         ho, False, np.zeros(4))


def test_haplotype_offsets_are_validated():
    d, ro, ho = _good()
    bad = ho.copy()
    bad[-1] += 1
    with pytest.raises(F.InvalidArgument, match="must end at the array length"):
        _call(d, ro, bad)


def test_output_of_the_wrong_length_is_rejected():  # lib.rs:325-335
    d, ro, ho = _good()
    with pytest.raises(F.InvalidArgument, match="output has 19 elements but reads x haplotypes is 20"):
        _call(d, ro, ho, out=np.zeros(19))


def test_empty_read_and_haplotype_are_rejected():  # lib.rs:54-58, 279-281
    d, ro, ho = _good()
    ro2 = ro.copy()
    ro2[1] = 0  # read 0 empty
    with pytest.raises(F.InvalidArgument, match="read 0 has no bases"):
        _call(d, ro2, ho)
    ho2 = ho.copy()
    ho2[2] = ho2[1]  # haplotype 1 empty
    with pytest.raises(F.InvalidArgument, match="haplotype 1 has no bases"):
        _call(d, ro, ho2)


def test_pd_flags_of_another_length_are_rejected():  # crates/jni/src/lib.rs:232-237
    d = PdRegion.generate(4, 5, 50, 4, 80)
    with pytest.raises(F.InvalidArgument, match="haplotype bases and PD flags differ in total length"):
        F.pd_compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, d.read_offsets.astype(np.int32),
                            d.hap_bases, d.hap_flags[:-1], d.hap_offsets.astype(np.int32), False,
                            np.zeros(d.n_pairs()))


def test_empty_inputs_produce_no_output_without_touching_a_gpu():  # FgklPairHmmTest.java:163-170
    d, ro, ho = _good()
    F.compute_native(np.zeros(0, np.uint8), np.zeros(0, np.uint8), np.zeros(0, np.uint8),
                     np.zeros(0, np.uint8), np.zeros(0, np.uint8), np.array([0], np.int32),
                     d.hap_bases, ho, False, np.zeros(0))
    F.compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, ro, np.zeros(0, np.uint8),
                     np.array([0], np.int32), False, np.zeros(0))


@pytest.mark.skipif(F.device_count() > 0, reason="checks the no-GPU host")
def test_long_reads_and_haplotypes_pass_validation():
    """lib.rs:55-67 validates only emptiness and matching lengths: reads of more than 511 bases
    and haplotypes of thousands are valid (without a GPU the call then fails for want of one)."""
    for n, m in ((513, 600), (2000, 2100), (150, 5000)):
        b = np.full(n, ord("A"), np.uint8)
        h = np.full(m, ord("A"), np.uint8)
        with pytest.raises(F.NativeError, match="no HIP device"):
            F.compute_native(b, b, b, b, b, np.array([0, n], np.int32), h,
                             np.array([0, m], np.int32), False, np.zeros(1))


@pytest.mark.skipif(F.device_count() > 0, reason="checks the no-GPU host")
def test_without_a_gpu_the_engine_fails_loudly():
    d, ro, ho = _good()
    with pytest.raises(F.NativeError, match="no HIP device"):
        _call(d, ro, ho)
    with pytest.raises(F.BackendUnavailable):
        F.PairHmm()


def test_pooled_batches_survive_the_threads_that_used_them():
    """The per-call entry points check reusable batches out of a process-wide pool for one call
    each.  Sixteen threads of 50 checkouts each run and exit: nothing is destroyed while they run
    or at their exit (no HIP call at thread exit), every checkout is returned, at most one object
    per thread was ever created, later checkouts reuse idle ones, and release_idle frees them."""
    st = F.selftest_pool(16, 50)
    assert st["destroyed_while_threads_ran"] == 0
    assert st["outstanding"] == 0
    assert 1 <= st["created"] <= 16
    assert st["idle"] == st["created"]
    assert st["reused"] == 16 * 50 - st["created"]
    # the product's pools start empty on a host that never ran a call
    if F.device_count() == 0:
        assert F.pool_stats(F.POOL_PAIRHMM) == (0, 0, 0)
        assert F.release_idle() == 0
    with pytest.raises(F.InvalidArgument):
        F.pool_stats(7)


def test_aggregated_calls_validate_and_fail_like_single_calls():
    """With aggregation on, each call is still checked on its own thread before it is queued, and a
    batch that cannot run hands every caller the error of its own region."""
    d, ro, ho = _good()
    with pytest.raises(F.InvalidArgument, match="non-negative"):
        F.aggregate(-1)
    F.aggregate(4, 2000)
    try:
        with pytest.raises(F.InvalidArgument, match="output has 19 elements"):
            _call(d, ro, ho, out=np.zeros(19))
        F.compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, ro, np.zeros(0, np.uint8),
                         np.array([0], np.int32), False, np.zeros(0))
        if F.device_count() == 0:
            b0, r0 = F.aggregate_stats()
            errors = []

            def worker():
                try:
                    _call(d, ro, ho)
                except F.NativeError as e:
                    errors.append(str(e))

            threads = [threading.Thread(target=worker) for _ in range(3)]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
            assert len(errors) == 3 and all("no HIP device" in e for e in errors), errors
            b1, r1 = F.aggregate_stats()
            assert r1 - r0 == 3 and 1 <= b1 - b0 <= 3
    finally:
        F.aggregate(0)


@pytest.mark.skipif(F.device_count() > 0, reason="checks the no-GPU host")
def test_the_environment_turns_aggregation_on_for_unchanged_callers():
    """FGKL_AGGREGATE=<regions>[:<wait us>] routes per-region calls through the aggregator in a
    process that never calls fgkl_aggregate_config (a JVM with the unchanged bindings)."""
    import os
    import subprocess
    import sys

    code = (
        "import numpy as np, fgkl_amd as F\n"
        "from fgkl_amd.synthetic import Region\n"
        "d = Region.generate(3, 4, 30, 2, 60)\n"
        "ro, ho = d.read_offsets.astype(np.int32), d.hap_offsets.astype(np.int32)\n"
        "try:\n"
        "    F.compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, ro, d.hap_bases, ho, False,"
        " np.zeros(d.n_pairs()))\n"
        "except F.NativeError as e:\n"
        "    assert 'no HIP device' in str(e), e\n"
        "print(F.aggregate_stats())\n")
    for env, want in (("4:100", "(1, 1)"), ("", "(0, 0)")):
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                             env={**os.environ, "FGKL_AGGREGATE": env}, timeout=120)
        assert out.returncode == 0, out.stderr
        assert out.stdout.strip().splitlines()[-1] == want, (env, out.stdout)


# ---- Python mirror validation (lib.rs:1070-1096, pd.rs:438-464) ----------------------------------
def test_mirror_validates_like_the_reference_before_any_device_work():
    good = F.ReadRef(b"ACGT", [30] * 4, [45] * 4, [45] * 4, [10] * 4)
    short = F.ReadRef(b"ACGT", [30] * 3, [45] * 4, [45] * 4, [10] * 4)
    with pytest.raises(F.MismatchedReadArrays) as e:
        short.validate(0)
    assert str(e.value) == "read 0: bases, qualities and penalties differ in length"
    with pytest.raises(F.EmptyRead):
        F.ReadRef(b"", [], [], [], []).validate(3)
    assert str(F.OutputLength(2, 1)) == "output has 1 elements but reads x haplotypes is 2"
    assert str(F.EmptyHaplotype(0)) == "haplotype 0 has no bases"
    assert str(F.MismatchedHaplotypeArrays(1)) == "haplotype 1: bases and flags differ in length"
    good.validate(0)


# ---- synthetic data ------------------------------------------------------------------------------
def test_python_and_engine_splitmix_agree():
    r = Rng(1234567)
    assert [r.next_u64() for _ in range(20)] == splitmix(1234567, 20)


def test_same_seed_gives_the_same_region():  # synthetic.rs:348-355
    a = Region.generate(5, 20, 50, 4, 80)
    b = Region.generate(5, 20, 50, 4, 80)
    assert a.region_haps_list() == b.region_haps_list()
    assert a.read(0).bases == b.read(0).bases and a.read(19).quals == b.read(19).quals


def test_pd_region_reads_are_valid_and_haplotypes_carry_flags():  # synthetic.rs:340-346
    r = PdRegion.generate(31, 70, 120, 24, 300)
    assert all(len(x.bases) > 0 for x in r.region_reads_list())
    assert any(any(f) for f in (r.flags(h) for h in range(r.n_haps)))
    assert r.cells() > 0


def test_region_shapes_follow_the_generator():
    d = Region.generate(11, 70, 120, 24, 300)
    assert d.n_reads == 70 and d.n_haps == 24
    lens = np.diff(d.read_offsets)
    assert lens.max() <= 120 and lens.min() >= 1
    haps = d.region_haps_list()
    assert len(haps[0]) == 300 and haps[0] != haps[1]


def test_workloads_have_their_stated_shapes():
    w0 = workload(0, seed=3)
    assert (w0.n_reads, w0.n_haps) == (32, 8)
    assert set(np.diff(w0.read_offsets).tolist()) == {151}
    assert 300 <= len(w0.hap(0)) <= 350
    w2 = workload(2, seed=3, n_regions=2)
    assert w2.n_reads == 128 and w2.n_haps == 16 and len(w2.hap(0)) == 600
    assert set(np.diff(w2.read_offsets).tolist()) == {250}
    w1 = workload(1, seed=3, n_regions=20)
    for g in range(20):
        r, h = w1.region_shape(g)
        assert 1 <= r <= 400 and 2 <= h <= 16
    lens = np.diff(w1.read_offsets)
    assert lens.min() >= 50 and lens.max() <= 250
    w3 = workload(3, seed=3, n_regions=4)
    frac = np.count_nonzero(w3.hap_flags) / len(w3.hap_flags)
    assert 0.01 < frac < 0.1
    w4 = workload(4, seed=3, n_regions=20)
    for g in range(20):
        r, h = w4.region_shape(g)
        assert 1 <= r <= 200 and 2 <= h <= 8
