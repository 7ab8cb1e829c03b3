"""The JNI entry points of libfgkl.so through a mock JNIEnv (tests/jni/mock_jni.c): the
contract of crates/jni/src/lib.rs:149-313, FgklPairHmmTest.java:124-190 and
FgklSmithWatermanTest.java:73-122 without a JVM."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

import fgkl_amd as F
from fgkl_amd import _lib
from fgkl_amd.synthetic import PdRegion, Region
from _oracle import pd_reference_all, reference_all

HERE = Path(__file__).resolve().parent
PAIR = "Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_computeNative"
PD = "Java_com_fulcrumgenomics_fgkl_pdhmm_FgklPdHmm_computeNative"
SW = "Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_alignNative"
SW_BACKEND = "Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_backendNative"


@pytest.fixture(scope="module")
def mock(tmp_path_factory):
    out = tmp_path_factory.mktemp("jni") / "libmockjni.so"
    subprocess.run(["gcc", "-O1", "-shared", "-fPIC", "-o", str(out), str(HERE / "jni" / "mock_jni.c")],
                   check=True)
    m = C.CDLL(str(out))
    m.mock_compute.restype = C.c_int
    m.mock_sw_align.restype = C.c_char_p
    m.mock_backend.restype = C.c_char_p
    m.mock_exception_class.restype = C.c_char_p
    m.mock_exception_message.restype = C.c_char_p
    return m


def entry(name):
    return C.cast(getattr(_lib.lib(), name), C.c_void_p)


def p8(a):
    return np.ascontiguousarray(a, dtype=np.int8).ctypes.data_as(C.POINTER(C.c_int8))


def call(m, data, pd=False, dp=False, out=None, nulls=0, **over):
    arrs = {"bases": data.bases, "quals": data.quals, "ins": data.ins, "dels": data.dels,
            "gcp": data.gcp, "ro": data.read_offsets.astype(np.int32), "hb": data.hap_bases,
            "hf": data.hap_flags, "ho": data.hap_offsets.astype(np.int32)}
    arrs.update(over)
    keep = {k: (np.ascontiguousarray(v, dtype=np.int32) if k in ("ro", "ho")
                else np.ascontiguousarray(v).view(np.int8)) for k, v in arrs.items()}
    if out is None:
        out = np.full(data.n_pairs(), -7.0)
    i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    rc = m.mock_compute(entry(PD if pd else PAIR), int(pd),
                        p8(keep["bases"]), len(keep["bases"]), p8(keep["quals"]), len(keep["quals"]),
                        p8(keep["ins"]), len(keep["ins"]), p8(keep["dels"]), len(keep["dels"]),
                        p8(keep["gcp"]), len(keep["gcp"]), i32(keep["ro"]), len(keep["ro"]),
                        p8(keep["hb"]), len(keep["hb"]), p8(keep["hf"]), len(keep["hf"]),
                        i32(keep["ho"]), len(keep["ho"]), int(dp),
                        out.ctypes.data_as(C.POINTER(C.c_double)), len(out), nulls)
    exc = (m.mock_exception_class().decode(), m.mock_exception_message().decode()) if rc else None
    return out, exc


def test_backend_names_the_engine(mock):
    for name in ("Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_backendNative",
                 "Java_com_fulcrumgenomics_fgkl_pdhmm_FgklPdHmm_backendNative"):
        assert mock.mock_backend(entry(name)).decode() == "hip-gfx950"


def test_malformed_inputs_throw_illegal_argument(mock):  # FgklPairHmmTest.java:172-190
    d = Region.generate(5, 5, 50, 4, 80)
    _, exc = call(mock, d, out=np.zeros(19))
    assert exc == ("java/lang/IllegalArgumentException",
                   "output has 19 elements but reads x haplotypes is 20")
    _, exc = call(mock, d, quals=d.quals[:-1])
    assert exc[0] == "java/lang/IllegalArgumentException" and "differ in total length" in exc[1]
    ro = d.read_offsets.astype(np.int32).copy()
    ro[2] = ro[1]
    _, exc = call(mock, d, ro=ro)
    assert exc == ("java/lang/IllegalArgumentException", "read 1 has no bases")
    ho = d.hap_offsets.astype(np.int32).copy()
    ho[-1] -= 1
    _, exc = call(mock, d, ho=ho)
    assert exc == ("java/lang/IllegalArgumentException", "offset table must end at the array length")
    pdd = PdRegion.generate(5, 5, 50, 4, 80)
    _, exc = call(mock, pdd, pd=True, hf=pdd.hap_flags[:-1])
    assert exc == ("java/lang/IllegalArgumentException",
                   "haplotype bases and PD flags differ in total length")


def test_null_arrays_are_a_jni_error(mock):
    d = Region.generate(4, 5, 50, 4, 80)
    for bit in (0, 5, 6, 8, 9):
        _, exc = call(mock, d, nulls=1 << bit)
        assert exc == ("java/lang/RuntimeException", "JNI error: null array"), bit


def test_empty_inputs_write_nothing_and_throw_nothing(mock):  # FgklPairHmmTest.java:163-170
    d = Region.generate(3, 5, 50, 4, 80)
    empty = np.zeros(0, np.uint8)
    _, exc = call(mock, d, bases=empty, quals=empty, ins=empty, dels=empty, gcp=empty,
                  ro=np.zeros(1, np.int32), out=np.zeros(0))
    assert exc is None


@pytest.mark.skipif(F.device_count() > 0, reason="checks the no-GPU host")
def test_without_a_gpu_the_call_throws_runtime_exception(mock):
    d = Region.generate(4, 5, 50, 4, 80)
    _, exc = call(mock, d)
    assert exc[0] == "java/lang/RuntimeException" and "no HIP device" in exc[1]


@pytest.mark.gpu
def test_jni_entries_match_the_oracle(gpu, mock):  # FgklPairHmmTest.java:129-140
    d = Region.generate(1, 60, 120, 20, 300)
    expected = reference_all(d.region_reads_list(), d.region_haps_list())
    out, exc = call(mock, d, dp=True)
    assert exc is None
    assert np.all(np.abs(out - expected) <= 1e-12 * np.maximum(1.0, np.abs(expected)))
    out, exc = call(mock, d, dp=False)
    assert exc is None and np.max(np.abs(out - expected)) <= 1e-4
    pdd = PdRegion.generate(31, 40, 120, 12, 300)
    expected = pd_reference_all(pdd.region_reads_list(), pdd.region_pd_haps_list())
    out, exc = call(mock, pdd, pd=True, dp=True)
    assert exc is None
    assert np.all(np.abs(out - expected) <= 1e-12 * np.maximum(1.0, np.abs(expected)))


def sw_call(m, ref: bytes, alt: bytes, params=(200, -150, -260, -11), strategy=1, n_offset=1,
            nulls=0):
    r = np.frombuffer(ref, np.int8).copy() if ref else np.zeros(0, np.int8)
    a = np.frombuffer(alt, np.int8).copy() if alt else np.zeros(0, np.int8)
    w = (C.c_int32 * 4)(*params)
    off = np.full(max(n_offset, 1), -9, np.int32)
    res = m.mock_sw_align(entry(SW), p8(r), len(r), p8(a), len(a), w, strategy,
                          off.ctypes.data_as(C.POINTER(C.c_int32)), n_offset, nulls)
    exc = (m.mock_exception_class().decode(), m.mock_exception_message().decode())
    return (res.decode() if res is not None else None), int(off[0]), (exc if exc[1] else None)


def test_sw_backend_names_the_engine(mock):
    assert mock.mock_backend(entry(SW_BACKEND)).decode() == "hip-gfx950"


def test_sw_bad_strategy_and_empty_sequences_throw_illegal_argument(mock):
    # crates/jni/src/lib.rs:283-291; FgklSmithWatermanTest.java:113-122
    cig, _, exc = sw_call(mock, b"ACGT", b"ACGT", strategy=4)
    assert cig is None and exc == ("java/lang/IllegalArgumentException",
                                   "unknown overhang strategy code 4")
    cig, _, exc = sw_call(mock, b"", b"ACGT")
    assert cig is None and exc[0] == "java/lang/IllegalArgumentException"
    assert "non-empty" in exc[1]
    cig, _, exc = sw_call(mock, b"ACGT", b"")
    assert cig is None and exc[0] == "java/lang/IllegalArgumentException"


def test_sw_null_arrays_are_a_jni_error(mock):
    for bit in (0, 1):
        cig, _, exc = sw_call(mock, b"ACGT", b"ACGT", nulls=1 << bit)
        assert cig is None and exc == ("java/lang/RuntimeException", "JNI error: null array")


@pytest.mark.skipif(F.device_count() > 0, reason="checks the no-GPU host")
def test_sw_without_a_gpu_throws_runtime_exception(mock):
    cig, _, exc = sw_call(mock, b"ACGTTT", b"ACGT")
    assert cig is None and exc[0] == "java/lang/RuntimeException" and "no HIP device" in exc[1]


@pytest.mark.gpu
def test_sw_jni_entry_matches_the_oracle(gpu, mock):  # FgklSmithWatermanTest.java:82-111
    from _oracle import sw_align
    from _sw_cases import HAP_TO_REF, READ_TO_HAP, STRATEGIES, related_pair
    from fgkl_amd.synthetic import Rng

    assert sw_call(mock, b"AAACGTACGTAAA", b"ACGTACGT", strategy=0)[:2] == ("8M", 2)
    assert sw_call(mock, b"ACGTACGTAGGCCTTAGCA", b"GGGGACGTACGTAGGCCTTAGCAGGGG",
                   params=READ_TO_HAP, strategy=0)[:2] == ("4S19M4S", 0)
    rng = Rng(17)
    for i in range(40):
        ref, alt = (related_pair(rng, 40 + rng.below(300), 20 + rng.below(200)) if i % 2 == 0
                    else related_pair(rng, 5 + rng.below(30), 3 + rng.below(30)))
        for params in (HAP_TO_REF, READ_TO_HAP):
            for code, strategy in enumerate(STRATEGIES):
                cig, off, exc = sw_call(mock, ref, alt, params=params, strategy=code)
                assert exc is None
                assert (cig, off) == sw_align(ref, alt, params, strategy), (i, strategy)
    # the offset array must hold the offset (the JVM throws on an empty one)
    cig, _, exc = sw_call(mock, b"ACGTTT", b"ACGT", n_offset=0)
    assert cig is None and exc[0] == "java/lang/ArrayIndexOutOfBoundsException"
