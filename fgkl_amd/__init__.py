"""fgkl_amd: PairHMM and partially determined PairHMM likelihoods on one AMD MI355X (gfx950).

The Python surface mirrors the reference's ``fgkl_pairhmm`` crate API
(crates/pairhmm/src/lib.rs:30-344, pd.rs:13-110): ``Precision``, ``Config``, ``ReadRef``,
``PairHmm.compute_log10_likelihoods``, ``PdHaplotype``, ``PdPairHmm``, ``fallback_pairs`` and
the ``Error`` variants with the same messages.  All computation runs in the HIP engine
(``libfgkl.so``) through its C-ABI; there is no CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import lib, last_error

__all__ = [
    "Precision", "Config", "ReadRef", "PdHaplotype", "PairHmm", "PdPairHmm",
    "Error", "EmptyRead", "MismatchedReadArrays", "EmptyHaplotype",
    "MismatchedHaplotypeArrays", "OutputLength", "BackendUnavailable", "InvalidArgument",
    "NativeError", "SNP", "DEL_START", "DEL_END", "ALT_A", "ALT_C", "ALT_G", "ALT_T",
    "compute_native", "pd_compute_native", "aggregate", "aggregate_stats", "Batch",
    "backend_name", "device_count",
]

# pd.rs:15-27
SNP, DEL_START, DEL_END, ALT_A, ALT_C, ALT_G, ALT_T = 1, 2, 4, 8, 16, 32, 64

FGKL_OK, FGKL_INVALID_ARGUMENT, FGKL_INTERNAL_ERROR = 0, 1, 2
MODE_FP32_RESCUE, MODE_FP64, MODE_FP64_PER_REGION, MODE_FP32_NO_RESCUE = 0, 1, 2, 3


class Precision(enum.Enum):
    """lib.rs:72-105"""
    FLOAT = "float"
    DOUBLE = "double"


@dataclass
class Config:
    """lib.rs:213-238: the backend is always the HIP engine; suffix sharing is not offered."""
    precision: Precision = Precision.FLOAT
    double_fallback: bool = True


# ---- errors (lib.rs:305-344) -----------------------------------------------------------------
class Error(Exception):
    """A refused computation; ``str()`` is the reference's Display text."""


class EmptyRead(Error):
    def __init__(self, index: int):
        self.index = index
        super().__init__(f"read {index} has no bases")


class MismatchedReadArrays(Error):
    def __init__(self, index: int):
        self.index = index
        super().__init__(f"read {index}: bases, qualities and penalties differ in length")


class EmptyHaplotype(Error):
    def __init__(self, index: int):
        self.index = index
        super().__init__(f"haplotype {index} has no bases")


class MismatchedHaplotypeArrays(Error):
    def __init__(self, index: int):
        self.index = index
        super().__init__(f"haplotype {index}: bases and flags differ in length")


class OutputLength(Error):
    def __init__(self, expected: int, actual: int):
        self.expected, self.actual = expected, actual
        super().__init__(f"output has {actual} elements but reads x haplotypes is {expected}")


class BackendUnavailable(Error):
    pass


class InvalidArgument(Error):
    """FGKL_INVALID_ARGUMENT from the C-ABI: the JNI shim throws IllegalArgumentException."""


class NativeError(RuntimeError):
    """FGKL_INTERNAL_ERROR from the C-ABI: the JNI shim throws RuntimeException."""


def _check(rc: int) -> None:
    if rc == FGKL_OK:
        return
    msg = last_error()
    if rc == FGKL_INVALID_ARGUMENT:
        raise InvalidArgument(msg)
    raise NativeError(msg)


def backend_name() -> str:
    return lib().fgkl_backend_name().decode()


def device_count() -> int:
    return int(lib().fgkl_device_count())


# ---- array helpers -----------------------------------------------------------------------------
def _u8(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray)):
        return np.frombuffer(bytes(a), dtype=np.uint8)
    return np.ascontiguousarray(a, dtype=np.uint8)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class ReadRef:
    """lib.rs:30-68: one read's bases and per-base penalties, all of one length."""
    bases: bytes
    quals: Sequence[int]
    ins_gop: Sequence[int]
    del_gop: Sequence[int]
    gcp: Sequence[int]

    def __len__(self) -> int:
        return len(self.bases)

    def validate(self, index: int) -> None:
        n = len(self.bases)
        if n == 0:
            raise EmptyRead(index)
        if any(len(x) != n for x in (self.quals, self.ins_gop, self.del_gop, self.gcp)):
            raise MismatchedReadArrays(index)


@dataclass
class PdHaplotype:
    """pd.rs:29-37"""
    bases: bytes
    flags: Sequence[int]


def _concat_reads(reads: Sequence[ReadRef]):
    offs = np.zeros(len(reads) + 1, dtype=np.int64)
    for i, r in enumerate(reads):
        offs[i + 1] = offs[i] + len(r.bases)
    cat = [np.concatenate([_u8(getattr(r, f)) for r in reads]) if reads else np.zeros(0, np.uint8)
           for f in ("bases", "quals", "ins_gop", "del_gop", "gcp")]
    return offs, cat


def _concat(seqs) -> tuple[np.ndarray, np.ndarray]:
    offs = np.zeros(len(seqs) + 1, dtype=np.int64)
    for i, s in enumerate(seqs):
        offs[i + 1] = offs[i] + len(s)
    data = np.concatenate([_u8(s) for s in seqs]) if seqs else np.zeros(0, np.uint8)
    return offs, data


class _Engine:
    _pd = False

    def __init__(self, config: Config | None = None):
        self.config = config or Config()
        h = lib().fgkl_pairhmm_new(int(self._pd), int(self.config.precision is Precision.DOUBLE),
                                   int(self.config.double_fallback))
        if not h:
            raise BackendUnavailable(last_error())
        self._h = C.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().fgkl_pairhmm_free(h)
            self._h = None

    def backend(self) -> str:
        return backend_name()

    def precision(self) -> Precision:
        return self.config.precision

    def fallback_pairs(self) -> int:
        """lib.rs:263-268"""
        return int(lib().fgkl_pairhmm_fallback_pairs(self._h))

    def _run(self, reads, hap_bases, hap_flags, out):
        expected = len(reads) * len(hap_bases)
        if out is None:
            out = np.zeros(expected, dtype=np.float64)
        elif len(out) != expected:
            raise OutputLength(expected, len(out))
        if expected == 0:
            return out
        roff, (b, q, i, d, g) = _concat_reads(reads)
        hoff, hb = _concat(hap_bases)
        hf = _concat(hap_flags)[1] if hap_flags is not None else None
        res = np.zeros(expected, dtype=np.float64)
        _check(lib().fgkl_pairhmm_compute(
            self._h, len(reads), _ptr(roff, C.c_int64), _ptr(b, C.c_uint8), _ptr(q, C.c_uint8),
            _ptr(i, C.c_uint8), _ptr(d, C.c_uint8), _ptr(g, C.c_uint8), len(hap_bases),
            _ptr(hoff, C.c_int64), _ptr(hb, C.c_uint8),
            _ptr(hf, C.c_uint8) if hf is not None else None, _ptr(res, C.c_double), expected))
        out[:] = res
        return out


class PairHmm(_Engine):
    """lib.rs:240-294: a configured PairHMM on the MI355X engine."""

    def compute_log10_likelihoods(self, reads: Sequence[ReadRef], haplotypes: Sequence[bytes],
                                  out: np.ndarray | None = None) -> np.ndarray:
        """``out[r * len(haplotypes) + h]`` is log10 P(read r | haplotype h) (lib.rs:270-293)."""
        for idx, r in enumerate(reads):
            r.validate(idx)
        for idx, h in enumerate(haplotypes):
            if len(h) == 0:
                raise EmptyHaplotype(idx)
        return self._run(reads, list(haplotypes), None, out)


class PdPairHmm(_Engine):
    """pd.rs:39-96: the partially determined PairHMM on the MI355X engine."""
    _pd = True

    def compute_log10_likelihoods(self, reads: Sequence[ReadRef],
                                  haplotypes: Sequence[PdHaplotype],
                                  out: np.ndarray | None = None) -> np.ndarray:
        for idx, r in enumerate(reads):
            r.validate(idx)
        for idx, h in enumerate(haplotypes):
            if len(h.bases) == 0:
                raise EmptyHaplotype(idx)
            if len(h.flags) != len(h.bases):
                raise MismatchedHaplotypeArrays(idx)
        return self._run(reads, [h.bases for h in haplotypes], [h.flags for h in haplotypes], out)


# ---- the JNI-shaped entry points (crates/jni/src/lib.rs:159-254) ------------------------------
def _native(pd, bases, quals, ins, dels, gcp, read_offsets, hap_bases, hap_flags, hap_offsets,
            double_precision, likelihoods, noalloc=False):
    arrs = [_u8(x) for x in (bases, quals, ins, dels, gcp)]
    ro = np.ascontiguousarray(read_offsets, dtype=np.int32)
    ho = np.ascontiguousarray(hap_offsets, dtype=np.int32)
    hb = _u8(hap_bases)
    args = []
    for a in arrs:
        args += [_ptr(a, C.c_uint8), len(a)]
    args += [_ptr(ro, C.c_int32), len(ro), _ptr(hb, C.c_uint8), len(hb)]
    if pd:
        hf = _u8(hap_flags)
        args += [_ptr(hf, C.c_uint8), len(hf)]
    args += [_ptr(ho, C.c_int32), len(ho), int(bool(double_precision)),
             likelihoods.ctypes.data_as(C.POINTER(C.c_double)), len(likelihoods)]
    if noalloc:
        fn = lib().fgkl_pairhmm_compute_native_noalloc
    else:
        fn = lib().fgkl_pdhmm_compute_native if pd else lib().fgkl_pairhmm_compute_native
    _check(fn(*args))


def compute_native(bases, quals, ins, dels, gcp, read_offsets, hap_bases, hap_offsets,
                   double_precision, likelihoods: np.ndarray) -> None:
    """FgklPairHmm.computeNative (crates/jni/src/lib.rs:159-195)."""
    _native(False, bases, quals, ins, dels, gcp, read_offsets, hap_bases, None, hap_offsets,
            double_precision, likelihoods)


def compute_native_noalloc(bases, quals, ins, dels, gcp, read_offsets, hap_bases, hap_offsets,
                           double_precision, likelihoods: np.ndarray) -> None:
    """compute_native without allocating device memory (include/fgkl.h,
    fgkl_pairhmm_compute_native_noalloc)."""
    _native(False, bases, quals, ins, dels, gcp, read_offsets, hap_bases, None, hap_offsets,
            double_precision, likelihoods, noalloc=True)


def pd_compute_native(bases, quals, ins, dels, gcp, read_offsets, hap_bases, hap_flags,
                      hap_offsets, double_precision, likelihoods: np.ndarray) -> None:
    """FgklPdHmm.computeNative (crates/jni/src/lib.rs:207-254)."""
    _native(True, bases, quals, ins, dels, gcp, read_offsets, hap_bases, hap_flags, hap_offsets,
            double_precision, likelihoods)


def aggregate(max_regions: int, max_wait_us: int = 1000) -> None:
    """Gathers concurrent compute_native / pd_compute_native calls (up to `max_regions` arriving
    within `max_wait_us` of the first) into one batched run; 0 turns it off (include/fgkl.h,
    fgkl_aggregate_config)."""
    _check(lib().fgkl_aggregate_config(int(max_regions), int(max_wait_us)))


def aggregate_stats() -> tuple[int, int]:
    """(batches run, regions they held) since the library loaded."""
    b, r = C.c_int64(), C.c_int64()
    lib().fgkl_aggregate_stats(C.byref(b), C.byref(r))
    return b.value, r.value


def release_idle() -> int:
    """Frees the idle reusable batches of the per-call entry points and of aggregation; returns
    how many (include/fgkl.h, fgkl_release_idle)."""
    n = lib().fgkl_release_idle()
    if n < 0:
        raise NativeError(last_error())
    return int(n)


VALU_FP32_FMA, VALU_FP64_FMA, VALU_U32 = 0, 1, 2


def measure_valu_rate(kind: int) -> tuple[float, float]:
    """(lane operations per second, best launch ms) of a full grid issuing one kind of vector ALU
    instruction (include/fgkl.h, fgkl_measure_valu_rate)."""
    rate, ms = C.c_double(), C.c_double()
    _check(lib().fgkl_measure_valu_rate(int(kind), C.byref(rate), C.byref(ms)))
    return rate.value, ms.value


POOL_PAIRHMM, POOL_PDHMM, POOL_SMITH_WATERMAN = 0, 1, 2


def pool_stats(kind: int = POOL_PAIRHMM) -> tuple[int, int, int]:
    """(idle batches, checkouts not yet returned, checkouts served by an idle batch) of one
    per-call pool."""
    idle, out, reused = C.c_int64(), C.c_int64(), C.c_int64()
    _check(lib().fgkl_pool_stats(int(kind), C.byref(idle), C.byref(out), C.byref(reused)))
    return idle.value, out.value, reused.value


def selftest_pool(threads: int, rounds: int) -> dict:
    """Runs the pool on the CPU with `threads` threads of `rounds` checkouts each (all of which
    exit before it reports) and returns what it saw."""
    out = (C.c_int64 * 5)()
    _check(lib().fgkl_selftest_pool(int(threads), int(rounds), out))
    keys = ("created", "destroyed_while_threads_ran", "idle", "outstanding", "reused")
    return dict(zip(keys, (int(v) for v in out)))


def plan_snapshots(data: "BatchData", pd: bool | None = None) -> dict:
    """Prefix-sharing snapshot records of a batch's plan per FP32 pass, summed over reads, planned
    on the host alone (fgkl_plan_snapshots): written to HBM slots, kept in a wave's LDS record,
    and sweeps that resume from a record."""
    d = data
    out = (C.c_int64 * 3)()
    _check(lib().fgkl_plan_snapshots(
        int(d.pd if pd is None else pd), d.n_regions, _ptr(d.region_reads, C.c_int64),
        _ptr(d.region_haps, C.c_int64), _ptr(d.read_offsets, C.c_int64),
        _ptr(d.hap_offsets, C.c_int64), _ptr(d.hap_bases, C.c_uint8),
        _ptr(d.hap_flags, C.c_uint8), out))
    return dict(zip(("hbm", "lds", "restarts"), (int(x) for x in out)))


# ---- batched engine --------------------------------------------------------------------------
class Batch:
    """Many regions resident in device memory (include/fgkl.h, fgkl_batch_*)."""

    def __init__(self, data: "BatchData", pd: bool | None = None):
        self.data = data
        self.pd = data.pd if pd is None else pd
        h = C.c_void_p()
        d = data
        _check(lib().fgkl_batch_create(
            C.byref(h), int(self.pd), d.n_regions, _ptr(d.region_reads, C.c_int64),
            _ptr(d.region_haps, C.c_int64), _ptr(d.read_offsets, C.c_int64),
            _ptr(d.bases, C.c_uint8), _ptr(d.quals, C.c_uint8), _ptr(d.ins, C.c_uint8),
            _ptr(d.dels, C.c_uint8), _ptr(d.gcp, C.c_uint8), _ptr(d.hap_offsets, C.c_int64),
            _ptr(d.hap_bases, C.c_uint8), _ptr(d.hap_flags, C.c_uint8)))
        self._h = h

    def run(self, mode: int = MODE_FP32_RESCUE, share: bool = True, graph: bool = False):
        st = _lib.RunStats()
        _check(lib().fgkl_batch_run(self._h, mode, int(share), int(graph), C.byref(st)))
        return st

    def run_pairs(self, mode: int = MODE_FP32_RESCUE):
        """Every (read, haplotype) pair in its own wave (include/fgkl.h, fgkl_batch_run_pairs)."""
        st = _lib.RunStats()
        _check(lib().fgkl_batch_run_pairs(self._h, mode, C.byref(st)))
        return st

    def work(self) -> dict:
        """DP cells of the plan: nominal, prefix sharing alone, and forward / backward with
        suffix sharing (fgkl_batch_work)."""
        w = (C.c_int64 * 4)()
        _check(lib().fgkl_batch_work(self._h, w))
        return dict(zip(("nominal", "prefix", "forward", "backward"), (int(x) for x in w)))

    def set_uniform_gaps(self, enable: bool):
        """Runs the items whose reads have uniform gap qualities with the uniform-gap kernel build
        (the default) or every item with the general build (fgkl_batch_set_uniform_gaps); the
        bits are the same."""
        _check(lib().fgkl_batch_set_uniform_gaps(self._h, int(bool(enable))))

    def set_lds_records(self, enable: bool):
        """Lets prefix-sharing launches keep a snapshot record in each wave's LDS (the default)
        or sends every record to HBM (fgkl_batch_set_lds_records); the bits are the same."""
        _check(lib().fgkl_batch_set_lds_records(self._h, int(bool(enable))))

    def snapshots(self) -> dict:
        """Snapshot records of the plan per FP32 pass (HBM, LDS, restarts) and the sweep launches
        so far that kept a record in LDS (fgkl_batch_snapshots)."""
        out = (C.c_int64 * 4)()
        _check(lib().fgkl_batch_snapshots(self._h, out))
        return dict(zip(("hbm", "lds", "restarts", "lds_launches"), (int(x) for x in out)))

    def tune(self, share: bool = True):
        """Times each FP32 bucket's launch configurations and keeps the fastest
        (fgkl_batch_tune); returns the report."""
        rep = _lib.TuneReport()
        _check(lib().fgkl_batch_tune(self._h, int(share), C.byref(rep)))
        return rep

    def set_launch(self, bucket_rows: int, rows: int = 0, block: int = 256, boost: bool = False):
        _check(lib().fgkl_batch_set_launch(self._h, bucket_rows, rows, block, int(boost)))

    def get_launch(self, bucket_rows: int) -> tuple[int, int, bool]:
        """(rows per lane, threads per workgroup, occupancy-raised build) of bucket R."""
        r, b, o = C.c_int(), C.c_int(), C.c_int()
        _check(lib().fgkl_batch_get_launch(self._h, bucket_rows, C.byref(r), C.byref(b),
                                           C.byref(o)))
        return r.value, b.value, bool(o.value)

    def results(self) -> np.ndarray:
        n = int(lib().fgkl_batch_pairs(self._h))
        out = np.empty(n, dtype=np.float64)
        _check(lib().fgkl_batch_results(self._h, _ptr(out, C.c_double), n))
        return out

    def close(self):
        if self._h:
            lib().fgkl_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from .synthetic import BatchData  # noqa: E402  (re-export)
