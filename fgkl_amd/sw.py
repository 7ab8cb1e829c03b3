"""Smith-Waterman on the device: the Python face of fgkl's aligner (crates/smithwaterman/src/lib.rs)
and GATK's SWAlignerNativeBinding (FgklSmithWaterman.java), over the C-ABI of include/fgkl.h.

``Aligner.align`` mirrors ``Aligner::align`` (lib.rs:330-360): one pair, GATK's result (CIGAR and
reference offset), ``EmptySequence`` for an empty input.  ``SwBatch`` aligns many pairs per launch
with the pairs resident on the device.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import InvalidArgument, _check, _ptr, _u8
from ._lib import SwStats, lib

_OPS = "MIDNSHP=X"


class OverhangStrategy(enum.IntEnum):
    """GATK's SWOverhangStrategy in FgklSmithWaterman.strategyCode's numbering (java:49-57)."""
    SOFTCLIP = 0
    INDEL = 1
    LEADING_INDEL = 2
    IGNORE = 3


@dataclass(frozen=True)
class SwParameters:
    """GATK's SWParameters (lib.rs:41-64): a positive match value and negative penalties."""
    match_value: int
    mismatch_penalty: int
    gap_open_penalty: int
    gap_extend_penalty: int

    def as_array(self) -> np.ndarray:
        return np.array([self.match_value, self.mismatch_penalty, self.gap_open_penalty,
                         self.gap_extend_penalty], dtype=np.int32)


# GATK's SmithWatermanAlignmentConstants
NEW_SW_PARAMETERS = SwParameters(200, -150, -260, -11)
ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS = SwParameters(10, -15, -30, -5)
STANDARD_NGS = SwParameters(25, -50, -110, -6)
ORIGINAL_DEFAULT = SwParameters(3, -1, -4, -3)


class EmptySequence(InvalidArgument):
    """Error::EmptySequence (lib.rs:145-153): the reference or the alternate has no bases."""


@dataclass
class Alignment:
    """A CIGAR as BAM encodes it, (length << 4) | op, and the offset on the reference."""
    cigar: list[int]
    offset: int

    def cigar_string(self) -> str:
        return "".join(f"{e >> 4}{_OPS[e & 15]}" for e in self.cigar)


def parse_cigar(text: str) -> list[int]:
    out, num = [], ""
    for ch in text:
        if ch.isdigit():
            num += ch
        else:
            out.append((int(num) << 4) | _OPS.index(ch))
            num = ""
    return out


def backend_name() -> str:
    """FgklSmithWaterman.backend() (crates/jni/src/lib.rs:256-268)."""
    return lib().fgkl_sw_backend_name().decode()


def _params(params) -> np.ndarray:
    if isinstance(params, SwParameters):
        return params.as_array()
    return np.ascontiguousarray(params, dtype=np.int32).reshape(4)


def align_native(reference, alternate, params, strategy) -> tuple[str, int]:
    """fgkl_sw_align_native, the C-ABI behind FgklSmithWaterman.alignNative: (CIGAR, offset)."""
    r, a, w = _u8(reference), _u8(alternate), _params(params)
    cap = 2 * (len(r) + len(a)) + 64
    buf = C.create_string_buffer(cap)
    off = C.c_int32()
    rc = lib().fgkl_sw_align_native(_ptr(r, C.c_uint8), len(r), _ptr(a, C.c_uint8), len(a),
                                    int(w[0]), int(w[1]), int(w[2]), int(w[3]), int(strategy),
                                    C.byref(off), buf, cap)
    _check(rc)
    return buf.value.decode(), off.value


class Aligner:
    """fgkl's Aligner (lib.rs:280-360) on the device, one pair per call."""

    def backend(self) -> str:
        return backend_name()

    def align(self, reference, alternate, params, strategy) -> Alignment:
        if len(reference) == 0 or len(alternate) == 0:
            raise EmptySequence("sequences must be non-empty")
        text, offset = align_native(reference, alternate, params, strategy)
        return Alignment(parse_cigar(text), offset)


class SwBatch:
    """Many pairs resident on the device, aligned per launch (fgkl_sw_batch_*)."""

    def __init__(self, pairs: Sequence[tuple[bytes, bytes]] | None = None, *, ref=None,
                 ref_offsets=None, alt=None, alt_offsets=None):
        if pairs is not None:
            refs = [_u8(r) for r, _ in pairs]
            alts = [_u8(a) for _, a in pairs]
            ref_offsets = np.zeros(len(pairs) + 1, np.int64)
            alt_offsets = np.zeros(len(pairs) + 1, np.int64)
            ref_offsets[1:] = np.cumsum([len(x) for x in refs])
            alt_offsets[1:] = np.cumsum([len(x) for x in alts])
            ref = np.concatenate(refs) if refs else np.zeros(0, np.uint8)
            alt = np.concatenate(alts) if alts else np.zeros(0, np.uint8)
        self.ref = _u8(ref)
        self.alt = _u8(alt)
        self.ref_offsets = np.ascontiguousarray(ref_offsets, np.int64)
        self.alt_offsets = np.ascontiguousarray(alt_offsets, np.int64)
        self.n = len(self.ref_offsets) - 1
        h = C.c_void_p()
        rc = lib().fgkl_sw_batch_create(C.byref(h), self.n, _ptr(self.ref_offsets, C.c_int64),
                                        _ptr(self.ref, C.c_uint8),
                                        _ptr(self.alt_offsets, C.c_int64),
                                        _ptr(self.alt, C.c_uint8))
        _check(rc)
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().fgkl_sw_batch_free(h)
            self._h = None

    @property
    def cells(self) -> int:
        return int(lib().fgkl_sw_batch_cells(self._h))

    def run(self, params, strategy) -> SwStats:
        w = _params(params)
        st = SwStats()
        _check(lib().fgkl_sw_batch_run(self._h, _ptr(w, C.c_int32), int(strategy), C.byref(st)))
        return st

    def results_arrays(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(offsets, cigar_starts, cigar elements) of the last run."""
        cap = int(lib().fgkl_sw_batch_cigar_bound(self._h))
        offs = np.zeros(self.n, np.int32)
        starts = np.zeros(self.n + 1, np.int64)
        cig = np.zeros(max(cap, 1), np.uint32)
        _check(lib().fgkl_sw_batch_results(self._h, _ptr(offs, C.c_int32),
                                           _ptr(starts, C.c_int64), _ptr(cig, C.c_uint32), cap))
        return offs, starts, cig[: starts[-1]]

    def results(self) -> list[Alignment]:
        offs, starts, cig = self.results_arrays()
        return [Alignment([int(e) for e in cig[starts[p]:starts[p + 1]]], int(offs[p]))
                for p in range(self.n)]

    def align(self, params, strategy) -> list[Alignment]:
        self.run(params, strategy)
        return self.results()


def bench_pairs(n_pairs: int = 2000, ref_len: int = 500, alt_len: int = 500, seed: int = 5):
    """The aligner benchmark's synthetic pairs (crates/smithwaterman/src/bin/bench.rs:41-53, the same
    SplitMix64 stream): a random reference of ref_len bases and its first alt_len bases with three
    substitutions and one deletion.  Returns (ref_offsets, ref, alt_offsets, alt) arrays."""
    from .synthetic import splitmix_array

    k = min(alt_len, ref_len)
    per = ref_len + 7
    u = splitmix_array(seed, n_pairs * per).reshape(n_pairs, per)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    ref = acgt[(u[:, :ref_len] % np.uint64(4)).astype(np.int64)]
    alt = ref[:, :k].copy()
    rows = np.arange(n_pairs)
    for e in range(3):
        pos = (u[:, ref_len + 2 * e] % np.uint64(k)).astype(np.int64)
        alt[rows, pos] = acgt[(u[:, ref_len + 2 * e + 1] % np.uint64(4)).astype(np.int64)]
    drop = (u[:, ref_len + 6] % np.uint64(k)).astype(np.int64)
    keep = np.ones_like(alt, dtype=bool)
    keep[rows, drop] = False
    alt = alt[keep].reshape(n_pairs, k - 1)
    ro = np.arange(n_pairs + 1, dtype=np.int64) * ref_len
    ao = np.arange(n_pairs + 1, dtype=np.int64) * (k - 1)
    return ro, ref.reshape(-1), ao, alt.reshape(-1)
