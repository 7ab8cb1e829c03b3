"""Replays GATK PairHMM / PD-HMM results dumps through the engine.

HaplotypeCaller writes every read x haplotype likelihood it computes with
``--pair-hmm-results-file`` (PD-HMM in DRAGEN-3.7.8 mode: ``--pdhmm-results-file``); fgkl replays
such dumps to check its kernels against GATK's numbers and to time them on real region shapes
(crates/pairhmm/src/bin/replay.rs, pdreplay.rs).  This module reads them with the library's
parser (fgkl_replay_load: regions rebuilt from the repeating haplotype lists), runs every region
through the batched engine in the requested modes, and reports per mode the device time, GCUPS,
the worst deviation from GATK's values, the pairs recomputed in FP64, and FNV-1a checksums of the
result bits (all pairs; pairs GATK kept above the FP32 threshold; those rounded to 1e-6) so two
runs or platforms can be compared, as replay.rs:267-310 does.  ``--per-region`` instead times one
synchronous call per region through the PairHmm API, as HaplotypeCaller calls.

    python -m fgkl_amd.replay DUMP [--pd] [--max-regions N] [--modes float-nofb,float,double]
                                   [--iters N] [--per-region] [--write-results DIR]
                                   [--parse-only]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

from ._lib import lib
from .synthetic import BatchData, _from_handle

MODES = {"float-nofb": 3, "float": 0, "double": 1}   # FGKL_MODE_* of include/fgkl.h
LOG10_F32_THRESHOLD = -64.0  # GATK recomputes in double below 1e-28 of the 2^120-scaled value


def load(path: str | Path, pd: bool = False, max_regions: int = 0):
    """(BatchData, expected log10 likelihoods, {lines, skipped_lines, dropped_regions})."""
    from . import _check
    h = C.c_void_p()
    _check(lib().fgkl_replay_load(str(path).encode(), int(pd), max_regions, C.byref(h)))
    n = int(lib().fgkl_synth_expected(h, None, 0))
    expected = np.zeros(n, dtype=np.float64)
    lib().fgkl_synth_expected(h, expected.ctypes.data_as(C.POINTER(C.c_double)), n)
    st = [C.c_int64() for _ in range(3)]
    lib().fgkl_replay_stats(h, *[C.byref(x) for x in st])
    stats = dict(zip(("lines", "skipped_lines", "dropped_regions"), (x.value for x in st)))
    data = _from_handle(h, pd)  # frees the handle
    return data, expected, stats


def write(path: str | Path, data: BatchData, values) -> None:
    """Writes `data` with one value per pair in the dump format (PD format when data.pd)."""
    from . import _check
    v = np.ascontiguousarray(values, dtype=np.float64)
    assert len(v) == data.n_pairs()
    p = lambda a, t: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))
    _check(lib().fgkl_replay_write(
        str(path).encode(), int(data.pd), data.n_regions, p(data.region_reads, C.c_int64),
        p(data.region_haps, C.c_int64), p(data.read_offsets, C.c_int64), p(data.bases, C.c_uint8),
        p(data.quals, C.c_uint8), p(data.ins, C.c_uint8), p(data.dels, C.c_uint8),
        p(data.gcp, C.c_uint8), p(data.hap_offsets, C.c_int64), p(data.hap_bases, C.c_uint8),
        p(data.hap_flags, C.c_uint8), v.ctypes.data_as(C.POINTER(C.c_double)), len(v)))


def fnv1a(values) -> int:
    """FNV-1a over 64-bit words (replay.rs:267-276)."""
    h = 0xCBF29CE484222325
    for w in np.asarray(values).view(np.uint64).tolist():
        h = ((h ^ w) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def checksums(out: np.ndarray, expected: np.ndarray) -> dict:
    kept = expected >= LOG10_F32_THRESHOLD
    rounded = np.round(out[kept] * 1e6).astype(np.int64)
    return {"all": f"{fnv1a(out):016x}", "above_threshold": f"{fnv1a(out[kept]):016x}",
            "rounded_1e-6": f"{fnv1a(rounded):016x}"}


def region_lists(data: BatchData):
    """Per region: ([ReadRef], [bytes or PdHaplotype]) for the per-call API."""
    from . import PdHaplotype, ReadRef
    out = []
    for g in range(data.n_regions):
        reads = []
        for r in range(int(data.region_reads[g]), int(data.region_reads[g + 1])):
            a, b = int(data.read_offsets[r]), int(data.read_offsets[r + 1])
            reads.append(ReadRef(bytes(data.bases[a:b]), data.quals[a:b], data.ins[a:b],
                                 data.dels[a:b], data.gcp[a:b]))
        haps = []
        for h in range(int(data.region_haps[g]), int(data.region_haps[g + 1])):
            a, b = int(data.hap_offsets[h]), int(data.hap_offsets[h + 1])
            hb = bytes(data.hap_bases[a:b])
            haps.append(PdHaplotype(hb, data.hap_flags[a:b]) if data.pd else hb)
        out.append((reads, haps))
    return out


def replay(data: BatchData, expected: np.ndarray, modes, iters: int = 2,
           per_region: bool = False, write_dir: str | None = None, stem: str = "dump") -> dict:
    """Runs every mode over the dump's regions; returns the report."""
    import fgkl_amd as F
    cells = data.cells()
    report = {"regions": data.n_regions, "pairs": data.n_pairs(), "cells": cells,
              "gatk_below_f32_threshold": int((expected < LOG10_F32_THRESHOLD).sum()),
              "modes": {}}
    batch = None if per_region else F.Batch(data)
    lists = region_lists(data) if per_region else None
    for name in modes:
        mode = MODES[name]
        best, out, fallback = float("inf"), None, 0
        for _ in range(iters):
            if per_region:
                precision = F.Precision.DOUBLE if name == "double" else F.Precision.FLOAT
                cfg = F.Config(precision=precision, double_fallback=name != "float-nofb")
                hmm = (F.PdPairHmm if data.pd else F.PairHmm)(cfg)
                parts = []
                t0 = time.perf_counter()
                for reads, haps in lists:
                    parts.append(hmm.compute_log10_likelihoods(reads, haps))
                dt = time.perf_counter() - t0
                out = np.concatenate(parts) if parts else np.zeros(0)
                fallback = int(hmm.fallback_pairs())
            else:
                st = batch.run(mode, share=True, graph=False)
                dt = st.total_ms * 1e-3
                out = batch.results()
                fallback = int(st.rescued_pairs)
            best = min(best, dt)
        finite = np.isfinite(out)
        err = np.abs(out - expected)[finite]
        report["modes"][name] = {
            "seconds": round(best, 6), "gcups": round(cells / best / 1e9, 3) if best > 0 else None,
            "max_abs_err_vs_gatk": float(err.max()) if err.size else None,
            "nonfinite": int((~finite).sum()), "fallback_pairs": fallback,
            "checksum": checksums(out, expected)}
        if write_dir:
            d = Path(write_dir)
            d.mkdir(parents=True, exist_ok=True)
            out.astype("<f8").tofile(d / f"{stem}.{name}.f64")
            expected.astype("<f8").tofile(d / f"{stem}.expected.f64")
    if batch is not None:
        batch.close()
    return report


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m fgkl_amd.replay", description=__doc__.split("\n\n")[0])
    ap.add_argument("dump")
    ap.add_argument("--pd", action="store_true", help="a --pdhmm-results-file dump")
    ap.add_argument("--max-regions", type=int, default=0)
    ap.add_argument("--modes", default="float-nofb,float,double")
    ap.add_argument("--iters", type=int, default=2)
    ap.add_argument("--per-region", action="store_true",
                    help="one synchronous call per region through the PairHmm API")
    ap.add_argument("--write-results", metavar="DIR")
    ap.add_argument("--parse-only", action="store_true", help="report the regions and exit")
    args = ap.parse_args(argv)
    modes = [m for m in args.modes.split(",") if m]
    for m in modes:
        if m not in MODES:
            ap.error(f"unknown mode {m}")
    data, expected, stats = load(args.dump, args.pd, args.max_regions)
    if data.n_regions == 0:
        print(f"no complete regions found in {args.dump}", file=sys.stderr)
        return 1
    summary = {"dump": str(args.dump), **stats, "regions": data.n_regions,
               "pairs": data.n_pairs(), "cells": data.cells(),
               "mean_haps_per_region": round(data.n_haps / data.n_regions, 2),
               "mean_reads_per_region": round(data.n_reads / data.n_regions, 2)}
    if not args.parse_only:
        summary.update(replay(data, expected, modes, args.iters, args.per_region,
                              args.write_results, Path(args.dump).stem))
    print(json.dumps(summary, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
