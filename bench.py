"""Benchmark: batched FP32 + FP64-rescue PairHMM GCUPS on one MI355X.

A step is one pass of the engine over the whole batch (FP32 sweeps of every read against its
region's haplotypes with prefix sharing, then the FP64 recomputation of every pair below 1e-28),
replayed from one captured hipGraph with inputs resident in HBM.  Workload: 4096 uneven regions
(R~U(1,400) reads of U(50,250) bp, H~U(2,16) haplotypes of read_max+U(50,400) bp), synthetic,
seeded.  GCUPS counts nominal cells, sum over regions of (sum of read lengths) x (sum of
haplotype lengths), as the reference counts them (synthetic.rs:166-170).

Alongside: the plain per-call FP64 GPU engine on the same inputs (one pair per wave, one
synchronous launch set per region), parity of every pair against it, the roofline of the
dominant sweep kernel, an end-to-end call from host arrays, the drop-in path (one JNI-shaped call
per region), the gradient engine, a rescue-stress and a PD-HMM workload, Smith-Waterman on the
reference's aligner benchmark, and fgkl's lane-batched CPU backend (restated in C) timed on the
host cores over a bounded sample (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "PairHMM GCUPS, batched FP32+FP64-rescue vs plain FP64, 1 GPU, |Δlog10L|≤1e-4"
FLOPS_PER_CELL = 12          # 4 mul + 4 FMA per cell (reference.rs:49-54)
PEAK_FP32_TFLOPS = 157.3     # MI355X published vector FP32 peak (it counts packed FP32 math)
PEAK_FP64_TFLOPS = 78.6      # MI355X published vector FP64 peak
ADJOINT_FLOPS_PER_CELL = 36  # forward recompute 12 + reverse recurrence 12 + gradient terms 12


# The measured vector issue rates (valu_peak), filled in by main before any workload runs.
MEASURED: dict = {}


def vs_measured(achieved_tflops: float, key: str) -> dict:
    """The achieved rate against the measured issue rate of the same instruction kind."""
    peak = MEASURED.get(key)
    if not peak:
        return {}
    return {"peak_measured": peak, "frac_of_measured": round(achieved_tflops / peak, 4)}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


PROFILE = ROOT / "profiles" / "sweep_kernel_batched_uneven_4096_traffic.txt"


def measured_traffic(launch: str):
    """HBM bytes per launch of one bucket's sweep on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (scaled by a copy kernel of known size), when they were taken
    with the same launch configuration `launch` (R:ROWS:BLOCK:BOOST); otherwise None."""
    if not PROFILE.exists():
        return None
    m = re.search(rf"^traffic_bytes launch={re.escape(launch)} read=(\d+) write=(\d+)$",
                  PROFILE.read_text(), re.M)
    return None if m is None else int(m.group(1)) + int(m.group(2))


def first_regions(data, max_regions: int):
    """The first regions of a BatchData, as a BatchData."""
    from fgkl_amd.synthetic import BatchData

    g = min(max_regions, data.n_regions)
    r1, h1 = int(data.region_reads[g]), int(data.region_haps[g])
    rb, hb = int(data.read_offsets[r1]), int(data.hap_offsets[h1])
    sub = BatchData(g, data.region_reads[: g + 1].copy(), data.region_haps[: g + 1].copy(),
                    data.read_offsets[: r1 + 1].copy(), data.bases[:rb], data.quals[:rb],
                    data.ins[:rb], data.dels[:rb], data.gcp[:rb], data.hap_offsets[: h1 + 1].copy(),
                    data.hap_bases[:hb], data.hap_flags[:hb], data.pd)
    return sub


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(data, args, fp32_res) -> dict:
    """fgkl's lane-batched CPU backend, restated (oracle/port.c), timed on this host's cores: one
    region per worker thread, best of three runs over the first regions of the workload, sized
    by a probe to about --cpu-seconds per run.  Also the scalar FP64 oracle on a few regions, as
    the parity check of the GPU results."""
    sys.path.insert(0, str(ROOT / "tests"))
    from _oracle import batch_reference, port_backend, port_batch

    try:
        threads = min(len(os.sched_getaffinity(0)), 16)
    except AttributeError:
        threads = min(os.cpu_count() or 1, 16)
    probe = first_regions(data, max(args.cpu_regions, threads))
    t0 = time.perf_counter()
    port_batch(probe, nthreads=threads)
    per_region = (time.perf_counter() - t0) / probe.n_regions
    want = int(min(data.n_regions, max(probe.n_regions, args.cpu_seconds / max(per_region, 1e-9))))
    sub = first_regions(data, want)
    times, got = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        got = port_batch(sub, nthreads=threads)
        times.append(time.perf_counter() - t0)
    dt = min(times)
    n_sub = sub.n_pairs()
    d_port = np.abs(fp32_res[:n_sub] - got)
    small = first_regions(data, args.cpu_regions)
    ref = batch_reference(small, nthreads=threads)
    d_oracle = np.abs(fp32_res[: small.n_pairs()] - ref)
    # The reference's pairhmm-bench region on one thread, as that benchmark runs one call.
    from fgkl_amd.synthetic import Region
    region = Region.generate(*REFERENCE_BENCH_REGION)
    rb_times = []
    for _ in range(2):
        t0 = time.perf_counter()
        port_batch(region, nthreads=1)
        rb_times.append(time.perf_counter() - t0)
    rb = {"cells": region.cells(), "threads": 1, "seconds": [round(t, 3) for t in rb_times],
          "mcells_per_s": round(region.cells() / min(rb_times) / 1e6, 1),
          "note": "pairhmm-bench's default region (REFERENCE_BENCH_REGION) through the port, "
                  "float with FP64 rescue, best of 2"}
    log(f"cpu baseline on the pairhmm-bench region, 1 thread: {rb['mcells_per_s']:.0f} Mcells/s")
    log(f"cpu baseline ({port_backend()}): {sub.n_regions} regions, best of 3 {dt:.2f}s on "
        f"{threads} threads = {sub.cells() / dt / 1e9:.2f} GCUPS; gpu vs port max |d| "
        f"{d_port.max():.3g}, gpu vs oracle max |d| {d_oracle.max():.3g}")
    return {
        "value": round(sub.cells() / dt / 1e9, 3), "unit": "GCUPS", "cores": threads,
        "kind": "port",
        "sample": f"first {sub.n_regions} regions of the same workload ({n_sub} pairs, "
                  f"{sub.cells():.3e} cells), best of 3",
        "implementation": f"fgkl's lane-batched FP32 backend restated in C, {port_backend()}, "
                          "row sweep with prefix sharing, FTZ/DAZ, lost pairs recomputed in "
                          "FP64; one region per thread",
        "cpu_model": cpu_model(), "seconds": [round(t, 3) for t in times],
        "max_abs_dlog10L_gpu_vs_port": float(d_port.max()),
        "max_abs_dlog10L_gpu_vs_oracle": float(d_oracle.max()),
        "oracle_sample_pairs": small.n_pairs(),
        "reference_bench_region": rb,
    }


def valu_peak() -> dict:
    """The vector ALU's measured issue rate on this device (fgkl_measure_valu_rate): unpacked
    FP32 FMAs, FP64 FMAs and 32-bit integer ops from a full grid of dependence-free chains.  It
    is the one issue model every roofline here is also quoted against, beside the published
    peaks: the library's kernels run with packed FP32 disabled."""
    import fgkl_amd as F

    f32, f32_ms = F.measure_valu_rate(F.VALU_FP32_FMA)
    f64, f64_ms = F.measure_valu_rate(F.VALU_FP64_FMA)
    u32, u32_ms = F.measure_valu_rate(F.VALU_U32)
    cus = 256
    out = {
        "fp32_fma_tflops": round(2 * f32 / 1e12, 2),
        "fp64_fma_tflops": round(2 * f64 / 1e12, 2),
        "u32_tops": round(u32 / 1e12, 2),
        "fp32_wave_instructions_per_s": float(f"{f32 / 64:.4g}"),
        "cycles_per_wave64_fp32_at_2400mhz": round(cus * 4 * 2.4e9 / (f32 / 64), 3),
        "cycles_per_wave64_fp64_at_2400mhz": round(cus * 4 * 2.4e9 / (f64 / 64), 3),
        "cycles_per_wave64_u32_at_2400mhz": round(cus * 4 * 2.4e9 / (u32 / 64), 3),
        "launch_ms": {"fp32": round(f32_ms, 3), "fp64": round(f64_ms, 3), "u32": round(u32_ms, 3)},
        "method": "256 CUs x 8 waves per SIMD, 16 independent chains per lane, packed FP32 off",
    }
    log(f"valu peak: fp32 fma {out['fp32_fma_tflops']} TFLOP/s, fp64 fma "
        f"{out['fp64_fma_tflops']} TFLOP/s, u32 {out['u32_tops']} Tops/s")
    return out


def _adjoint_run(prob, e, di, dd, g):
    prob.backward(e, di, dd, g)  # warm-up
    runs = [prob.backward(e, di, dd, g, timing=True) for _ in range(2)]
    ms = min(r[-1] for r in runs)
    for a, b in zip(runs[0][:5], runs[1][:5]):
        assert np.array_equal(a, b), "adjoint runs differ"
    fwd_ms = min(prob.forward(e, di, dd, g, timing=True)[1] for _ in range(2))
    return runs[-1][:5], ms, fwd_ms


def adjoint_bench(args) -> dict:
    """The gradient engine on the adjoint workload, timed by HIP events around its kernels: the
    FP32 mode (forward recompute, tape and reverse sweep in FP32, gradient terms accumulated in
    FP64, FP64 rerun of pairs below 1e-28) and the FP64 mode, and the FP32 mode's distance from
    the FP64 mode's gradients."""
    from fgkl_amd.grad import GradProblem, probabilities
    from fgkl_amd.synthetic import workload

    data = workload(4, seed=args.seed, n_regions=args.adjoint_regions)
    cells = data.cells()
    e, di, dd, g = probabilities(data)
    out = {}
    for precision in ("float", "double"):
        prob = GradProblem(data, precision=precision)
        res, ms, fwd_ms = _adjoint_run(prob, e, di, dd, g)
        prob.close()
        achieved = ADJOINT_FLOPS_PER_CELL * cells / (ms * 1e-3) / 1e12
        peak = PEAK_FP32_TFLOPS if precision == "float" else PEAK_FP64_TFLOPS
        key = "fp32_fma_tflops" if precision == "float" else "fp64_fma_tflops"
        log(f"adjoint {precision}: {data.n_regions} regions, {data.n_pairs()} pairs, "
            f"{cells:.3e} cells; backward {ms:.1f} ms ({cells / ms / 1e6:.1f} GCUPS), "
            f"forward {fwd_ms:.1f} ms")
        out[precision] = (res, {
            "dtype": "f32 (f64 accumulation, f64 rescue)" if precision == "float" else "f64",
            "backward_ms": round(ms, 3), "gcups": round(cells / ms / 1e6, 2),
            "forward_ms": round(fwd_ms, 3), "forward_gcups": round(cells / fwd_ms / 1e6, 2),
            "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": peak,
                         "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                         "frac_of_fp64_peak": round(achieved / PEAK_FP64_TFLOPS, 4),
                         "flops_per_cell": ADJOINT_FLOPS_PER_CELL, **vs_measured(achieved, key)},
            "finite": bool(all(np.isfinite(x).all() for x in res[1:4])),
        })
    (r32, f32), (r64, f64) = out["float"], out["double"]
    dist = {"max_abs_dlog10L": float(np.max(np.abs(r32[0] - r64[0]))) if len(r64[0]) else 0.0}
    for name, a, b in (("e", r32[1], r64[1]), ("d_ins", r32[2], r64[2]), ("d_del", r32[3], r64[3])):
        dist[f"max_rel_d_grad_{name}"] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    dist["rel_d_grad_g"] = float(abs(r32[4] - r64[4]) / max(abs(r64[4]), 1e-300))
    log(f"adjoint float vs double: {dist}")
    return {
        "workload": "adjoint regions: R~U(1,200) reads of U(50,250) bp, H~U(2,8) haplotypes of "
                    "read_max+U(50,400) bp",
        "regions": data.n_regions, "pairs": data.n_pairs(), "cells": cells,
        **f32, "fp64": f64, "fp32_vs_fp64": dist,
    }


SIDE_WORKLOADS = {
    2: ("rescue_stress", "1024 regions x 64 reads of 250 bp x 8 haplotypes of 600 bp, quals "
                         "U(30,41), 30% of reads from unrelated sequence"),
    3: ("pd", "PD-HMM, 1024 regions of the headline shapes, flags at 5% of haplotype positions "
              "(undetermined SNPs with an alternate, deletion runs of 1-8)"),
}


def side_bench(args, config: int) -> dict:
    """A secondary workload through the same engine: graph-replayed FP32 + FP64-rescue GCUPS, the
    rescued fraction and the FP64 rescue's share of device time (eager run, HIP events), and every
    pair against the engine's FP64 mode."""
    import fgkl_amd as F
    from fgkl_amd.synthetic import workload

    name, desc = SIDE_WORKLOADS[config]
    data = workload(config, seed=args.seed)
    cells, n_pairs = data.cells(), data.n_pairs()
    b = F.Batch(data)
    tune = None if args.no_tune else b.tune(share=True)  # per bucket, as for the headline
    st = b.run(F.MODE_FP32_RESCUE, share=True, graph=False)
    eager = b.results()
    primary_ms, rescue_ms, total_ms = st.primary_ms, st.rescue_ms, st.total_ms
    rescued = int(st.rescued_pairs)
    b.run(F.MODE_FP32_RESCUE, share=True, graph=True)
    t0 = time.perf_counter()
    steps = 3
    for _ in range(steps):
        b.run(F.MODE_FP32_RESCUE, share=True, graph=True)
    elapsed = time.perf_counter() - t0
    assert np.array_equal(b.results(), eager), f"{name}: graph replay differs from the eager run"
    st64 = b.run(F.MODE_FP64, share=True)
    d = np.abs(eager - b.results())
    b.close()
    gcups = cells * steps / elapsed / 1e9
    achieved = FLOPS_PER_CELL * cells / (primary_ms * 1e-3) / 1e12
    log(f"{name}: {cells:.3e} cells, {gcups:.1f} GCUPS, rescued {rescued}/{n_pairs}, "
        f"fp64 rescue {rescue_ms:.1f} of {total_ms:.1f} ms; max |d| vs fp64 {d.max():.3g}")
    return {
        "workload": desc, "regions": data.n_regions, "pairs": n_pairs, "cells": cells,
        "gcups": round(gcups, 2), "ms_per_step": round(elapsed * 1e3 / steps, 3),
        "fp32_sweep_ms": round(primary_ms, 3), "fp64_rescue_ms": round(rescue_ms, 3),
        "rescue": {"pairs": rescued, "fraction": round(rescued / n_pairs, 6),
                   "fp64_time_share": round(rescue_ms / total_ms, 4) if total_ms else None},
        "fp64_mode_gcups": round(cells / (st64.total_ms * 1e-3) / 1e9, 2),
        "roofline_fp32_sweeps": {"bound": "valu", "achieved": round(achieved, 3),
                                 "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                 "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                                 **vs_measured(achieved, "fp32_fma_tflops")},
        "parity": {"max_abs_dlog10L_vs_fp64": float(d.max()), "pairs_over_1e-4": int((d > 1e-4).sum()),
                   "nonfinite": int((~np.isfinite(eager)).sum())},
        "launches": None if tune is None else ",".join(
            f"{R}:{tune.rows[R - 1]}:{tune.block[R - 1]}:{int(tune.boost[R - 1])}"
            for R in range(1, 9) if tune.rows[R - 1]),
    }


def per_call_bench(data, fp64_res, n_regions: int) -> dict:
    """The drop-in path: one JNI-shaped call per region (fgkl_pairhmm_compute_native, the C-ABI
    behind FgklPairHmm.computeNative) over the first regions of the workload, from host arrays to
    host results, as HaplotypeCaller makes them; every pair against the batch's FP64 result."""
    import fgkl_amd as F

    g = min(n_regions, data.n_regions)
    regs = []
    for k in range(g):
        r0, r1 = int(data.region_reads[k]), int(data.region_reads[k + 1])
        h0, h1 = int(data.region_haps[k]), int(data.region_haps[k + 1])
        o0, o1 = int(data.read_offsets[r0]), int(data.read_offsets[r1])
        q0, q1 = int(data.hap_offsets[h0]), int(data.hap_offsets[h1])
        regs.append(([np.ascontiguousarray(a[o0:o1]) for a in
                      (data.bases, data.quals, data.ins, data.dels, data.gcp)],
                     (data.read_offsets[r0:r1 + 1] - o0).astype(np.int32),
                     np.ascontiguousarray(data.hap_bases[q0:q1]),
                     (data.hap_offsets[h0:h1 + 1] - q0).astype(np.int32),
                     np.zeros((r1 - r0) * (h1 - h0))))
    cells = first_regions(data, g).cells()
    out = {"regions": g, "cells": cells,
           "entry": "fgkl_pairhmm_compute_native (FgklPairHmm.computeNative), one call per region"}
    for name, dp in (("fp32", 0), ("fp64", 1)):
        arrs, ro, hb, ho, res = regs[0]
        F.compute_native(*arrs, ro, hb, ho, dp, res)
        t0 = time.perf_counter()
        for arrs, ro, hb, ho, res in regs:
            F.compute_native(*arrs, ro, hb, ho, dp, res)
        dt = time.perf_counter() - t0
        got = np.concatenate([r[4] for r in regs])
        ref = fp64_res[: len(got)]
        d = np.abs(got - ref)
        out[name] = {"ms_per_region": round(dt * 1e3 / g, 3), "gcups": round(cells / dt / 1e9, 2),
                     "max_abs_dlog10L_vs_batch_fp64": float(d.max())}
        log(f"per-call {name}: {dt * 1e3 / g:.3f} ms per region, {cells / dt / 1e9:.1f} GCUPS, "
            f"max |d| vs batch fp64 {d.max():.3g}")
    out["threads"] = threaded_calls(regs, cells, fp64_res)
    return out


def threaded_calls(regs, cells, fp64_res, thread_counts=(1, 8, 32)) -> dict:
    """The same per-region FP32 calls from several caller threads at once, each thread taking every
    T-th region: on their own (each thread's region on its own stream) and gathered into shared
    batches by the aggregator (fgkl_aggregate_config with up to T/2 regions a batch, so that one
    batch gathers while the other runs)."""
    import threading

    import fgkl_amd as F

    def run(T):
        def worker(t):
            for arrs, ro, hb, ho, res in regs[t::T]:
                F.compute_native(*arrs, ro, hb, ho, 0, res)

        threads = [threading.Thread(target=worker, args=(t,)) for t in range(T)]
        t0 = time.perf_counter()
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        return time.perf_counter() - t0

    out = {}
    for T in thread_counts:
        row = {}
        for label, agg in (("alone", 0), ("aggregated", max(T // 2, 2))):
            F.aggregate(agg, 2000)
            try:
                run(T)
                b0, r0 = F.aggregate_stats()
                dt = run(T)
                b1, r1 = F.aggregate_stats()
            finally:
                F.aggregate(0)
            got = np.concatenate([r[4] for r in regs])
            d = np.abs(got - fp64_res[: len(got)])
            row[label] = {"gcups": round(cells / dt / 1e9, 2),
                          "ms_per_region": round(dt * 1e3 / len(regs), 3),
                          "max_abs_dlog10L_vs_batch_fp64": float(d.max())}
            if agg:
                row[label]["regions_per_batch"] = round((r1 - r0) / max(b1 - b0, 1), 2)
            log(f"per-call fp32 x{T} threads {label}: {cells / dt / 1e9:.1f} GCUPS, "
                f"max |d| vs batch fp64 {d.max():.3g}")
        out[f"threads_{T}"] = row
    return out


REFERENCE_BENCH_REGION = (7, 2000, 150, 64, 500)  # bench.rs:27-33 / pdbench.rs:27-33 defaults
HC_REGION = (7, 46, 150, 4, 300)  # a typical HaplotypeCaller region (docs/design.md:25)


def reference_bench() -> dict:
    """The reference's own throughput benchmarks (pairhmm-bench, bench.rs:54-107; pdhmm-bench,
    pdbench.rs:56-134) at their default region: one region of 2000 reads of 150 bp against 64
    haplotypes of 500 bp, seed 7, one compute call per iteration, best of 3 wall times, cells /
    second, in float and double.  Here the call is the per-region C-ABI entry
    (fgkl_pairhmm_compute_native / fgkl_pdhmm_compute_native) on host arrays; float results are
    compared with the engine's double results (the double path is pinned to the oracle by the
    tests)."""
    import fgkl_amd as F
    from fgkl_amd.synthetic import PdRegion, Region

    out = {"region": "seed 7, 2000 reads x 150 bp, 64 haplotypes x 500 bp (the benchmarks' "
                     "defaults)", "entry": "one per-region call, best of 3"}
    for name, gen, pd in (("pairhmm", Region.generate, False), ("pdhmm", PdRegion.generate, True)):
        d = gen(*REFERENCE_BENCH_REGION)
        ro, ho = d.read_offsets.astype(np.int32), d.hap_offsets.astype(np.int32)
        cells = d.cells()
        row, res = {"cells": cells}, {}
        for prec, dp in (("double", True), ("float", False)):
            o = np.zeros(d.n_pairs())

            def call():
                if pd:
                    F.pd_compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, ro, d.hap_bases,
                                        d.hap_flags, ho, dp, o)
                else:
                    F.compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, ro, d.hap_bases, ho,
                                     dp, o)

            call()
            best = float("inf")
            for _ in range(3):
                t0 = time.perf_counter()
                call()
                best = min(best, time.perf_counter() - t0)
            res[prec] = o.copy()
            row[prec] = {"ms": round(best * 1e3, 3), "mcells_per_s": round(cells / best / 1e6, 1),
                         "gcups": round(cells / best / 1e9, 2)}
        row["max_abs_dlog10L_float_vs_double"] = float(np.abs(res["float"] - res["double"]).max())
        out[name] = row
        log(f"reference bench {name}: float {row['float']['ms']} ms "
            f"({row['float']['gcups']:.0f} GCUPS), double {row['double']['ms']} ms "
            f"({row['double']['gcups']:.0f} GCUPS), max |float - double| "
            f"{row['max_abs_dlog10L_float_vs_double']:.3g}")
    # A region of HaplotypeCaller's typical shape (46 reads x 3-4 haplotypes on real data): the
    # latency of one per-region call, mean over 1000 calls.
    d = Region.generate(*HC_REGION)
    ro, ho = d.read_offsets.astype(np.int32), d.hap_offsets.astype(np.int32)
    o = np.zeros(d.n_pairs())
    small = {"region": "seed 7, 46 reads x 150 bp, 4 haplotypes x 300 bp", "cells": d.cells()}
    for prec, dp in (("double", True), ("float", False)):
        for _ in range(20):
            F.compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, ro, d.hap_bases, ho, dp, o)
        t0 = time.perf_counter()
        for _ in range(1000):
            F.compute_native(d.bases, d.quals, d.ins, d.dels, d.gcp, ro, d.hap_bases, ho, dp, o)
        us = (time.perf_counter() - t0) * 1e3
        small[prec] = {"us_per_call": round(us, 1), "gcups": round(d.cells() / us / 1e3, 2)}
    out["hc_region_per_call"] = small
    log(f"per-call latency, 46 x 4 region: float {small['float']['us_per_call']} us, double "
        f"{small['double']['us_per_call']} us")
    return out


SW_PEAK_INT32_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # 256 CUs x 4 SIMD16 x 2.4 GHz, 32-bit lane ops
SW_VALU_PER_CELL = 24.4  # VALU instructions per cell of the 8-rows-per-lane steady-state step


def sw_bench(args) -> dict:
    """Smith-Waterman: the reference's aligner benchmark workload (sw-bench, bench.rs:41-53: pairs
    of a 500-base reference and its first 500 bases with three substitutions and a deletion,
    haplotype-to-reference weights, INDEL strategy), here as batches resident on the device,
    timed with HIP events; a larger batch beside it; parity of every pair against the SIMD CPU
    baseline and of a sample against the scalar oracle; the baseline timed on the host cores over
    the larger batch."""
    import fgkl_amd.sw as S

    params, strategy = S.NEW_SW_PARAMETERS, S.OverhangStrategy.INDEL
    out = {"workload": "sw-bench: 500-base references, alternates = first 500 bases with 3 "
                       "substitutions and 1 deletion (499 bases), weights 200/-150/-260/-11, "
                       "INDEL", "dtype": "int32"}
    results = {}
    sizes = [("pairs_2000", 2000), ("pairs_large", args.sw_large_pairs)]
    if args.sw_large_only:
        sizes = sizes[1:]
    for label, n_pairs in sizes:
        ro, ref, ao, alt = S.bench_pairs(n_pairs, 500, 500, seed=5)
        batch = S.SwBatch(ref=ref, ref_offsets=ro, alt=alt, alt_offsets=ao)
        cells = batch.cells
        for _ in range(max(args.warmup, 1)):
            batch.run(params, strategy)
        dev, t0 = [], time.perf_counter()
        for _ in range(args.steps):
            dev.append(batch.run(params, strategy).total_ms)
        wall = (time.perf_counter() - t0) / args.steps
        offs, starts, cig = batch.results_arrays()
        results[label] = (ro, ref, ao, alt, offs, starts, cig)
        ms = min(dev)
        out[label] = {"pairs": n_pairs, "cells": cells, "device_ms": round(ms, 3),
                      "gcups": round(cells / (ms * 1e-3) / 1e9, 2),
                      "gcups_wall": round(cells / wall / 1e9, 2)}
        log(f"smith-waterman {n_pairs} pairs: {ms:.2f} ms device, "
            f"{cells / (ms * 1e-3) / 1e9:.1f} GCUPS")
        del batch
    best = max(out[label]["gcups"] for label, _ in sizes)
    achieved = best * 1e9 * SW_VALU_PER_CELL / 1e12  # lane ops: one cell per lane per row
    out["roofline"] = {"bound": "valu", "achieved": round(achieved, 3), "peak": SW_PEAK_INT32_TOPS,
                       "unit": "Tops/s (32-bit lane ops)", "frac": round(achieved / SW_PEAK_INT32_TOPS, 4),
                       "valu_per_cell": SW_VALU_PER_CELL, **vs_measured(achieved, "u32_tops")}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        from _oracle import cigar_text, sw_align, sw_port_backend, sw_port_raw

        try:
            threads = min(len(os.sched_getaffinity(0)), 16)
        except AttributeError:
            threads = min(os.cpu_count() or 1, 16)
        ro, ref, ao, alt, offs, starts, cig = results["pairs_large"]
        times, got = [], None
        for _ in range(3):
            t0 = time.perf_counter()
            got = sw_port_raw(ro, ref, ao, alt, params.as_array().tolist(), int(strategy),
                              nthreads=threads)
            times.append(time.perf_counter() - t0)
        dt = min(times)
        p_offs, p_start, p_cig, p_n = got
        mismatch = 0
        for p in range(len(ro) - 1):
            a = cig[starts[p]:starts[p + 1]]
            b = p_cig[p_start[p]:p_start[p] + p_n[p]]
            if offs[p] != p_offs[p] or len(a) != len(b) or not np.array_equal(a, b):
                mismatch += 1
        sample = 16
        oracle_mismatch = 0
        for p in range(sample):
            want = sw_align(bytes(ref[ro[p]:ro[p + 1]]), bytes(alt[ao[p]:ao[p + 1]]),
                            params.as_array().tolist(), "INDEL")
            if (cigar_text(cig[starts[p]:starts[p + 1]]), int(offs[p])) != want:
                oracle_mismatch += 1
        cells = out["pairs_large"]["cells"]
        out["parity"] = {"pairs_differing_from_cpu_baseline": mismatch,
                         "oracle_sample_pairs": sample, "pairs_differing_from_oracle": oracle_mismatch}
        out["cpu_baseline"] = {
            "value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"the {len(ro) - 1} pairs of the larger batch (sw-bench shapes), best of 3",
            "implementation": f"GATK's aligner vectorised across pairs in C ({sw_port_backend()}), "
                              "one trace byte per cell, traceback per pair; fgkl's anti-diagonal "
                              "16-bit SIMD backend is Rust and not built here",
            "cpu_model": cpu_model(), "seconds": [round(t, 3) for t in times]}
        log(f"smith-waterman cpu baseline ({sw_port_backend()}, {threads} threads): "
            f"{cells / dt / 1e9:.2f} GCUPS; pairs differing from it: {mismatch}, "
            f"from the oracle sample: {oracle_mismatch}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--regions", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-regions", type=int, default=16,
                    help="regions in the CPU probe and in the oracle parity sample")
    ap.add_argument("--cpu-seconds", type=float, default=5.0,
                    help="target time of one CPU baseline run (three are timed)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-baseline", action="store_true", help="skip the FP64 per-call run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient workload")
    ap.add_argument("--adjoint-regions", type=int, default=4096)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launches")
    ap.add_argument("--launch", default="", metavar="R:ROWS:BLOCK:BOOST,...",
                    help="set these bucket launch configurations instead of tuning")
    ap.add_argument("--no-lds-records", action="store_true",
                    help="write every prefix-sharing snapshot record to HBM (same bits)")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the rescue-stress and PD workloads")
    ap.add_argument("--no-sw", action="store_true", help="skip the Smith-Waterman workload")
    ap.add_argument("--no-per-call", action="store_true", help="skip the per-region call leg")
    ap.add_argument("--per-call-regions", type=int, default=256)
    ap.add_argument("--sw-only", action="store_true",
                    help="run only the Smith-Waterman workload and print its JSON")
    ap.add_argument("--sw-large-pairs", type=int, default=16384)
    ap.add_argument("--sw-large-only", action="store_true",
                    help="time only the larger Smith-Waterman batch (profiling runs)")
    ap.add_argument("--pd-only", action="store_true",
                    help="run only the PD-HMM workload and print its JSON (profiling runs)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the gradient workload and print its JSON (profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "launches": 3}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X; the path shards by region with no exchange "
              "between GPUs; run with --gpus 1", file=sys.stderr)
        return 2
    MEASURED.update(valu_peak())
    if args.sw_only:
        print(json.dumps({"smith_waterman": sw_bench(args)}), flush=True)
        return 0
    if args.pd_only:
        print(json.dumps({"pd": side_bench(args, 3)}), flush=True)
        return 0
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(args)}), flush=True)
        return 0

    import fgkl_amd as F
    from fgkl_amd.synthetic import workload

    t0 = time.perf_counter()
    data = workload(1, seed=args.seed, n_regions=args.regions)
    cells = data.cells()
    n_pairs = data.n_pairs()
    log(f"generated {data.n_regions} regions, {data.n_reads} reads, {n_pairs} pairs, "
        f"{cells:.3e} cells in {time.perf_counter() - t0:.1f}s")

    t0 = time.perf_counter()
    batch = F.Batch(data)
    prep_s = time.perf_counter() - t0
    log(f"planned + uploaded in {prep_s:.2f}s")
    if args.no_lds_records:
        batch.set_lds_records(False)

    # Launch autotuning per rows-per-lane bucket (untimed, like planning), or the configurations
    # given with --launch (so a profiling run can repeat a tuned run's kernels).
    tune = None
    if args.launch:
        for item in args.launch.split(","):
            R, rows, block, boost = (int(x) for x in item.split(":"))
            batch.set_launch(R, rows, block, bool(boost))
    elif not args.no_tune:
        tune = batch.tune(share=True)
    if tune is not None:
        chosen = {f"R={R}": {"rows": tune.rows[R - 1], "block": tune.block[R - 1],
                             "boost": bool(tune.boost[R - 1]),
                             "default_ms": round(tune.default_ms[R - 1], 3),
                             "tuned_ms": round(tune.tuned_ms[R - 1], 3)}
                  for R in range(1, 9) if tune.rows[R - 1]}
        log(f"autotune: {tune.candidates} configurations in {tune.seconds:.2f}s; " +
            ", ".join(f"{k}: rows {v['rows']} block {v['block']} boost {int(v['boost'])} "
                      f"{v['default_ms']}->{v['tuned_ms']} ms" for k, v in chosen.items()))

    # Eager run: per-bucket kernel times (HIP events on the engine's stream) and results.
    st = batch.run(F.MODE_FP32_RESCUE, share=True, graph=False)
    eager_ms = st.total_ms
    bucket_ms = list(st.bucket_ms)
    bc = [int(x) for x in st.bucket_cells]  # nominal cells each rows-per-lane launch covers
    bw = [int(x) for x in st.bucket_work]   # cells it computes (prefix + suffix sharing)
    primary_ms, rescue_ms, rescued = st.primary_ms, st.rescue_ms, int(st.rescued_pairs)
    fp32_res = batch.results()
    log(f"eager: total {eager_ms:.1f} ms, fp32 {primary_ms:.1f} ms, fp64 rescue {rescue_ms:.1f} ms, "
        f"rescued {rescued}")

    for _ in range(args.warmup):
        batch.run(F.MODE_FP32_RESCUE, share=True, graph=True)
    import ctypes as C
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = batch.run(F.MODE_FP32_RESCUE, share=True, graph=True)
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    gcups = cells * args.steps / elapsed / 1e9
    graph_res = batch.results()
    log(f"graph steps: {ms_per_step:.1f} ms/step, {gcups:.1f} GCUPS")

    result = {
        "metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1,
        "valu_peak_measured": dict(MEASURED),
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 rescue)", "data": "synthetic",
        "config": {"workload": "batched uneven PairHMM, 4096 regions: R~U(1,400) reads of "
                               "U(50,250) bp, H~U(2,16) haplotypes of read_max+U(50,400) bp",
                   "regions": data.n_regions, "reads": data.n_reads, "pairs": n_pairs,
                   "cells": cells, "seed": args.seed, "prefix_sharing": True, "graph": True},
    }

    # Roofline: the dominant sweep kernel.
    # The dominant launch is the bucket with the most cells to compute.  Buckets with under 2% of
    # the cells run on a second stream beside the large ones, so their event times span the
    # large launches they overlap and are not their own cost.
    dom = int(np.argmax(bw))
    achieved = FLOPS_PER_CELL * bc[dom] / (bucket_ms[dom] * 1e-3) / 1e12
    d_rows, d_block, d_boost = batch.get_launch(dom + 1)
    # Reads whose insertion, deletion and gap-continuation qualities are each constant run with
    # the uniform-gap build (every read of this workload).
    starts = data.read_offsets[:-1].astype(np.int64)
    uniform = all(bool(np.all(np.maximum.reduceat(a, starts) == np.minimum.reduceat(a, starts)))
                  for a in (data.ins, data.dels, data.gcp))
    launches = ",".join(f"{R}:" + ":".join(str(int(x)) for x in batch.get_launch(R))
                        for R in range(1, 9) if bc[R - 1])
    result["roofline"] = {
        "bound": "valu",
        "kernel": f"sweep_kernel<float, {d_rows}, false, {str(d_boost).lower()}, true, "
                  f"{str(uniform).lower()}> (bucket R={dom + 1}, {d_block} threads per workgroup"
                  + (", uniform-gap build)" if uniform else ")"),
        "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": measured_traffic(
            f"{dom + 1}:{d_rows}:{d_block}:{int(d_boost)}"),
        "traffic_source": f"{PROFILE.relative_to(ROOT)} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
        "algorithmic_flops_per_launch": int(FLOPS_PER_CELL * bc[dom]),
        "launch_ms": round(bucket_ms[dom], 3),
        # The same launch counted by the cells it computes (prefix and suffix sharing skip the
        # rest): the vector ALU's actual share of its FP32 peak.
        "computed_cells_per_launch": bw[dom],
        "achieved_computed": round(FLOPS_PER_CELL * bw[dom] / (bucket_ms[dom] * 1e-3) / 1e12, 3),
        "frac_computed": round(FLOPS_PER_CELL * bw[dom] / (bucket_ms[dom] * 1e-3) / 1e12
                               / PEAK_FP32_TFLOPS, 4),
        **vs_measured(achieved, "fp32_fma_tflops"),
        "frac_computed_of_measured": round(
            FLOPS_PER_CELL * bw[dom] / (bucket_ms[dom] * 1e-3) / 1e12
            / MEASURED["fp32_fma_tflops"], 4) if MEASURED.get("fp32_fma_tflops") else None,
        "launches": launches,
    }
    if tune is not None:
        result["autotune"] = {"seconds": round(tune.seconds, 3), "candidates": tune.candidates,
                              "buckets": chosen}
    result["kernels"] = {
        "fp32_sweep_ms": round(primary_ms, 3), "fp64_rescue_ms": round(rescue_ms, 3),
        "bucket_ms": [round(x, 3) for x in bucket_ms],
        "side_stream_buckets": [r + 1 for r in range(8) if 0 < bw[r] and bw[r] * 50 < sum(bw)],
        "note": "side-stream buckets run beside the large ones; their bucket_ms span the overlap",
        "bucket_cells": bc,
        "bucket_computed_cells": bw,
        "fp32_sweep_gcups": round(cells / (primary_ms * 1e-3) / 1e9, 2),
    }
    result["rescue"] = {"pairs": rescued, "fraction": round(rescued / n_pairs, 6),
                        "fp64_time_share": round(rescue_ms / eager_ms, 4) if eager_ms else None}
    assert np.array_equal(graph_res, fp32_res), "graph replay differs from the eager run"

    # The batched FP64 mode on the same inputs (every pair in FP64, prefix sharing, one launch
    # set): it separates the FP32 gain from the sharing and batching gains in the speedup over
    # the per-call FP64 engine.
    st64 = batch.run(F.MODE_FP64, share=True)
    fp64_batched = batch.results()
    result["fp64_batched"] = {
        "gcups": round(cells / (st64.total_ms * 1e-3) / 1e9, 2), "ms": round(st64.total_ms, 3),
        "definition": "FP64 engine, every pair, prefix and suffix sharing (joins below 1e24 "
                      "rerun unshared), one launch per rows-per-lane bucket (HIP events)",
        "max_abs_dlog10L_fp32_rescue_vs_fp64_batched": float(np.abs(fp32_res - fp64_batched).max())}
    result["speedup_vs_fp64_batched"] = round(gcups / result["fp64_batched"]["gcups"], 2)
    log(f"fp64 batched: {st64.total_ms:.1f} ms, {result['fp64_batched']['gcups']:.1f} GCUPS")

    if not args.no_baseline:
        st = batch.run(F.MODE_FP64_PER_REGION)
        base_ms = st.total_ms
        fp64 = batch.results()
        base_gcups = cells / (base_ms * 1e-3) / 1e9
        d = np.abs(fp32_res - fp64)
        result["baseline_fp64_per_call"] = {
            "gcups": round(base_gcups, 2), "ms": round(base_ms, 1),
            "definition": "FP64 engine, one pair per wave, no prefix sharing, one synchronous "
                          "launch set per region (as HaplotypeCaller calls per region)"}
        result["speedup_vs_fp64_per_call"] = round(gcups / base_gcups, 2)
        result["parity"] = {"max_abs_dlog10L_vs_fp64": float(d.max()),
                            "pairs_over_1e-4": int((d > 1e-4).sum()),
                            "nonfinite": int((~np.isfinite(fp32_res)).sum())}
        log(f"fp64 per-call: {base_ms:.1f} ms, {base_gcups:.1f} GCUPS; max |d| {d.max():.3g}")

    # End to end: one call from host arrays to host results (plan, upload, one graph run, copy
    # back), PCIe included.  Not the headline, which times resident inputs.
    t0 = time.perf_counter()
    e2e = F.Batch(data)
    e2e.run(F.MODE_FP32_RESCUE, share=True, graph=True)
    e2e_res = e2e.results()
    e2e_s = time.perf_counter() - t0
    e2e.close()
    assert np.array_equal(e2e_res, fp32_res), "a fresh batch gave different results"
    work = batch.work()
    result["work"] = {**work, "prefix_fraction": round(work["prefix"] / work["nominal"], 4),
                      "suffix_fraction": round((work["forward"] + work["backward"])
                                               / work["nominal"], 4),
                      "note": "cells computed per step: FP32 pass with prefix and suffix sharing "
                              "(forward + backward); value counts nominal cells"}
    snaps = batch.snapshots()
    result["snapshots"] = {
        "hbm_records": snaps["hbm"], "lds_records": snaps["lds"], "restarts": snaps["restarts"],
        "lds_record_launches": snaps["lds_launches"], "lds_records_enabled": not args.no_lds_records,
        "note": "prefix-sharing snapshot records per FP32 pass (one per read and recorded "
                "column): a sweep records a column only when it is the last to pass it before a "
                "haplotype resumes there; one record per wave stays in LDS when its lifetime "
                "allows and the launch holds it at the same occupancy, the rest go to HBM"}
    result["end_to_end"] = {"seconds": round(e2e_s, 4), "gcups": round(cells / e2e_s / 1e9, 2),
                            "includes": "plan on the host, upload, graph build + run, download"}
    log(f"end to end: {e2e_s * 1e3:.1f} ms, {cells / e2e_s / 1e9:.1f} GCUPS")

    if not args.no_per_call:
        result["per_call"] = per_call_bench(data, fp32_res if args.no_baseline else fp64,
                                            args.per_call_regions)

    result["reference_bench"] = reference_bench()

    if not args.no_adjoint:
        result["adjoint"] = adjoint_bench(args)
    if not args.no_side:
        for config in SIDE_WORKLOADS:
            result[SIDE_WORKLOADS[config][0]] = side_bench(args, config)

    if not args.no_sw:
        result["smith_waterman"] = sw_bench(args)

    if not args.no_cpu:
        cb = cpu_baseline(data, args, fp32_res)
        result["cpu_baseline"] = cb
        result.setdefault("parity", {})["max_abs_dlog10L_vs_oracle_sample"] = \
            cb["max_abs_dlog10L_gpu_vs_oracle"]

    print(json.dumps(result), flush=True)
    batch.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
