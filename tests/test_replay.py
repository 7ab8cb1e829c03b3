"""GATK results-dump replay (crates/pairhmm/src/bin/replay.rs, pdreplay.rs).

The fixtures (tests/golden/replay_*.txt, written by make_replay_fixtures.py) are dumps in GATK's
two formats whose recorded likelihoods are the CPU oracle's; their *.layout.json gives the region layout
the reconstruction must produce.  CPU tests pin the parser (layout, skipped and dropped lines,
pairing of every line with its read and haplotype, round trips); GPU tests replay the dumps
through the engine in every mode."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import fgkl_amd as F
import fgkl_amd.replay as R
from fgkl_amd.synthetic import workload
from _oracle import batch_reference

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = [("replay_plain", False), ("replay_pd", True)]


def _load(name, pd):
    data, expected, stats = R.load(GOLDEN / f"{name}.txt", pd)
    return data, expected, stats, json.loads((GOLDEN / f"{name}.layout.json").read_text())


@pytest.mark.parametrize("name,pd", FIXTURES)
def test_dump_regions_are_rebuilt_from_the_repeating_haplotype_lists(name, pd):
    data, expected, stats, meta = _load(name, pd)
    shapes = [data.region_shape(g) for g in range(data.n_regions)]
    assert shapes == [(m["reads"], m["haps"]) for m in meta["regions"]]
    assert stats["skipped_lines"] == meta["skipped_lines"]
    assert stats["dropped_regions"] == meta["dropped_regions"]
    assert len(expected) == data.n_pairs() and data.pd == pd


@pytest.mark.parametrize("name,pd", FIXTURES)
def test_every_recorded_value_belongs_to_its_read_and_haplotype(name, pd):
    # The oracle recomputes each pair from the parsed arrays: a line paired with the wrong read,
    # haplotype or flags would miss by far more than the recorded digits allow.
    data, expected, _, meta = _load(name, pd)
    ref = batch_reference(data)
    base = data.out_base()
    for g, m in enumerate(meta["regions"]):
        sl = slice(int(base[g]), int(base[g + 1]))
        rel = 1e-6 if m.get("short_values") else 1e-12
        d = np.abs(ref[sl] - expected[sl]) / np.maximum(1.0, np.abs(ref[sl]))
        assert d.max() <= rel, (g, d.max())


@pytest.mark.parametrize("wl", [1, 3])
def test_write_then_load_round_trips(tmp_path, wl):
    data = workload(wl, seed=31 + wl, n_regions=4)
    values = -np.linspace(1.0, 90.0, data.n_pairs())
    path = tmp_path / "dump.txt"
    R.write(path, data, values)
    back, expected, stats = R.load(path, data.pd)
    assert stats == {"lines": data.n_pairs(), "skipped_lines": 0, "dropped_regions": 0}
    assert np.array_equal(expected, values)
    for f in ("region_reads", "region_haps", "read_offsets", "bases", "quals", "ins", "dels",
              "gcp", "hap_offsets", "hap_bases", "hap_flags"):
        assert np.array_equal(getattr(back, f), getattr(data, f)), f


def test_pd_flags_are_read_as_java_signed_bytes(tmp_path):
    line = "ACGT\t[0, -128, 127, 6]\tAC\t??\t??\t??\t++\t-3.5e+00\n"
    (tmp_path / "d.txt").write_text(line)
    data, expected, _ = R.load(tmp_path / "d.txt", pd=True)
    assert list(data.hap_flags) == [0, 128, 127, 6] and list(expected) == [-3.5]
    assert list(data.quals) == [30, 30] and list(data.gcp) == [10, 10]


def test_a_missing_dump_is_an_invalid_argument(tmp_path):
    with pytest.raises(F.InvalidArgument):
        R.load(tmp_path / "absent.txt")


def test_the_command_line_reports_the_regions_without_a_gpu():
    out = subprocess.run([sys.executable, "-m", "fgkl_amd.replay", str(GOLDEN / "replay_plain.txt"),
                          "--parse-only"], capture_output=True, text=True, check=True,
                         cwd=GOLDEN.parents[1])
    report = json.loads(out.stdout)
    assert report["regions"] == 5 and report["dropped_regions"] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd", FIXTURES)
def test_dumps_replay_through_the_engine(gpu, name, pd, tmp_path):
    data, expected, _, _ = _load(name, pd)
    ref = batch_reference(data)
    rep = R.replay(data, expected, ["float-nofb", "float", "double"], iters=1,
                   write_dir=str(tmp_path), stem="batch")
    assert rep["pairs"] == data.n_pairs()
    assert rep["modes"]["double"]["max_abs_err_vs_gatk"] <= 5e-5   # 7-digit %e values included
    assert rep["modes"]["float"]["max_abs_err_vs_gatk"] <= 1e-4
    assert rep["modes"]["float"]["nonfinite"] == 0
    b = F.Batch(data)
    b.run(F.MODE_FP64, share=True)
    d = np.abs(b.results() - ref) / np.maximum(1.0, np.abs(ref))
    assert d.max() <= 1e-12
    b.close()
    # One call per region through the PairHmm API: each pair in its own wave, without prefix
    # sharing, so within FP64 rounding of the batch (which starts haplotypes from rescaled
    # snapshots) and within the FP32 bound in float mode.
    per = R.replay(data, expected, ["float", "double"], iters=1, per_region=True,
                   write_dir=str(tmp_path), stem="per")
    for mode, tol in (("double", 1e-12), ("float", 1e-4)):
        a = np.fromfile(tmp_path / f"per.{mode}.f64", "<f8")
        b = np.fromfile(tmp_path / f"batch.{mode}.f64", "<f8")
        assert np.all(np.abs(a - b) <= tol * np.maximum(1.0, np.abs(b))), mode
        assert np.all(np.abs(a - ref) <= tol * np.maximum(1.0, np.abs(ref))), mode
