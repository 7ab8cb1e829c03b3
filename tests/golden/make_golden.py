"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

The reference keeps no data files (its tests compute expected values at run time from seeded
generators, synthetic.rs), so the fixtures are the oracle's log10 likelihoods on the reference's
own seeded test regions, regenerated draw for draw.  They pin the oracle and the generator
against later drift and give the GPU tests fixed expected values.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

from fgkl_amd.synthetic import PdRegion, Region, workload  # noqa: E402
from _oracle import batch_reference  # noqa: E402

CASES = {
    # name: (kind, args); regions of crates/pairhmm/src/lib.rs:617, :748 and pd.rs:158
    "region_11_70_120_24_300": ("plain", (11, 70, 120, 24, 300)),
    "region_21_90_150_40_260": ("plain", (21, 90, 150, 40, 260)),
    "pd_region_31_70_120_24_300": ("pd", (31, 70, 120, 24, 300)),
    "workload0_seed7": ("workload", (0, 7, 1)),
}


def build(kind, args):
    if kind == "plain":
        return Region.generate(*args)
    if kind == "pd":
        return PdRegion.generate(*args)
    return workload(*args)


def main() -> None:
    for name, (kind, args) in CASES.items():
        data = build(kind, args)
        values = batch_reference(data)
        fixture = {
            "generator": kind, "args": list(args), "n_reads": data.n_reads,
            "n_haps": data.n_haps, "cells": data.cells(),
            "read_bases_checksum": int(data.bases.astype("uint64").sum()),
            "hap_bases_checksum": int(data.hap_bases.astype("uint64").sum()),
            "log10_likelihoods": [float(v) for v in values],
        }
        (HERE / f"{name}.json").write_text(json.dumps(fixture) + "\n")
        print(name, len(values))


if __name__ == "__main__":
    main()
