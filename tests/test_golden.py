"""Golden fixtures (tests/golden/*.json, written by tests/golden/make_golden.py from the oracle on
the reference's seeded test regions): the generator and the oracle reproduce them on the CPU,
and the engine matches them on the GPU (FP64 within 1e-12 relative, FP32 within 1e-4)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import fgkl_amd as F
from _oracle import batch_reference
from fgkl_amd.synthetic import PdRegion, Region, workload

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.json") if not p.name.endswith(".layout.json"))


def load(name):
    fx = json.loads((GOLDEN / f"{name}.json").read_text())
    kind, args = fx["generator"], fx["args"]
    if kind == "plain":
        data = Region.generate(*args)
    elif kind == "pd":
        data = PdRegion.generate(*args)
    else:
        data = workload(*args)
    return fx, data


@pytest.mark.parametrize("name", NAMES)
def test_generator_and_oracle_reproduce_the_fixture(name):
    fx, data = load(name)
    assert (data.n_reads, data.n_haps, data.cells()) == (fx["n_reads"], fx["n_haps"], fx["cells"])
    assert int(data.bases.astype("uint64").sum()) == fx["read_bases_checksum"]
    assert int(data.hap_bases.astype("uint64").sum()) == fx["hap_bases_checksum"]
    assert np.array_equal(batch_reference(data), np.array(fx["log10_likelihoods"]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_engine_matches_the_fixture(gpu, name):
    fx, data = load(name)
    expected = np.array(fx["log10_likelihoods"])
    b = F.Batch(data)
    b.run(F.MODE_FP64)
    got = b.results()
    assert np.all(np.abs(got - expected) <= 1e-12 * np.maximum(1.0, np.abs(expected)))
    b.run(F.MODE_FP32_RESCUE)
    got = b.results()
    assert np.max(np.abs(got - expected)) <= 1e-4
    b.close()
