"""Writes the GATK-dump fixtures of tests/golden/ (replay_plain.txt, replay_pd.txt).

No GATK dump exists in this repository's environment, so these are dumps in GATK's formats
(--pair-hmm-results-file: space-separated; --pdhmm-results-file: tab-separated with
Arrays.toString flags) of small seeded regions, with the CPU oracle's FP64 log10 likelihoods in
the recorded-likelihood column.  Besides clean regions they hold the cases the region
reconstruction has to get right (crates/pairhmm/src/bin/replay.rs:83-159): duplicate reads back
to back, a region of one haplotype followed by another region, comment and malformed lines, a
region whose haplotype list breaks mid-read (dropped), and values in GATK's short %e form.
The layout each fixture must parse into is written next to it (replay_*.layout.json).

    python tests/golden/make_replay_fixtures.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

import fgkl_amd as F  # noqa: E402
from fgkl_amd.synthetic import PdRegion, Region  # noqa: E402
from _oracle import oracle, pd_reference_all, reference_all  # noqa: E402


def q33(a) -> str:
    return "".join(chr(int(x) + 33) for x in a)


def plain_line(hap: bytes, r: F.ReadRef, v: float, fmt: str = "%.17e") -> str:
    return " ".join([hap.decode(), bytes(r.bases).decode(), q33(r.quals), q33(r.ins_gop),
                     q33(r.del_gop), q33(r.gcp), fmt % v])


def pd_line(h: F.PdHaplotype, r: F.ReadRef, v: float, fmt: str = "%.17e") -> str:
    flags = "[" + ", ".join(str(int(np.int8(np.uint8(f)))) for f in h.flags) + "]"
    return "\t".join([bytes(h.bases).decode(), flags, bytes(r.bases).decode(), q33(r.quals),
                      q33(r.ins_gop), q33(r.del_gop), q33(r.gcp), fmt % v])


def region_lines(reads, haps, line, values, fmt="%.17e"):
    out = []
    for i, r in enumerate(reads):
        for j, h in enumerate(haps):
            out.append(line(h, r, values[i * len(haps) + j], fmt))
    return out


def build(pd: bool):
    gen = PdRegion if pd else Region
    line = pd_line if pd else plain_line
    ref_all = pd_reference_all if pd else reference_all
    lines, layout = ["# GATK results dump fixture"], []

    def add(reads, haps, fmt="%.17e"):
        vals = ref_all(reads, haps)
        lines.extend(region_lines(reads, haps, line, vals, fmt))
        layout.append({"reads": len(reads), "haps": len(haps)})

    # 1. a clean region with a duplicate read back to back
    r = gen.generate(101, 6, 90, 5, 160)
    reads, haps = r.region_reads_list(), (r.region_pd_haps_list() if pd else r.region_haps_list())
    reads = reads[:3] + [reads[2]] + reads[3:]
    add(reads, haps)
    # 2. a region of one haplotype, then
    r = gen.generate(102, 4, 60, 1, 120)
    add(r.region_reads_list(), r.region_pd_haps_list() if pd else r.region_haps_list())
    lines.append("# a comment between regions")
    # 3. a clean region written with GATK's %e (seven significant digits), a malformed line inside
    r = gen.generate(103, 5, 70, 4, 140)
    reads, haps = r.region_reads_list(), (r.region_pd_haps_list() if pd else r.region_haps_list())
    vals = ref_all(reads, haps)
    body = region_lines(reads, haps, line, vals, "%e")
    lines.extend(body[:3] + ["TOO FEW FIELDS"] + body[3:])
    layout.append({"reads": len(reads), "haps": len(haps), "short_values": True})
    # 4. a region broken mid-read: its lines are dropped, the breaking line starts a new region
    r = gen.generate(104, 3, 50, 3, 100)
    reads, haps = r.region_reads_list(), (r.region_pd_haps_list() if pd else r.region_haps_list())
    vals = ref_all(reads, haps)
    lines.extend(region_lines(reads[:1], haps, line, vals))  # first read complete
    lines.append(line(haps[0], reads[1], vals[3]))            # second read starts ...
    lines.append(line(haps[2], reads[1], vals[5]))            # ... and skips haps[1]
    # 5. the region that line starts, completed
    tail_haps = [haps[2], haps[0]]
    tail_reads = [reads[1], reads[2]]
    tv = ref_all(tail_reads, tail_haps)
    lines.append(line(tail_haps[1], tail_reads[0], tv[1]))
    lines.extend(region_lines(tail_reads[1:], tail_haps, line, tv[2:]))
    layout.append({"reads": 2, "haps": 2})
    # 6. a last clean region
    r = gen.generate(105, 7, 110, 6, 200)
    add(r.region_reads_list(), r.region_pd_haps_list() if pd else r.region_haps_list())
    return lines, layout


def main() -> None:
    oracle()
    for pd in (False, True):
        lines, layout = build(pd)
        name = "replay_pd" if pd else "replay_plain"
        (HERE / f"{name}.txt").write_text("\n".join(lines) + "\n")
        meta = {"regions": layout, "skipped_lines": 1, "dropped_regions": 1,
                "generator": "tests/golden/make_replay_fixtures.py"}
        (HERE / f"{name}.layout.json").write_text(json.dumps(meta, indent=1) + "\n")
        print(name, len(lines), "lines")


if __name__ == "__main__":
    main()
