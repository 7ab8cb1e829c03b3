"""Synthetic regions and benchmark workloads (synthetic.rs restated in the engine library).

``Region.generate`` / ``PdRegion.generate`` regenerate the reference's own seeded test regions
draw for draw; ``workload`` builds the five benchmark workloads.  ``Rng`` is SplitMix64
(synthetic.rs:10-40) in Python, for small hand-built test inputs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import lib

MASK64 = (1 << 64) - 1


class Rng:
    """SplitMix64 (synthetic.rs:10-40)."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        return self.next_u64() % n

    def chance(self, p: float) -> bool:
        return (self.next_u64() >> 11) / float(1 << 53) < p

    def base(self) -> int:
        return b"ACGT"[self.below(4)]


@dataclass
class BatchData:
    """Regions of reads and haplotypes, concatenated (the batch C-ABI layout)."""
    n_regions: int
    region_reads: np.ndarray
    region_haps: np.ndarray
    read_offsets: np.ndarray
    bases: np.ndarray
    quals: np.ndarray
    ins: np.ndarray
    dels: np.ndarray
    gcp: np.ndarray
    hap_offsets: np.ndarray
    hap_bases: np.ndarray
    hap_flags: np.ndarray
    pd: bool = False

    @property
    def n_reads(self) -> int:
        return len(self.read_offsets) - 1

    @property
    def n_haps(self) -> int:
        return len(self.hap_offsets) - 1

    def region_shape(self, g: int) -> tuple[int, int]:
        return (int(self.region_reads[g + 1] - self.region_reads[g]),
                int(self.region_haps[g + 1] - self.region_haps[g]))

    def n_pairs(self) -> int:
        return int(sum(r * h for r, h in (self.region_shape(g) for g in range(self.n_regions))))

    def out_base(self) -> np.ndarray:
        sizes = [r * h for r, h in (self.region_shape(g) for g in range(self.n_regions))]
        return np.concatenate([[0], np.cumsum(sizes, dtype=np.int64)]).astype(np.int64)

    def cells(self) -> int:
        """synthetic.rs:166-170 summed over regions: (sum of read lengths) x (sum of hap lengths)."""
        rr = self.read_offsets[self.region_reads]
        hh = self.hap_offsets[self.region_haps]
        return int(np.sum((rr[1:] - rr[:-1]).astype(np.int64) * (hh[1:] - hh[:-1]).astype(np.int64)))

    def read(self, r: int):
        from . import ReadRef
        lo, hi = int(self.read_offsets[r]), int(self.read_offsets[r + 1])
        return ReadRef(bytes(self.bases[lo:hi]), self.quals[lo:hi].tolist(),
                       self.ins[lo:hi].tolist(), self.dels[lo:hi].tolist(),
                       self.gcp[lo:hi].tolist())

    def hap(self, h: int) -> bytes:
        return bytes(self.hap_bases[int(self.hap_offsets[h]):int(self.hap_offsets[h + 1])])

    def flags(self, h: int) -> list[int]:
        return self.hap_flags[int(self.hap_offsets[h]):int(self.hap_offsets[h + 1])].tolist()

    def region_reads_list(self, g: int = 0):
        return [self.read(r) for r in range(int(self.region_reads[g]), int(self.region_reads[g + 1]))]

    def region_haps_list(self, g: int = 0):
        return [self.hap(h) for h in range(int(self.region_haps[g]), int(self.region_haps[g + 1]))]

    def region_pd_haps_list(self, g: int = 0):
        from . import PdHaplotype
        return [PdHaplotype(self.hap(h), self.flags(h))
                for h in range(int(self.region_haps[g]), int(self.region_haps[g + 1]))]


def batch_from_regions(regions, pd: bool = False) -> BatchData:
    """A BatchData of explicit regions: each is (reads, haplotypes) with reads as ReadRef and
    haplotypes as bytes (PdHaplotype when pd)."""
    rr, hh, ro, ho = [0], [0], [0], [0]
    cols = {k: [] for k in ("bases", "quals", "ins", "dels", "gcp", "hap", "flags")}
    for reads, haps in regions:
        for rd in reads:
            cols["bases"].append(np.frombuffer(bytes(rd.bases), np.uint8))
            cols["quals"].append(np.asarray(rd.quals, np.uint8))
            cols["ins"].append(np.asarray(rd.ins_gop, np.uint8))
            cols["dels"].append(np.asarray(rd.del_gop, np.uint8))
            cols["gcp"].append(np.asarray(rd.gcp, np.uint8))
            ro.append(ro[-1] + len(rd.bases))
        for h in haps:
            b = bytes(h.bases) if pd else bytes(h)
            cols["hap"].append(np.frombuffer(b, np.uint8))
            cols["flags"].append(np.asarray(h.flags if pd else [0] * len(b), np.uint8))
            ho.append(ho[-1] + len(b))
        rr.append(rr[-1] + len(reads))
        hh.append(hh[-1] + len(haps))
    cat = lambda k: np.concatenate(cols[k]).astype(np.uint8) if cols[k] else np.zeros(0, np.uint8)
    return BatchData(len(regions), np.array(rr, np.int64), np.array(hh, np.int64),
                     np.array(ro, np.int64), cat("bases"), cat("quals"), cat("ins"), cat("dels"),
                     cat("gcp"), np.array(ho, np.int64), cat("hap"), cat("flags"), pd)


def _from_handle(h, pd: bool) -> BatchData:
    L = lib()
    try:
        sizes = [C.c_int64() for _ in range(5)]
        L.fgkl_synth_sizes(h, *[C.byref(s) for s in sizes])
        g, nr, rb, nh, hb = (s.value for s in sizes)
        arr = lambda n, t: np.zeros(n, dtype=t)
        d = BatchData(g, arr(g + 1, np.int64), arr(g + 1, np.int64), arr(nr + 1, np.int64),
                      arr(rb, np.uint8), arr(rb, np.uint8), arr(rb, np.uint8), arr(rb, np.uint8),
                      arr(rb, np.uint8), arr(nh + 1, np.int64), arr(hb, np.uint8),
                      arr(hb, np.uint8), pd)
        p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        L.fgkl_synth_copy(h, p(d.region_reads, C.c_int64), p(d.region_haps, C.c_int64),
                          p(d.read_offsets, C.c_int64), p(d.bases, C.c_uint8),
                          p(d.quals, C.c_uint8), p(d.ins, C.c_uint8), p(d.dels, C.c_uint8),
                          p(d.gcp, C.c_uint8), p(d.hap_offsets, C.c_int64),
                          p(d.hap_bases, C.c_uint8), p(d.hap_flags, C.c_uint8))
        return d
    finally:
        L.fgkl_synth_free(h)


class Region:
    @staticmethod
    def generate(seed: int, num_reads: int, read_len: int, num_haplotypes: int,
                 hap_len: int) -> BatchData:
        """synthetic.rs:103-153"""
        return _from_handle(lib().fgkl_synth_region(seed, num_reads, read_len, num_haplotypes,
                                                    hap_len, 0), False)


class PdRegion:
    @staticmethod
    def generate(seed: int, num_reads: int, read_len: int, num_haplotypes: int,
                 hap_len: int) -> BatchData:
        """synthetic.rs:199-289"""
        return _from_handle(lib().fgkl_synth_region(seed, num_reads, read_len, num_haplotypes,
                                                    hap_len, 1), True)


WORKLOADS = {
    0: "plain FP64 region: 32 reads x 8 haplotypes, 151 bp reads, 300-350 bp haplotypes",
    1: "batched uneven regions: R~U(1,400) reads of U(50,250) bp, H~U(2,16) haplotypes",
    2: "underflow-heavy: 64 reads x 8 haplotypes, 250 bp reads, 600 bp haplotypes, 30% foreign",
    3: "PD regions (DRAGEN 3.7.8 flags at 5% of positions), shapes of workload 1",
    4: "adjoint regions: R~U(1,200), H~U(2,8), shapes of workload 1 otherwise",
}
DEFAULT_REGIONS = {0: 1, 1: 4096, 2: 1024, 3: 1024, 4: 4096}


def workload(config: int, seed: int = 7, n_regions: int = 0) -> BatchData:
    n = n_regions if n_regions > 0 else DEFAULT_REGIONS[config]
    return _from_handle(lib().fgkl_synth_workload(config, seed, n), config == 3)


def splitmix_array(seed: int, n: int) -> np.ndarray:
    """The first n values of SplitMix64 from `seed`, as uint64."""
    out = np.zeros(n, dtype=np.uint64)
    lib().fgkl_synth_splitmix(seed, n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def splitmix(seed: int, n: int) -> list[int]:
    out = np.zeros(n, dtype=np.uint64)
    lib().fgkl_synth_splitmix(seed, n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return [int(x) for x in out]
