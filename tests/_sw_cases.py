"""Smith-Waterman test inputs: the reference's generator of related pairs
(crates/smithwaterman/src/lib.rs:772-810, same SplitMix64 stream) and its weight sets
(lib.rs:741-743)."""
from __future__ import annotations

from fgkl_amd.synthetic import Rng

HAP_TO_REF = (200, -150, -260, -11)
READ_TO_HAP = (10, -15, -30, -5)
ORIGINAL = (3, -1, -4, -3)
DANGLING = (25, -50, -110, -6)   # lib.rs:885, odd match values
STRATEGIES = ("SOFTCLIP", "INDEL", "LEADING_INDEL", "IGNORE")


def random_seq(rng: Rng, n: int) -> bytes:
    return bytes(rng.base() for _ in range(n))


def related_pair(rng: Rng, ref_len: int, alt_len: int) -> tuple[bytes, bytes]:
    """A reference and an edited copy of a window of it, like a haplotype or a read."""
    ref = random_seq(rng, ref_len)
    start = rng.below(max(ref_len - alt_len, 0) + 1)
    alt = bytearray(ref[start:min(start + alt_len, ref_len)])
    for _ in range(rng.below(4)):
        if not alt:
            break
        pos = rng.below(len(alt))
        kind = rng.below(3)
        if kind == 0:
            alt[pos] = rng.base()
        elif kind == 1:
            alt[pos:pos] = random_seq(rng, 1 + rng.below(6))
        else:
            end = min(pos + 1 + rng.below(6), len(alt))
            if len(alt) > end - pos:
                del alt[pos:end]
    if rng.chance(0.3):
        extra = random_seq(rng, 1 + rng.below(20))
        alt = bytearray(extra) + alt if rng.chance(0.5) else alt + extra
    return ref, bytes(alt or b"A")
