"""Loader for the CPU oracle (oracle/liboracle.so): test infrastructure, used only as the checker."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

_o = None
u8p = C.POINTER(C.c_uint8)
f64p = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)


def oracle() -> C.CDLL:
    global _o
    if _o is None:
        if not ORACLE_LIB.exists():
            subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
        o = C.CDLL(str(ORACLE_LIB))
        o.fgklo_error_prob.restype = C.c_double
        o.fgklo_error_prob.argtypes = [C.c_uint8]
        o.fgklo_match_to_match.restype = C.c_double
        o.fgklo_match_to_match.argtypes = [C.c_uint8, C.c_uint8]
        o.fgklo_transitions.restype = None
        o.fgklo_transitions.argtypes = [C.c_uint8, C.c_uint8, C.c_uint8, f64p]
        o.fgklo_priors.restype = None
        o.fgklo_priors.argtypes = [C.c_uint8, f64p]
        o.fgklo_approx_log10_sum_log10.restype = C.c_double
        o.fgklo_approx_log10_sum_log10.argtypes = [C.c_double, C.c_double]
        o.fgklo_initial_condition.restype = C.c_double
        o.fgklo_initial_condition_log10.restype = C.c_double
        o.fgklo_log10_likelihood.restype = C.c_double
        o.fgklo_log10_likelihood.argtypes = [u8p, C.c_int64, u8p, u8p, u8p, u8p, u8p, C.c_int64]
        o.fgklo_port_batch.restype = C.c_int
        o.fgklo_port_batch.argtypes = [C.c_int64, i64p, i64p, u8p, u8p, u8p, u8p, u8p, i64p,
                                       i64p, u8p, f64p, C.c_int, C.c_int, C.c_int]
        o.fgklo_port_backend.restype = C.c_char_p
        o.fgklo_sw_align.restype = C.c_int
        o.fgklo_sw_align.argtypes = [u8p, C.c_int64, u8p, C.c_int64, C.POINTER(C.c_int32), C.c_int,
                                     C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_int32)]
        o.fgklo_sw_port_batch.restype = C.c_int
        o.fgklo_sw_port_batch.argtypes = [C.c_int64, i64p, u8p, i64p, u8p, C.POINTER(C.c_int32),
                                          C.c_int, C.POINTER(C.c_int32), i64p,
                                          C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.c_int,
                                          C.c_int]
        o.fgklo_sw_port_backend.restype = C.c_char_p
        o.fgklo_base_matches_pd.restype = C.c_int
        o.fgklo_base_matches_pd.argtypes = [C.c_uint8, C.c_uint8]
        o.fgklo_pd_log10_likelihood.restype = C.c_double
        o.fgklo_pd_log10_likelihood.argtypes = [u8p, u8p, C.c_int64, u8p, u8p, u8p, u8p, u8p,
                                                C.c_int64]
        o.fgklo_batch.restype = None
        o.fgklo_batch.argtypes = [C.c_int64, i64p, i64p, u8p, u8p, u8p, u8p, u8p, i64p, i64p,
                                  u8p, u8p, f64p, C.c_int]
        o.fgklo_log10_likelihood_cont.restype = C.c_double
        o.fgklo_log10_likelihood_cont.argtypes = [u8p, C.c_int64, u8p, C.c_int64, f64p, f64p,
                                                  f64p, C.c_double]
        o.fgklo_log10_likelihood_grad.restype = C.c_double
        o.fgklo_log10_likelihood_grad.argtypes = [u8p, C.c_int64, u8p, C.c_int64, f64p, f64p,
                                                  f64p, C.c_double, f64p, f64p, f64p, f64p]
        _o = o
    return _o


def _u8(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray)):
        return np.frombuffer(bytes(a), dtype=np.uint8).copy()
    return np.ascontiguousarray(a, dtype=np.uint8)


def _p(a, t=C.c_uint8):
    return a.ctypes.data_as(C.POINTER(t))


def log10_likelihood(hap, read) -> float:
    """reference.rs:20-64 via the oracle; `read` is a ReadRef."""
    h = _u8(hap)
    b, q, i, d, g = (_u8(x) for x in (read.bases, read.quals, read.ins_gop, read.del_gop, read.gcp))
    return oracle().fgklo_log10_likelihood(_p(h), len(h), _p(b), _p(q), _p(i), _p(d), _p(g), len(b))


def pd_log10_likelihood(hap, flags, read) -> float:
    """pd_reference.rs:29-118 via the oracle."""
    h, f = _u8(hap), _u8(flags)
    b, q, i, d, g = (_u8(x) for x in (read.bases, read.quals, read.ins_gop, read.del_gop, read.gcp))
    return oracle().fgklo_pd_log10_likelihood(_p(h), _p(f), len(h), _p(b), _p(q), _p(i), _p(d),
                                              _p(g), len(b))


def reference_all(reads, haps) -> np.ndarray:
    return np.array([log10_likelihood(h, r) for r in reads for h in haps])


def pd_reference_all(reads, haps) -> np.ndarray:
    return np.array([pd_log10_likelihood(h.bases, h.flags, r) for r in reads for h in haps])


def batch_reference(data, nthreads: int = 8) -> np.ndarray:
    """Every pair of a BatchData, region by region (out[base_g + r * H + h])."""
    out = np.zeros(data.n_pairs(), dtype=np.float64)
    if len(out) == 0:
        return out
    oracle().fgklo_batch(
        data.n_regions, _p(data.region_reads, C.c_int64), _p(data.read_offsets, C.c_int64),
        _p(data.bases), _p(data.quals), _p(data.ins), _p(data.dels), _p(data.gcp),
        _p(data.region_haps, C.c_int64), _p(data.hap_offsets, C.c_int64), _p(data.hap_bases),
        _p(data.hap_flags) if data.pd else None, _p(out, C.c_double), nthreads)
    return out


def port_batch(data, nthreads: int = 8, force_avx2: bool = False,
               share_suffixes: bool = True) -> np.ndarray:
    """The lane-batched CPU baseline (oracle/port.c) over a plain BatchData, same layout as
    batch_reference; suffix sharing on by default, as in fgkl's JNI configuration."""
    assert not data.pd, "the CPU baseline covers the plain PairHMM"
    out = np.full(data.n_pairs(), np.nan, dtype=np.float64)
    if len(out) == 0:
        return out
    rc = oracle().fgklo_port_batch(
        data.n_regions, _p(data.region_reads, C.c_int64), _p(data.read_offsets, C.c_int64),
        _p(data.bases), _p(data.quals), _p(data.ins), _p(data.dels), _p(data.gcp),
        _p(data.region_haps, C.c_int64), _p(data.hap_offsets, C.c_int64), _p(data.hap_bases),
        _p(out, C.c_double), nthreads, int(force_avx2), int(share_suffixes))
    if rc != 0:
        raise RuntimeError("the CPU baseline needs AVX-512 or AVX2")
    return out


SW_STRATEGIES = {"SOFTCLIP": 0, "INDEL": 1, "LEADING_INDEL": 2, "IGNORE": 3}


def cigar_text(elems) -> str:
    return "".join(f"{int(e) >> 4}{'MIDNSHP=X'[int(e) & 15]}" for e in elems)


def sw_align(ref: bytes, alt: bytes, params, strategy) -> tuple[str, int]:
    """GATK's aligner on one pair (oracle/sw.c): (CIGAR, reference offset)."""
    if isinstance(strategy, str):
        strategy = SW_STRATEGIES[strategy]
    cig = (C.c_uint32 * (len(ref) + len(alt) + 2))()
    n, off = C.c_int32(), C.c_int32()
    p = (C.c_int32 * 4)(*params)
    rb = np.frombuffer(bytes(ref), np.uint8)
    ab = np.frombuffer(bytes(alt), np.uint8)
    rc = oracle().fgklo_sw_align(_p(rb), len(ref), _p(ab), len(alt), p, strategy, C.byref(off),
                                 cig, C.byref(n))
    if rc != 0:
        raise ValueError("empty sequence")
    return cigar_text(cig[: n.value]), off.value


def sw_arrays(pairs):
    """Offset tables and concatenated bytes of (reference, alternate) pairs."""
    ro = np.zeros(len(pairs) + 1, np.int64)
    ao = np.zeros(len(pairs) + 1, np.int64)
    ro[1:] = np.cumsum([len(r) for r, _ in pairs])
    ao[1:] = np.cumsum([len(a) for _, a in pairs])
    ref = np.frombuffer(b"".join(r for r, _ in pairs), np.uint8).copy()
    alt = np.frombuffer(b"".join(a for _, a in pairs), np.uint8).copy()
    return ro, ref, ao, alt


def sw_port_raw(ro, ref, ao, alt, params, strategy, nthreads=1, force_avx2=False):
    """The lane-batched SIMD baseline (oracle/sw_port.c) on arrays: (offsets, cig_start, cigar,
    ncig)."""
    if isinstance(strategy, str):
        strategy = SW_STRATEGIES[strategy]
    n = len(ro) - 1
    caps = (ro[1:] - ro[:-1]) + (ao[1:] - ao[:-1]) + 4
    start = np.zeros(n + 1, np.int64)
    start[1:] = np.cumsum(caps)
    cig = np.zeros(max(int(start[-1]), 1), np.uint32)
    ncig = np.zeros(max(n, 1), np.int32)
    offs = np.zeros(max(n, 1), np.int32)
    w = (C.c_int32 * 4)(*params)
    rc = oracle().fgklo_sw_port_batch(n, _p(ro, C.c_int64), _p(ref), _p(ao, C.c_int64), _p(alt),
                                      w, strategy, _p(offs, C.c_int32), _p(start, C.c_int64),
                                      _p(cig, C.c_uint32), _p(ncig, C.c_int32), nthreads,
                                      int(force_avx2))
    if rc == -1:
        raise RuntimeError("the Smith-Waterman baseline needs AVX2")
    if rc == -2:
        raise ValueError("empty sequence")
    return offs[:n], start, cig, ncig[:n]


def sw_port(pairs, params, strategy, nthreads=1, force_avx2=False) -> list[tuple[str, int]]:
    offs, start, cig, ncig = sw_port_raw(*sw_arrays(pairs), params, strategy, nthreads, force_avx2)
    return [(cigar_text(cig[start[p]:start[p] + ncig[p]]), int(offs[p])) for p in range(len(pairs))]


def sw_port_backend() -> str:
    return oracle().fgklo_sw_port_backend().decode()


def port_backend() -> str:
    return oracle().fgklo_port_backend().decode()


def transitions(ins, dele, gcp) -> list[float]:
    t = (C.c_double * 6)()
    oracle().fgklo_transitions(ins, dele, gcp, t)
    return list(t)


def priors(q) -> tuple[float, float]:
    p = (C.c_double * 2)()
    oracle().fgklo_priors(q, p)
    return p[0], p[1]


def grad_pair(hap: bytes, bases: bytes, e, di, dd, g: float):
    """Gradient oracle of one pair: (log10 L, d/de, d/dd_ins, d/dd_del, d/dg)."""
    o = oracle()
    n = len(bases)
    e, di, dd = (np.ascontiguousarray(x, dtype=np.float64) for x in (e, di, dd))
    ge, gi, gd, gg = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(1)
    h, b = _u8(hap), _u8(bases)
    f = lambda a: a.ctypes.data_as(f64p)
    ll = o.fgklo_log10_likelihood_grad(_p(h), len(h), _p(b), n, f(e), f(di), f(dd), g, f(ge),
                                       f(gi), f(gd), f(gg))
    return ll, ge, gi, gd, float(gg[0])


def cont_pair(hap: bytes, bases: bytes, e, di, dd, g: float) -> float:
    """log10 L of the continuous-parameter forward (the gradient model)."""
    o = oracle()
    e, di, dd = (np.ascontiguousarray(x, dtype=np.float64) for x in (e, di, dd))
    h, b = _u8(hap), _u8(bases)
    f = lambda a: a.ctypes.data_as(f64p)
    return o.fgklo_log10_likelihood_cont(_p(h), len(h), _p(b), len(b), f(e), f(di), f(dd), g)


def grad_batch(data, e, di, dd, g: float, weights=None):
    """The gradient API's outputs from the oracle: per-pair log10 L and the gradients of
    sum_p w_p log10 L_p per read base and for g."""
    n_pairs = data.n_pairs()
    ll = np.zeros(n_pairs)
    ge, gi, gd = np.zeros(len(data.bases)), np.zeros(len(data.bases)), np.zeros(len(data.bases))
    gg = 0.0
    base = data.out_base()
    for gi_ in range(data.n_regions):
        r0, r1 = int(data.region_reads[gi_]), int(data.region_reads[gi_ + 1])
        h0, h1 = int(data.region_haps[gi_]), int(data.region_haps[gi_ + 1])
        H = h1 - h0
        for r in range(r0, r1):
            lo, hi = int(data.read_offsets[r]), int(data.read_offsets[r + 1])
            for k in range(H):
                p = int(base[gi_]) + (r - r0) * H + k
                w = 1.0 if weights is None else float(weights[p])
                l_, a, b, c, d = grad_pair(data.hap(h0 + k), bytes(data.bases[lo:hi]),
                                           e[lo:hi], di[lo:hi], dd[lo:hi], g)
                ll[p] = l_
                ge[lo:hi] += w * a
                gi[lo:hi] += w * b
                gd[lo:hi] += w * c
                gg += w * d
    return ll, ge, gi, gd, gg
