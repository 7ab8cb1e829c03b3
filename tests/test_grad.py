"""The gradient engine: the HIP adjoint kernel (include/fgkl.h, fgkl_grad_*) checked against

* the gradient oracle (fgklo_log10_likelihood_grad: reverse mode over the full FP64 matrices,
  itself pinned by central differences in test_oracle.py), per read base and for g, on regions
  whose reads cross the kernel's 64-row strips and whose haplotypes cross its checkpoint blocks;
* FP64 central differences with h = 1e-6 of the GPU's own forward, on 64 seeded pairs of the
  adjoint workload (reads U(50,250) bp, up to 200 per region, up to 8 haplotypes);
* the continuous-parameter forward of the oracle (log10 L per pair);
* linearity in the pair weights, bitwise determinism, and the torch.autograd.Function (CPU
  tensors: torch's wheel carries its own HIP runtime, so it stays off the device here);
* the FP32 mode (forward, tape and adjoint in FP32, FP64 accumulation, FP64 rerun of pairs below
  1e-28) against the same oracle with FP32 bounds, and its rescue: reads lost against every
  haplotype get the FP64 mode's bits, reads lost against some get the FP64 terms added.

The reference computes no gradient; its model is the FP64 recurrence of reference.rs:20-64.
"""
from __future__ import annotations

import numpy as np
import pytest

import fgkl_amd as F
from fgkl_amd.synthetic import Region, Rng, batch_from_regions, workload

import _oracle as O


def _grad_module():
    from fgkl_amd import grad
    return grad


# ---- CPU --------------------------------------------------------------------------------------
def test_probabilities_follow_the_phred_map():  # model.rs:53-84
    G = _grad_module()
    d = Region.generate(3, 4, 30, 2, 50)
    e, di, dd, g = G.probabilities(d)
    ep = O.oracle().fgklo_error_prob
    for got, q in ((e, d.quals), (di, d.ins), (dd, d.dels)):
        assert np.allclose(got, [ep(int(x)) for x in q], rtol=1e-15, atol=0)
    assert abs(g - ep(10)) < 1e-15


def test_create_validates_before_any_device_work():
    G = _grad_module()
    rd = F.ReadRef(b"", [], [], [], [])
    d = batch_from_regions([([rd], [b"ACGT"])])
    with pytest.raises(F.InvalidArgument, match="read 0 has no bases"):
        G.GradProblem(d)
    rd = F.ReadRef(b"ACGT", [30] * 4, [45] * 4, [45] * 4, [10] * 4)
    d = batch_from_regions([([rd], [b"ACGT", b""])])
    with pytest.raises(F.InvalidArgument, match="haplotype 1 has no bases"):
        G.GradProblem(d)


@pytest.mark.skipif(F.device_count() > 0, reason="checks the no-GPU host")
def test_without_a_gpu_the_gradient_engine_fails_loudly():
    G = _grad_module()
    with pytest.raises(F.NativeError, match="no HIP device"):
        G.GradProblem(Region.generate(3, 4, 30, 2, 50))


# ---- GPU --------------------------------------------------------------------------------------
def _gpu_grad(data, weights=None, g=None, precision="double"):
    G = _grad_module()
    e, di, dd, g0 = G.probabilities(data)
    g = g0 if g is None else g
    P = G.GradProblem(data, precision=precision)
    ll, ge, gi, gd, gg = P.backward(e, di, dd, g, weights=weights)
    P.close()
    return (ll, ge, gi, gd, np.array([gg])), (e, di, dd, g)


def _read(rng: Rng, n: int, alphabet=b"ACGT") -> F.ReadRef:
    bases = bytes(alphabet[rng.below(len(alphabet))] for _ in range(n))
    quals = [2 + rng.below(40) for _ in range(n)]
    ins = [10 + rng.below(36) for _ in range(n)]
    dels = [10 + rng.below(36) for _ in range(n)]
    return F.ReadRef(bases, quals, ins, dels, [10] * n)


def _mutated(rng: Rng, hap: bytes, n: int) -> bytes:
    """A read-sized window of `hap` with a few substitutions, so likelihoods are realistic."""
    start = rng.below(max(1, len(hap) - n + 1))
    s = bytearray(hap[start:start + n])
    while len(s) < n:
        s.append(b"ACGT"[rng.below(4)])
    for _ in range(max(1, n // 30)):
        s[rng.below(n)] = b"ACGT"[rng.below(4)]
    return bytes(s)


# FP32 mode bounds.  log10 L: the engine's FP32 bound everywhere (1e-4 absolute).  Gradients:
# every forward and adjoint value is a product along a path of at most n + m <= 700 FP32
# roundings (2^-24 each), so a term's relative error is at most ~700 * 6e-8 = 4.2e-5, and the sum
# of terms of both signs is held to that times the largest component of the read set.
F32_LOG10_TOL = 1e-4
F32_RTOL = 4.2e-5


def _check_against_oracle(data, got, params, weights=None, rtol=1e-9, ll_tol=None):
    e, di, dd, g = params
    ll, ge, gi, gd, gg = got
    rl, re_, ri, rd, rg = O.grad_batch(data, e, di, dd, g, weights)
    if ll_tol is None:
        assert np.all(np.abs(ll - rl) <= 1e-12 * np.abs(rl) + 1e-12), np.max(np.abs(ll - rl))
    else:
        assert np.all(np.abs(ll - rl) <= ll_tol), np.max(np.abs(ll - rl))
    for name, a, b in (("e", ge, re_), ("d_ins", gi, ri), ("d_del", gd, rd)):
        # FP64 against FP64 in another summation order; terms of both signs are summed, so the
        # bound is relative to the largest gradient component of the read set.
        tol = rtol * np.abs(b).max() + 1e-12
        err = np.abs(a - b).max()
        assert err <= tol, (name, err, tol)
    assert abs(gg[0] - rg) <= rtol * max(1.0, abs(rg)), (gg[0], rg)


@pytest.mark.gpu
def test_dpp_wave_shl_moves_each_lane_down(gpu):
    import ctypes as C
    out = (C.c_int32 * 64)()
    assert F._lib.lib().fgkl_selftest_dpp_shl(out) == 0
    assert list(out) == [1000 + i + 1 for i in range(63)] + [-1]


def _bounds(precision):
    return {} if precision == "double" else {"rtol": F32_RTOL, "ll_tol": F32_LOG10_TOL}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_gradient_matches_the_oracle_on_seeded_regions(gpu, precision):
    d = Region.generate(11, 12, 150, 4, 220)   # 3 strips per read, 30 (FP64) / 15 (FP32) blocks
    got, params = _gpu_grad(d, precision=precision)
    _check_against_oracle(d, got, params, **_bounds(precision))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_gradient_matches_the_oracle_across_strip_and_block_edges(gpu, precision):
    rng = Rng(77)
    haps = [bytes(b"ACGT"[rng.below(4)] for _ in range(m)) for m in (1, 2, 7, 8, 9, 64, 151)]
    reads = []
    for n in (1, 2, 63, 64, 65, 127, 128, 129, 200):
        reads.append(_read(rng, n) if n < 8 else F.ReadRef(
            _mutated(rng, haps[-1], n), *(getattr(_read(rng, n), f)
                                          for f in ("quals", "ins_gop", "del_gop", "gcp"))))
    d = batch_from_regions([(reads, haps)])
    got, params = _gpu_grad(d, precision=precision)
    _check_against_oracle(d, got, params, **_bounds(precision))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_gradient_with_n_bases_matches_the_oracle(gpu, precision):
    rng = Rng(91)
    hap = bytes(b"ACGTN"[rng.below(5)] for _ in range(90))
    reads = [F.ReadRef(_mutated(rng, hap, n).replace(b"A", b"N", 2),
                       *(getattr(_read(rng, n), f) for f in ("quals", "ins_gop", "del_gop", "gcp")))
             for n in (40, 70, 100)]
    d = batch_from_regions([(reads, [hap, hap[5:], b"NNNN" + hap])])
    got, params = _gpu_grad(d, precision=precision)
    _check_against_oracle(d, got, params, **_bounds(precision))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_forward_matches_the_continuous_oracle(gpu, precision):
    G = _grad_module()
    d = Region.generate(21, 9, 120, 5, 260)
    e, di, dd, g = G.probabilities(d)
    P = G.GradProblem(d, precision=precision)
    ll = P.forward(e, di, dd, g)
    base = d.out_base()
    H = d.n_haps
    for r in range(d.n_reads):
        lo, hi = int(d.read_offsets[r]), int(d.read_offsets[r + 1])
        for k in range(H):
            ref = O.cont_pair(d.hap(k), bytes(d.bases[lo:hi]), e[lo:hi], di[lo:hi], dd[lo:hi], g)
            tol = 1e-12 * abs(ref) if precision == "double" else F32_LOG10_TOL
            assert abs(ll[int(base[0]) + r * H + k] - ref) <= tol, (r, k)


def _fd_pairs(n_pairs=64, seed=3):
    data = workload(4, seed=seed, n_regions=16)
    rng = Rng(seed * 7 + 1)
    picks = []
    while len(picks) < n_pairs:
        g = rng.below(data.n_regions)
        r0, r1 = int(data.region_reads[g]), int(data.region_reads[g + 1])
        h0, h1 = int(data.region_haps[g]), int(data.region_haps[g + 1])
        if r1 == r0:
            continue
        picks.append((r0 + rng.below(r1 - r0), h0 + rng.below(h1 - h0)))
    return data, picks


@pytest.mark.gpu
def test_gradient_matches_central_differences_on_64_pairs(gpu):
    """FP64 central differences, h = 1e-6, of the GPU forward (one region per perturbed pair).

    The difference quotient's own error is its truncation term h^2 f'''/6 (f = log10 L), which
    for a probability p is of relative size (h/p)^2: p >= 3.2e-5 here (Phred <= 45), so up to
    1e-3 for the smallest gap-open probabilities; plus rounding of ~1e-16 |f| / h ~ 1e-8.  The
    bound below is that estimate per parameter; the same gradients match the oracle to 1e-9.
    """
    G = _grad_module()
    data, picks = _fd_pairs()
    single = batch_from_regions([([data.read(r)], [data.hap(h)]) for r, h in picks])
    (ll, ge, gi, gd, gg), (e, di, dd, g) = _gpu_grad(single)
    # The same pairs through the oracle.
    _check_against_oracle(single, (ll, ge, gi, gd, gg), (e, di, dd, g))
    h = 1e-6
    rng = Rng(19)
    regions, plan = [], []
    for p, (r, hp) in enumerate(picks):
        n = int(single.read_offsets[p + 1] - single.read_offsets[p])
        for pos in sorted({0, n - 1, rng.below(n)}):
            for t in range(3):
                for sgn in (1.0, -1.0):
                    regions.append(([data.read(r)], [data.hap(hp)]))
                    plan.append((p, t, pos, sgn))
    fdb = batch_from_regions(regions)
    E, DI, DD, _ = G.probabilities(fdb)
    for j, (p, t, pos, sgn) in enumerate(plan):
        (E, DI, DD)[t][int(fdb.read_offsets[j]) + pos] += sgn * h
    P = G.GradProblem(fdb)
    f = P.forward(E, DI, DD, g)
    worst = 0.0
    for j in range(0, len(plan), 2):
        p, t, pos, _ = plan[j]
        fd = (f[j] - f[j + 1]) / (2 * h)
        idx = int(single.read_offsets[p]) + pos
        val = (e, di, dd)[t][idx]
        grad = (ge, gi, gd)[t][idx]
        bound = 1e-7 + 2.0 * (h / val) ** 2 * max(abs(grad), abs(fd)) + 1e-7 * abs(grad)
        worst = max(worst, abs(fd - grad) / bound)
        assert abs(fd - grad) <= bound, (p, t, pos, fd, grad, val)
    # g is shared by every pair: one difference of the summed log likelihoods.
    P1 = G.GradProblem(single)
    s = lambda gv: P1.forward(e, di, dd, gv).sum()
    fd = (s(g + h) - s(g - h)) / (2 * h)
    assert abs(fd - gg[0]) <= 1e-6 * max(1.0, abs(fd)) + 2.0 * (h / g) ** 2 * abs(fd), (fd, gg[0])
    print(f"worst |fd - grad| / bound = {worst:.3g}")


@pytest.mark.gpu
def test_weights_scale_the_gradient_and_zero_weights_remove_pairs(gpu):
    d = Region.generate(5, 10, 90, 3, 150)
    (ll1, ge1, gi1, gd1, gg1), params = _gpu_grad(d)
    (ll2, ge2, gi2, gd2, gg2), _ = _gpu_grad(d, weights=np.full(d.n_pairs(), 2.0))
    for a, b in ((ge1, ge2), (gi1, gi2), (gd1, gd2), (gg1, gg2)):
        # Doubling the seed doubles every adjoint exactly except those in the subnormal range
        # (adjoints are ~1/L per 2^1020 of forward mass), so the sums agree to rounding.
        assert np.abs(2.0 * a - b).max() <= 1e-13 * np.abs(b).max()
    assert np.array_equal(ll1, ll2)
    (_, ge0, gi0, gd0, gg0), _ = _gpu_grad(d, weights=np.zeros(d.n_pairs()))
    assert not ge0.any() and not gi0.any() and not gd0.any() and gg0[0] == 0.0
    w = np.random.default_rng(4).normal(size=d.n_pairs())
    got, params = _gpu_grad(d, weights=w)
    _check_against_oracle(d, got, params, weights=w)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_backward_is_bitwise_deterministic(gpu, precision):
    d = workload(4, seed=9, n_regions=6)
    a, _ = _gpu_grad(d, precision=precision)
    b, _ = _gpu_grad(d, precision=precision)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "float"])
def test_a_paired_read_matches_its_lone_run(gpu, precision):
    """Two reads of a region whose last strips fit in one run as one item (the second read's rows
    follow the first's, starting in the lane below its last row).  Each read gets its lone-run
    log10 L bits (the forward is unchanged by the pairing) and its lone-run gradients to rounding
    (a paired item runs more checkpoint blocks in the shared strip, and the reverse sweep's
    partial sums can round differently there: measured at most 4e-16 of the largest component in
    FP64).  Pairings: remainders mod 64 of 22 + 22, 63 + 1 (the second read starting in lane 63),
    8 + 56 (filling the strip), 20 + 44 and 20 + 30, and a 1-row second read."""
    rng = Rng(5)
    hap = bytes(b"ACGT"[rng.below(4)] for _ in range(300))
    hap2 = hap[:140] + b"T" + hap[141:]
    rtol = 1e-14 if precision == "double" else F32_RTOL
    for na, nb in ((150, 150), (127, 65), (200, 120), (212, 108), (212, 30), (86, 1), (63, 1)):
        ra = F.ReadRef(_mutated(rng, hap, na), *(getattr(_read(rng, na), f)
                                                 for f in ("quals", "ins_gop", "del_gop", "gcp")))
        rb = F.ReadRef(_mutated(rng, hap, nb), *(getattr(_read(rng, nb), f)
                                                 for f in ("quals", "ins_gop", "del_gop", "gcp")))
        both = batch_from_regions([([ra, rb], [hap, hap2])])
        (l2, e2, i2, d2, g2), params = _gpu_grad(both, precision=precision)
        _check_against_oracle(both, (l2, e2, i2, d2, g2), params, **_bounds(precision))
        for k, rd in enumerate((ra, rb)):
            alone = batch_from_regions([([rd], [hap, hap2])])
            (l1, e1, i1, d1, _), _p = _gpu_grad(alone, precision=precision)
            lo, hi = int(both.read_offsets[k]), int(both.read_offsets[k + 1])
            assert np.array_equal(l1, l2[2 * k:2 * k + 2]), (na, nb, k)
            for x, y in ((e1, e2), (i1, i2), (d1, d2)):
                assert np.abs(x - y[lo:hi]).max() <= rtol * np.abs(x).max(), (na, nb, k)


def _riddled(rng: Rng, hap: bytes, start: int, n: int, every: int) -> F.ReadRef:
    """A window of `hap` with a substitution every `every` bases at Phred 35: log10 L falls by
    about 4 per substitution, below the FP32 threshold (scaled L < 1e-28, log10 L < -64.1) for
    n / every > 20, while staying far inside FP64's range."""
    b = bytearray(hap[start:start + n])
    for j in range(0, n, every):
        b[j] = ord("A") if b[j] != ord("A") else ord("C")
    return F.ReadRef(bytes(b), [35] * n, [40] * n, [40] * n, [10] * n)


@pytest.mark.gpu
def test_float_mode_reruns_lost_pairs_in_double(gpu):
    """Reads lost against every haplotype: the FP32 pass leaves their gradients at zero and the
    FP64 rerun adds every term, so their log10 L are the FP64 mode's bits and their gradients
    the FP64 mode's to FP64 rounding (the rerun runs each read alone, where the FP64 mode may
    pair it with another read).  A read lost against one haplotype only gets that pair's FP64
    terms added to its FP32 ones and matches the oracle at the FP32 bounds."""
    rng = Rng(41)
    hap = bytes(b"ACGT"[rng.below(4)] for _ in range(420))
    alt = bytearray(hap)
    for j in range(100, 300, 9):  # alt differs from hap every 9 bases over the span 100..300
        alt[j] = ord("G") if alt[j] != ord("G") else ord("T")
    alt = bytes(alt)
    ok = [F.ReadRef(_mutated(rng, hap, n), *(getattr(_read(rng, n), f)
                                             for f in ("quals", "ins_gop", "del_gop", "gcp")))
          for n in (90, 150)]
    lost_both = [_riddled(rng, hap, 20, 200, 8), _riddled(rng, hap, 150, 180, 7)]
    half = F.ReadRef(hap[110:290], [35] * 180, [40] * 180, [40] * 180, [10] * 180)
    reads = ok + lost_both + [half]
    d = batch_from_regions([(reads, [hap, alt])])
    (l64, e64, i64, d64, g64), params = _gpu_grad(d)
    (l32, e32, i32, d32, g32), _ = _gpu_grad(d, precision="float")
    assert np.all(np.isfinite(l64))
    H = 2
    assert np.all(l64[2 * H:4 * H] < -70), l64  # both riddled reads, both haplotypes
    assert l64[4 * H] > -10 and l64[4 * H + 1] < -70, l64[4 * H:]  # half: lost against alt only
    assert np.array_equal(l32[2 * H:5 * H][[0, 1, 2, 3, 5]], l64[2 * H:5 * H][[0, 1, 2, 3, 5]])
    for r in (2, 3):
        lo, hi = int(d.read_offsets[r]), int(d.read_offsets[r + 1])
        for a, b in ((e32, e64), (i32, i64), (d32, d64)):
            assert np.abs(a[lo:hi] - b[lo:hi]).max() <= 1e-14 * np.abs(b[lo:hi]).max(), r
    _check_against_oracle(d, (l32, e32, i32, d32, g32), params, rtol=F32_RTOL,
                          ll_tol=F32_LOG10_TOL)


@pytest.mark.gpu
def test_autograd_function_runs_the_adjoint(gpu):
    import torch
    G = _grad_module()
    d = Region.generate(8, 7, 80, 3, 140)
    e, di, dd, g = G.probabilities(d)
    P = G.GradProblem(d)
    t = lambda x: torch.tensor(np.asarray(x, np.float64), requires_grad=True)
    E, DI, DD, GG = t(e), t(di), t(dd), t([g])
    ll = G.log10_likelihoods(P, E, DI, DD, GG)
    assert np.array_equal(ll.detach().numpy(), P.forward(e, di, dd, g))
    w = torch.linspace(0.5, 1.5, ll.numel(), dtype=torch.float64)
    (ll * w).sum().backward()
    _, ge, gi, gd, gg = P.backward(e, di, dd, g, weights=w.numpy())
    for a, b in ((E.grad, ge), (DI.grad, gi), (DD.grad, gd), (GG.grad, np.array([gg]))):
        assert np.array_equal(a.numpy(), b)
