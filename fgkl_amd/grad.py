"""Gradients of PairHMM log10 likelihoods on the GPU, for calibrating the per-base error model.

The model is fgkl's FP64 forward recurrence (crates/pairhmm/src/reference.rs:20-64) with the
Phred-derived probabilities made continuous (crates/pairhmm/src/model.rs:66-84): per read base
an error probability e (priors 1 - e / e/3), gap-open probabilities d_ins and d_del
(M->M = 1 - d_ins - d_del), and one global gap-continuation probability g.  The reference has no
gradient; ``GradProblem.backward`` runs the HIP adjoint kernel (reverse anti-diagonal sweep over
checkpointed forward blocks) and ``log10_likelihoods`` wraps it as a ``torch.autograd.Function``.
``precision="double"`` (the default) runs the forward and the adjoint in FP64;
``precision="float"`` runs them in FP32 with FP64 accumulation of the gradient terms and an FP64
rerun of every pair whose FP32 likelihood falls below 1e-28.

Arrays cross the boundary in host memory (NumPy, or CPU torch tensors in the Function): the
engine links ROCm's HIP runtime, and torch's wheel carries a HIP runtime of its own; one process
can drive the GPU through only one of them, so torch never touches the device here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _ptr
from ._lib import lib

__all__ = ["GradProblem", "probabilities", "PairHmmLikelihood", "log10_likelihoods"]


def probabilities(data):
    """Per-base (e, d_ins, d_del) and the global g of a BatchData, from its Phred scores as
    model.rs:53-84 maps them (scores clamped to 254).  g comes from the gap-continuation score,
    which must be the same for every base."""
    def p(q):
        return 10.0 ** (-np.minimum(q.astype(np.float64), 254.0) / 10.0)

    gcp = np.unique(data.gcp)
    if len(gcp) > 1:
        raise ValueError("the gradient model has one global gap-continuation probability")
    g = float(p(gcp)[0]) if len(gcp) else 0.1
    return p(data.quals), p(data.ins), p(data.dels), g


class GradProblem:
    """Regions of reads and haplotypes resident on the GPU (include/fgkl.h, fgkl_grad_*).

    Pairs are ordered region after region, ``[r * H + h]`` inside each, as the batch engine
    orders them; per-base arrays follow the reads' bases.
    """

    def __init__(self, data, precision: str = "double"):
        if precision not in ("double", "float"):
            raise ValueError("precision must be 'double' or 'float'")
        h = C.c_void_p()
        d = data
        _check(lib().fgkl_grad_create(
            C.byref(h), d.n_regions, _ptr(d.region_reads, C.c_int64),
            _ptr(d.region_haps, C.c_int64), _ptr(d.read_offsets, C.c_int64),
            _ptr(d.bases, C.c_uint8), _ptr(d.hap_offsets, C.c_int64),
            _ptr(d.hap_bases, C.c_uint8)))
        self._h = h
        self.precision = precision
        if precision == "float":
            _check(lib().fgkl_grad_set_precision(h, 1))
        self.n_pairs = int(lib().fgkl_grad_pairs(h))
        self.n_read_bases = int(lib().fgkl_grad_read_bases(h))

    def _params(self, err, ins, dels):
        out = []
        for name, x in (("err", err), ("ins", ins), ("dels", dels)):
            x = np.ascontiguousarray(x, dtype=np.float64)
            if x.shape != (self.n_read_bases,):
                raise ValueError(f"{name} must hold {self.n_read_bases} values")
            out.append(x)
        return out

    def forward(self, err, ins, dels, gcp: float, timing: bool = False):
        """log10 likelihood of every pair; with timing, also the kernels' device time in ms."""
        e, i, d = self._params(err, ins, dels)
        out = np.empty(self.n_pairs, dtype=np.float64)
        ms = C.c_double(0.0)
        _check(lib().fgkl_grad_forward(self._h, _ptr(e, C.c_double), _ptr(i, C.c_double),
                                       _ptr(d, C.c_double), float(gcp), _ptr(out, C.c_double),
                                       C.byref(ms)))
        return (out, ms.value) if timing else out

    def backward(self, err, ins, dels, gcp: float, weights=None, timing: bool = False):
        """(log10 L, d/d e, d/d d_ins, d/d d_del, d/d g) of sum_p w_p log10 L_p; with timing,
        also the kernels' device time in ms."""
        e, i, d = self._params(err, ins, dels)
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float64)
            if w.shape != (self.n_pairs,):
                raise ValueError(f"weights must hold {self.n_pairs} values")
        ll = np.empty(self.n_pairs, dtype=np.float64)
        ge, gi, gd = (np.empty(self.n_read_bases, dtype=np.float64) for _ in range(3))
        gg = C.c_double(0.0)
        ms = C.c_double(0.0)
        _check(lib().fgkl_grad_backward(
            self._h, _ptr(e, C.c_double), _ptr(i, C.c_double), _ptr(d, C.c_double), float(gcp),
            None if w is None else _ptr(w, C.c_double), _ptr(ll, C.c_double),
            _ptr(ge, C.c_double), _ptr(gi, C.c_double), _ptr(gd, C.c_double), C.byref(gg),
            C.byref(ms)))
        out = (ll, ge, gi, gd, gg.value)
        return out + (ms.value,) if timing else out

    def close(self):
        if self._h:
            lib().fgkl_grad_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _function():
    import torch

    class PairHmmLikelihood(torch.autograd.Function):
        """log10 L of every pair of a GradProblem as a differentiable function of the per-base
        probabilities and g (CPU float64 tensors; g a one-element tensor).  The backward pass
        runs the adjoint kernel with grad_output as the pair weights."""

        @staticmethod
        def forward(ctx, err, ins, dels, gcp, problem):
            ctx.problem = problem
            ctx.save_for_backward(err, ins, dels, gcp)
            np_ = lambda t: t.detach().cpu().numpy()
            ll = problem.forward(np_(err), np_(ins), np_(dels), float(gcp.reshape(-1)[0]))
            return torch.from_numpy(ll)

        @staticmethod
        def backward(ctx, grad_out):
            err, ins, dels, gcp = ctx.saved_tensors
            np_ = lambda t: t.detach().cpu().numpy()
            _, ge, gi, gd, gg = ctx.problem.backward(
                np_(err), np_(ins), np_(dels), float(gcp.reshape(-1)[0]),
                weights=np_(grad_out.to(torch.float64)))
            t = lambda a: torch.from_numpy(a)
            return t(ge), t(gi), t(gd), torch.tensor([gg], dtype=torch.float64).reshape(gcp.shape), None

    return PairHmmLikelihood


_FN = None


def __getattr__(name):
    global _FN
    if name == "PairHmmLikelihood":
        if _FN is None:
            _FN = _function()
        return _FN
    raise AttributeError(name)


def log10_likelihoods(problem: GradProblem, err, ins, dels, gcp):
    """Differentiable log10 likelihoods of every pair of `problem` (CPU float64 tensors)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(err, ins, dels, gcp, problem)
