"""Sinkhorn divergence with gradients (libculamp ``culamp_sinkhorn_divergence``).

The reference has no gradients; this is the differentiable wrapper a Julia
caller would expose as a ChainRulesCore ``rrule``.  Here the host language is
Python, so the same role is played by :class:`SinkhornDivergence`, a
``torch.autograd.Function`` (forward: the value; backward: the cotangent
times the gradients computed on the device).

Cost: squared Euclidean, C_ij = |x_i - y_j|^2, fixed eps (default
eta_p * max_ij C_ij of the (x, y) pair, held fixed under differentiation).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import (FP64, MIXED, Context, check, culamp_divergence_opts, culamp_result, load, ptr)
from .solvers import default_context


def divergence_opts(eps: float | None = None, eta_p: float = 1e-2, tol: float = 1e-6,
                    itermax: int = 100000, frequency: int = 10, unroll: int = 64,
                    checkpoint: int = 16, precision: str = "mixed",
                    polish_tol: float = 1e-9, tmax: float = math.inf) -> culamp_divergence_opts:
    o = culamp_divergence_opts()
    load().culamp_divergence_opts_default(C.byref(o))
    o.eps = 0.0 if eps is None else float(eps)
    o.eta_p = float(eta_p)
    o.tol = float(tol)
    o.itermax = int(itermax)
    o.frequency = int(frequency)
    o.unroll = int(unroll)
    o.checkpoint = int(checkpoint)
    o.precision = {"fp64": FP64, "mixed": MIXED}[precision]
    o.polish_tol = float(polish_tol)
    o.tmax = float(tmax)
    return o


def _pts(a, name):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"{name} must be an (n, 3) array")
    return a


def _w(a, n, name):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size != n:
        raise ValueError(f"{name} must have {n} entries")
    return a


def sinkhorn_divergence(x, a, y, b, *, grad: bool = True, ctx: Context | None = None,
                        return_terms: bool = False, **opts):
    """S(a, x; b, y) and, with ``grad``, (dS/dx, dS/da, dS/dy, dS/db).

    Returns ``value`` or ``(value, gx, ga, gy, gb)``; with ``return_terms``
    also the three OT solves' summaries."""
    ctx = ctx or default_context()
    x, y = _pts(x, "x"), _pts(y, "y")
    n, m = x.shape[0], y.shape[0]
    a, b = _w(a, n, "a"), _w(b, m, "b")
    o = divergence_opts(**opts)
    val = C.c_double()
    terms = (culamp_result * 3)()
    gx = np.empty((n, 3)) if grad else None
    ga = np.empty(n) if grad else None
    gy = np.empty((m, 3)) if grad else None
    gb = np.empty(m) if grad else None
    check(load().culamp_sinkhorn_divergence(ctx.handle, ptr(x), n, ptr(a), ptr(y), m, ptr(b),
                                            C.byref(o), C.byref(val), ptr(gx), ptr(ga), ptr(gy),
                                            ptr(gb), terms))
    out = (val.value, gx, ga, gy, gb) if grad else val.value
    if return_terms:
        return out, [t.as_dict() for t in terms]
    return out


def _autograd_function():
    import torch

    class SinkhornDivergence(torch.autograd.Function):
        """S = SinkhornDivergence.apply(x, a, y, b, opts) on CPU float64
        tensors; backward returns the cotangent times dS/dx, dS/da, dS/dy,
        dS/db (computed together with the value on the device)."""

        @staticmethod
        def forward(fctx, x, a, y, b, opts=None):
            opts = dict(opts or {})
            ctx = opts.pop("ctx", None)
            val, gx, ga, gy, gb = sinkhorn_divergence(
                x.detach().cpu().numpy(), a.detach().cpu().numpy(), y.detach().cpu().numpy(),
                b.detach().cpu().numpy(), grad=True, ctx=ctx, **opts)
            fctx.save_for_backward(*(torch.from_numpy(g) for g in (gx, ga, gy, gb)))
            return torch.tensor(val, dtype=torch.float64)

        @staticmethod
        def backward(fctx, dS):
            gx, ga, gy, gb = fctx.saved_tensors
            return dS * gx, dS * ga, dS * gy, dS * gb, None

    return SinkhornDivergence


_SD = None


def __getattr__(name):
    # built on first use so that importing the package does not import torch
    global _SD
    if name == "SinkhornDivergence":
        if _SD is None:
            _SD = _autograd_function()
        return _SD
    raise AttributeError(name)
