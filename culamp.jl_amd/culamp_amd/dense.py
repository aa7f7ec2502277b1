"""Dense cost-matrix passes (the reference's plain dense GPU path,
src/solvers/MirrorProx.jl:236-333), kept as the HBM-streaming comparator of
the on-the-fly passes: the row log-sum-exp and column marginal of LAMP
computed from a stored n x m matrix W instead of from the points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FP64, MIXED, Context, DeviceArray, check, load
from .solvers import default_context


class DenseProblem:
    """W (n x m) and its transpose resident on the device, FP64 or FP32
    ("mixed": FP32 costs and exponents, FP64 sums)."""

    def __init__(self, W, precision: str = "fp64", ctx: Context | None = None):
        self.ctx = ctx or default_context()
        w = np.ascontiguousarray(W, dtype=np.float64)
        if w.ndim != 2:
            raise ValueError("W must be a 2-D array")
        self.n, self.m = w.shape
        code = {"fp64": FP64, "mixed": MIXED}[precision]
        h = C.c_void_p()
        check(load().culamp_dense_create(self.ctx.handle, w.ctypes.data, self.n, self.m, code,
                                         C.byref(h)))
        self.handle = h
        self.ctx.adopt(self)

    def close(self) -> None:
        if getattr(self, "handle", None) and self.ctx.handle:
            load().culamp_dense_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lse_rows(self, out: DeviceArray, theta: DeviceArray, reg: float, st: float,
                 winf: float) -> None:
        """warp_logsumexp_fused!: out_i = LSE_j(-(st W_ij/(2 W_inf) + theta_j)/reg)."""
        check(load().culamp_dense_lse_rows(self.ctx.handle, self.handle, theta.ptr, float(reg),
                                           float(st), float(winf), out.ptr))

    def col_marginal(self, out: DeviceArray, r: DeviceArray, theta: DeviceArray,
                     logZ: DeviceArray, reg: float, st: float, winf: float) -> None:
        """residual_c!: out_j = sum_i r_i exp(-(st W_ij/(2 W_inf) + theta_j)/reg - logZ_i)."""
        check(load().culamp_dense_col_marginal(self.ctx.handle, self.handle, r.ptr, theta.ptr,
                                               logZ.ptr, float(reg), float(st), float(winf),
                                               out.ptr))


__all__ = ["DenseProblem"]
