"""Dense cost-matrix passes (the reference's plain dense GPU path,
src/solvers/MirrorProx.jl:236-333), kept as the HBM-streaming comparator of
the on-the-fly passes: the row log-sum-exp and column marginal of LAMP
computed from a stored n x m matrix W instead of from the points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (FP64, LOG_FN, MIXED, Context, DeviceArray, check, culamp_result, load, ptr)
from .args import EOTArgs
from .solvers import _stdout_log, default_context


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

    def lamp(self, r, c, args: EOTArgs, frequency: int = 50, theta0=None, plan: bool = True,
             log=_stdout_log):
        """LAMP(r, c, W::CuArray, args, frequency; _θ) (MirrorProx.jl:397-503)
        -> (plan or None, theta, info)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        if r.size != self.n or c.size != self.m:
            raise ValueError("r and c must match W's rows and columns")
        t0 = None if theta0 is None else np.ascontiguousarray(theta0, dtype=np.float64)
        theta = np.empty(self.m)
        P = np.empty((self.n, self.m)) if plan else None
        res, a = culamp_result(), args.to_c()
        cb = LOG_FN(0) if log is None else LOG_FN(lambda _u, line: log(line.decode()))
        check(load().culamp_dense_lamp_solve(self.ctx.handle, self.handle, ptr(r), ptr(c),
                                             C.byref(a), int(frequency), ptr(t0), cb, None,
                                             ptr(theta), ptr(P), C.byref(res)))
        return P, theta, res.as_dict()


def dense_lamp(r, c, W, args: EOTArgs, frequency: int = 50, *, theta0=None,
               precision: str = "fp64", log=_stdout_log, ctx: Context | None = None):
    """LAMP(r, c, W, args, frequency) on the device (MirrorProx.jl:397-503)
    -> (r .* p, θ)."""
    dm = DenseProblem(W, precision, ctx)
    try:
        P, theta, _ = dm.lamp(r, c, args, frequency, theta0, True, log)
    finally:
        dm.close()
    return P, theta


__all__ = ["DenseProblem", "dense_lamp"]
