"""Cell-similarity (feature-cost) solvers with the reference's names,
arguments and returns.

* ``extragradient_cellsim``      -- src/solvers/MirrorProxCellSim.jl:94-188
* ``sinkhorn_cellsim``           -- src/solvers/SinkhornCellSim.jl:240-310
* ``annealed_sinkhorn_cellsim``  -- src/solvers/AnnealedSinkhornCellSim.jl:1-63

``data`` is a (features x cells) array -- the reference's ``data`` matrix,
whose rows are the transported points -- passed to libculamp column-major.
``metric`` is a string or symbol name of src/CellUtilities.jl:5 ("pearson" is
"correlation"); ``scale`` is the reference's ``kernel.scale``
(``_cost_scale``, CellUtilities.jl:78-84).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import LOG_FN, Context, check, culamp_result, load, ptr
from .args import EOTArgs
from .solvers import _stdout_log, _vec, default_context

CELL_COST_METRICS = ("l1", "l2", "cosine", "pearson", "correlation")
_METRIC_ID = {"l1": 0, "l2": 1, "cosine": 2, "correlation": 3}


def normalize_cell_cost_metric(metric) -> str:
    """normalize_cell_cost_metric (src/CellUtilities.jl:25-30)."""
    m = str(metric).lstrip(":").lower()
    if m == "pearson":
        return "correlation"
    if m not in CELL_COST_METRICS:
        raise ValueError(f"Unsupported cell cost metric: {metric}")
    return m


def cost_scale(metric) -> float:
    """_cost_scale (src/CellUtilities.jl:78-84)."""
    m = normalize_cell_cost_metric(metric)
    return {"l1": 2.0, "l2": math.sqrt(2.0), "cosine": 2.0, "correlation": 2.0}[m]


class FeatureProblem:
    """Device-resident feature problem: rows of data1 (n x d) against rows of
    data2 (m x d, or data1 itself when data2 is None)."""

    def __init__(self, data1, data2, scale, metric, r, c, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.metric = normalize_cell_cost_metric(metric)
        self.x = np.asfortranarray(data1, dtype=np.float64)
        self.self_cost = data2 is None or data2 is data1
        self.y = self.x if self.self_cost else np.asfortranarray(data2, dtype=np.float64)
        if self.x.ndim != 2 or self.y.ndim != 2 or self.x.shape[1] != self.y.shape[1]:
            raise ValueError("data must be (points x features) matrices with equal widths")
        self.n, self.d = self.x.shape
        self.m = self.y.shape[0]
        self.r = _vec(r, self.n, "marginal1")
        self.c = _vec(c, self.m, "marginal2")
        h = C.c_void_p()
        check(load().culamp_feature_problem_create(
            self.ctx.handle, self.x.ctypes.data, self.n,
            None if self.self_cost else self.y.ctypes.data, self.m, self.d,
            _METRIC_ID[self.metric], float(scale), ptr(self.r), ptr(self.c), C.byref(h)))
        self.handle = h
        self.ctx.adopt(self)

    def close(self) -> None:
        # the context destroys its problems itself once it is closed
        if getattr(self, "handle", None) and self.ctx.handle:
            load().culamp_problem_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _logger(log):
        return LOG_FN(0) if log is None else LOG_FN(lambda _u, line: log(line.decode()))

    def lamp(self, args: EOTArgs, frequency: int = 100, log=_stdout_log):
        nu, phi, psi = np.empty(self.m), np.empty(self.n), np.empty(self.m)
        res, a, cb = culamp_result(), args.to_c(), self._logger(log)
        check(load().culamp_lamp_cellsim_solve(self.ctx.handle, self.handle, C.byref(a),
                                               int(frequency), cb, None, ptr(nu), ptr(phi),
                                               ptr(psi), C.byref(res)))
        return nu, phi, psi, res.as_dict()

    def sinkhorn(self, args: EOTArgs, frequency: int = 100, phi0=None, psi0=None, log=_stdout_log):
        phi, psi = np.empty(self.n), np.empty(self.m)
        p0 = None if phi0 is None else _vec(phi0, self.n, "φ0")
        q0 = None if psi0 is None else _vec(psi0, self.m, "ψ0")
        res, a, cb = culamp_result(), args.to_c(), self._logger(log)
        check(load().culamp_sinkhorn_cellsim_solve(self.ctx.handle, self.handle, C.byref(a),
                                                   int(frequency), ptr(p0), ptr(q0), cb, None,
                                                   ptr(phi), ptr(psi), C.byref(res)))
        return phi, psi, res.as_dict()

    def annealed_sinkhorn(self, args: EOTArgs, frequency: int = 100, log=_stdout_log):
        phi, psi = np.empty(self.n), np.empty(self.m)
        res, a, cb = culamp_result(), args.to_c(), self._logger(log)
        check(load().culamp_annealed_sinkhorn_cellsim_solve(self.ctx.handle, self.handle,
                                                            C.byref(a), int(frequency), cb, None,
                                                            ptr(phi), ptr(psi), C.byref(res)))
        return phi, psi, res.as_dict()


def extragradient_cellsim(data, scale, metric, marginal1, marginal2, args: EOTArgs,
                          frequency: int = 100, log_output: bool = False, *, log=_stdout_log,
                          ctx: Context | None = None, return_info: bool = False):
    """MirrorProxCellSim.jl:94-188 -> (ν, φ, ψ, objective)."""
    pb = FeatureProblem(data, None, scale, metric, marginal1, marginal2, ctx)
    try:
        nu, phi, psi, info = pb.lamp(args, frequency, log)
    finally:
        pb.close()
    out = (nu, phi, psi, info["objective"])
    return out + (info,) if return_info else out


def sinkhorn_cellsim(data1, data2, scale, metric, marginal1, marginal2, args: EOTArgs,
                     frequency: int = 100, *, return_cuda: bool = False, φ0=None, ψ0=None,
                     log=_stdout_log, ctx: Context | None = None, return_info: bool = False):
    """SinkhornCellSim.jl:240-310 -> (φ, ψ, objective, residual) (+ num_iter
    with return_cuda=True)."""
    pb = FeatureProblem(data1, data2, scale, metric, marginal1, marginal2, ctx)
    try:
        phi, psi, info = pb.sinkhorn(args, frequency, φ0, ψ0, log)
    finally:
        pb.close()
    out = (phi, psi, info["objective"], info["residual"])
    if return_cuda:
        out = out + (info["iterations"],)
    return out + (info,) if return_info else out


def annealed_sinkhorn_cellsim(data1, data2, scale, metric, marginal1, marginal2,
                              args: EOTArgs, frequency: int = 100, *, log=_stdout_log,
                              ctx: Context | None = None, return_info: bool = False):
    """AnnealedSinkhornCellSim.jl:1-63 -> (φ, ψ, objective)."""
    pb = FeatureProblem(data1, data2, scale, metric, marginal1, marginal2, ctx)
    try:
        phi, psi, info = pb.annealed_sinkhorn(args, frequency, log)
    finally:
        pb.close()
    out = (phi, psi, info["objective"])
    return out + (info,) if return_info else out
