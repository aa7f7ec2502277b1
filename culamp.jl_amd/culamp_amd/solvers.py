"""Solver entry points with the reference's names, arguments and returns.

* ``extragradient_color_transfer`` -- src/solvers/MirrorProxKernel.jl:1019-1132
* ``extragradient_euclidean``      -- MirrorProxKernel.jl:1143-1185
* ``sinkhorn_color_transfer``      -- src/solvers/SinkhornKernel.jl:477-559
* ``sinkhorn_euclidean``           -- SinkhornKernel.jl:579-608
* ``annealed_sinkhorn_euclidean``  -- src/solvers/AnnealedSinkhorn.jl:65-177

Each call uploads the problem into a device-resident ``culamp_problem`` and
runs the solver loop inside libculamp (C++ host loop over gfx950 kernels).
Log lines are the reference's CSV lines, passed to ``log`` (default: stdout,
where the reference's ``println`` / ``@printf`` write).  ``Problem`` keeps the
inputs resident for repeated solves (benchmarks).
"""
from __future__ import annotations

import ctypes as C
import sys
from typing import Callable

import numpy as np

from ._lib import (LOG_FN, Context, check, culamp_result, load, ptr)
from .args import EOTArgs

_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _points(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"point clouds are (n, 3) arrays, got shape {a.shape}")
    return a


def _vec(a, n: int, name: str) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if a.size != n:
        raise ValueError(f"{name} has {a.size} entries, expected {n}")
    return a


def _stdout_log(line: str) -> None:
    sys.stdout.write(line)


class Problem:
    """Device-resident inputs (x: (n,3), y: (m,3), r: n, c: m)."""

    def __init__(self, x, y, r, c, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.x = _points(x)
        self.y = _points(y)
        self.n, self.m = self.x.shape[0], self.y.shape[0]
        self.r = _vec(r, self.n, "marginal1")
        self.c = _vec(c, self.m, "marginal2")
        h = C.c_void_p()
        check(load().culamp_problem_create(self.ctx.handle, ptr(self.x), self.n, ptr(self.y),
                                           self.m, ptr(self.r), ptr(self.c), C.byref(h)))
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

    # -- raw solver calls on the resident problem -------------------------------
    def _logger(self, log):
        if log is None:
            return LOG_FN(0)
        cb = LOG_FN(lambda _u, line: log(line.decode()))
        return cb

    def lamp(self, args: EOTArgs, frequency: int = 100, p: float = 2.0,
             log: Callable[[str], None] | None = _stdout_log, assignments: bool = True):
        nu = np.empty(self.m)
        phi = np.empty(self.n)
        psi = np.empty(self.m)
        img1 = np.empty((self.n, 3)) if assignments else None
        img2 = np.empty((self.m, 3)) if assignments else None
        res = culamp_result()
        a = args.to_c()
        cb = self._logger(log)
        check(load().culamp_lamp_solve(self.ctx.handle, self.handle, C.byref(a), int(frequency),
                                       float(p), cb, None, ptr(nu), ptr(phi), ptr(psi),
                                       ptr(img1), ptr(img2), C.byref(res)))
        return nu, phi, psi, img1, img2, res.as_dict()

    def sinkhorn(self, args: EOTArgs, frequency: int = 100, p: float = 2.0, phi0=None, psi0=None,
                 log: Callable[[str], None] | None = _stdout_log, assignments: bool = True):
        phi = np.empty(self.n)
        psi = np.empty(self.m)
        img1 = np.empty((self.n, 3)) if assignments else None
        img2 = np.empty((self.m, 3)) if assignments else None
        p0 = None if phi0 is None else _vec(phi0, self.n, "φ0")
        q0 = None if psi0 is None else _vec(psi0, self.m, "ψ0")
        res = culamp_result()
        a = args.to_c()
        cb = self._logger(log)
        check(load().culamp_sinkhorn_solve(self.ctx.handle, self.handle, C.byref(a), int(frequency),
                                           float(p), ptr(p0), ptr(q0), cb, None, ptr(phi), ptr(psi),
                                           ptr(img1), ptr(img2), C.byref(res)))
        return phi, psi, img1, img2, res.as_dict()

    def annealed_sinkhorn(self, args: EOTArgs, frequency: int = 100, p: float = 2.0,
                          log: Callable[[str], None] | None = _stdout_log, assignments: bool = True):
        phi = np.empty(self.n)
        psi = np.empty(self.m)
        img1 = np.empty((self.n, 3)) if assignments else None
        img2 = np.empty((self.m, 3)) if assignments else None
        res = culamp_result()
        a = args.to_c()
        cb = self._logger(log)
        check(load().culamp_annealed_sinkhorn_solve(self.ctx.handle, self.handle, C.byref(a),
                                                    int(frequency), float(p), cb, None, ptr(phi),
                                                    ptr(psi), ptr(img1), ptr(img2), C.byref(res)))
        return phi, psi, img1, img2, res.as_dict()


def extragradient_color_transfer(img1, img2, marginal1, marginal2, args: EOTArgs,
                                 frequency: int = 100, p: float = 2.0, *,
                                 log=_stdout_log, ctx: Context | None = None,
                                 return_info: bool = False):
    """MirrorProxKernel.jl:1019-1132 -> (ν, φ, ψ, img1_new, img2_new)."""
    pb = Problem(img1, img2, marginal1, marginal2, ctx)
    try:
        nu, phi, psi, a1, a2, info = pb.lamp(args, frequency, p, log)
    finally:
        pb.close()
    out = (nu, phi, psi, a1, a2)
    return out + (info,) if return_info else out


def sinkhorn_color_transfer(img1, img2, marginal1, marginal2, args: EOTArgs,
                            frequency: int = 100, p: float = 2, *, return_cuda: bool = False,
                            return_assignments: bool = True, φ0=None, ψ0=None,
                            log=_stdout_log, ctx: Context | None = None,
                            return_info: bool = False):
    """SinkhornKernel.jl:477-559.

    return_assignments=True  -> (φ, ψ, img1_new, img2_new)
    return_assignments=False -> (φ, ψ, objective, residual) (+ num_iter when
    return_cuda=True, the reference's device-array branch)."""
    pb = Problem(img1, img2, marginal1, marginal2, ctx)
    try:
        phi, psi, a1, a2, info = pb.sinkhorn(args, frequency, p, φ0, ψ0, log, return_assignments)
    finally:
        pb.close()
    if return_assignments:
        out = (phi, psi, a1, a2)
    elif return_cuda:
        out = (phi, psi, info["objective"], info["residual"], info["iterations"])
    else:
        out = (phi, psi, info["objective"], info["residual"])
    return out + (info,) if return_info else out


def julia_float_str(v: float) -> str:
    """Julia's `println(::Float64)` text (shortest round-trip, `1.0e-5` style)."""
    v = float(v)
    if v != v:
        return "NaN"
    if v in (float("inf"), float("-inf")):
        return "Inf" if v > 0 else "-Inf"
    if v == 0.0:
        return "-0.0" if np.signbit(v) else "0.0"
    s = repr(v)
    a = abs(v)
    if 1e-4 <= a < 1e6:
        if "e" in s:
            s = np.format_float_positional(v, trim="0")
        if "." not in s:
            s += ".0"
        elif s.endswith("."):
            s += "0"
        return s
    mant, _, exp = np.format_float_scientific(v, trim="0").partition("e")
    if "." not in mant:
        mant += ".0"
    elif mant.endswith("."):
        mant += "0"
    return f"{mant}e{int(exp)}"


def _write_vec(path: str, v) -> None:
    with open(path, "w") as f:
        for x in v:
            f.write(julia_float_str(x) + "\n")


def julia_matrix_str(a) -> str:
    """Julia's string of a Float64 matrix (``"$(A)"``): rows joined by "; ",
    entries by " ", ";;" closing an n x 1 matrix (Julia >= 1.7)."""
    a = np.asarray(a, dtype=np.float64)
    rows = ["; ".join(" ".join(julia_float_str(v) for v in row) for row in a)]
    tail = ";;" if a.shape[1] == 1 else ""
    return "[" + rows[0] + tail + "]"


def _write_assignments(path: str, a) -> None:
    """The reference's assignment files (MirrorProxKernel.jl:1171-1184,
    SinkhornKernel.jl:594-607, AnnealedSinkhorn.jl:158-172): for every linear
    index k of the 3 x N barycentric-image matrix, one line "<matrix>,<k>",
    the whole matrix interpolated each time -- 3N lines of O(N) characters,
    reproduced as written there."""
    text = julia_matrix_str(np.asarray(a, dtype=np.float64).T)
    n = np.asarray(a).size
    with open(path, "w") as f:
        for k in range(1, n + 1):
            f.write(f"{text},{k}\n")


def extragradient_euclidean(marginal1, marginal2, location1, location2, out1: str, out2: str,
                            outmu: str, args: EOTArgs, frequency: int, p: float, **kw):
    """MirrorProxKernel.jl:1143-1185: solve and write potentials
    (`outmu`.spp_col / .row / .col, one value per line)."""
    nu, phi, psi, a1, a2 = extragradient_color_transfer(location1, location2, marginal1,
                                                        marginal2, args, frequency, p, **kw)[:5]
    if outmu:
        _write_vec(outmu + ".spp_col", nu)
        _write_vec(outmu + ".row", phi)
        _write_vec(outmu + ".col", psi)
    if out1:
        _write_assignments(out1, a1)
    if out2:
        _write_assignments(out2, a2)
    return nu, phi, psi, a1, a2


def sinkhorn_euclidean(marginal1, marginal2, location1, location2, out1: str, out2: str,
                       potentials: str, args: EOTArgs, frequency: int, p: float, **kw):
    """SinkhornKernel.jl:579-608: solve and write `potentials`.row / .col."""
    phi, psi, a1, a2 = sinkhorn_color_transfer(location1, location2, marginal1, marginal2,
                                               args, frequency, p, **kw)[:4]
    if potentials:
        _write_vec(potentials + ".row", phi)
        _write_vec(potentials + ".col", psi)
    if out1:
        _write_assignments(out1, a1)
    if out2:
        _write_assignments(out2, a2)
    return phi, psi, a1, a2


def annealed_sinkhorn_euclidean(marginal1, marginal2, location1, location2, out1: str,
                                out2: str, potentials: str, args: EOTArgs, frequency: int,
                                p: float, *, log=_stdout_log, ctx: Context | None = None,
                                return_info: bool = False):
    """AnnealedSinkhorn.jl:65-177 -> (φ, ψ, objective)."""
    pb = Problem(location1, location2, marginal1, marginal2, ctx)
    try:
        phi, psi, a1, a2, info = pb.annealed_sinkhorn(args, frequency, p, log, True)
    finally:
        pb.close()
    if potentials:
        _write_vec(potentials + ".row", phi)
        _write_vec(potentials + ".col", psi)
    if out1:
        _write_assignments(out1, a1)
    if out2:
        _write_assignments(out2, a2)
    out = (phi, psi, info["objective"])
    return out + (info,) if return_info else out
