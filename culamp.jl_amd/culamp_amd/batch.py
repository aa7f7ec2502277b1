"""Batched solves: many independent instances advanced together on one
device (libculamp batches, include/culamp.h ``culamp_batch_*``).

Per instance the loop, stop test and outputs are those of the single-instance
solvers -- ``sinkhorn_color_transfer`` / ``extragradient_color_transfer``
(src/solvers/SinkhornKernel.jl:477-559, MirrorProxKernel.jl:1019-1132) for
point clouds and ``sinkhorn_cellsim`` / ``extragradient_cellsim``
(src/solvers/SinkhornCellSim.jl:240-310, MirrorProxCellSim.jl:94-188) for
feature instances -- while every pass of every instance shares one kernel
launch and the iterations between checks replay a captured HIP graph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (Context, check, culamp_batch_tuning, culamp_instance, culamp_result, load,
                   ptr)
from .args import _PRECISIONS, EOTArgs
from .cellsim import _METRIC_ID, normalize_cell_cost_metric
from .solvers import _vec, default_context


class Batch:
    """Instances ``(x, y, r, c)``: point clouds as (n, 3) / (m, 3) arrays, or,
    with ``metric`` given, feature rows (n, d) / (m, d) (``y`` None for the
    self cost) with the cell-similarity ``metric`` and ``scale``."""

    def __init__(self, instances, ctx: Context | None = None, metric=None, scale=None):
        self.ctx = ctx or default_context()
        self.feature = metric is not None
        self.metric = normalize_cell_cost_metric(metric) if self.feature else None
        keep, descs = [], []
        d = 0
        # each distinct input array is converted once, so instances that pass
        # the same array share one device copy of it (culamp_batch_create
        # matches them by pointer)
        conv = {}

        def once(a, fn):
            key = id(a)
            if key not in conv:
                conv[key] = (a, fn(a))
            return conv[key][1]

        fortran = lambda a: np.asfortranarray(a, dtype=np.float64)  # noqa: E731
        contig = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        for k, (x, y, r, c) in enumerate(instances):
            if self.feature:
                xa = once(x, fortran)
                ya = None if (y is None or y is x) else once(y, fortran)
                if xa.ndim != 2 or (ya is not None and ya.shape[1] != xa.shape[1]):
                    raise ValueError(f"instance {k}: feature rows must be (points x d) with equal d")
                if d and xa.shape[1] != d:
                    raise ValueError("every instance of a feature batch needs the same d")
                d = xa.shape[1]
                n, m = xa.shape[0], (xa.shape[0] if ya is None else ya.shape[0])
            else:
                xa = once(x, contig)
                ya = once(y, contig)
                if xa.ndim != 2 or xa.shape[1] != 3 or ya.ndim != 2 or ya.shape[1] != 3:
                    raise ValueError(f"instance {k}: point clouds must be (n, 3) arrays")
                n, m = xa.shape[0], ya.shape[0]
            ra, ca = _vec(r, n, "r"), _vec(c, m, "c")
            keep.append((xa, ya, ra, ca))
            descs.append(culamp_instance(n, m, xa.ctypes.data,
                                         None if ya is None else ya.ctypes.data,
                                         ra.ctypes.data, ca.ctypes.data))
        if not descs:
            raise ValueError("empty batch")
        self.n = np.array([dd.n for dd in descs], dtype=np.int64)
        self.m = np.array([dd.m for dd in descs], dtype=np.int64)
        self.off_n = np.concatenate([[0], np.cumsum(self.n)])
        self.off_m = np.concatenate([[0], np.cumsum(self.m)])
        arr = (culamp_instance * len(descs))(*descs)
        h = C.c_void_p()
        scale = 1.0 if scale is None else float(scale)
        check(load().culamp_batch_create(self.ctx.handle, arr, len(descs), int(d),
                                         _METRIC_ID[self.metric] if self.feature else -1, scale,
                                         C.byref(h)))
        self.handle = h
        self.count = len(descs)
        self.ctx.adopt(self)

    def close(self) -> None:
        if getattr(self, "handle", None) and self.ctx.handle:
            load().culamp_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tuning(self) -> dict:
        t = culamp_batch_tuning()
        check(load().culamp_batch_get_tuning(self.handle, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        cur = self.tuning
        cur.update(kw)
        t = culamp_batch_tuning(**cur)
        check(load().culamp_batch_set_tuning(self.handle, C.byref(t)))

    def autotune(self, precision: str = "mixed", p: float = 2.0, reps: int = 5):
        """Time candidate launch shapes of the row log-sum-exp pass; keeps the
        fastest.  Returns (tuning dict, ms per pass)."""
        t, ms = culamp_batch_tuning(), C.c_double()
        check(load().culamp_batch_autotune(self.ctx.handle, self.handle, _PRECISIONS[precision],
                                           float(p), int(reps), C.byref(t), C.byref(ms)))
        return t.as_dict(), ms.value

    def _split(self, v, offs):
        return [v[offs[k]:offs[k + 1]] for k in range(self.count)]

    def time_pass(self, precision: str = "mixed", p: float = 2.0, reps: int = 20):
        """Mean time (ms) of the row log-sum-exp pass over all instances at the
        current tuning, from HIP events on the context's stream, and the kernel
        launches per pass (2 when a split merge follows)."""
        ms, nl = C.c_double(), C.c_int32()
        check(load().culamp_batch_time_pass(self.ctx.handle, self.handle, _PRECISIONS[precision],
                                            float(p), int(reps), C.byref(ms), C.byref(nl)))
        return ms.value, nl.value

    def sinkhorn(self, args: EOTArgs, frequency: int = 100, p: float = 2.0, phi0=None, psi0=None):
        """Returns a list of (phi, psi, info) per instance."""
        N, M = int(self.off_n[-1]), int(self.off_m[-1])
        phi, psi = np.empty(N), np.empty(M)
        p0 = None if phi0 is None else np.ascontiguousarray(np.concatenate(phi0), dtype=np.float64)
        q0 = None if psi0 is None else np.ascontiguousarray(np.concatenate(psi0), dtype=np.float64)
        if p0 is not None and p0.size != N or q0 is not None and q0.size != M:
            raise ValueError("warm starts must match the instances' sizes")
        res = (culamp_result * self.count)()
        a = args.to_c()
        check(load().culamp_batch_sinkhorn_solve(self.ctx.handle, self.handle, C.byref(a),
                                                 int(frequency), float(p), ptr(p0), ptr(q0),
                                                 ptr(phi), ptr(psi), res))
        return list(zip(self._split(phi, self.off_n), self._split(psi, self.off_m),
                        [r.as_dict() for r in res]))

    def lamp(self, args: EOTArgs, frequency: int = 100, p: float = 2.0):
        """Returns a list of (nu, phi, psi, info) per instance."""
        N, M = int(self.off_n[-1]), int(self.off_m[-1])
        nu, phi, psi = np.empty(M), np.empty(N), np.empty(M)
        res = (culamp_result * self.count)()
        a = args.to_c()
        check(load().culamp_batch_lamp_solve(self.ctx.handle, self.handle, C.byref(a),
                                             int(frequency), float(p), ptr(nu), ptr(phi), ptr(psi),
                                             res))
        return list(zip(self._split(nu, self.off_m), self._split(phi, self.off_n),
                        self._split(psi, self.off_m), [r.as_dict() for r in res]))
