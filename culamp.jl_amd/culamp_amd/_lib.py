"""ctypes binding of libculamp (include/culamp.h).

The shared library is loaded from the package directory (``culamp.jl_amd/
libculamp.so``, built by ``culamp.jl_amd/build.py``).  There is no fallback:
if the library is missing or has no device, calls raise ``CulampError``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
import weakref

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("CULAMP_LIB", os.path.join(PKG_DIR, "libculamp.so"))

CULAMP_OK = 0
CULAMP_ERR_INVALID_ARG = 1
CULAMP_ERR_HIP = 2
CULAMP_ERR_NONFINITE = 3
CULAMP_ERR_NOT_CONVERGED = 4
CULAMP_ERR_NO_DEVICE = 5
CULAMP_ERR_ALLOC = 6

FP64 = 0
MIXED = 1
BF16 = 2
BF16X3 = 3


class CulampError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libculamp error {code}: {msg}")
        self.code = code


class culamp_args(C.Structure):
    _fields_ = [
        ("eta_p", C.c_double), ("eta_mu", C.c_double), ("tau_p", C.c_double),
        ("tau_mu", C.c_double), ("alpha", C.c_double), ("B", C.c_double),
        ("epsilon", C.c_double), ("anneal_mult", C.c_double),
        ("itermax", C.c_int64), ("inner_iter", C.c_int64), ("tmax", C.c_double),
        ("verbose", C.c_int32), ("precision", C.c_int32),
        ("stop", C.c_int32), ("reserved", C.c_int32),
    ]


class culamp_result(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64), ("objective", C.c_double),
        ("residual", C.c_double), ("primal", C.c_double), ("dual", C.c_double),
        ("winf", C.c_double), ("eta_t", C.c_double), ("elapsed_s", C.c_double),
        ("pairs", C.c_double), ("passes", C.c_int64),
        ("converged", C.c_int32), ("nonfinite", C.c_int32), ("fp64_from", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class culamp_instance(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("x", C.c_void_p), ("y", C.c_void_p),
                ("r", C.c_void_p), ("c", C.c_void_p)]


class culamp_batch_tuning(C.Structure):
    _fields_ = [("rows_per_thread", C.c_int32), ("graph", C.c_int32),
                ("max_split", C.c_int64), ("min_items", C.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class culamp_divergence_opts(C.Structure):
    _fields_ = [("eps", C.c_double), ("eta_p", C.c_double), ("tol", C.c_double),
                ("polish_tol", C.c_double), ("itermax", C.c_int64), ("frequency", C.c_int64),
                ("unroll", C.c_int32), ("checkpoint", C.c_int32), ("precision", C.c_int32),
                ("reserved", C.c_int32), ("tmax", C.c_double)]


LOG_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)

_P = C.c_void_p
_D = C.c_double
_I64 = C.c_int64
_I = C.c_int

# name -> (restype, argtypes); mirrors include/culamp.h
SIGNATURES = {
    "culamp_abi_version": (_I, []),
    "culamp_last_error": (C.c_char_p, []),
    "culamp_args_default": (None, [C.POINTER(culamp_args)]),
    "culamp_device_count": (_I, [C.POINTER(_I)]),
    "culamp_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "culamp_ctx_destroy": (_I, [_P]),
    "culamp_ctx_synchronize": (_I, [_P]),
    "culamp_ctx_set_launch": (_I, [_P, _I, _I]),
    "culamp_ctx_set_layout": (_I, [_P, _I]),
    "culamp_ctx_set_source_records": (_I, [_P, _I]),
    "culamp_ctx_set_fp64_fallback": (_I, [_P, _I]),
    "culamp_probe_rates": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "culamp_probe_mfma": (_I, [_P, C.POINTER(C.c_double)]),
    "culamp_malloc": (_I, [_P, C.c_size_t, C.POINTER(_P)]),
    "culamp_free": (_I, [_P, _P]),
    "culamp_memcpy_htod": (_I, [_P, _P, _P, C.c_size_t]),
    "culamp_memcpy_dtoh": (_I, [_P, _P, _P, C.c_size_t]),
    "culamp_memcpy_dtod": (_I, [_P, _P, _P, C.c_size_t]),
    "culamp_timer_start": (_I, [_P]),
    "culamp_timer_stop": (_I, [_P, C.POINTER(_D)]),
    "culamp_dense_create": (_I, [_P, _P, C.c_int64, C.c_int64, _I, C.POINTER(_P)]),
    "culamp_dense_destroy": (_I, [_P]),
    "culamp_dense_lse_rows": (_I, [_P, _P, _P, _D, _D, _D, _P]),
    "culamp_dense_col_marginal": (_I, [_P, _P, _P, _P, _P, _D, _D, _D, _P]),
    "culamp_dense_lamp_solve": (_I, [_P, _P, _P, _P, C.POINTER(culamp_args), C.c_int64, _P,
                                     LOG_FN, _P, _P, _P, C.POINTER(culamp_result)]),
    "culamp_capture_begin": (_I, [_P]),
    "culamp_capture_end": (_I, [_P, C.POINTER(_P)]),
    "culamp_graph_launch": (_I, [_P, _P]),
    "culamp_graph_destroy": (_I, [_P]),
    "culamp_max_cost": (_I, [_P, _P, _I64, _P, _I64, _D, _P]),
    "culamp_lse_rows": (_I, [_P, _P, _I64, _P, _I64, _P, _D, _D, _D, _D, _I, _P]),
    "culamp_lse_rows_noalloc": (_I, [_P, _P, _I64, _P, _I64, _P, _D, _D, _D, _D, _I, _P]),
    "culamp_col_marginal_cost": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _D, _I, _P, _P]),
    "culamp_minplus_rows": (_I, [_P, _P, _I64, _P, _I64, _P, _D, _D, _P]),
    "culamp_lamp_theta_update": (_I, [_P, _P, _P, _P, _P, _P, _I64, _D, _I, _D, _D, _D]),
    "culamp_sinkhorn_half_step": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _D, _D, _D, _I, _P]),
    "culamp_row_residual_cost": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _I, _P, _P]),
    "culamp_barycentric_sinkhorn": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _D, _D, _D, _P]),
    "culamp_barycentric_lamp_rows": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _P]),
    "culamp_barycentric_lamp_cols": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _P]),
    "culamp_problem_create": (_I, [_P, _P, _I64, _P, _I64, _P, _P, C.POINTER(_P)]),
    "culamp_problem_destroy": (_I, [_P]),
    "culamp_lamp_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, _D, LOG_FN, _P,
                               _P, _P, _P, _P, _P, C.POINTER(culamp_result)]),
    "culamp_sinkhorn_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, _D, _P, _P, LOG_FN, _P,
                                   _P, _P, _P, _P, C.POINTER(culamp_result)]),
    "culamp_annealed_sinkhorn_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, _D, LOG_FN, _P,
                                            _P, _P, _P, _P, C.POINTER(culamp_result)]),
    "culamp_feature_problem_create": (_I, [_P, _P, _I64, _P, _I64, _I64, _I, _D, _P, _P,
                                           C.POINTER(_P)]),
    "culamp_feature_lse_rows": (_I, [_P, _P, _P, _D, _D, _D, _I, _P]),
    "culamp_feature_col_marginal_cost": (_I, [_P, _P, _P, _P, _P, _D, _D, _D, _I, _P, _P]),
    "culamp_feature_sinkhorn_half_step": (_I, [_P, _P, _I, _P, _P, _D, _D, _I, _P]),
    "culamp_feature_row_residual_cost": (_I, [_P, _P, _P, _P, _P, _D, _D, _I, _P, _P]),
    "culamp_feature_max_cost": (_I, [_P, _P, _P]),
    "culamp_feature_minplus_rows": (_I, [_P, _P, _P, _D, _P]),
    "culamp_lamp_cellsim_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, LOG_FN, _P, _P, _P,
                                       _P, C.POINTER(culamp_result)]),
    "culamp_sinkhorn_cellsim_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, _P, _P, LOG_FN,
                                           _P, _P, _P, C.POINTER(culamp_result)]),
    "culamp_annealed_sinkhorn_cellsim_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, LOG_FN,
                                                    _P, _P, _P, C.POINTER(culamp_result)]),
    "culamp_batch_create": (_I, [_P, C.POINTER(culamp_instance), C.c_int32, _I64, _I, _D,
                                 C.POINTER(_P)]),
    "culamp_batch_destroy": (_I, [_P]),
    "culamp_batch_set_tuning": (_I, [_P, C.POINTER(culamp_batch_tuning)]),
    "culamp_batch_get_tuning": (_I, [_P, C.POINTER(culamp_batch_tuning)]),
    "culamp_batch_time_pass": (_I, [_P, _P, _I, _D, _I, C.POINTER(_D), C.POINTER(C.c_int32)]),
    "culamp_batch_autotune": (_I, [_P, _P, _I, _D, _I, C.POINTER(culamp_batch_tuning),
                                   C.POINTER(_D)]),
    "culamp_batch_sinkhorn_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, _D, _P, _P, _P, _P,
                                         C.POINTER(culamp_result)]),
    "culamp_batch_lamp_solve": (_I, [_P, _P, C.POINTER(culamp_args), _I64, _D, _P, _P, _P,
                                     C.POINTER(culamp_result)]),
    "culamp_divergence_opts_default": (None, [C.POINTER(culamp_divergence_opts)]),
    "culamp_sinkhorn_divergence": (_I, [_P, _P, _I64, _P, _P, _I64, _P,
                                        C.POINTER(culamp_divergence_opts), C.POINTER(_D), _P, _P,
                                        _P, _P, _P]),
    "culamp_sinkhorn_divergence_vjp": (_I, [_P, _P, _I64, _P, _P, _I64, _P,
                                            C.POINTER(culamp_divergence_opts), _D, _P, _P, _P, _P]),
    "culamp_lse_adjoint": (_I, [_P, _P, _I64, _P, _I64, _D, _P, _D, _P, _P, _D, _I, _P, _D, _D,
                                _D, _P, _P]),
}

_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load libculamp once; raise CulampError when it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise CulampError(CULAMP_ERR_NO_DEVICE,
                              f"{LIB_PATH} not found: build it with culamp.jl_amd/build.py")
        lib = C.CDLL(LIB_PATH, mode=os.RTLD_NOW | os.RTLD_LOCAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    if rc != CULAMP_OK:
        msg = load().culamp_last_error()
        raise CulampError(rc, msg.decode() if msg else "")


def ptr(a: np.ndarray | None) -> int | None:
    """Host pointer of a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected C-contiguous float64"
    return a.ctypes.data


def device_count() -> int:
    n = C.c_int(0)
    check(load().culamp_device_count(C.byref(n)))
    return n.value


class Context:
    """One libculamp context (one HIP stream on one device)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.culamp_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device
        self._children = weakref.WeakSet()

    def adopt(self, child) -> None:
        """Register an object holding device resources of this context; it is
        closed before the context is destroyed."""
        self._children.add(child)

    def close(self) -> None:
        if self.handle:
            for ch in list(self._children):
                try:
                    ch.close()
                except Exception:
                    pass
            self.lib.culamp_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        check(self.lib.culamp_ctx_synchronize(self.handle))

    def set_launch(self, rows_per_thread: int = 0, splits: int = 0) -> None:
        check(self.lib.culamp_ctx_set_launch(self.handle, rows_per_thread, splits))

    def set_layout(self, layout: str = "auto") -> None:
        """Source-split layout of point passes: "auto", "split" (over
        workgroups, merged by a second launch) or "wide" (inside one
        workgroup, merged in LDS)."""
        codes = {"auto": 0, "split": 1, "wide": 2}
        check(self.lib.culamp_ctx_set_layout(self.handle, codes[layout]))

    def set_source_records(self, enable: bool = True) -> None:
        """Whether the mixed-precision LAMP loop passes staged FP32 source
        records from each pass's epilogue to the next pass (identical results
        either way)."""
        check(self.lib.culamp_ctx_set_source_records(self.handle, 1 if enable else 0))

    def set_fp64_fallback(self, enable: bool = True) -> None:
        """Whether reduced-precision solves hand their remaining passes to
        FP64 when the infeasibility stalls above the stop tolerance (10 checks
        without reaching 0.97 of its best); the result's ``fp64_from`` gives
        the iteration of the switch (0: none)."""
        check(self.lib.culamp_ctx_set_fp64_fallback(self.handle, 1 if enable else 0))

    def probe_rates(self) -> dict:
        """Measured FP32 exponential and fused multiply-add throughput of the
        device, lane operations per second."""
        e, f = C.c_double(), C.c_double()
        check(self.lib.culamp_probe_rates(self.handle, C.byref(e), C.byref(f)))
        return {"exp_per_s": e.value, "fma_per_s": f.value}

    def probe_mfma(self) -> float:
        """Measured sustained dense bf16 MFMA throughput of the device, flop/s
        (32x32x16 MFMA chains at three waves per SIMD)."""
        r = C.c_double()
        check(self.lib.culamp_probe_mfma(self.handle, C.byref(r)))
        return r.value

    def capture(self, fn) -> "Graph":
        """Records the kernel-level calls made by fn() into a replayable graph."""
        check(self.lib.culamp_capture_begin(self.handle))
        try:
            fn()
        finally:
            g = C.c_void_p()
            rc = self.lib.culamp_capture_end(self.handle, C.byref(g))
        check(rc)
        return Graph(self, g)

    def timer_start(self) -> None:
        check(self.lib.culamp_timer_start(self.handle))

    def timer_stop(self) -> float:
        ms = C.c_double(0.0)
        check(self.lib.culamp_timer_stop(self.handle, C.byref(ms)))
        return ms.value


class Graph:
    """A captured sequence of kernel-level calls (Context.capture)."""

    def __init__(self, ctx: Context, handle):
        self.ctx, self.handle = ctx, handle
        ctx.adopt(self)

    def launch(self) -> None:
        check(self.ctx.lib.culamp_graph_launch(self.ctx.handle, self.handle))

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.culamp_graph_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float64 device buffer owned through libculamp (culamp_malloc)."""

    def __init__(self, ctx: Context, n: int):
        self.ctx = ctx
        self.n = int(n)
        p = C.c_void_p()
        check(ctx.lib.culamp_malloc(ctx.handle, max(self.n, 1) * 8, C.byref(p)))
        self.ptr = p.value
        ctx.adopt(self)

    @classmethod
    def from_host(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(ctx, a.size)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        assert a.size == self.n
        check(self.ctx.lib.culamp_memcpy_htod(self.ctx.handle, self.ptr, a.ctypes.data, a.size * 8))

    def get(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        check(self.ctx.lib.culamp_memcpy_dtoh(self.ctx.handle, out.ctypes.data, self.ptr, self.n * 8))
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            check(self.ctx.lib.culamp_free(self.ctx.handle, self.ptr))
        self.ptr = None

    close = free

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.ctx.lib.culamp_free(self.ctx.handle, self.ptr)
        except Exception:
            pass


def dptr(a: DeviceArray | None):
    return None if a is None else a.ptr
