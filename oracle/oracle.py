"""ctypes loader for the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results.  The product library never
loads it.

Arrays follow the oracle's C conventions: points (n, 3) C-contiguous float64
(= 3 x n column-major), vectors float64.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class or_args(C.Structure):
    _fields_ = [
        ("eta_p", C.c_double), ("eta_mu", C.c_double), ("tau_p", C.c_double),
        ("tau_mu", C.c_double), ("alpha", C.c_double), ("B", C.c_double),
        ("epsilon", C.c_double), ("anneal_mult", C.c_double),
        ("itermax", C.c_int64), ("inner_iter", C.c_int64), ("tmax", C.c_double),
        ("verbose", C.c_int32), ("stop", C.c_int32),
    ]


class or_result(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64), ("objective", C.c_double), ("residual", C.c_double),
        ("primal", C.c_double), ("dual", C.c_double), ("winf", C.c_double),
        ("eta_t", C.c_double), ("converged", C.c_int32), ("_pad", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "_pad"}


LOG_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)
_P, _D, _I64, _I = C.c_void_p, C.c_double, C.c_int64, C.c_int

_SIG = {
    "or_set_threads": (None, [_I]),
    "or_get_threads": (_I, []),
    "or_cost3": (_D, [_P, _P, _D]),
    "or_max_cost_rows": (None, [_P, _I64, _P, _I64, _D, _P]),
    "or_lse_rows": (None, [_P, _I64, _P, _I64, _P, _D, _D, _D, _D, _P]),
    "or_col_marginal_cost": (None, [_P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _D, _P, _P]),
    "or_minplus_rows": (None, [_P, _I64, _P, _I64, _P, _D, _D, _P]),
    "or_theta_update": (None, [_P, _P, _P, _P, _P, _I64, _D, _I, _D, _D, _D]),
    "or_sinkhorn_half_step": (None, [_P, _I64, _P, _I64, _P, _P, _D, _D, _D, _P]),
    "or_row_residual_cost": (None, [_P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _P, _P]),
    "or_barycentric_sinkhorn": (None, [_P, _I64, _P, _I64, _P, _P, _D, _D, _D, _P]),
    "or_barycentric_lamp_rows": (None, [_P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _D, _P]),
    "or_barycentric_lamp_cols": (None, [_P, _I64, _P, _I64, _P, _P, _P, _D, _D, _D, _D, _P]),
    "or_lamp_solve": (_I, [_P, _I64, _P, _I64, _P, _P, C.POINTER(or_args), _I64, _D, LOG_FN, _P,
                           _P, _P, _P, _P, _P, C.POINTER(or_result)]),
    "or_sinkhorn_solve": (_I, [_P, _I64, _P, _I64, _P, _P, C.POINTER(or_args), _I64, _D, _P, _P,
                               LOG_FN, _P, _P, _P, _P, _P, C.POINTER(or_result)]),
    "or_annealed_sinkhorn_solve": (_I, [_P, _I64, _P, _I64, _P, _P, C.POINTER(or_args), _I64, _D,
                                        LOG_FN, _P, _P, _P, _P, _P, C.POINTER(or_result)]),
    "or_sinkhorn_log_dense": (_I, [_P, _I64, _P, _I64, _P, C.POINTER(or_args), _I64, LOG_FN, _P,
                                   _P, _P, _P, C.POINTER(or_result)]),
    "or_grid_cost_matrix": (None, [_I64, _I64, _D, _P]),
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for k, (res, args) in _SIG.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def make_args(eta_p=0.0, eta_mu=0.0, tau_p=1.0, tau_mu=1.0, alpha=0.01, B=1.0, epsilon=1e-4,
              anneal_mult=0.95, itermax=10000, inner_iter=10, tmax=float("inf"), verbose=True,
              stop="reference", **_ignored) -> or_args:
    a = or_args()
    a.eta_p, a.eta_mu, a.tau_p, a.tau_mu = eta_p, eta_mu, tau_p, tau_mu
    a.alpha, a.B, a.epsilon, a.anneal_mult = alpha, B, epsilon, anneal_mult
    a.itermax, a.inner_iter, a.tmax, a.verbose = int(itermax), int(inner_iter), tmax, int(bool(verbose))
    a.stop = {"reference": 0, "infeas": 1}[stop]
    return a


def cost3(a, b, p) -> float:
    a, b = _f(a), _f(b)
    return lib().or_cost3(_p(a), _p(b), float(p))


def max_cost_rows(x, y, p):
    x, y = _f(x), _f(y)
    out = np.empty(x.shape[0])
    lib().or_max_cost_rows(_p(x), x.shape[0], _p(y), y.shape[0], float(p), _p(out))
    return out


def lse_rows(x, y, theta, reg, st, winf, p):
    x, y, theta = _f(x), _f(y), _f(theta)
    out = np.empty(x.shape[0])
    lib().or_lse_rows(_p(x), x.shape[0], _p(y), y.shape[0], _p(theta), reg, st, winf, p, _p(out))
    return out


def col_marginal_cost(x, y, r, theta, logZ, reg, st, winf, p):
    x, y, r, theta, logZ = map(_f, (x, y, r, theta, logZ))
    res = np.empty(y.shape[0])
    cost = np.empty(y.shape[0])
    lib().or_col_marginal_cost(_p(x), x.shape[0], _p(y), y.shape[0], _p(r), _p(theta), _p(logZ),
                               reg, st, winf, p, _p(res), _p(cost))
    return res, cost


def minplus_rows(x, y, theta, winf, p):
    x, y, theta = map(_f, (x, y, theta))
    out = np.empty(x.shape[0])
    lib().or_minplus_rows(_p(x), x.shape[0], _p(y), y.shape[0], _p(theta), winf, p, _p(out))
    return out


def theta_update(theta0, res, c, eta_mu_i, eta_mu, adjust, minv, maxv, winf):
    theta0, res, c, eta_mu_i = map(_f, (theta0, res, c, eta_mu_i))
    out = np.empty(theta0.size)
    lib().or_theta_update(_p(out), _p(theta0), _p(res), _p(c), _p(eta_mu_i), theta0.size, eta_mu,
                          int(adjust), minv, maxv, winf)
    return out


def sinkhorn_half_step(x, y, marg, psi, reg, winf, p):
    x, y, marg, psi = map(_f, (x, y, marg, psi))
    out = np.empty(x.shape[0])
    lib().or_sinkhorn_half_step(_p(x), x.shape[0], _p(y), y.shape[0], _p(marg), _p(psi), reg, winf,
                                p, _p(out))
    return out


def row_residual_cost(x, y, marg, phi, psi, reg, winf, p):
    x, y, marg, phi, psi = map(_f, (x, y, marg, phi, psi))
    res = np.empty(x.shape[0])
    cost = np.empty(x.shape[0])
    lib().or_row_residual_cost(_p(x), x.shape[0], _p(y), y.shape[0], _p(marg), _p(phi), _p(psi),
                               reg, winf, p, _p(res), _p(cost))
    return res, cost


def barycentric_sinkhorn(x, y, phi, psi, reg, winf, p):
    x, y, phi, psi = map(_f, (x, y, phi, psi))
    out = np.empty((x.shape[0], 3))
    lib().or_barycentric_sinkhorn(_p(x), x.shape[0], _p(y), y.shape[0], _p(phi), _p(psi), reg,
                                  winf, p, _p(out))
    return out


def barycentric_lamp_rows(x, y, nu, r, logZ, reg, st, winf, p):
    x, y, nu, r, logZ = map(_f, (x, y, nu, r, logZ))
    out = np.empty((x.shape[0], 3))
    lib().or_barycentric_lamp_rows(_p(x), x.shape[0], _p(y), y.shape[0], _p(nu), _p(r), _p(logZ),
                                   reg, st, winf, p, _p(out))
    return out


def barycentric_lamp_cols(x, y, r, nu, logZ, reg, st, winf, p):
    x, y, r, nu, logZ = map(_f, (x, y, r, nu, logZ))
    out = np.empty((y.shape[0], 3))
    lib().or_barycentric_lamp_cols(_p(x), x.shape[0], _p(y), y.shape[0], _p(r), _p(nu), _p(logZ),
                                   reg, st, winf, p, _p(out))
    return out


class _Log:
    def __init__(self):
        self.lines: list[str] = []
        self.cb = LOG_FN(lambda _u, s: self.lines.append(s.decode()))


def lamp_solve(x, y, r, c, frequency=100, p=2.0, **args):
    x, y, r, c = map(_f, (x, y, r, c))
    n, m = x.shape[0], y.shape[0]
    nu, phi, psi = np.empty(m), np.empty(n), np.empty(m)
    img1, img2 = np.empty((n, 3)), np.empty((m, 3))
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = lib().or_lamp_solve(_p(x), n, _p(y), m, _p(r), _p(c), C.byref(a), int(frequency),
                             float(p), lg.cb, None, _p(nu), _p(phi), _p(psi), _p(img1), _p(img2),
                             C.byref(res))
    assert rc == 0
    return dict(nu=nu, phi=phi, psi=psi, img1=img1, img2=img2, log=lg.lines, **res.as_dict())


def sinkhorn_solve(x, y, r, c, frequency=100, p=2.0, phi0=None, psi0=None, **args):
    x, y, r, c = map(_f, (x, y, r, c))
    n, m = x.shape[0], y.shape[0]
    phi, psi = np.empty(n), np.empty(m)
    img1, img2 = np.empty((n, 3)), np.empty((m, 3))
    p0 = None if phi0 is None else _f(phi0)
    q0 = None if psi0 is None else _f(psi0)
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = lib().or_sinkhorn_solve(_p(x), n, _p(y), m, _p(r), _p(c), C.byref(a), int(frequency),
                                 float(p), _p(p0), _p(q0), lg.cb, None, _p(phi), _p(psi),
                                 _p(img1), _p(img2), C.byref(res))
    assert rc == 0
    return dict(phi=phi, psi=psi, img1=img1, img2=img2, log=lg.lines, **res.as_dict())


def annealed_sinkhorn_solve(x, y, r, c, frequency=100, p=2.0, **args):
    x, y, r, c = map(_f, (x, y, r, c))
    n, m = x.shape[0], y.shape[0]
    phi, psi = np.empty(n), np.empty(m)
    img1, img2 = np.empty((n, 3)), np.empty((m, 3))
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = lib().or_annealed_sinkhorn_solve(_p(x), n, _p(y), m, _p(r), _p(c), C.byref(a),
                                          int(frequency), float(p), lg.cb, None, _p(phi), _p(psi),
                                          _p(img1), _p(img2), C.byref(res))
    assert rc == 0
    return dict(phi=phi, psi=psi, img1=img1, img2=img2, log=lg.lines, **res.as_dict())


def sinkhorn_log_dense(r, c, W, frequency=50, **args):
    r, c, W = map(_f, (r, c, W))
    n, m = r.size, c.size
    phi, psi, P = np.empty(n), np.empty(m), np.empty((n, m))
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = lib().or_sinkhorn_log_dense(_p(r), n, _p(c), m, _p(W), C.byref(a), int(frequency), lg.cb,
                                     None, _p(phi), _p(psi), _p(P), C.byref(res))
    assert rc == 0
    return dict(phi=phi, psi=psi, P=P, log=lg.lines, **res.as_dict())


def grid_cost_matrix(h, w, p):
    W = np.empty((h * w, h * w))
    lib().or_grid_cost_matrix(h, w, float(p), _p(W))
    return W


# ---- feature (cell-similarity) costs: data are (n, d) arrays, passed to C
# column-major as the reference's `data::Matrix` ---------------------------------
METRICS = {"l1": 0, "l2": 1, "cosine": 2, "correlation": 3, "pearson": 3}
SCALES = {"l1": 2.0, "l2": 2.0 ** 0.5, "cosine": 2.0, "correlation": 2.0, "pearson": 2.0}

_SIG_F = {
    "or_feature_stats": (None, [_P, _I64, _I64, _P, _P, _P, _P]),
    "or_feature_cost_matrix": (None, [_I, _P, _I64, _P, _I64, _I64, _D, _P]),
    "or_lamp_cellsim_solve": (_I, [_P, _I64, _I64, _D, _I, _P, _P, C.POINTER(or_args), _I64, LOG_FN,
                                   _P, _P, _P, _P, C.POINTER(or_result)]),
    "or_cpu_lamp_solve": (_I, [_P, _P, _I64, _P, C.POINTER(or_args), _I64, LOG_FN, _P, _P, _P,
                               C.POINTER(or_result)]),
    "or_lamp_cellsim2_solve": (_I, [_P, _I64, _P, _I64, _I64, _D, _I, _P, _P, C.POINTER(or_args),
                                    _I64, LOG_FN, _P, _P, _P, _P, C.POINTER(or_result)]),
    "or_sinkhorn_cellsim_solve": (_I, [_P, _I64, _P, _I64, _I64, _D, _I, _P, _P, C.POINTER(or_args),
                                       _I64, _P, _P, LOG_FN, _P, _P, _P, C.POINTER(or_result)]),
    "or_annealed_sinkhorn_cellsim_solve": (_I, [_P, _I64, _P, _I64, _I64, _D, _I, _P, _P,
                                                C.POINTER(or_args), _I64, LOG_FN, _P, _P, _P,
                                                C.POINTER(or_result)]),
}


def _flib():
    L = lib()
    if not getattr(L, "_f_ready", False):
        for k, (res, args) in _SIG_F.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        L._f_ready = True
    return L


def _colmajor(a):
    return np.asfortranarray(a, dtype=np.float64)


def _fp(a):
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"]
    return a.ctypes.data


def feature_stats(x):
    x = _colmajor(x)
    n, d = x.shape
    s = [np.empty(n) for _ in range(4)]
    _flib().or_feature_stats(_fp(x), n, d, *[_p(v) for v in s])
    return s


def feature_cost_matrix(x, y, metric, scale=None):
    x, y = _colmajor(x), _colmajor(y)
    scale = SCALES[metric] if scale is None else scale
    C_ = np.empty((x.shape[0], y.shape[0]))
    _flib().or_feature_cost_matrix(METRICS[metric], _fp(x), x.shape[0], _fp(y), y.shape[0],
                                   x.shape[1], float(scale), _p(C_))
    return C_


def lamp_cellsim_solve(data, r, c, metric="cosine", scale=None, frequency=100, **args):
    data = _colmajor(data)
    r, c = _f(r), _f(c)
    n, d = data.shape
    scale = SCALES[metric] if scale is None else scale
    nu, phi, psi = np.empty(n), np.empty(n), np.empty(n)
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = _flib().or_lamp_cellsim_solve(_fp(data), n, d, float(scale), METRICS[metric], _p(r), _p(c),
                                       C.byref(a), int(frequency), lg.cb, None, _p(nu), _p(phi),
                                       _p(psi), C.byref(res))
    assert rc == 0
    return dict(nu=nu, phi=phi, psi=psi, log=lg.lines, **res.as_dict())


def cpu_lamp_solve(r, c, W, frequency=50, **args):
    """LAMP(r, c, W::AbstractMatrix, args, frequency) (src/solvers/MirrorProx.jl:
    506-575), the reference's CPU path, on a dense n x n W: returns the plan
    r .* p, theta, the log lines and the result fields (iterations is the
    loop index at exit)."""
    r, c, W = _f(r), _f(c), np.ascontiguousarray(W, dtype=np.float64)
    n = r.size
    assert W.shape == (n, n) and c.size == n
    plan, theta = np.empty((n, n)), np.empty(n)
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = _flib().or_cpu_lamp_solve(_p(r), _p(c), n, _p(W), C.byref(a), int(frequency), lg.cb, None,
                                   _p(plan), _p(theta), C.byref(res))
    assert rc == 0
    return dict(plan=plan, theta=theta, log=lg.lines, **res.as_dict())


def lamp_cellsim2_solve(data1, data2, r, c, metric="cosine", scale=None, frequency=100, **args):
    """extragradient_cellsim with the cost between rows of data1 and rows of
    data2 (lamp_cellsim_solve is the data1 == data2 case)."""
    data1, data2 = _colmajor(data1), _colmajor(data2)
    r, c = _f(r), _f(c)
    n, d = data1.shape
    m = data2.shape[0]
    scale = SCALES[metric] if scale is None else scale
    nu, phi, psi = np.empty(m), np.empty(n), np.empty(m)
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = _flib().or_lamp_cellsim2_solve(_fp(data1), n, _fp(data2), m, d, float(scale),
                                        METRICS[metric], _p(r), _p(c), C.byref(a), int(frequency),
                                        lg.cb, None, _p(nu), _p(phi), _p(psi), C.byref(res))
    assert rc == 0
    return dict(nu=nu, phi=phi, psi=psi, log=lg.lines, **res.as_dict())


def sinkhorn_cellsim_solve(data1, data2, r, c, metric="cosine", scale=None, frequency=100,
                           phi0=None, psi0=None, **args):
    x, y = _colmajor(data1), _colmajor(data2)
    r, c = _f(r), _f(c)
    n, d = x.shape
    m = y.shape[0]
    scale = SCALES[metric] if scale is None else scale
    phi, psi = np.empty(n), np.empty(m)
    p0 = None if phi0 is None else _f(phi0)
    q0 = None if psi0 is None else _f(psi0)
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = _flib().or_sinkhorn_cellsim_solve(_fp(x), n, _fp(y), m, d, float(scale), METRICS[metric],
                                           _p(r), _p(c), C.byref(a), int(frequency), _p(p0), _p(q0),
                                           lg.cb, None, _p(phi), _p(psi), C.byref(res))
    assert rc == 0
    return dict(phi=phi, psi=psi, log=lg.lines, **res.as_dict())


def annealed_sinkhorn_cellsim_solve(data1, data2, r, c, metric="cosine", scale=None,
                                    frequency=100, **args):
    x, y = _colmajor(data1), _colmajor(data2)
    r, c = _f(r), _f(c)
    n, d = x.shape
    m = y.shape[0]
    scale = SCALES[metric] if scale is None else scale
    phi, psi = np.empty(n), np.empty(m)
    a, res, lg = make_args(**args), or_result(), _Log()
    rc = _flib().or_annealed_sinkhorn_cellsim_solve(_fp(x), n, _fp(y), m, d, float(scale),
                                                    METRICS[metric], _p(r), _p(c), C.byref(a),
                                                    int(frequency), lg.cb, None, _p(phi), _p(psi),
                                                    C.byref(res))
    assert rc == 0
    return dict(phi=phi, psi=psi, log=lg.lines, **res.as_dict())


# ---- dense-matrix LAMP (numpy restatement; small sizes) ---------------------

def _softmax_rows(x):
    """softmax(x; norm_dims=2) (src/Utilities.jl:93-105)."""
    mx = x.max(axis=1, keepdims=True)
    v1 = np.exp(x - mx)
    return v1 / v1.sum(axis=1, keepdims=True)


def _lse_rows_np(x):
    """logsumexp(x, 2) (src/Utilities.jl:34-41)."""
    mx = x.max(axis=1, keepdims=True)
    return (np.log(np.exp(x - mx).sum(axis=1, keepdims=True)) + mx)[:, 0]


def _round_plan(g, mu, nu):
    """round(gamma, mu, nu) (src/Utilities.jl:158-165)."""
    gp = g * np.minimum(mu / g.sum(axis=1), 1.0)[:, None]
    gpp = gp * np.minimum(nu / gp.sum(axis=0), 1.0)[None, :]
    rmu = mu - gpp.sum(axis=1)
    rnu = nu - gpp.sum(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return gpp + np.outer(rmu, rnu) / np.abs(rmu).sum()


def _neg_entropy_rows(p):
    """neg_entropy(p, dims=2) (src/Utilities.jl:137-143)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(p > 0.0, p * np.log(np.where(p > 0.0, p, 1.0)), 0.0)
    return t.sum(axis=1)


def dense_lamp_solve(r, c, W, frequency=50, eta_p=0.0, eta_mu=0.0, tau_p=1.0, tau_mu=1.0,
                     alpha=0.01, B=1.0, epsilon=1e-4, itermax=10000, verbose=True, theta0=None,
                     **_ignored):
    """LAMP(r, c, W::CuArray, args, frequency) (src/solvers/MirrorProx.jl:397-503)
    restated in numpy, FP64: returns dict(plan, theta, log, iterations).
    W is len(r) x len(c)."""
    r, c, W = np.asarray(r, float), np.asarray(c, float), np.asarray(W, float)
    winf = np.abs(W).max()                                       # 404
    n = c.size
    theta = np.zeros(n) if theta0 is None else np.array(theta0, float)
    etap = eta_p / 2 / winf                                      # 409
    nu, nub, thetab = theta.copy(), theta.copy(), theta.copy()
    eta_mu_v = (c + alpha / n) / tau_mu                           # 417
    minv, maxv = np.tanh(-B / 2), np.tanh(B / 2)

    def infeas(th, eta):                                          # 425-428
        a = 1.0 / (2 * eta * winf)
        sumvals = _lse_rows_np(-(a * W + (th / eta)[None, :]))
        val = -(0.5 * W / winf + th[None, :]) / eta - sumvals[:, None]
        return (r[:, None] * np.exp(val)).sum(axis=0)

    def update(th0, res, em, adjust):                              # MirrorProx.jl:159-180
        diff = 2.0 * (res - c) / eta_mu_v
        mx = np.maximum(diff, -diff)
        expv = np.exp(diff - mx) * ((th0 + 1) / (1 - th0)) ** (1 - em)
        out = (expv - np.exp(-mx)) / (expv + np.exp(-mx))
        return np.clip(out, minv, maxv) if adjust else out

    def primalv(p):                                               # 28-32
        rp = r[:, None] * p
        return 2 * (0.5 * (W * rp).sum() + (winf * np.abs(rp.sum(axis=0) - c)).sum()
                    + etap * r @ _neg_entropy_rows(p) + etap * r @ np.log(r))

    def dualv(th):                                                # 7-13
        if etap == 0:
            return -(c @ (2 * winf * th)) + r @ (W + 2 * winf * th[None, :]).min(axis=1)
        lse = _lse_rows_np(-(0.5 * W + (winf * th)[None, :]) / etap)
        return 2 * ((-winf * (c @ th) - etap * (r @ lse)) + etap * (r @ np.log(r)))

    etat = np.inf
    with np.errstate(divide="ignore", invalid="ignore"):
        res = infeas(nu, etat)
    log = []
    it_done = 0
    for i in range(1, itermax + 1):
        if (i - 1) % frequency == 0:                              # 444-461
            with np.errstate(divide="ignore", invalid="ignore"):
                p = _softmax_rows(-(W / (2 * winf) + nu[None, :]) / etat)
            obj = (W * _round_plan(r[:, None] * p, r, c)).sum()
            pv, dv = primalv(p), dualv(theta)
            feas = np.abs(c - (r[:, None] * p).sum(axis=0)).sum()
            log.append((i, feas, obj, pv, dv))
            if feas < epsilon / 2 and pv - dv < epsilon / 2:
                break
        thetab = update(theta, res, tau_mu * eta_mu, False)        # 463
        etat = 1 / (tau_p + (1 / etat) * (1 - etap))              # 465
        nub = tau_p * etat * theta + (1 - tau_p * etat) * nu       # 467
        res = infeas(nub, etat)                                    # 475
        theta = update(theta, res, eta_mu, False)                  # 476
        nu = (1 - tau_p * etat) * nu + tau_p * etat * thetab       # 477
        theta = np.clip(theta, minv, maxv)                         # 484
        res = infeas(nu, etat)                                     # 487
        it_done = i
    p = _softmax_rows(-(W / (2 * winf) + nu[None, :]) / etat)     # 501
    return {"plan": r[:, None] * p, "theta": theta, "log": log, "iterations": it_done,
            "eta_t": etat, "winf": winf}
