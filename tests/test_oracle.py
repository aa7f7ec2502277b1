"""The FP64 oracle pinned against (a) the mpmath golden vectors and (b)
closed forms / invariants of the reference's algorithms.  CPU only."""
import math

import numpy as np
import pytest
from scipy.optimize import linprog

from conftest import pval, unhex


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_golden_lamp_kernels(golden, oracle):
    n_cases = 0
    for c in golden:
        if c["kind"] != "lamp":
            continue
        x, y = unhex(c["x"], (-1, 3)), unhex(c["y"], (-1, 3))
        th, r = unhex(c["theta"]), unhex(c["r"])
        p, reg = pval(c["p"]), pval(c["reg"])
        logZ = oracle.lse_rows(x, y, th, reg, c["st"], c["winf"], p)
        assert rel(logZ, c["logZ"]) < 1e-13, c["p"]
        res, cost = oracle.col_marginal_cost(x, y, r, th, unhex(c["logZ_in"]), reg, c["st"], c["winf"], p)
        assert rel(res, c["res"]) < 1e-12
        assert rel(cost, c["cost"]) < 1e-12
        assert rel(oracle.minplus_rows(x, y, th, c["winf"], p), c["minplus"]) < 1e-13
        assert rel(oracle.max_cost_rows(x, y, p), c["maxcost"]) < 1e-14
        if "bary_rows" in c:
            b1 = oracle.barycentric_lamp_rows(x, y, th, r, unhex(c["logZ_in"]), reg, 1.0, c["winf"], p)
            b2 = oracle.barycentric_lamp_cols(x, y, r, th, unhex(c["logZ_in"]), reg, 1.0, c["winf"], p)
            assert rel(b1, c["bary_rows"]) < 1e-12
            assert rel(b2, c["bary_cols"]) < 1e-12
        n_cases += 1
    assert n_cases == 12


def test_golden_sinkhorn_kernels(golden, oracle):
    n_cases = 0
    for c in golden:
        if c["kind"] != "sinkhorn":
            continue
        x, y = unhex(c["x"], (-1, 3)), unhex(c["y"], (-1, 3))
        marg, phi, psi = unhex(c["marg"]), unhex(c["phi"]), unhex(c["psi"])
        p = pval(c["p"])
        half = oracle.sinkhorn_half_step(x, y, marg, psi, c["reg"], c["winf"], p)
        assert rel(half, c["half"]) < 1e-13
        res, cost = oracle.row_residual_cost(x, y, marg, phi, psi, c["reg"], c["winf"], p)
        assert np.max(np.abs(res - np.array(c["resid"]))) < 1e-12 * max(1.0, np.max(np.abs(c["resid"])))
        assert rel(cost, c["rcost"]) < 1e-12
        assert rel(oracle.barycentric_sinkhorn(x, y, phi, psi, c["reg"], c["winf"], p), c["bary"]) < 1e-12
        n_cases += 1
    assert n_cases == 8


def test_golden_theta_update(golden, oracle):
    for c in golden:
        if c["kind"] != "theta":
            continue
        out = oracle.theta_update(unhex(c["theta0"]), unhex(c["res"]), unhex(c["c"]), unhex(c["etamu"]),
                                  c["eta_mu"], c["adjust"], c["minv"], c["maxv"], 1.0)
        assert rel(out, c["out"]) < 1e-13


# ---- closed forms / invariants ------------------------------------------------

def grid(g, scale=1.0):
    k = np.arange(g * g)
    return np.stack([(k // g) * scale, (k % g) * scale, np.zeros(g * g)], axis=1).astype(float)


@pytest.mark.parametrize("p,expect", [(2.0, 2 * 31 ** 2), (1.0, 2 * 31), (math.inf, 31.0)])
def test_grid_winf_closed_form(oracle, p, expect):
    """W_inf on the 32x32 integer grid: 2(g-1)^2, 2(g-1), g-1."""
    x = grid(32)
    assert oracle.max_cost_rows(x, x, p).max() == expect
    # dense get_euclidean_distance agrees (src/Utilities.jl:145-157, p >= 10 is Chebyshev)
    W = oracle.grid_cost_matrix(32, 32, p if p != math.inf else 10.0)
    assert W.max() == expect


def test_lse_identities(oracle):
    rng = np.random.default_rng(1)
    x, y = rng.uniform(size=(6, 3)), rng.uniform(size=(11, 3))
    th = rng.uniform(-1, 1, 11)
    base = oracle.lse_rows(x, y, th, 0.07, 1.0, 1.3, 2.0)
    # theta + kappa shifts the result by -kappa / reg
    sh = oracle.lse_rows(x, y, th + 0.25, 0.07, 1.0, 1.3, 2.0)
    assert np.allclose(sh, base - 0.25 / 0.07, rtol=0, atol=1e-12)
    # st = 0 removes the cost: log sum_j exp(-theta_j / reg)
    z = oracle.lse_rows(x, y, th, 0.07, 0.0, 1.3, 2.0)
    ref = np.log(np.sum(np.exp(-th / 0.07)))
    assert np.allclose(z, ref, rtol=1e-14)
    # reg = inf (first LAMP iteration): log m
    assert np.allclose(oracle.lse_rows(x, y, th, math.inf, 1.0, 1.3, 2.0), np.log(11), rtol=1e-15)
    # m = 1: the exponent itself
    one = oracle.lse_rows(x, y[:1], th[:1], 0.07, 1.0, 1.3, 2.0)
    c0 = np.sum((x - y[0]) ** 2, axis=1)
    assert np.allclose(one, -(c0 / 2.6 + th[0]) / 0.07, rtol=1e-13)


def test_col_marginal_mass_invariant(oracle):
    """With logZ the row LSE, sum_j res_j = sum_i r_i (rows normalised)."""
    rng = np.random.default_rng(2)
    x, y = rng.uniform(size=(40, 3)), rng.uniform(size=(33, 3))
    th = rng.uniform(-1, 1, 33)
    r = rng.uniform(0.1, 1, 40)
    r /= r.sum()
    for p in (1.0, 2.0, math.inf):
        logZ = oracle.lse_rows(x, y, th, 0.03, 1.0, 2.0, p)
        res, cost = oracle.col_marginal_cost(x, y, r, th, logZ, 0.03, 1.0, 2.0, p)
        assert abs(res.sum() - 1.0) < 1e-13
        assert np.all(cost >= 0)


def test_sinkhorn_matches_dense_and_lp(oracle):
    """Kernel Sinkhorn (on-the-fly, relative reg) and dense sinkhorn_log on the
    normalised W reach the same unique entropic fixed point; small reg
    approaches the LP optimum (scipy HiGHS)."""
    g = 4
    x = grid(g)
    rng = np.random.default_rng(3)
    r = rng.uniform(size=g * g) + 1e-3
    r /= r.sum()
    c = rng.uniform(size=g * g) + 1e-3
    c /= c.sum()
    eta = 0.05
    sk = oracle.sinkhorn_solve(x, x, r, c, frequency=1, p=2.0, eta_p=eta, epsilon=6e-12,
                               itermax=200000, verbose=False)
    W = oracle.grid_cost_matrix(g, g, 2.0)
    dense = oracle.sinkhorn_log_dense(r, c, W / W.max(), frequency=1, eta_p=eta, epsilon=6e-12,
                                      itermax=200000, verbose=False)
    P_k = np.exp(-W / W.max() / eta + sk["phi"][:, None] + sk["psi"][None, :])
    assert np.allclose(P_k, dense["P"], atol=1e-10)
    assert np.allclose(P_k.sum(1), r, atol=1e-10) and np.allclose(P_k.sum(0), c, atol=1e-12)
    # transport cost in raw units vs the LP optimum
    n = g * g
    A = np.zeros((2 * n, n * n))
    for i in range(n):
        A[i, i * n:(i + 1) * n] = 1
        A[n + i, i::n] = 1
    lp = linprog(W.ravel(), A_eq=A, b_eq=np.concatenate([r, c]), bounds=(0, None), method="highs")
    sk_small = oracle.sinkhorn_solve(x, x, r, c, frequency=10, p=2.0, eta_p=2e-3, epsilon=1e-9,
                                     itermax=400000, verbose=False)
    assert abs(sk_small["objective"] - lp.fun) / lp.fun < 0.05
    assert sk_small["objective"] >= lp.fun * (1 - 1e-9)


def test_lamp_converges_to_lp(oracle):
    """LAMP (MirrorProxKernel.jl:1019-1132) with small eta_p reaches the stop
    test and its transport cost approaches the LP optimum."""
    g = 4
    x = grid(g, 1.0 / (g - 1))
    rng = np.random.default_rng(4)
    r = rng.uniform(size=g * g) + 1e-3
    r /= r.sum()
    c = rng.uniform(size=g * g) + 1e-3
    c /= c.sum()
    # eta_p = 0: the reference's own stop test (primal - dual and infeasibility)
    out = oracle.lamp_solve(x, x, r, c, frequency=10, p=2.0, eta_p=0.0, epsilon=6e-6, B=1.1,
                            itermax=200000, verbose=True)
    assert out["converged"] == 1
    assert out["residual"] < 1e-6 and out["primal"] - out["dual"] < 1e-6
    # eta_p > 0: infeasibility stop
    out2 = oracle.lamp_solve(x, x, r, c, frequency=10, p=2.0, eta_p=1e-3, epsilon=6e-6, B=1.1,
                             itermax=200000, verbose=False, stop="infeas")
    assert out2["converged"] == 1 and out2["residual"] < 1e-6
    n = g * g
    W = np.sum((x[:, None, :] - x[None, :, :]) ** 2, axis=2)
    A = np.zeros((2 * n, n * n))
    for i in range(n):
        A[i, i * n:(i + 1) * n] = 1
        A[n + i, i::n] = 1
    lp = linprog(W.ravel(), A_eq=A, b_eq=np.concatenate([r, c]), bounds=(0, None), method="highs")
    assert abs(out["objective"] - lp.fun) / lp.fun < 1e-3
    assert abs(out2["objective"] - lp.fun) / lp.fun < 0.05
    # log format: header + rows with 7 fields ending in lamp_kernel
    assert out["log"][0] == "time(s),iter,infeas,ot_objective,primal,dual,solver\n"
    assert all(l.rstrip("\n").split(",")[-1] == "lamp_kernel" and len(l.split(",")) == 7 for l in out["log"][1:])


def test_theta_update_clamp_bounds(oracle):
    m = 50
    rng = np.random.default_rng(5)
    out = oracle.theta_update(rng.uniform(-0.99, 0.99, m), rng.uniform(0, 10, m), np.zeros(m),
                              np.full(m, 1e-3), 0.0, 1, math.tanh(-0.55), math.tanh(0.55), 1.0)
    assert np.all(out <= math.tanh(0.55)) and np.all(out >= math.tanh(-0.55))


def test_annealing_schedule(oracle):
    """eta_k = max(q^k eta_0, eps): stages and final eta (AnnealedSinkhorn.jl:79,128)."""
    g = 3
    x = grid(g)
    r = np.full(g * g, 1.0 / (g * g))
    out = oracle.annealed_sinkhorn_solve(x, x, r, r, frequency=5, p=2.0, eta_p=0.1, epsilon=0.02,
                                         anneal_mult=0.5, itermax=10000, verbose=True)
    etas = [float(l.split(",")[4]) for l in out["log"][1:]]
    assert etas == [0.1, 0.05, 0.025, 0.02]
    assert out["eta_t"] == 0.02


def test_thread_count_invariance(oracle):
    rng = np.random.default_rng(6)
    x, y = rng.uniform(size=(300, 3)), rng.uniform(size=(200, 3))
    th = rng.uniform(-1, 1, 200)
    oracle.set_threads(1)
    a = oracle.lse_rows(x, y, th, 0.01, 1.0, 3.0, 2.0)
    oracle.set_threads(4)
    b = oracle.lse_rows(x, y, th, 0.01, 1.0, 3.0, 2.0)
    oracle.set_threads(1)
    assert np.array_equal(a, b)


def test_golden_feature_costs(golden, oracle):
    k = 0
    for c in golden:
        if c["kind"] != "feature":
            continue
        x = unhex(c["x"], (c["n"], c["d"]))
        y = unhex(c["y"], (c["m"], c["d"]))
        C = oracle.feature_cost_matrix(x, y, c["metric"])
        # l2 is sqrt(|x|^2 + |y|^2 - 2 x.y) as in the reference: an FP64
        # cancellation of size eps becomes sqrt(eps) ~ 1.5e-8 after the root
        tol = 5e-8 if c["metric"] == "l2" else 1e-14
        assert np.max(np.abs(C - np.array(c["C"]))) < tol, c["metric"]
        k += 1
    assert k == 4


def test_feature_cost_closed_forms(oracle):
    """cosine: identical direction -> 0, orthogonal -> 1/2; correlation:
    perfectly (anti)correlated -> 0 (1); l1/l2 of identical rows 0."""
    x = np.array([[1.0, 2.0, 3.0, 0.0], [0.0, 0.0, 0.0, 5.0]])
    y = np.array([[2.0, 4.0, 6.0, 0.0], [3.0, 2.0, 1.0, 4.0]])
    C = oracle.feature_cost_matrix(x, y, "cosine")
    assert abs(C[0, 0]) < 1e-15 and abs(C[1, 0] - 0.5) < 1e-15
    a = np.array([[1.0, 2.0, 3.0, 4.0]])
    b = np.array([[2.0, 4.0, 6.0, 8.0], [4.0, 3.0, 2.0, 1.0]])
    Cc = oracle.feature_cost_matrix(a, b, "correlation")
    assert abs(Cc[0, 0]) < 1e-15 and abs(Cc[0, 1] - 1.0) < 1e-14
    # Pearson correlation of the row-sum-normalised rows equals that of the raw rows
    from scipy.stats import pearsonr
    rho = pearsonr(a[0], b[1])[0]
    assert abs(Cc[0, 1] - (1 - rho) / 2) < 1e-12
    assert np.allclose(np.diag(oracle.feature_cost_matrix(b, b, "l1")), 0.0)
    assert np.allclose(np.diag(oracle.feature_cost_matrix(b, b, "l2")), 0.0, atol=1e-7)


def test_cellsim_sinkhorn_matches_point_machinery(oracle):
    """With the cosine cost precomputed, the cellsim Sinkhorn reaches the same
    unique fixed point as the dense sinkhorn_log on C / W_inf."""
    rng = np.random.default_rng(21)
    X = np.abs(rng.normal(size=(30, 12))) * (rng.uniform(size=(30, 12)) < 0.5) + 1e-3
    r, c = rng.dirichlet(np.ones(30)), rng.dirichlet(np.ones(30))
    eta = 0.05
    o = oracle.sinkhorn_cellsim_solve(X, X, r, c, "cosine", frequency=1, eta_p=eta, epsilon=6e-12,
                                      itermax=100000, verbose=False)
    C = oracle.feature_cost_matrix(X, X, "cosine")
    W = max(C.max(), 0.0)
    dense = oracle.sinkhorn_log_dense(r, c, C / W, frequency=1, eta_p=eta, epsilon=6e-12,
                                      itermax=100000, verbose=False)
    P = np.exp(-C / (eta * W) + o["phi"][:, None] + o["psi"][None, :])
    assert np.allclose(P, dense["P"], atol=1e-10)


def test_cellsim_lamp_stop_and_annealing(oracle):
    rng = np.random.default_rng(22)
    X = np.abs(rng.normal(size=(25, 10))) + 0.1
    r, c = rng.dirichlet(np.ones(25)), rng.dirichlet(np.ones(25))
    o = oracle.lamp_cellsim_solve(X, r, c, "correlation", frequency=10, eta_p=1e-3, epsilon=6e-6,
                                  B=1.1, itermax=50000, verbose=False, stop="infeas")
    assert o["converged"] == 1 and o["residual"] < 1e-6
    a = oracle.annealed_sinkhorn_cellsim_solve(X, X, r, c, "l1", frequency=5, eta_p=0.2,
                                               epsilon=0.05, anneal_mult=0.5, verbose=True)
    etas = [float(l.split(",")[4]) for l in a["log"][1:]]
    assert etas == [0.2, 0.1, 0.05]
    assert a["log"][1].rstrip().endswith("annealed_sinkhorn_cellsim")


def test_lamp_infeas_stop_skips_product_plan_check(oracle):
    """With uniform target weights the plan of the first check (eta_t = inf:
    r (x) 1/m) is exactly feasible; the infeasibility stop must not end there.
    At eta_p > 0 it then stops where the reference's gap-and-infeasibility
    stop (MirrorProxKernel.jl:1081-1101) does, with the same state."""
    from culamp_amd import workloads
    x, y, r, c = workloads.gmm_clouds(150, 150)
    assert np.allclose(c, 1.0 / 150)
    kw = dict(frequency=10, p=2.0, eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000)
    inf = oracle.lamp_solve(x, y, r, c, verbose=False, stop="infeas", **kw)
    ref = oracle.lamp_solve(x, y, r, c, verbose=True, stop="reference", **kw)
    prod = float(r @ ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1).mean(1))
    assert inf["converged"] == 1 and inf["iterations"] > 0 and inf["residual"] < 1e-6
    assert inf["objective"] < 0.6 * prod
    assert ref["converged"] == 1 and ref["iterations"] == inf["iterations"]
    assert ref["objective"] == inf["objective"]
    assert np.array_equal(ref["nu"], inf["nu"])
    # eta_p = 0: the reference's stop does not fire (its gap closes slowly);
    # the infeasibility stop does, after real iterations
    z = oracle.lamp_solve(x, y, r, c, verbose=False, stop="infeas", **{**kw, "eta_p": 0.0})
    assert z["converged"] == 1 and z["iterations"] > 0 and z["objective"] < 0.6 * prod


def _feature_instance(seed, n, d, metric):
    rng = np.random.default_rng(seed)
    X = np.abs(rng.normal(size=(n, d))) * (rng.uniform(size=(n, d)) < 0.5) + 1e-3
    r, c = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
    return X, r, c


def test_cellsim_lamp_converges_to_entropic_plan_at_2winf_eta(oracle):
    """extragradient_cellsim keeps eta = eta_p (MirrorProxCellSim.jl:118) on
    the cost C / (2 W_inf) (281-282), so at convergence (eta_t -> eta_p,
    infeasibility 1e-11) its plan r_i softmax_j(-(C_ij / (2 W_inf) + nu_j) /
    eta_t) is the entropic OT plan of the raw cost at regularisation
    2 W_inf eta_p, which an independent dense log-domain Sinkhorn gives.  A
    restatement with the point-cloud convention eta = eta_p / (2 W_inf) would
    converge to the plan at regularisation eta_p instead (1e-3 away here)."""
    from scipy.special import logsumexp
    X, r, c = _feature_instance(7, 30, 10, "cosine")
    C = oracle.feature_cost_matrix(X, X, "cosine", 2.0)
    W = C.max()
    for eta_p in (0.05, 0.2):
        o = oracle.lamp_cellsim_solve(X, r, c, "cosine", 2.0, frequency=10, eta_p=eta_p,
                                      epsilon=6e-11, B=1.1, itermax=400000, verbose=False,
                                      stop="infeas")
        assert o["converged"] == 1 and o["winf"] == W
        assert abs(o["eta_t"] - eta_p) < 1e-8 * eta_p
        L = -(C / (2 * W) + o["nu"][None, :]) / o["eta_t"]
        P = r[:, None] * np.exp(L - logsumexp(L, axis=1, keepdims=True))
        ref = oracle.sinkhorn_log_dense(r, c, C, frequency=10, eta_p=2 * W * eta_p, epsilon=6e-14,
                                        itermax=200000, verbose=False)["P"]
        assert np.abs(P - ref).max() < 1e-8 * ref.max()
        other = oracle.sinkhorn_log_dense(r, c, C, frequency=10, eta_p=eta_p, epsilon=6e-14,
                                          itermax=200000, verbose=False)["P"]
        assert np.abs(P - other).max() > 1e-4 * other.max()
        # psi = -2 W_inf nu / eta_p (MirrorProxCellSim.jl:184)
        assert np.allclose(o["psi"], -2 * W * o["nu"] / eta_p, rtol=0, atol=1e-12 * W / eta_p)


def test_cellsim_lamp_eta0_reaches_lp_with_weak_duality(oracle):
    """eta_p = 0: the cellsim LAMP's cost tends to the LP optimum (HiGHS), and
    at every logged check its dual (the averaged-gamma min-plus form,
    MirrorProxCellSim.jl:140, 157-159, 179) and primal bracket it."""
    from scipy.optimize import linprog
    rng = np.random.default_rng(8)
    n = 12
    X = np.abs(rng.normal(size=(n, 6))) + 0.05
    r, c = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
    C = oracle.feature_cost_matrix(X, X, "l1", 2.0)
    A = np.zeros((2 * n, n * n))
    for i in range(n):
        A[i, i * n:(i + 1) * n] = 1
        A[n + i, i::n] = 1
    lp = linprog(C.ravel(), A_eq=A, b_eq=np.concatenate([r, c]), bounds=(0, None),
                 method="highs").fun
    o = oracle.lamp_cellsim_solve(X, r, c, "l1", 2.0, frequency=10, eta_p=0.0, epsilon=6e-4,
                                  B=1.1, itermax=400000, verbose=True)
    assert o["converged"] == 1
    assert abs(o["objective"] - lp) < 1e-6 * lp
    rows = [l.split(",") for l in o["log"][1:]]
    primal = np.array([float(q[4]) for q in rows])
    dual = np.array([float(q[5]) for q in rows])
    assert np.all(dual <= lp + 1e-12) and np.all(primal >= lp - 1e-12)
    assert primal[-1] - dual[-1] < 1e-4


@pytest.mark.parametrize("kind", ["points", "cellsim"])
def test_annealed_sinkhorn_ends_at_the_eps_fixed_point(oracle, kind):
    """Annealed Sinkhorn (AnnealedSinkhorn.jl:65-177, AnnealedSinkhornCellSim.jl:
    1-63) ends with a Sinkhorn stage at eta = epsilon stopped at infeasibility
    epsilon / 6: its plan is within a few times that residual (L1) of the
    fully converged Sinkhorn plan at epsilon, and far from the one at
    2 epsilon (so a schedule that stopped a stage early would fail)."""
    rng = np.random.default_rng(8)
    eps = 1e-2
    if kind == "points":
        x, y = rng.uniform(size=(40, 3)), rng.uniform(size=(35, 3))
        a, b = rng.dirichlet(np.ones(40)), rng.dirichlet(np.ones(35))
        C = ((x[:, None] - y[None]) ** 2).sum(-1)
        an = oracle.annealed_sinkhorn_solve(x, y, a, b, frequency=5, p=2.0, eta_p=0.5, epsilon=eps,
                                            anneal_mult=0.5, verbose=False)
        direct = lambda e: oracle.sinkhorn_solve(x, y, a, b, frequency=5, p=2.0, eta_p=e,  # noqa: E731
                                                 epsilon=6e-14, itermax=400000, verbose=False)
    else:
        X = np.abs(rng.normal(size=(12, 6))) + 0.05
        a, b = rng.dirichlet(np.ones(12)), rng.dirichlet(np.ones(12))
        C = oracle.feature_cost_matrix(X, X, "l1", 2.0)
        an = oracle.annealed_sinkhorn_cellsim_solve(X, X, a, b, "l1", 2.0, frequency=5, eta_p=0.5,
                                                    epsilon=eps, anneal_mult=0.5, verbose=False)
        direct = lambda e: oracle.sinkhorn_cellsim_solve(X, X, a, b, "l1", 2.0, frequency=5,  # noqa: E731
                                                         eta_p=e, epsilon=6e-14, itermax=400000,
                                                         verbose=False)
    W = C.max()

    def plan(e, phi, psi):
        return np.exp(-C / (e * W) + phi[:, None] + psi[None, :])

    assert an["residual"] <= eps / 6
    pa = plan(eps, an["phi"], an["psi"])
    d1 = direct(eps)
    d2 = direct(2 * eps)
    l1 = np.abs(pa - plan(eps, d1["phi"], d1["psi"])).sum()
    l2 = np.abs(pa - plan(2 * eps, d2["phi"], d2["psi"])).sum()
    assert l1 <= 4 * an["residual"], (l1, an["residual"])
    assert l1 < 0.1 * l2, (l1, l2)


def test_cpu_lamp_restatement_lp_limit_and_entropic_fixed_point(oracle):
    """The reference's CPU LAMP (src/solvers/MirrorProx.jl:506-575), restated
    in the oracle.  eta_p = 0: its own stop (gap and infeasibility below
    epsilon/6, tested at every check) ends the solve with the LP cost (HiGHS)
    and every logged primal / dual bracket the optimum.  eta_p > 0: eta_t
    tends to eta_p / (2 W_inf) and the plan to the dense Sinkhorn plan at
    regularisation eta_p (the exponent is -(W / 2 W_inf + nu) / eta_t)."""
    from scipy.optimize import linprog
    rng = np.random.default_rng(3)
    n = 16
    x = rng.uniform(size=(n, 2))
    W = ((x[:, None] - x[None]) ** 2).sum(-1)
    r, c = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
    A = np.zeros((2 * n, n * n))
    for i in range(n):
        A[i, i * n:(i + 1) * n] = 1
        A[n + i, i::n] = 1
    lp = linprog(W.ravel(), A_eq=A, b_eq=np.concatenate([r, c]), bounds=(0, None),
                 method="highs").fun
    o = oracle.cpu_lamp_solve(r, c, W, frequency=10, eta_p=0.0, epsilon=6e-6, B=1.1,
                              itermax=300000, verbose=True)
    assert o["converged"] == 1
    assert abs(o["objective"] - lp) < 1e-5 * lp
    rows = [l.split(",") for l in o["log"][1:]]
    assert all(q[-1].strip() == "lamp" and len(q) == 7 for q in rows)
    assert o["log"][0] == "time(s),iter,infeas,ot_objective,primal,dual,solver\n"
    primal = np.array([float(q[4]) for q in rows])
    dual = np.array([float(q[5]) for q in rows])
    assert np.all(dual <= lp + 1e-12) and np.all(primal >= lp - 1e-12)
    eta = 0.05
    o = oracle.cpu_lamp_solve(r, c, W, frequency=10, eta_p=eta, epsilon=6e-6, B=1.1,
                              itermax=3000, verbose=False)
    assert abs(o["eta_t"] - eta / (2 * W.max())) < 1e-12 * eta
    d = oracle.sinkhorn_log_dense(r, c, W, frequency=10, eta_p=eta, epsilon=6e-14, itermax=200000,
                                  verbose=False)
    assert np.abs(o["plan"] - d["P"]).max() < 1e-10 * d["P"].max()
    assert abs(o["plan"].sum(0) - c).sum() < 1e-12
