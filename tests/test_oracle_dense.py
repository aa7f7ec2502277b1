"""Pins of the oracle's numpy restatement of the dense-matrix LAMP
(oracle.dense_lamp_solve, src/solvers/MirrorProx.jl:397-503) without a GPU:
the one-point closed form, weak duality at every check, the plan's row sums,
and agreement at convergence with the oracle's on-the-fly LAMP (C
restatement of MirrorProxKernel.jl), which solves the same entropic problem."""
import numpy as np


def test_one_point_closed_form(oracle):
    o = oracle.dense_lamp_solve([1.0], [1.0], [[3.0]], frequency=5, eta_p=0.1, epsilon=1e-6,
                                itermax=50)
    it, feas, obj, pv, dv = o["log"][0]
    # the plan is exactly feasible, so round()'s correction divides 0 by
    # |r_mu|_1 = 0: the reference's ot_objective is NaN here
    assert it == 1 and feas == 0.0 and np.isnan(obj)
    assert abs(pv - 3.0) < 1e-14 and abs(dv - 3.0) < 1e-14
    assert o["iterations"] == 0  # stopped at the first check


def test_weak_duality_and_rows(oracle):
    from culamp_amd import workloads
    x, y, r, c = workloads.uniform_grid_pair(6)
    W = oracle.grid_cost_matrix(6, 6, 2.0)
    for eta in (1e-2, 0.0):
        o = oracle.dense_lamp_solve(r, c, W, frequency=10, eta_p=eta, epsilon=1e-9, B=1.1,
                                    itermax=300)
        for it, feas, obj, pv, dv in o["log"]:
            assert pv - dv >= -1e-12 * max(1.0, abs(pv)), (eta, it, pv, dv)
            assert feas >= 0.0 and obj > 0.0
        np.testing.assert_allclose(o["plan"].sum(axis=1), r, rtol=1e-13)


def test_converged_plan_matches_on_the_fly_lamp(oracle):
    """Both LAMP forms reach the unique entropic optimum of the same problem
    (eta = eta_p / (2 W_inf) on C = W): the transport costs agree."""
    from culamp_amd import workloads
    x, y, r, c = workloads.uniform_grid_pair(6, unit=False)
    W = oracle.grid_cost_matrix(6, 6, 2.0)
    d = oracle.dense_lamp_solve(r, c, W, frequency=10, eta_p=1e-1, epsilon=1e-12, B=1.1,
                                itermax=6000)
    k = oracle.lamp_solve(x, y, r, c, frequency=10, p=2.0, eta_p=1e-1, epsilon=1e-11, B=1.1,
                          itermax=20000, verbose=False, stop="infeas")
    dense_cost = float((W * d["plan"]).sum())
    assert d["log"][-1][1] < 1e-9  # infeasibility at the last check
    assert abs(dense_cost - k["objective"]) < 1e-6 * k["objective"], (dense_cost, k["objective"])
