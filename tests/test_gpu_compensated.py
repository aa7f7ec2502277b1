"""Compensated FP64 sums on the GPU.

Marginal and cost sums of the pair kernels and the check reductions (the
infeasibility, cost and dual sums) add with TwoSum and carry the rounding
error in a second FP64 word.  These cases are built so that every term is
exact, the exact sum is known (math.fsum of the same doubles) and plain FP64
summation of the same terms is measurably off."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ULP = 2.0 ** -52


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_fp64_column_marginal_is_compensated(gpu_ctx):
    """All points coincide (C = 0), theta = 0, logZ = 0, reg = 1: the FP64
    column pass's w_ij = r_i exactly, so every column marginal is sum_i r_i.
    The weights r = 1e8 (g - mean g) over 65536 points cancel to a sum about
    1e-12 of their magnitude; plain FP64 summation (sequential, pairwise, or
    strided lanes plus a tree) is off by percents of it, the compensated pass
    (every term TwoSum-folded, compensated lane and split merges) gives it
    to 1e-9 relative in every column.  The cost sums are exact zeros."""
    import culamp_amd
    from culamp_amd import kernels as K
    rng = np.random.default_rng(11)
    n, m = 65536, 96
    g = rng.standard_normal(n)
    r = 1e8 * (g - g.mean())
    exact = math.fsum(r)
    seq = float(np.add.accumulate(r)[-1])
    pairwise = float(np.sum(r))
    say("exact", repr(exact), "plain sequential rel err", (seq - exact) / exact,
        "pairwise rel err", (pairwise - exact) / exact)
    assert abs(seq - exact) > 1e-4 * abs(exact) and abs(pairwise - exact) > 1e-4 * abs(exact)
    x = np.zeros((n, 3))
    y = np.zeros((m, 3))
    dev = lambda a: culamp_amd.DeviceArray.from_host(gpu_ctx, np.ascontiguousarray(a, float))
    dx, dy, dr = dev(x), dev(y), dev(r)
    th, lz = dev(np.zeros(m)), dev(np.zeros(n))
    res = culamp_amd.DeviceArray(gpu_ctx, m)
    cost = culamp_amd.DeviceArray(gpu_ctx, m)
    K.residual_spp_c(gpu_ctx, res, cost, dx, dy, dr, th, lz, 1.0, 1.0, 1.0, 2.0,
                     precision=culamp_amd.FP64)
    got = res.get()
    err = float(np.max(np.abs(got - exact))) / abs(exact)
    say("gpu column marginal max rel err", err)
    assert err <= 1e-9
    assert np.all(cost.get() == 0.0)


def test_dual_reduction_is_compensated(gpu_ctx):
    """The dual <r, phi> + <c, psi> of a Sinkhorn solve run for zero
    iterations from a warm start: r = 2^-12 (n = 4096), phi = K (g - mean g)
    with K = 1e8, psi = 0.  Each product r_i phi_i is exact and the terms
    cancel to a sum many orders below their size; the compensated reduction
    (FMA-exact products, TwoSum folds and butterflies) returns it within a few
    ulps of the exact sum, where plain FP64 summation is off by orders of
    magnitude more."""
    import culamp_amd
    from culamp_amd import workloads
    rng = np.random.default_rng(3)
    n = m = 4096
    x, y, _, _ = workloads.grid_histograms(64)
    r = np.full(n, 2.0 ** -12)
    c = np.full(m, 2.0 ** -12)
    g = rng.standard_normal(n)
    phi0 = 1e8 * (g - g.mean())
    psi0 = np.zeros(m)
    exact = math.fsum(phi0) * 2.0 ** -12  # (power-of-two scaling: exact)
    terms = phi0 * 2.0 ** -12
    plain = float(np.add.accumulate(terms)[-1])
    say("exact", repr(exact), "plain sequential err", plain - exact,
        "pairwise err", float(np.sum(terms)) - exact)
    assert abs(plain - exact) > 1e4 * max(abs(exact), 1e-300) * ULP
    pb = culamp_amd.Problem(x, y, r, c, gpu_ctx)
    args = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, itermax=0, verbose=False)
    _, _, _, _, info = pb.sinkhorn(args, 10, 2.0, phi0=phi0, psi0=psi0, log=None,
                                   assignments=False)
    pb.close()
    say("gpu dual", repr(info["dual"]), "err", info["dual"] - exact)
    assert info["iterations"] == 0
    assert abs(info["dual"] - exact) <= 4 * abs(exact) * ULP + 1e-300
