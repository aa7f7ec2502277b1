"""The differentiable Sinkhorn divergence at its benchmark size: two clouds
of n = m = 65536 points in 3-D (x ~ N(0, I), y a 4-component Gaussian
mixture, weights Uniform(0.5, 1.5) normalised), squared Euclidean cost,
eps = 1e-2 max C.  The gradient (mixed forward to infeasibility 1e-6, FP64
polish to 1e-9, adjoint unrolled over the last 64 iterations with
checkpoints every 16) against central finite differences of FP64 forwards
to infeasibility 1e-13: 32 seeded coordinates of x (h = 1e-3) and 8
mass-preserving pairs of weights (a_i + h, a_j - h, h = 1e-7: the weights are
about 1.5e-5, and the central difference's truncation error from the a log a
terms grows as h^2 / a^2)."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_divergence_gradient_65536_vs_finite_differences(gpu_ctx):
    import culamp_amd
    from culamp_amd import workloads
    x, y, a, b = workloads.divergence_clouds()
    n = len(x)
    t0 = time.perf_counter()
    (val, gx, ga, gy, gb), terms = culamp_amd.sinkhorn_divergence(
        x, a, y, b, ctx=gpu_ctx, eta_p=1e-2, tol=1e-6, polish_tol=1e-9, unroll=64, checkpoint=16,
        precision="mixed", return_terms=True, tmax=60.0)
    say("value", val, "eps", terms[0]["eta_t"], "forward+backward",
        f"{time.perf_counter() - t0:.2f}s", [(t["iterations"], t["residual"]) for t in terms])
    assert all(t["converged"] for t in terms)
    eps = terms[0]["eta_t"]
    kw = dict(grad=False, ctx=gpu_ctx, eps=eps, tol=1e-13, precision="fp64", itermax=20000,
              tmax=60.0)
    rng = np.random.default_rng(503)
    coords = [(int(rng.integers(n)), int(rng.integers(3))) for _ in range(32)]
    h = 1e-3
    fd_x = []
    for q, (i, k) in enumerate(coords):
        xp, xm = x.copy(), x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        fd = (culamp_amd.sinkhorn_divergence(xp, a, y, b, **kw) -
              culamp_amd.sinkhorn_divergence(xm, a, y, b, **kw)) / (2 * h)
        fd_x.append(fd)
        say(f"  x[{i},{k}] fd {fd:.10e} grad {gx[i, k]:.10e} "
            f"rel {abs(fd - gx[i, k]) / max(abs(fd), 1e-300):.2e} ({time.perf_counter() - t0:.0f}s)")
    fd_x = np.array(fd_x)
    g_x = np.array([gx[i, k] for i, k in coords])
    ex = np.max(np.abs(g_x - fd_x)) / np.max(np.abs(fd_x))
    h_a = 1e-7
    pairs = [tuple(int(v) for v in rng.choice(n, 2, replace=False)) for _ in range(8)]
    fd_a = []
    for i, j in pairs:
        ap, am = a.copy(), a.copy()
        ap[i] += h_a; ap[j] -= h_a
        am[i] -= h_a; am[j] += h_a
        fd = (culamp_amd.sinkhorn_divergence(x, ap, y, b, **kw) -
              culamp_amd.sinkhorn_divergence(x, am, y, b, **kw)) / (2 * h_a)
        fd_a.append(fd)
        say(f"  a[{i}]-a[{j}] fd {fd:.10e} grad {ga[i] - ga[j]:.10e} "
            f"({time.perf_counter() - t0:.0f}s)")
    fd_a = np.array(fd_a)
    g_a = np.array([ga[i] - ga[j] for i, j in pairs])
    ea = np.max(np.abs(g_a - fd_a)) / np.max(np.abs(fd_a))
    say("relative error: x", ex, "a", ea, f"total {time.perf_counter() - t0:.0f}s")
    assert ex < 1e-4 and ea < 1e-4
