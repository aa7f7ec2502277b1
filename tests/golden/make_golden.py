"""Write golden vectors for the hot-path kernels, computed independently of
the oracle with mpmath at 50 significant digits (brute force over all pairs,
no max shift, no streaming).

    python tests/golden/make_golden.py      # -> tests/golden/kernels.json

Each case stores its seeded float64 inputs exactly (as hex floats) and the
expected outputs rounded to float64.  The formulas are the reference
kernels' definitions (file:line in the case's "ref" field).
"""
from __future__ import annotations

import json
import math
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def hx(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64).ravel()]


def cost(a, b, p):
    d = [abs(mp.mpf(float(a[k])) - mp.mpf(float(b[k]))) for k in range(3)]
    if p == 1:
        return d[0] + d[1] + d[2]
    if p == 2:
        return d[0] ** 2 + d[1] ** 2 + d[2] ** 2
    if p == math.inf:
        return max(d)
    return sum(mp.power(v, p) for v in d)


def lse(vals):
    return mp.log(mp.fsum(mp.exp(v) for v in vals))


def pts(rng, n, scale=1.0):
    return rng.uniform(0.0, scale, size=(n, 3))


def main():
    rng = np.random.default_rng(20261203)
    cases = []
    for p in (1.0, 2.0, math.inf, 3.0):
        for (n, m, reg) in ((5, 7, 0.5), (6, 4, 0.02), (3, 9, math.inf)):
            x, y = pts(rng, n), pts(rng, m)
            theta = rng.uniform(-1, 1, m)
            r = rng.uniform(0.1, 1.0, n)
            r /= r.sum()
            C = [[cost(x[i], y[j], p) for j in range(m)] for i in range(n)]
            winf = float(max(max(row) for row in C))
            st = 1.0
            # MirrorProxKernel.jl:833-938
            invreg = mp.mpf(0) if reg == math.inf else 1 / mp.mpf(reg)
            c1 = mp.mpf(st) / (2 * mp.mpf(winf))
            V = [[-(C[i][j] * c1 + mp.mpf(float(theta[j]))) * invreg for j in range(m)] for i in range(n)]
            logZ = [lse(V[i]) for i in range(n)]
            # MirrorProxKernel.jl:7-94
            res = [mp.fsum(mp.exp(V[i][j] - logZ[i]) * mp.mpf(float(r[i])) for i in range(n)) for j in range(m)]
            cst = [mp.fsum(mp.exp(V[i][j] - logZ[i]) * C[i][j] * mp.mpf(float(r[i])) for i in range(n)) for j in range(m)]
            # MirrorProxKernel.jl:941-1016
            mins = [min(C[i][j] + 2 * mp.mpf(winf) * mp.mpf(float(theta[j])) for j in range(m)) for i in range(n)]
            maxs = [max(C[i]) for i in range(n)]
            logZf = np.array([float(v) for v in logZ])
            case = dict(kind="lamp", ref="src/solvers/MirrorProxKernel.jl:7-94,280-326,833-938,941-1016",
                        p="inf" if p == math.inf else p, reg="inf" if reg == math.inf else reg,
                        st=st, winf=winf, x=hx(x), y=hx(y), theta=hx(theta), r=hx(r),
                        logZ=[float(v) for v in logZ], res=[float(v) for v in res],
                        cost=[float(v) for v in cst], minplus=[float(v) for v in mins],
                        maxcost=[float(v) for v in maxs], logZ_in=hx(logZf))
            if reg != math.inf:
                # MirrorProxKernel.jl:97-189 barycentric maps, weights from the
                # float64 logZ handed to the kernel
                W = [[mp.exp(V[i][j] - mp.mpf(logZf[i])) * mp.mpf(float(r[i])) for j in range(m)] for i in range(n)]
                b1 = [[float(mp.fsum(W[i][j] * mp.mpf(float(y[j][k])) for j in range(m)) / mp.fsum(W[i])) for k in range(3)] for i in range(n)]
                b2 = [[float(mp.fsum(W[i][j] * mp.mpf(float(x[i][k])) for i in range(n)) / mp.fsum(W[i][j] for i in range(n))) for k in range(3)] for j in range(m)]
                case["bary_rows"] = b1
                case["bary_cols"] = b2
            cases.append(case)
    for p in (1.0, 2.0, math.inf, 3.0):
        for (n, m, reg) in ((5, 7, 0.3), (8, 3, 0.05)):
            x, y = pts(rng, n), pts(rng, m)
            marg = rng.uniform(0.1, 1.0, n)
            marg /= marg.sum()
            psi = rng.normal(0, 2, m)
            phi = rng.normal(0, 2, n)
            C = [[cost(x[i], y[j], p) for j in range(m)] for i in range(n)]
            winf = float(max(max(row) for row in C))
            k = -1 / (mp.mpf(reg) * mp.mpf(winf))
            # SinkhornKernel.jl:341-444
            half = [mp.log(mp.mpf(float(marg[i]))) - lse([C[i][j] * k + mp.mpf(float(psi[j])) for j in range(m)]) for i in range(n)]
            # SinkhornKernel.jl:9-81
            E = [[mp.exp(C[i][j] * k + mp.mpf(float(phi[i])) + mp.mpf(float(psi[j]))) for j in range(m)] for i in range(n)]
            resid = [mp.mpf(float(marg[i])) - mp.fsum(E[i]) for i in range(n)]
            rcost = [mp.fsum(E[i][j] * C[i][j] for j in range(m)) for i in range(n)]
            # SinkhornKernel.jl:124-162
            bary = [[float(mp.fsum(E[i][j] * mp.mpf(float(y[j][kk])) for j in range(m)) / mp.fsum(E[i])) for kk in range(3)] for i in range(n)]
            cases.append(dict(kind="sinkhorn", ref="src/solvers/SinkhornKernel.jl:9-81,124-162,341-444",
                              p="inf" if p == math.inf else p, reg=reg, winf=winf, x=hx(x), y=hx(y),
                              marg=hx(marg), phi=hx(phi), psi=hx(psi),
                              half=[float(v) for v in half], resid=[float(v) for v in resid],
                              rcost=[float(v) for v in rcost], bary=bary))
    # update_theta_residual, src/solvers/MirrorProx.jl:159-180: closed form
    # tanh((res - c) W / eta_mu_i + (1 - eta_mu) atanh(theta0)), clamped.
    for eta_mu in (0.0, 0.3):
        for adjust in (0, 1):
            m = 9
            theta0 = rng.uniform(-0.9, 0.9, m)
            res = rng.uniform(0, 0.3, m)
            c = rng.uniform(0, 0.3, m)
            etamu = rng.uniform(0.05, 0.4, m)
            B = 1.1
            minv, maxv = math.tanh(-B / 2), math.tanh(B / 2)
            out = []
            for j in range(m):
                t = mp.tanh((mp.mpf(float(res[j])) - mp.mpf(float(c[j]))) / mp.mpf(float(etamu[j]))
                            + (1 - mp.mpf(eta_mu)) * mp.atanh(mp.mpf(float(theta0[j]))))
                if adjust:
                    t = min(max(t, mp.mpf(minv)), mp.mpf(maxv))
                out.append(float(t))
            cases.append(dict(kind="theta", ref="src/solvers/MirrorProx.jl:159-180", eta_mu=eta_mu,
                              adjust=adjust, minv=minv, maxv=maxv, theta0=hx(theta0), res=hx(res),
                              c=hx(c), etamu=hx(etamu), out=out))
    # feature costs (src/CellUtilities.jl:187-301), scales 78-84, stats 304-311
    scales = {"l1": 2, "l2": mp.sqrt(2), "cosine": 2, "correlation": 2}
    for metric in ("l1", "l2", "cosine", "correlation"):
        n, m, d = 6, 5, 9
        x = np.abs(rng.normal(size=(n, d))) * (rng.uniform(size=(n, d)) < 0.6)
        y = np.abs(rng.normal(size=(m, d))) * (rng.uniform(size=(m, d)) < 0.6)
        x[0, :] = 0.0
        x[0, 2] = 1.0  # a one-hot row (zero centred norm for correlation)
        y[1] = x[1] * 2.5  # proportional rows: cosine / correlation 1, l1 / l2 0
        def stats(a):
            out = []
            for row in a:
                v = [mp.mpf(float(t)) for t in row]
                s = mp.fsum(v)
                mu = s / len(v)
                out.append((s, mp.fsum(t * t for t in v), mu, mp.fsum((t - mu) ** 2 for t in v), v))
            return out
        sx, sy = stats(x), stats(y)
        sc = scales[metric]
        C = []
        for i in range(n):
            row = []
            for j in range(m):
                si, qi, mi, ci, vi = sx[i]
                sj, qj, mj, cj, vj = sy[j]
                if metric == "l1":
                    val = mp.fsum(abs(a / si - b / sj) for a, b in zip(vi, vj)) / sc
                else:
                    dot = mp.fsum(a * b for a, b in zip(vi, vj)) / (si * sj)
                    if metric == "l2":
                        val = mp.sqrt(max(qi / si ** 2 + qj / sj ** 2 - 2 * dot, 0)) / sc
                    elif metric == "cosine":
                        den = mp.sqrt((qi / si ** 2) * (qj / sj ** 2))
                        val = 0 if den <= 0 else (1 - max(-1, min(1, dot / den))) / sc
                    else:
                        mui, muj = mi / si, mj / sj
                        den = mp.sqrt((ci / si ** 2) * (cj / sj ** 2))
                        if den <= 0:
                            val = 0 if abs(mui - muj) <= mp.sqrt(mp.mpf(2) ** -52) else 1 / sc
                        else:
                            val = (1 - max(-1, min(1, (dot - d * mui * muj) / den))) / sc
                row.append(float(val))
            C.append(row)
        cases.append(dict(kind="feature", ref="src/CellUtilities.jl:187-311", metric=metric,
                          x=hx(x), y=hx(y), n=n, m=m, d=d, C=C))
    with open(os.path.join(HERE, "kernels.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "dps": 50, "cases": cases}, f)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
