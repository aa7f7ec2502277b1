"""Benchmark: fused on-the-fly cost + LSE on one MI355X, batched and plain.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one batched LAMP solve, to L1 infeasibility 1e-6 (stop rule
"infeas", epsilon = 6e-6), of 256 independent cell-similarity instances
(n, m log-uniform in [256, 4096], d = 128 sparse unit feature rows, cosine
cost / 2, Dirichlet(1) weights, eta_p = 5e-2): one launch per pass covers
every active instance, the cost tile is a bf16 MFMA x.y^T with FP32
accumulation, potentials and sums are FP64, the tile shape is autotuned and
the iterations between checks replay a captured HIP graph.  Inputs are
resident in HBM.

Printed as one JSON line (rank 0):
  value        dense-equivalent cost-LSE GB/s over the timed solves: per pass
               and instance n*m*4 B (the FP32 cost matrix a dense path would
               stream) + the O(n + m) row and vector bytes, / time;
  roofline     the all-instances batched row-LSE pass (MFMA kernel) timed
               with HIP events on the library's stream: 256 bf16 flops per
               pair / average launch time, against 2.5 PFLOP/s, with the
               measured HBM traffic and the VALU / exponential bound beside it;
  cpu_baseline the FP64 oracle (dense restatement of the reference's cellsim
               LAMP loop) on a time-capped sample of the same instances.
Extra fields: batched Sinkhorn; the same instances one after the other on
the single-instance path in bf16 (what batching buys) and in FP64 (the plain
path); the 64x64-grid LAMP at eta_p = 1e-2 and 1e-4 (p = 1, 2, inf); the
16384-point RGB clouds; the 65536-point Sinkhorn divergence with gradients;
the C-index over all cell pairs; the dense cost-matrix comparator.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "culamp.jl_amd"))

METRIC = "cost-LSE GB/s + solves/s to 1e-6 infeasibility, batched/diff vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0


def pass_bytes(n: int, m: int, b_c: int, b_x: int = 8) -> float:
    """Algorithmic bytes of one O(n*m) pass: n*m*b_c + (n+m)*3*b_x + 8(n+m) + 8n."""
    return n * m * b_c + (n + m) * 3 * b_x + 8 * (n + m) + 8 * n


MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense bf16
TRAFFIC_BATCH_FILE = os.path.join(ROOT, "profiles", "traffic_mfma_batch_cellsim256.json")
# VALU issue slots per pair in the batched bf16 row-LSE loop (gfx950 assembly
# of the offset-folded fast path: per 16 pairs of a lane 16 v_fma, 16 v_exp
# (half rate: two slots each), 16 v_add and 4 more, plus the loop head's
# address arithmetic): the pass's own VALU bound is the measured FMA issue
# rate / this
BATCH_LOOP_SLOTS_PER_PAIR = (16 + 2 * 16 + 16 + 4 + 4) / 16


def cell_pass_bytes(n: int, m: int, d: int = 128, b_x: int = 2) -> float:
    """Dense-equivalent bytes of one feature pass: the n*m FP32 cost matrix
    plus the (n + m) bf16 feature rows and the FP64 vectors."""
    return n * m * 4 + (n + m) * d * b_x + 8 * (n + m) + 8 * n


def batched_headline(ctx, a, log):
    """The headline workload: `batch_count` cell-similarity instances solved
    by one batched LAMP to L1 infeasibility 1e-6 per step.  Returns the batch,
    its instances and the timed record."""
    import numpy as np
    import culamp_amd
    from culamp_amd import workloads
    insts = workloads.cell_batch(a.batch_count)
    b = culamp_amd.Batch(insts, ctx, metric="cosine", scale=2.0)
    prec = a.batch_precision
    tune, _ = b.autotune(prec, reps=3)
    lp = culamp_amd.EOTArgs(eta_p=5e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False,
                            precision=prec, stop="infeas")
    for _ in range(a.warmup):
        b.lamp(lp, a.frequency)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = b.lamp(lp, a.frequency)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    infos = [r[-1] for r in res]
    gbytes = sum(i["passes"] * cell_pass_bytes(x.shape[0], y.shape[0])
                 for i, (x, y, _, _) in zip(infos, insts)) / 1e9
    pairs = sum(i["pairs"] for i in infos)
    rec = {"value": a.steps * gbytes / dt, "ms_per_step": dt / a.steps * 1e3,
           "solves_per_s": a.steps * len(insts) / dt, "tuning": tune, "precision": prec,
           "converged": int(sum(i["converged"] for i in infos)),
           "max_infeasibility": max(i["residual"] for i in infos),
           "iterations_mean": float(np.mean([i["iterations"] for i in infos])),
           "pairs_per_step": pairs,
           "pairs_all_active_pass": float(sum(x.shape[0] * y.shape[0] for x, y, _, _ in insts)),
           "mfma_tflops": a.steps * pairs * 256 * (3 if prec == "bf16x3" else 1) / 1e12 / dt,
           "objectives": [i["objective"] for i in infos], "args": lp}
    log(f"batched lamp ({prec}): {a.steps} x {len(insts)} solves in {dt:.3f}s, "
        f"{rec['converged']} converged, max infeas {rec['max_infeasibility']:.2e}, "
        f"{rec['value']:.0f} GB/s dense-eq")
    return b, insts, rec


def batch_roofline(ctx, b, insts, a, log) -> dict:
    """The dominant kernel: the batched row-LSE pass over all 256 instances
    (mfma_feature_pass_batch), `kernel_launches` launches between HIP events
    on the library's stream."""
    prec = a.batch_precision
    ms, launches = b.time_pass(prec, reps=a.kernel_launches)
    pairs = sum(x.shape[0] * y.shape[0] for x, y, _, _ in insts)
    flops = pairs * 256 * (3 if prec == "bf16x3" else 1)
    achieved = flops / (ms * 1e-3) / 1e12
    dense_bytes = sum(cell_pass_bytes(x.shape[0], y.shape[0]) for x, y, _, _ in insts)
    traffic = None
    if os.path.exists(TRAFFIC_BATCH_FILE):
        with open(TRAFFIC_BATCH_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    rates = ctx.probe_rates()
    loop_bound = rates["fma_per_s"] / BATCH_LOOP_SLOTS_PER_PAIR
    pps = pairs / (ms * 1e-3)
    log(f"batched row-LSE pass: {ms * 1e3:.1f} us/launch ({launches} kernel launch(es) per pass), "
        f"{achieved:.0f} TFLOP/s bf16, {pps / 1e12:.2f} T pairs/s")
    return {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / MFMA_PEAK_TFLOPS, 4), "traffic": traffic,
            "kernel": f"mfma_feature_pass_batch<KIND_LSE, 1, 8> ({prec} MFMA cosine, d = 128), "
                      "all 256 instances active",
            "avg_launch_us": round(ms * 1e3, 3), "launches_per_pass": launches,
            "flops_per_launch": flops, "pairs_per_launch": pairs,
            "dense_eq": {"bytes_per_launch": dense_bytes,
                         "GBps": dense_bytes / (ms * 1e-3) / 1e9,
                         "frac_of_8TBps": dense_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "compute": {"bound": "v_exp_f32 + VALU", "unit": "pairs/s", "achieved": pps,
                        "exp_per_s_measured": rates["exp_per_s"],
                        "fma_per_s_measured": rates["fma_per_s"],
                        "loop_slots_per_pair": BATCH_LOOP_SLOTS_PER_PAIR,
                        "loop_bound_pairs_per_s": loop_bound,
                        "frac_of_loop_bound": pps / loop_bound}}


def batched_extras(ctx, b, insts, head, a, log) -> dict:
    """Batched Sinkhorn, and the same instances solved one after the other on
    the single-instance path in the batch's precision (what batching buys)
    and in FP64 (the plain path: host-checked stop every `frequency`
    iterations, no batching, no graph), each time-capped."""
    import culamp_amd
    prec = a.batch_precision
    lp = head["args"]
    sk = culamp_amd.EOTArgs(eta_p=5e-2, epsilon=6e-6, itermax=20000, verbose=False, precision=prec)
    b.sinkhorn(sk, a.frequency)
    ctx.synchronize()
    t0 = time.perf_counter()
    res = b.sinkhorn(sk, a.frequency)
    ctx.synchronize()
    dsk = time.perf_counter() - t0
    sinfos = [r[-1] for r in res]
    out = {"sinkhorn": {"solves_per_s": len(insts) / dsk, "seconds": dsk,
                        "converged": int(sum(i["converged"] for i in sinfos)),
                        "max_infeasibility": max(i["residual"] for i in sinfos)}}
    log(f"batched sinkhorn ({prec}): {len(insts)} solves in {dsk * 1e3:.1f} ms")
    sobj = [i["objective"] for i in sinfos]
    pbs = [culamp_amd.FeatureProblem(x, y, 2.0, "cosine", r, c, ctx) for x, y, r, c in insts]

    def sequential(name, args, cap):
        objs, done = [], 0
        t1 = time.perf_counter()
        for pb in pbs:
            if name == "sinkhorn":
                _, _, info = pb.sinkhorn(args, a.frequency, log=None)
            else:
                _, _, _, info = pb.lamp(args, a.frequency, log=None)
            objs.append(info["objective"])
            done += 1
            if time.perf_counter() - t1 > cap:
                break
        return done / (time.perf_counter() - t1), done, objs

    for name, args, bobj, bsolves in (("lamp", lp, head["objectives"], head["solves_per_s"]),
                                      ("sinkhorn", sk, sobj, out["sinkhorn"]["solves_per_s"])):
        rec = out.setdefault(name, {})
        rate_s, n_s, o_s = sequential(name, args, a.plain_seconds / 4)
        rate_f, n_f, o_f = sequential(name, args.copy(precision="fp64"), a.plain_seconds)
        rec.update({
            f"sequential_{prec}_solves_per_s": rate_s, f"sequential_{prec}_instances_timed": n_s,
            "batching_speedup_same_precision": bsolves / rate_s,
            "sequential_same_precision_max_rel_diff": max(abs(u - v) / abs(v) for u, v in
                                                          zip(bobj[:n_s], o_s)),
            "plain_fp64_solves_per_s": rate_f, "plain_fp64_instances_timed": n_f,
            "speedup_vs_plain_fp64": bsolves / rate_f,
            "objective_max_rel_err_vs_fp64": max(abs(u - v) / abs(v) for u, v in
                                                 zip(bobj[:n_f], o_f))})
        log(f"{name}: sequential {prec} {rate_s:.1f} solves/s ({n_s} timed), plain fp64 "
            f"{rate_f:.1f} solves/s ({n_f} timed); batched {bsolves:.0f} solves/s = "
            f"{rec['batching_speedup_same_precision']:.1f}x / {rec['speedup_vs_plain_fp64']:.1f}x, "
            f"cost rel err vs fp64 {rec['objective_max_rel_err_vs_fp64']:.1e}")
    for pb in pbs:
        pb.close()
    return out


def cpu_baseline_batched(insts, head, a, log) -> dict:
    """The FP64 oracle (dense restatement of the reference's cellsim LAMP loop,
    src/solvers/MirrorProxCellSim.jl:94-188, between two point sets) on the
    headline's instances in order until `cpu_seconds` have passed, at the
    host-thread share and at one thread; GB/s with the headline's byte
    accounting, so the ratio of the two values is the speedup."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the FP64 CPU oracle, timed here as the CPU baseline
    nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    kw = dict(eta_p=5e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False, stop="infeas")

    def run(threads, seconds):
        oracle.set_threads(threads)
        done, gbytes, t0 = 0, 0.0, time.perf_counter()
        for x, y, r, c in insts:
            o = oracle.lamp_cellsim2_solve(x, y, r, c, "cosine", 2.0, frequency=a.frequency, **kw)
            it = max(o["iterations"], 1)
            checks = (it + a.frequency - 1) // a.frequency
            gbytes += (4 * it + 3 * checks) * cell_pass_bytes(x.shape[0], y.shape[0]) / 1e9
            done += 1
            if time.perf_counter() - t0 > seconds:
                break
        dt = time.perf_counter() - t0
        return gbytes / dt, done / dt, done

    v, sps, done = run(share, a.cpu_seconds)
    v1, sps1, done1 = run(1, a.cpu_seconds / 2)
    log(f"cpu oracle: {share} threads {done} instances, {sps:.2f} solves/s ({v:.2f} GB/s dense-eq); "
        f"1 thread {done1} instances, {sps1:.3f} solves/s")
    return {"value": round(v, 3), "unit": "GB/s", "cores": share, "kind": "port",
            "sample": f"FP64 oracle cellsim LAMP (dense cost matrix) on the first {done} of the same "
                      f"{len(insts)} instances (time cap {a.cpu_seconds:.0f} s), bytes counted as "
                      "for value",
            "solves_per_s": sps, "nproc": nproc,
            "one_thread": {"value": round(v1, 3), "solves_per_s": sps1, "instances": done1}}


def bench_cindex(ctx, a, log) -> dict:
    """Cell-similarity C-index (the reference's multi-instance workload,
    src/CellSimilarity.jl:164-255): every pair of `cindex_cells` cells of
    synthetic clustered data with 96 features, cosine cost between features,
    LAMP distances to L1 infeasibility 1e-6 (eta_p = 5e-2); all pairs in
    batches on one shared device copy of the features (bf16 MFMA), against
    the plain FP64 single-instance path pair by pair (time-capped sample)."""
    import numpy as np
    import culamp_amd
    from culamp_amd import cindex
    cd = cindex.synthetic_cell_data(96, a.cindex_cells, 4, seed=3)
    kern = cindex.CellCostKernel.build(cd.data, "cosine")
    kw = dict(eta_p=5e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False, stop="infeas")
    npairs = a.cindex_cells * (a.cindex_cells - 1) // 2
    args = culamp_amd.EOTArgs(precision=a.batch_precision, **kw)
    cindex.pair_distances(kern, cindex._pair_order(a.cindex_cells, 64), args, a.frequency, "lamp",
                          64, ctx)  # warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    out = cindex.streaming_c_index(cd, kern, args, npairs, frequency=a.frequency, solver="lamp",
                                   batch_size=a.cindex_batch, ctx=ctx)
    dt = time.perf_counter() - t0
    pairs = cindex._pair_order(a.cindex_cells, npairs)
    f64 = culamp_amd.EOTArgs(precision="fp64", **kw)
    sums = cd.data.sum(axis=0)
    data = np.asfortranarray(cd.data)
    done, t1 = 0, time.perf_counter()
    for i, j in pairs:
        cindex.extragradient_cell_distance(data, cindex.cell_marginal(cd.data, i, sums),
                                           cindex.cell_marginal(cd.data, j, sums), kern.scale,
                                           "cosine", f64, a.frequency, ctx)
        done += 1
        if time.perf_counter() - t1 > a.plain_seconds / 4:
            break
    plain_rate = done / (time.perf_counter() - t1)
    log(f"c-index: {npairs} pairs in {dt:.2f}s batched ({npairs / dt:.0f} pairs/s), "
        f"plain fp64 {plain_rate:.1f} pairs/s, c_index {out['c_index']:.4f}")
    return {"workload": f"C-index over all {npairs} pairs of {a.cindex_cells} synthetic cells "
                        "(96 features, 4 clusters), cosine/2 cost between features, LAMP to L1 "
                        "infeasibility 1e-6, eta_p=5e-2",
            "precision": a.batch_precision, "pairs": npairs, "seconds": dt,
            "pairs_per_s": npairs / dt, "c_index": out["c_index"],
            "plain_fp64_pairs_per_s": plain_rate, "plain_fp64_pairs_timed": done,
            "speedup_vs_plain_fp64": (npairs / dt) / plain_rate}


def _sq_cost_matrix(x, y):
    import numpy as np
    W = np.empty((len(x), len(y)))
    for i0 in range(0, len(x), 512):
        d = x[i0:i0 + 512, None, :] - y[None, :, :]
        W[i0:i0 + 512] = (d * d).sum(-1)
    return W


def bench_dense(ctx, a, log) -> dict:
    """The reference's plain dense GPU path (row LSE and column marginal over
    a stored cost matrix, MirrorProx.jl:236-333) as the HBM-streaming
    comparator: the same passes as the on-the-fly kernels on g x g grids
    (squared Euclidean), the FP32 matrix and its transpose resident in HBM,
    each pass timed from graph replays; achieved = matrix bytes + vector
    bytes per pass / time."""
    import numpy as np
    import culamp_amd
    from culamp_amd import MIXED, workloads
    from culamp_amd import kernels as K

    def timed(fn, reps=20):
        fn()
        fn()
        g = ctx.capture(lambda: [fn() for _ in range(reps)])
        g.launch()
        ctx.timer_start()
        for _ in range(5):
            g.launch()
        ms = ctx.timer_stop() / (5 * reps)
        g.close()
        return ms * 1e3

    out = []
    for grid in a.dense_grids:
        x, y, r, c = workloads.grid_histograms(grid)
        n = len(x)
        W = _sq_cost_matrix(x, y)
        winf = float(W.max())
        th = np.random.default_rng(0).uniform(-1e-3, 1e-3, n)
        dev = lambda v: culamp_amd.DeviceArray.from_host(ctx, v)  # noqa: E731
        dth, dr, dx, dy = dev(th), dev(r), dev(x), dev(y)
        o, res = culamp_amd.DeviceArray(ctx, n), culamp_amd.DeviceArray(ctx, n)
        dm = culamp_amd.DenseProblem(W, "mixed", ctx)
        del W
        reg = a.eta / 4.0
        t_lse = timed(lambda: dm.lse_rows(o, dth, reg, 1.0, winf))
        lz = dev(o.get())
        t_col = timed(lambda: dm.col_marginal(res, dr, dth, lz, reg, 1.0, winf))
        t_otf = timed(lambda: K.warp_logsumexp_spp_ct_fused(ctx, o, dx, dy, dth, reg, 1.0, winf,
                                                            2.0, MIXED))
        # the reference's dense LAMP loop over the same matrix: iterations/s
        # (4 dense passes per iteration, one check at the first iteration)
        lamp_args = culamp_amd.EOTArgs(eta_p=a.eta, epsilon=1e-30, B=1.1, itermax=200,
                                       verbose=False)
        dm.lamp(r, c, lamp_args.copy(itermax=5), 10 ** 9, plan=False, log=None)
        ctx.synchronize()
        t1 = time.perf_counter()
        _, _, dl = dm.lamp(r, c, lamp_args, 10 ** 9, plan=False, log=None)
        dense_it_s = dl["iterations"] / (time.perf_counter() - t1)
        dm.close()
        stream = n * n * 4 + 3 * 8 * n
        rec = {"n": n, "m": n, "matrix_bytes": n * n * 4,
               "dense_lse_us": t_lse, "dense_col_us": t_col, "on_the_fly_lse_us": t_otf,
               "dense_lse_gbs": stream / (t_lse * 1e-6) / 1e9,
               "dense_col_gbs": (stream + 8 * n) / (t_col * 1e-6) / 1e9,
               "dense_lse_frac": stream / (t_lse * 1e-6) / 1e9 / HBM_PEAK_GBS,
               "on_the_fly_speedup_vs_dense": t_lse / t_otf,
               "dense_lamp_iterations_per_s": dense_it_s}
        log(f"dense comparator n={n}: dense LSE {t_lse:.1f} us ({rec['dense_lse_gbs']:.0f} GB/s), "
            f"dense column {t_col:.1f} us, on-the-fly LSE {t_otf:.1f} us "
            f"({rec['on_the_fly_speedup_vs_dense']:.2f}x); dense LAMP {dense_it_s:.0f} it/s")
        out.append(rec)
    return {"workload": "row LSE / column marginal over a stored FP32 cost matrix (and its "
                        "transpose) on g x g grids, squared Euclidean, against the on-the-fly "
                        "mixed row LSE on the same points",
            "sizes": out}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_lse_mixed_grid64.json")


def copy_ceiling(ctx, log, mib: int = 512, reps: int = 10) -> dict:
    """Measured HBM stream ceiling on the box: a device-to-device copy of
    `mib` MiB (read + write) timed with HIP events on the library's stream."""
    import culamp_amd
    n = (mib << 20) // 8
    src, dst = culamp_amd.DeviceArray(ctx, n), culamp_amd.DeviceArray(ctx, n)
    for _ in range(2):
        culamp_amd._lib.check(ctx.lib.culamp_memcpy_dtod(ctx.handle, dst.ptr, src.ptr, n * 8))
    ctx.timer_start()
    for _ in range(reps):
        culamp_amd._lib.check(ctx.lib.culamp_memcpy_dtod(ctx.handle, dst.ptr, src.ptr, n * 8))
    ms = ctx.timer_stop() / reps
    gbs = 2 * n * 8 / (ms * 1e-3) / 1e9
    log(f"copy ceiling: {mib} MiB device-to-device in {ms * 1e3:.1f} us, {gbs:.0f} GB/s read + write")
    return {"bytes_moved": 2 * n * 8, "us": ms * 1e3, "GBps": gbs, "frac_of_8TBps": gbs / HBM_PEAK_GBS}


def mfma_ceiling(ctx, roofline: dict, log) -> dict:
    """Measured dense bf16 MFMA ceiling on the box (32x32x16 MFMA chains at
    three waves per SIMD, the pair pass's occupancy), and the batched pass's
    fraction of it."""
    tf = ctx.probe_mfma() / 1e12
    log(f"mfma ceiling: {tf:.0f} TFLOP/s bf16 sustained ({tf / MFMA_PEAK_TFLOPS:.2f} of the published peak)")
    return {"TFLOPs": round(tf, 1), "frac_of_peak": round(tf / MFMA_PEAK_TFLOPS, 4),
            "pass_frac_of_ceiling": round(roofline["achieved"] / tf, 4) if tf > 0 else None}


def bench_shard(ctx, a, log) -> dict:
    """configs[4] at one GPU's share of an 8-GPU row sharding: the row LSE of
    8192 of the 65,536 source points against all 65,536 targets (the pass each
    GPU would run per half step; the all-gather of the other side's 65,536
    FP64 potentials, 512 KiB, is not run), mixed and FP64, graph-replayed and
    timed with HIP events."""
    import numpy as np
    import culamp_amd
    from culamp_amd import kernels as K
    from culamp_amd import workloads
    x, y, _, _ = workloads.divergence_clouds(65536, 65536)
    rows = 65536 // 8
    xs = np.ascontiguousarray(x[:rows])
    C = ((x[:64, None, :] - y[None, :4096, :]) ** 2).sum(-1)
    reg = 1e-2  # of the cost scale below
    winf = float(C.max())
    th = np.zeros(len(y))
    dx = culamp_amd.DeviceArray.from_host(ctx, xs)
    dy = culamp_amd.DeviceArray.from_host(ctx, y)
    dth = culamp_amd.DeviceArray.from_host(ctx, th)
    out = culamp_amd.DeviceArray(ctx, rows)
    rec = {"workload": f"row LSE of {rows} of the 65,536 divergence source points (N(0, I3)) against "
                       "all 65,536 targets (4-component GMM), squared Euclidean, reg = 1e-2 of the "
                       "sampled max cost: one GPU's pass of an 8-GPU row sharding",
           "rows": rows, "cols": len(y), "allgather_bytes_per_half_step": len(y) * 8}
    for prec, name in ((culamp_amd.MIXED, "mixed"), (culamp_amd.FP64, "fp64")):
        def one():
            K.warp_logsumexp_spp_ct_fused(ctx, out, dx, dy, dth, reg, 1.0, winf, 2.0, prec)
        for _ in range(3):
            one()
        g = ctx.capture(lambda: [one() for _ in range(5)])
        g.launch()
        ctx.timer_start()
        for _ in range(4):
            g.launch()
        ms = ctx.timer_stop() / 20
        g.close()
        rec[name] = {"us": ms * 1e3, "pairs_per_s": rows * len(y) / (ms * 1e-3)}
        log(f"shard {rows}x{len(y)} {name}: {ms * 1e3:.1f} us per row LSE, "
            f"{rows * len(y) / (ms * 1e-3) / 1e12:.2f} T pairs/s")
    return rec


def bench_divergence(ctx, a, log) -> dict:
    """Differentiable Sinkhorn divergence: x ~ N(0, I_3) (n points), y a
    4-component Gaussian mixture, Uniform(0.5, 1.5) weights; squared
    Euclidean cost, eps = eta_p * max C (eta_p = 1e-2); forward to L1
    infeasibility 1e-6 (mixed, then FP64 to 1e-9), backward = envelope +
    adjoint of the last 64 iterations of each solve, checkpoints every 16."""
    import numpy as np
    import culamp_amd
    from culamp_amd import workloads
    x, y, wa, wb = workloads.divergence_clouds(a.diff_n, a.diff_n)
    kw = dict(ctx=ctx, eta_p=1e-2, tol=1e-6, unroll=64, checkpoint=16, precision="mixed",
              frequency=10)
    t0 = time.perf_counter()
    (val, gx, ga, gy, gb), terms = culamp_amd.sinkhorn_divergence(x, wa, y, wb, return_terms=True,
                                                                  **kw)
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    val0, terms0 = culamp_amd.sinkhorn_divergence(x, wa, y, wb, grad=False, return_terms=True, **kw)
    dtf = time.perf_counter() - t1
    pairs = sum(t["pairs"] for t in terms)
    log(f"divergence n={a.diff_n}: value {val:.6e}, forward+backward {dt:.2f}s, forward {dtf:.2f}s")
    return {"workload": f"Sinkhorn divergence, n = m = {a.diff_n}, x ~ N(0, I3), y 4-component GMM, "
                        "squared Euclidean, eps = 1e-2 * max C, forward to infeasibility 1e-6 "
                        "(mixed + FP64 polish to 1e-9), unrolled adjoint over 64 iterations, "
                        "checkpoints every 16",
            "value": val, "seconds_forward_backward": dt, "seconds_forward": dtf,
            "eps": terms[0]["eta_t"],
            "terms": [{k: t[k] for k in ("iterations", "residual", "objective", "converged")}
                      for t in terms],
            "forward_pairs": pairs, "grad_x_norm": float(np.linalg.norm(gx)),
            "grad_a_centred_norm": float(np.linalg.norm(ga - ga.mean()))}


def bench_clouds(ctx, a, log) -> dict:
    """Colour-transfer-shaped clouds: n = m = 16384 RGB points, 5- and
    8-component Gaussian mixtures, uniform weights, squared Euclidean,
    eta_p = 1e-2; Sinkhorn and LAMP to L1 infeasibility 1e-6 (checked every
    `frequency` iterations), mixed precision against the plain FP64 path."""
    import culamp_amd
    from culamp_amd import workloads
    x, y, r, c = workloads.gmm_clouds(a.clouds_n, a.clouds_n)
    n, m = len(x), len(y)
    pb = culamp_amd.Problem(x, y, r, c, ctx)
    base = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, B=1.1, itermax=20000, verbose=False,
                              stop="infeas")
    out = {"workload": f"two Gaussian-mixture RGB clouds (5 / 8 components), n = m = {n}, "
                       "uniform weights, squared Euclidean, eta_p = 1e-2; Sinkhorn and LAMP to L1 "
                       "infeasibility 1e-6"}
    for solver in ("sinkhorn", "lamp"):
        rec = {}
        for prec in ("mixed", "fp64"):
            args = base.copy(precision=prec)
            fn = (lambda: pb.sinkhorn(args, a.frequency, 2.0, log=None, assignments=False)[-1]) \
                if solver == "sinkhorn" else \
                (lambda: pb.lamp(args, a.frequency, 2.0, log=None, assignments=False)[-1])
            if prec == "mixed":
                fn()
            ctx.synchronize()
            t0 = time.perf_counter()
            info = fn()
            ctx.synchronize()
            dt = time.perf_counter() - t0
            b_pair = 4 if prec == "mixed" else 8
            gbs = (info["pairs"] * b_pair + info["passes"] * pass_bytes(n, m, 0)) / 1e9 / dt
            rec[prec] = {"seconds": dt, "iterations": info["iterations"],
                         "iterations_per_s": info["iterations"] / dt,
                         "pairs_per_s": info["pairs"] / dt,
                         "residual": info["residual"], "objective": info["objective"],
                         "converged": info["converged"], "dense_eq_GBps": gbs}
        rec["speedup_mixed_vs_fp64"] = rec["fp64"]["seconds"] / rec["mixed"]["seconds"]
        # (LAMP's cost still moves between the two precisions' stop checks)
        om, of = rec["mixed"]["objective"], rec["fp64"]["objective"]
        rec["objective_rel_diff"] = abs(om - of) / abs(of)
        if solver == "lamp":
            # both precisions after the FP64 solve's iteration count (epsilon 0
            # never stops; the check at it = K + 1 reports the state after K)
            k = rec["fp64"]["iterations"]
            fixed = base.copy(stop="reference", verbose=True, epsilon=0.0, itermax=k + 1)
            same = {p: pb.lamp(fixed.copy(precision=p), a.frequency, 2.0, log=None,
                               assignments=False)[-1] for p in ("mixed", "fp64")}
            rec["same_iterations"] = {
                "iterations": k,
                "objective_rel_diff": abs(same["mixed"]["objective"] - same["fp64"]["objective"])
                / abs(same["fp64"]["objective"]),
                "infeasibility_mixed": same["mixed"]["residual"],
                "infeasibility_fp64": same["fp64"]["residual"]}
        log(f"clouds n={n} {solver}: mixed {rec['mixed']['seconds'] * 1e3:.1f} ms "
            f"({rec['mixed']['iterations']} it, {rec['mixed']['dense_eq_GBps']:.0f} GB/s dense-eq), "
            f"fp64 {rec['fp64']['seconds'] * 1e3:.1f} ms ({rec['fp64']['iterations']} it); "
            f"objective rel diff {rec['objective_rel_diff']}")
        out[solver] = rec
    pb.close()
    return out


def lse_roofline(ctx, x, y, theta, a, log) -> dict:
    """Roofline of the dominant kernel: the mixed row LSE (pair_pass + split
    merge, culamp_lse_rows) timed with HIP events on the library's stream."""
    import culamp_amd
    from culamp_amd import kernels as K
    n, m = len(x), len(y)
    dx = culamp_amd.DeviceArray.from_host(ctx, x)
    dy = culamp_amd.DeviceArray.from_host(ctx, y)
    dth = culamp_amd.DeviceArray.from_host(ctx, theta)
    dout = culamp_amd.DeviceArray(ctx, n)
    reg = a.eta / 4.0  # eta_t at convergence = eta_p / (2 W_inf), W_inf = 2 on the unit grid

    def one():
        K.warp_logsumexp_spp_ct_fused(ctx, dout, dx, dy, dth, reg, 1.0, 2.0, 2.0, culamp_amd.MIXED)

    for _ in range(5):
        one()
    # the launches are replayed from a captured graph so that host-side call
    # overhead cannot leave the GPU idle between them
    per_graph = 25
    graph = ctx.capture(lambda: [one() for _ in range(per_graph)])
    graph.launch()
    reps = max(1, a.kernel_launches // per_graph)
    ctx.timer_start()
    for _ in range(reps):
        graph.launch()
    ms = ctx.timer_stop() / (reps * per_graph)
    graph.close()
    bpl = pass_bytes(n, m, 4)
    achieved = bpl / (ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    log(f"row-LSE kernel: {ms * 1e3:.2f} us/launch, {achieved:.0f} GB/s dense-eq")
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "culamp_lse_rows, mixed, sq-Euclidean: pair_pass_wide (16 source splits merged in LDS)",
            "avg_launch_us": round(ms * 1e3, 3), "bytes_per_launch": bpl,
            "pairs_per_s": n * m / (ms * 1e-3)}


def kernel_only(a) -> None:
    """Launches only the roofline kernel (and, for counter calibration, a
    device-to-device copy of 64 MiB) so that a profiler run sees them alone."""
    import numpy as np
    import culamp_amd
    from culamp_amd import workloads

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    ctx = culamp_amd.Context(0)
    x, y, r, c = workloads.grid_histograms(a.grid)
    theta = np.random.default_rng(0).uniform(-1e-3, 1e-3, len(y))
    lse_roofline(ctx, x, y, theta, a, log)
    nbytes = 64 << 20
    src, dst = culamp_amd.DeviceArray(ctx, nbytes // 8), culamp_amd.DeviceArray(ctx, nbytes // 8)
    for _ in range(10):
        culamp_amd._lib.check(ctx.lib.culamp_memcpy_dtod(ctx.handle, dst.ptr, src.ptr, nbytes))
    ctx.synchronize()
    log(f"copy calibration: 10 x {nbytes} B device-to-device")


def solve_only(a) -> None:
    """Only the headline LAMP solves (mixed, 64 x 64 grids), warm-up and
    timed, so that a profiler run sees the solve's kernels alone."""
    import culamp_amd
    from culamp_amd import workloads

    ctx = culamp_amd.Context(0)
    x, y, r, c = workloads.grid_histograms(a.grid)
    pb = culamp_amd.Problem(x, y, r, c, ctx)
    mixed = culamp_amd.EOTArgs(eta_p=a.eta, epsilon=6e-6, B=1.1, itermax=200000, verbose=False,
                               stop="infeas", precision="mixed")
    for _ in range(a.warmup):
        pb.lamp(mixed, a.frequency, 2.0, log=None, assignments=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        info = pb.lamp(mixed, a.frequency, 2.0, log=None, assignments=False)[-1]
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    print(f"lamp solve: {dt * 1e3:.2f} ms, {info['iterations']} iterations, {info['passes']} passes, "
          f"infeasibility {info['residual']:.3e}", file=sys.stderr, flush=True)


def batch_kernel_only(a) -> None:
    """The batched row-LSE pass over all instances of the cell-similarity
    batch (the headline's dominant kernel), autotuned and then launched
    `kernel_launches` times between HIP events, plus a 64 MiB device-to-device
    copy for counter calibration, so that a profiler run sees them alone."""
    import culamp_amd
    from culamp_amd import workloads

    ctx = culamp_amd.Context(0)
    b = culamp_amd.Batch(workloads.cell_batch(a.batch_count), ctx, metric="cosine", scale=2.0)
    tune, _ = b.autotune(a.batch_precision, reps=3)
    ms, launches = b.time_pass(a.batch_precision, reps=a.kernel_launches)
    print(f"batched row-LSE pass ({a.batch_precision}, tuning {tune}): {ms * 1e3:.1f} us, "
          f"{launches} launch(es) per pass", file=sys.stderr, flush=True)
    nbytes = 64 << 20
    src, dst = culamp_amd.DeviceArray(ctx, nbytes // 8), culamp_amd.DeviceArray(ctx, nbytes // 8)
    for _ in range(10):
        culamp_amd._lib.check(ctx.lib.culamp_memcpy_dtod(ctx.handle, dst.ptr, src.ptr, nbytes))
    ctx.synchronize()
    print(f"copy calibration: 10 x {nbytes} B device-to-device", file=sys.stderr, flush=True)
    b.close()


def bench_grid64(ctx, a, log) -> dict:
    """The 64x64-grid histograms (n = m = 4096, half Uniform / half Gaussian
    bumps): mixed-precision LAMP to L1 infeasibility 1e-6 at eta_p = 1e-2
    (p = 2) against the plain FP64 path, the row-LSE kernel's figures, and
    eta_p = 1e-4 for p in {1, 2, inf}: FP64, mixed with its FP64 takeover,
    and the pure-mixed cost after the FP64 solve's iteration count."""
    import math
    import numpy as np
    import culamp_amd
    from culamp_amd import workloads
    x, y, r, c = workloads.grid_histograms(a.grid)
    n, m = len(x), len(y)
    pb = culamp_amd.Problem(x, y, r, c, ctx)
    mixed = culamp_amd.EOTArgs(eta_p=a.eta, epsilon=6e-6, B=1.1, itermax=200000, verbose=False,
                               stop="infeas", precision="mixed")
    pb.lamp(mixed, a.frequency, 2.0, log=None, assignments=False)
    ctx.synchronize()
    infos = []
    t0 = time.perf_counter()
    for _ in range(a.grid_steps):
        nu, _, _, _, _, info = pb.lamp(mixed, a.frequency, 2.0, log=None, assignments=False)
        infos.append(info)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    gbytes = sum(i["pairs"] * 4 + i["passes"] * pass_bytes(n, m, 0) for i in infos) / 1e9
    t1 = time.perf_counter()
    i64 = pb.lamp(mixed.copy(precision="fp64"), a.frequency, 2.0, log=None, assignments=False)[-1]
    d64 = time.perf_counter() - t1
    out = {"workload": f"LAMP, {a.grid}x{a.grid} grid histograms on [0,1]^2 (n = m = {n}), p = 2, "
                       f"eta_p = {a.eta}, to L1 infeasibility 1e-6",
           "mixed": {"solves_per_s": a.grid_steps / dt, "GBps_dense_eq": gbytes / dt,
                     "iterations": infos[-1]["iterations"],
                     "infeasibility": infos[-1]["residual"]},
           "fp64_plain": {"solves_per_s": 1.0 / d64, "iterations": i64["iterations"],
                          "infeasibility": i64["residual"],
                          "GBps_dense_eq": (i64["pairs"] * 8 + i64["passes"] * pass_bytes(n, m, 0))
                          / 1e9 / d64},
           "mixed_vs_fp64_objective_rel_err": abs(infos[-1]["objective"] - i64["objective"])
           / abs(i64["objective"])}
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64_plain"]["solves_per_s"]
    log(f"grid{a.grid} eta {a.eta}: mixed {dt / a.grid_steps * 1e3:.1f} ms/solve "
        f"({infos[-1]['iterations']} it), fp64 {d64 * 1e3:.1f} ms ({i64['iterations']} it)")
    lse = lse_roofline(ctx, x, y, nu, a, log)
    out["row_lse_kernel"] = {k: lse[k] for k in ("avg_launch_us", "achieved", "frac", "traffic",
                                                 "pairs_per_s", "kernel")}
    small = []
    base = culamp_amd.EOTArgs(eta_p=1e-4, epsilon=6e-6, B=1.1, itermax=30000, verbose=False,
                              stop="infeas")
    for p in (1.0, 2.0, math.inf):
        t1 = time.perf_counter()
        f = pb.lamp(base.copy(precision="fp64"), a.frequency, p, log=None, assignments=False)[-1]
        tf = time.perf_counter() - t1
        t1 = time.perf_counter()
        g = pb.lamp(base.copy(precision="mixed"), a.frequency, p, log=None, assignments=False)[-1]
        tg = time.perf_counter() - t1
        k = f["iterations"]
        ctx.set_fp64_fallback(False)
        h = pb.lamp(base.copy(precision="mixed", stop="reference", verbose=True, epsilon=0.0,
                              itermax=k + 1), a.frequency, p, log=None, assignments=False)[-1]
        ctx.set_fp64_fallback(True)
        rec = {"p": "inf" if math.isinf(p) else p,
               "fp64": {"iterations": k, "infeasibility": f["residual"], "objective": f["objective"],
                        "seconds": tf},
               "mixed": {"iterations": g["iterations"], "infeasibility": g["residual"],
                         "objective": g["objective"], "fp64_from": g["fp64_from"], "seconds": tg,
                         "converged": g["converged"]},
               "pure_mixed_after_fp64_iterations": {
                   "infeasibility": h["residual"],
                   "objective_rel_err_vs_fp64": abs(h["objective"] - f["objective"])
                   / abs(f["objective"])}}
        small.append(rec)
        log(f"grid{a.grid} eta 1e-4 p={rec['p']}: fp64 {k} it {tf * 1e3:.0f} ms; mixed "
            f"{g['iterations']} it (fp64 from {g['fp64_from']}) {tg * 1e3:.0f} ms, infeas "
            f"{g['residual']:.2e}; pure-mixed cost err {rec['pure_mixed_after_fp64_iterations']['objective_rel_err_vs_fp64']:.1e}")
    out["eta_1e-4"] = small
    pb.close()
    return out


# The reference's Table-1 medians of warp_logsumexp_spp_ct_fused! (ms; FP64,
# random points in [0,1]^3, theta in [-1, 1], reg = 1e-4, st = 0.5, W_inf = 1,
# p = 2), data_archive/data/kernel_timings_output.csv:3, 8, 13, 18 (GPU model
# unstated)
TABLE1_REFERENCE_MS = {1024: 0.120141, 4096: 1.312046, 16384: 17.260767, 65536: 252.7594105}


def _graph_time_us(ctx, fn, reps=20, launches=5):
    """Microseconds per call of `fn`, replayed from a captured graph."""
    fn()
    fn()
    g = ctx.capture(lambda: [fn() for _ in range(reps)])
    g.launch()
    ctx.timer_start()
    for _ in range(launches):
        g.launch()
    ms = ctx.timer_stop() / (launches * reps)
    g.close()
    return ms * 1e3


def bench_table1(ctx, a, log) -> dict:
    """The reference's Table-1 kernel benchmark (experiments/table_1_kernel_bench.jl:
    7-81) on the fused row LSE: N = M random points in [0,1]^3, theta uniform
    in [-1, 1], reg = 1e-4, st = 0.5, W_inf = 1, p = 2; our culamp_lse_rows in
    FP64 and in mixed precision against the reference's published medians."""
    import numpy as np
    import culamp_amd
    from culamp_amd import FP64, MIXED
    from culamp_amd import kernels as K
    out = []
    for n in a.table1_sizes:
        rng = np.random.default_rng(700 + n)
        x, y = rng.uniform(size=(n, 3)), rng.uniform(size=(n, 3))
        th = 2 * rng.uniform(size=n) - 1
        dev = lambda v: culamp_amd.DeviceArray.from_host(ctx, v)  # noqa: E731
        dx, dy, dth = dev(x), dev(y), dev(th)
        o64, o32 = culamp_amd.DeviceArray(ctx, n), culamp_amd.DeviceArray(ctx, n)
        rec = {"N": n}
        for name, prec, o in (("fp64", FP64, o64), ("mixed", MIXED, o32)):
            us = _graph_time_us(ctx, lambda: K.warp_logsumexp_spp_ct_fused(
                ctx, o, dx, dy, dth, 1e-4, 0.5, 1.0, 2.0, prec), reps=max(1, 20 * 1024 // n))
            rec[name] = {"ms": us / 1e3, "Gpairs_per_s": n * n / (us * 1e-6) / 1e9}
        ref = TABLE1_REFERENCE_MS.get(n)
        v64, v32 = o64.get(), o32.get()
        rec["mixed_max_abs_diff_vs_fp64"] = float(np.max(np.abs(v64 - v32)))
        rec["mixed_max_rel_diff_vs_fp64"] = float(np.max(np.abs(v64 - v32) / np.abs(v64)))
        if ref:
            rec["reference_fused_warp_ms"] = ref
            rec["fp64_speedup_vs_reference"] = ref / rec["fp64"]["ms"]
            rec["mixed_speedup_vs_reference"] = ref / rec["mixed"]["ms"]
        log(f"table-1 N={n}: fp64 {rec['fp64']['ms']:.4f} ms, mixed {rec['mixed']['ms']:.4f} ms"
            + (f" (reference {ref} ms)" if ref else ""))
        out.append(rec)
    return {"workload": "Table-1 fused row LSE (experiments/table_1_kernel_bench.jl): N = M random "
                        "points in [0,1]^3, theta in [-1, 1], reg = 1e-4, st = 0.5, W_inf = 1, p = 2",
            "reference_source": "data_archive/data/kernel_timings_output.csv:3,8,13,18 "
                                "(fused_warp medians, FP64, GPU model unstated)",
            "sizes": out}


def bench_cpu_reference(ctx, a, log) -> dict:
    """The reference's CPU path on the 32 x 32 uniform grids (n = m = 1024,
    p = 2, W normalised by its largest entry as run_dot does,
    src/CmdLineUtils.jl:195-198): the dense log-domain sinkhorn_log
    (src/solvers/Sinkhorn.jl:8-84) to infeasibility 1e-6 and the CPU LAMP
    (src/solvers/MirrorProx.jl:506-575, whose gap test never closes at
    eta_p > 0, so it is timed over a fixed number of iterations), both as the
    oracle's FP64 restatements at one thread and at the host-thread share;
    beside them the GPU's on-the-fly FP64 and mixed Sinkhorn on the same
    problem (the same fixed point)."""
    import culamp_amd
    from culamp_amd import workloads
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    x, y, r, c = workloads.uniform_grid_pair(32)
    W = oracle.grid_cost_matrix(32, 32, 2.0)
    W = W / W.max()
    nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    out = {"workload": "32x32 grid histograms (n = m = 1024), p = 2, eta_p = 1e-2, epsilon = 6e-6 "
                       "(stop at L1 infeasibility 1e-6), FP64", "nproc": nproc}
    for threads in (1, share):
        oracle.set_threads(threads)
        t0 = time.perf_counter()
        s = oracle.sinkhorn_log_dense(r, c, W, frequency=10, eta_p=1e-2, epsilon=6e-6,
                                      itermax=100000, verbose=False)
        ds = time.perf_counter() - t0
        t0 = time.perf_counter()
        lp = oracle.cpu_lamp_solve(r, c, W, frequency=10, eta_p=1e-2, epsilon=6e-6, B=1.1,
                                   itermax=a.cpu_lamp_iters, tmax=a.cpu_seconds, verbose=False)
        dl = time.perf_counter() - t0
        out[f"threads_{threads}"] = {
            "sinkhorn_log": {"seconds": ds, "iterations": s["iterations"],
                             "infeasibility": s["residual"], "converged": s["converged"]},
            "lamp": {"seconds": dl, "iterations": lp["iterations"],
                     "iterations_per_s": lp["iterations"] / dl}}
        log(f"cpu reference path ({threads} threads): sinkhorn_log {ds:.3f}s ({s['iterations']} it), "
            f"cpu LAMP {lp['iterations'] / dl:.1f} it/s")
    pb = culamp_amd.Problem(x, y, r, c, ctx)
    for prec in ("fp64", "mixed"):
        args = culamp_amd.EOTArgs(eta_p=1e-2, epsilon=6e-6, itermax=100000, verbose=False,
                                  precision=prec)
        pb.sinkhorn(args, 10, 2.0, log=None, assignments=False)
        t0 = time.perf_counter()
        info = pb.sinkhorn(args, 10, 2.0, log=None, assignments=False)[-1]
        dg = time.perf_counter() - t0
        out[f"gpu_sinkhorn_{prec}"] = {"seconds": dg, "iterations": info["iterations"],
                                       "infeasibility": info["residual"]}
        out[f"gpu_{prec}_speedup_vs_cpu_{share}_threads"] = out[f"threads_{share}"]["sinkhorn_log"][
            "seconds"] / dg
    pb.close()
    log(f"gpu on-the-fly sinkhorn: fp64 {out['gpu_sinkhorn_fp64']['seconds'] * 1e3:.2f} ms, "
        f"mixed {out['gpu_sinkhorn_mixed']['seconds'] * 1e3:.2f} ms")
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--grid-steps", type=int, default=3)
    ap.add_argument("--eta", type=float, default=1e-2)
    ap.add_argument("--frequency", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-lamp-iters", type=int, default=2000)
    ap.add_argument("--table1-sizes", type=int, nargs="*", default=[1024, 4096, 16384, 65536])
    ap.add_argument("--no-table1", action="store_true")
    ap.add_argument("--kernel-launches", type=int, default=200)
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-precision", default="bf16")
    ap.add_argument("--plain-seconds", type=float, default=30.0,
                    help="time budget of each sequential FP64 single-instance comparison")
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--dense-grids", type=int, nargs="*", default=[64, 128])
    ap.add_argument("--no-dense", action="store_true")
    ap.add_argument("--dense-only", action="store_true",
                    help="only the dense-matrix comparator (profiling)")
    ap.add_argument("--cindex-cells", type=int, default=64)
    ap.add_argument("--cindex-batch", type=int, default=2016)
    ap.add_argument("--no-cindex", action="store_true")
    ap.add_argument("--clouds-n", type=int, default=16384)
    ap.add_argument("--no-clouds", action="store_true")
    ap.add_argument("--clouds-only", action="store_true",
                    help="only the RGB-cloud Sinkhorn / LAMP comparison (profiling)")
    ap.add_argument("--diff-n", type=int, default=65536)
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--kernel-only", action="store_true",
                    help="launch only the 4096-grid row-LSE kernel and a copy of known size "
                         "(profiling)")
    ap.add_argument("--solve-only", action="store_true",
                    help="only the 4096-grid LAMP solves (profiling)")
    ap.add_argument("--batch-kernel-only", action="store_true",
                    help="only the batched all-instances row-LSE pass and a copy of known size "
                         "(profiling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: the library runs on one GPU (asked for --gpus {a.gpus}); "
              "sharding over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    if a.kernel_only:
        kernel_only(a)
        return
    if a.batch_kernel_only:
        batch_kernel_only(a)
        return
    if a.solve_only:
        solve_only(a)
        return
    if a.dense_only:
        import culamp_amd
        bench_dense(culamp_amd.Context(0), a, lambda *x: print(*x, file=sys.stderr, flush=True))
        return
    if a.clouds_only:
        import culamp_amd
        bench_clouds(culamp_amd.Context(0), a, lambda *x: print(*x, file=sys.stderr, flush=True))
        return

    import culamp_amd

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    ctx = culamp_amd.Context(0)
    b, insts, head = batched_headline(ctx, a, log)
    roofline = batch_roofline(ctx, b, insts, a, log)
    extras = batched_extras(ctx, b, insts, head, a, log)
    b.close()
    cpu = None if a.no_cpu else cpu_baseline_batched(insts, head, a, log)
    cpu_ref = None if a.no_cpu else bench_cpu_reference(ctx, a, log)
    table1 = None if a.no_table1 else bench_table1(ctx, a, log)
    grid = None if a.no_grid else bench_grid64(ctx, a, log)
    clouds = None if a.no_clouds else bench_clouds(ctx, a, log)
    divergence = None if a.no_diff else bench_divergence(ctx, a, log)
    shard = None if a.no_diff else bench_shard(ctx, a, log)
    roofline["copy_ceiling"] = copy_ceiling(ctx, log)
    roofline["mfma_ceiling"] = mfma_ceiling(ctx, roofline, log)
    cidx = None if a.no_cindex else bench_cindex(ctx, a, log)
    dense = None if a.no_dense else bench_dense(ctx, a, log)

    rec = {
        "metric": METRIC, "value": round(head["value"], 2), "unit": "GB/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"{len(insts)} independent cell-similarity OT instances, n, m "
                               "log-uniform in [256, 4096], d = 128 |N(0,1)|*Bernoulli(0.3) "
                               "L2-normalised feature rows, cosine cost / 2 as a bf16 MFMA x.y^T "
                               "tile with FP32 accumulation, Dirichlet(1) weights, eta_p = 5e-2, "
                               "one batched LAMP solve per step to L1 infeasibility 1e-6, FP64 "
                               "potentials",
                   "instances": len(insts), "d": 128, "eta_p": 5e-2, "frequency": a.frequency,
                   "tuning": head["tuning"]},
        "solves_per_s": head["solves_per_s"], "converged": head["converged"],
        "max_infeasibility": head["max_infeasibility"],
        "iterations_mean": head["iterations_mean"], "mfma_tflops_whole_solve": head["mfma_tflops"],
        "roofline": roofline, "cpu_baseline": cpu,
        "batched": extras,
    }
    for k, v in (("cpu_reference_path", cpu_ref), ("table1", table1), ("grid64", grid),
                 ("clouds", clouds), ("divergence", divergence), ("shard_one_gpu", shard),
                 ("cindex", cidx), ("dense_comparator", dense)):
        if v is not None:
            rec[k] = v
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
