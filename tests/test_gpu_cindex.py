"""Cell-similarity C-index on the GPU (culamp_amd.cindex over batched
solves): every pair distance, and so the index, against the FP64 oracle's
single solves; batched distances against the single-instance GPU path; the
bf16 MFMA batch within its stated tolerance."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _expected(cd, pairs, dist):
    from culamp_amd import cindex
    ext = cindex._Extremes(cindex.within_cluster_pairs(cd.clusters))
    sw = 0.0
    for (i, j), v in zip(pairs, dist):
        if cd.clusters[i] == cd.clusters[j]:
            sw += v
        ext.push(float(v))
    return sw, ext.smin, ext.smax


@pytest.mark.parametrize("solver", ["lamp", "sinkhorn"])
def test_c_index_fp64_vs_oracle(gpu_ctx, oracle, solver):
    import culamp_amd
    from culamp_amd import cindex
    say("c-index fp64", solver)
    cd = cindex.synthetic_cell_data(48, 10, 3, seed=7)
    kern = cindex.CellCostKernel.build(cd.data, "cosine")
    kw = dict(eta_p=5e-2, epsilon=6e-6, B=1.1, itermax=5000, verbose=False, stop="infeas")
    args = culamp_amd.EOTArgs(precision="fp64", **kw)
    out = cindex.streaming_c_index(cd, kern, args, 10 ** 6, frequency=10, solver=solver,
                                   batch_size=16, ctx=gpu_ctx)
    pairs = cindex._pair_order(10, 10 ** 6)
    sums = cd.data.sum(axis=0)
    dist = []
    for i, j in pairs:
        r, c = cindex.cell_marginal(cd.data, i, sums), cindex.cell_marginal(cd.data, j, sums)
        if solver == "lamp":
            o = oracle.lamp_cellsim_solve(cd.data, r, c, "cosine", 2.0, frequency=10, **kw)
        else:
            o = oracle.sinkhorn_cellsim_solve(cd.data, cd.data, r, c, "cosine", 2.0, frequency=10,
                                              **kw)
        dist.append(o["objective"])
    sw, smin, smax = _expected(cd, pairs, dist)
    say("c_index", out["c_index"], "oracle", (sw - smin) / (smax - smin))
    assert out["Nw"] == cindex.within_cluster_pairs(cd.clusters) and out["pairs"] == 45
    for key, ref in (("Sw", sw), ("Smin", smin), ("Smax", smax)):
        assert abs(out[key] - ref) <= 1e-9 * abs(ref), (key, out[key], ref)
    assert abs(out["c_index"] - (sw - smin) / (smax - smin)) <= 1e-8
    assert 0.02 < out["c_index"] < 0.9  # clusters overlap: the index is not trivially 0
    # batched distances against the single-instance GPU path, pair by pair
    batched = cindex.pair_distances(kern, pairs[:6], args, 10, solver, 4, gpu_ctx)
    fn = (cindex.extragradient_cell_distance if solver == "lamp"
          else cindex.sinkhorn_cell_distance)
    for q, (i, j) in enumerate(pairs[:6]):
        r, c = cindex.cell_marginal(cd.data, i, sums), cindex.cell_marginal(cd.data, j, sums)
        single = fn(np.asfortranarray(cd.data), r, c, 2.0, "cosine", args, 10, gpu_ctx)
        assert abs(batched[q] - single) <= 1e-11 * abs(single), (q, batched[q], single)


def test_c_index_bf16_mfma_batch(gpu_ctx):
    """Cosine cost through the bf16 MFMA batch: distances within 1e-2
    relative of the FP64 batch (bf16 similarity error 2^-7 on unit rows) and
    a C-index within 0.05 of the one the FP64 distances give."""
    import culamp_amd
    from culamp_amd import cindex
    cd = cindex.synthetic_cell_data(96, 24, 3, seed=11)
    kern = cindex.CellCostKernel.build(cd.data, "cosine")
    kw = dict(eta_p=5e-2, epsilon=6e-6, itermax=5000, verbose=False)
    pairs = cindex._pair_order(24, 10 ** 6)
    d64 = cindex.pair_distances(kern, pairs, culamp_amd.EOTArgs(precision="fp64", **kw), 10,
                                "sinkhorn", 128, gpu_ctx)
    d16 = cindex.pair_distances(kern, pairs, culamp_amd.EOTArgs(precision="bf16", **kw), 10,
                                "sinkhorn", 128, gpu_ctx)
    err = float(np.max(np.abs(d16 - d64) / np.abs(d64)))
    a = cindex.streaming_c_index(cd, kern, culamp_amd.EOTArgs(precision="bf16", **kw), 10 ** 6,
                                 frequency=10, solver="sinkhorn", batch_size=128, ctx=gpu_ctx)
    sw, smin, smax = _expected(cd, pairs, d64)
    ref = (sw - smin) / (smax - smin)
    say("bf16 c-index distances max rel err", err, "c_index", a["c_index"], "fp64", ref)
    assert err < 1e-2  # (observed 3.4e-3 to 4.2e-3: near-identical cells give small distances)
    assert 0.02 < ref < 0.9  # clusters overlap: the index is not trivially 0
    assert abs(a["c_index"] - ref) < 0.05
