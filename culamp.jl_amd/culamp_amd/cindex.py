"""Cell-similarity C-index over optimal-transport distances between cells
(src/CellSimilarity.jl:1-65, 124-255; src/CellUtilities.jl:7-120, 175-186).

Each cell is a distribution over features (its expression column divided by
the column sum, + 1e-7, renormalised); the distance of two cells is the OT
objective between their distributions under the feature-to-feature cost of
``CellCostKernel`` (rows of the features x cells matrix compared over cells).
``streaming_c_index`` walks the cell pairs in the reference's order (i = 1..n,
j = 1..i-1, stopping after ``num_subproblems`` pairs) and forms

    C = (S_w - S_min) / (S_max - S_min)

from the sum S_w of within-cluster distances and the sums of the N_w smallest
and largest distances (N_w = number of within-cluster pairs).

Every pair shares the feature cost, so the pairs are solved as batches of
instances on one device copy of the feature rows (``Batch``; instances that
pass the same array share it), ``batch_size`` pairs per batch; the distances,
and therefore every sum, are consumed in the reference's pair order.  The
annealed Sinkhorn solver has no batched form and runs pair by pair.

Deviation: ``read_otscomics_cell_data`` draws its random cell subset with
numpy's generator, not Julia's Xoshiro, so a given seed selects different
cells than the reference.
"""
from __future__ import annotations

import heapq
import re
import sys
from dataclasses import dataclass

import numpy as np

from ._lib import Context
from .args import EOTArgs
from .batch import Batch
from .cellsim import FeatureProblem, cost_scale, normalize_cell_cost_metric


@dataclass
class OTScOmicsCellData:
    """OTScOmicsCellData (src/CellUtilities.jl:7-12): ``data`` is features x
    cells."""
    data: np.ndarray
    feature_names: list
    cell_names: list
    clusters: list


@dataclass
class CellCostKernel:
    """CellCostKernel (src/CellUtilities.jl:14-23, 86-112)."""
    data: np.ndarray
    metric: str
    normalize_features: bool
    row_sums: np.ndarray
    row_sqnorms: np.ndarray
    row_means: np.ndarray
    row_centered_sqnorms: np.ndarray
    scale: float

    @classmethod
    def build(cls, data, metric: str = "correlation", normalize_features: bool = True):
        data = np.asarray(data, dtype=np.float64)
        m = normalize_cell_cost_metric(metric)
        sums = data.sum(axis=1)
        if normalize_features:
            sums = np.where(sums > 0.0, sums, 1.0)
        else:
            sums = np.ones(data.shape[0])
        means = data.mean(axis=1)
        centred = ((data - means[:, None]) ** 2).sum(axis=1)
        return cls(data, m, normalize_features, sums, (data ** 2).sum(axis=1), means, centred,
                   cost_scale(m))


def infer_cell_cluster(name: str) -> str:
    """infer_cell_cluster (src/CellUtilities.jl:31-36)."""
    hit = re.match(r"^Cell_\d+_(.+)$", name)
    if hit is not None:
        return hit.group(1)
    parts = name.split("_")
    return parts[-1] if parts else name


def infer_cell_clusters(cell_names) -> list:
    return [infer_cell_cluster(n) for n in cell_names]


def read_otscomics_cell_data(fpath: str, num_features: int, numcells: int, seed: int):
    """read_otscomics_cell_data (src/CellUtilities.jl:50-64): a table whose
    first column names the features and whose other columns are cells (.csv
    or .csv.gz); ``numcells`` random cells, then the ``num_features`` features
    of largest standard deviation (stable order for ties)."""
    import pandas as pd
    table = pd.read_csv(fpath)
    if table.shape[1] < 2:
        raise ValueError(f"Expected at least one feature column and one cell column in {fpath}")
    feature_names = [str(v) for v in table.iloc[:, 0]]
    cell_names = [str(v) for v in table.columns[1:]]
    numcells = min(numcells, len(cell_names))
    pick = np.random.default_rng(seed).choice(len(cell_names), numcells, replace=False)
    data = table.iloc[:, 1:].to_numpy(dtype=np.float64)[:, pick]
    num_features = min(num_features, data.shape[0])
    order = np.argsort(-data.std(axis=1, ddof=1), kind="stable")[:num_features]
    names = [cell_names[i] for i in pick]
    return OTScOmicsCellData(data[order, :], [feature_names[i] for i in order], names,
                             infer_cell_clusters(names))


def fill_normalized_cell(data, column: int, column_sums) -> np.ndarray:
    """fill_normalized_cell! (src/CellUtilities.jl:175-186), column 0-based."""
    scale = column_sums[column]
    if scale <= 0.0:
        return np.full(data.shape[0], 1.0 / data.shape[0])
    return data[:, column] / scale


def cell_marginal(data, column: int, column_sums) -> np.ndarray:
    """A cell's distribution over features as streaming_c_index forms it
    (src/CellSimilarity.jl:196-198, 201-203): + 1e-7, then L1-normalised."""
    v = fill_normalized_cell(data, column, column_sums) + 1e-7
    return v / np.abs(v).sum()


def within_cluster_pairs(clusters) -> int:
    """within_cluster_pairs (src/CellSimilarity.jl:156-162)."""
    counts = {}
    for c in clusters:
        counts[c] = counts.get(c, 0) + 1
    return sum(k * (k - 1) // 2 for k in counts.values())


class _Extremes:
    """The running sums of the ``keep`` smallest and largest values seen
    (push_smallest! / push_largest!, src/CellSimilarity.jl:124-154): a value
    that displaces the current extreme updates the sum as
    sum + value - displaced, in arrival order, as the reference does."""

    def __init__(self, keep: int):
        self.keep = keep
        self.small = []  # max-heap (negated) of the smallest values
        self.large = []  # min-heap of the largest values
        self.smin = 0.0
        self.smax = 0.0

    def push(self, v: float) -> None:
        if self.keep == 0:
            return
        if len(self.small) < self.keep:
            heapq.heappush(self.small, -v)
            self.smin = self.smin + v
        elif v < -self.small[0]:
            self.smin = self.smin + v - (-self.small[0])
            heapq.heapreplace(self.small, -v)
        if len(self.large) < self.keep:
            heapq.heappush(self.large, v)
            self.smax = self.smax + v
        elif v > self.large[0]:
            self.smax = self.smax + v - self.large[0]
            heapq.heapreplace(self.large, v)


def _pair_order(ncells: int, num_subproblems: int):
    """(i, j), j < i, in the reference's loop order (0-based), at most
    num_subproblems of them."""
    out = []
    for i in range(ncells):
        for j in range(i):
            out.append((i, j))
            if len(out) >= num_subproblems:  # checked after each solve, as there
                return out
    return out


def extragradient_cell_distance(data, marginal1, marginal2, scale, metric, args: EOTArgs,
                                frequency: int, ctx: Context | None = None) -> float:
    """extragradient_cell_distance (src/CellSimilarity.jl:1-22): the LAMP
    objective of one pair."""
    pb = FeatureProblem(data, None, scale, metric, marginal1, marginal2, ctx)
    try:
        return float(pb.lamp(args, frequency, None)[-1]["objective"])
    finally:
        pb.close()


def sinkhorn_cell_distance(data, marginal1, marginal2, scale, metric, args: EOTArgs,
                           frequency: int, ctx: Context | None = None) -> float:
    """sinkhorn_cell_distance (src/CellSimilarity.jl:24-44)."""
    pb = FeatureProblem(data, data, scale, metric, marginal1, marginal2, ctx)
    try:
        return float(pb.sinkhorn(args, frequency, log=None)[-1]["objective"])
    finally:
        pb.close()


def annealed_sinkhorn_cell_distance(data, marginal1, marginal2, scale, metric, args: EOTArgs,
                                    frequency: int, ctx: Context | None = None) -> float:
    """annealed_sinkhorn_cell_distance (src/CellSimilarity.jl:46-65)."""
    pb = FeatureProblem(data, data, scale, metric, marginal1, marginal2, ctx)
    try:
        return float(pb.annealed_sinkhorn(args, frequency, None)[-1]["objective"])
    finally:
        pb.close()


_BATCHED = {extragradient_cell_distance: "lamp", sinkhorn_cell_distance: "sinkhorn",
            "lamp": "lamp", "sinkhorn": "sinkhorn"}


def pair_distances(kernel: CellCostKernel, pairs, args: EOTArgs, frequency: int = 100,
                   solver=extragradient_cell_distance, batch_size: int = 512,
                   ctx: Context | None = None, log=None) -> np.ndarray:
    """OT distances of the given (i, j) cell pairs, in order."""
    data = np.asfortranarray(kernel.data, dtype=np.float64)
    column_sums = kernel.data.sum(axis=0)
    marg = {}

    def cell(k):
        if k not in marg:
            marg[k] = cell_marginal(kernel.data, k, column_sums)
        return marg[k]

    out = np.empty(len(pairs))
    mode = _BATCHED.get(solver)
    if mode is None:
        fn = {"annealed_sinkhorn": annealed_sinkhorn_cell_distance}.get(solver, solver)
        for q, (i, j) in enumerate(pairs):
            out[q] = fn(data, cell(i), cell(j), kernel.scale, kernel.metric, args, frequency, ctx)
        return out
    for q0 in range(0, len(pairs), batch_size):
        chunk = pairs[q0:q0 + batch_size]
        insts = [(data, None, cell(i), cell(j)) for i, j in chunk]
        b = Batch(insts, ctx, metric=kernel.metric, scale=kernel.scale)
        try:
            res = b.lamp(args, frequency) if mode == "lamp" else b.sinkhorn(args, frequency)
        finally:
            b.close()
        for q, r in enumerate(res):
            out[q0 + q] = r[-1]["objective"]
        if log is not None:
            log(f"processed_pairs={q0 + len(chunk)}/{len(pairs)}")
    return out


def streaming_c_index(cell_data: OTScOmicsCellData, kernel: CellCostKernel, args: EOTArgs,
                      num_subproblems: int, *, frequency: int = 100,
                      solver=extragradient_cell_distance, batch_size: int = 512,
                      ctx: Context | None = None, log=None) -> dict:
    """streaming_c_index (src/CellSimilarity.jl:164-255) -> dict(c_index, Sw,
    Smin, Smax, Nw, pairs)."""
    ncells = len(cell_data.cell_names)
    npairs = ncells * (ncells - 1) // 2
    nw = within_cluster_pairs(cell_data.clusters)
    pairs = _pair_order(ncells, num_subproblems)
    dist = pair_distances(kernel, pairs, args, frequency, solver, batch_size, ctx, log)
    ext = _Extremes(nw)
    sw = 0.0
    for (i, j), v in zip(pairs, dist):
        if cell_data.clusters[i] == cell_data.clusters[j]:
            sw += v
        ext.push(float(v))
    denominator = ext.smax - ext.smin
    if num_subproblems >= npairs:
        if not denominator > 0.0:
            raise ValueError("Degenerate C-index denominator; pairwise distances do not separate "
                             "within/between cluster pairs")
    elif denominator < 1e-10:
        denominator = 1.0
        print("WARNING: Insufficient pairs considered, disregard similarity output. "
              "Only use data for solver comparison", file=sys.stderr, flush=True)
    return {"c_index": (sw - ext.smin) / denominator, "Sw": sw, "Smin": ext.smin,
            "Smax": ext.smax, "Nw": nw, "pairs": npairs}


def format_summary(result: dict) -> str:
    """run_cellsim's summary line (src/CmdLineUtils.jl:254), Julia number
    formatting: "c_index=..,Sw=..,Smin=..,Smax=..,Nw=..,pairs=.."."""
    from .solvers import julia_float_str
    f = julia_float_str
    return (f"c_index={f(result['c_index'])},Sw={f(result['Sw'])},Smin={f(result['Smin'])},"
            f"Smax={f(result['Smax'])},Nw={int(result['Nw'])},pairs={int(result['pairs'])}")


def run_cellsim(file: str, num_features: int, num_subproblems: int, num_cells: int,
                args: EOTArgs, *, cost: str = "correlation", normalize_features: bool = True,
                frequency: int = 100, algorithm: str = "lamp", seed: int = 0, output: str = "",
                batch_size: int = 512, ctx: Context | None = None) -> dict:
    """run_cellsim (src/CmdLineUtils.jl:239-261): the C-index of an
    OT-scOmics table, printed (and written to ``output``) as the summary
    line."""
    solver = {"lamp": extragradient_cell_distance, "sinkhorn": sinkhorn_cell_distance,
              "annealed_sinkhorn": annealed_sinkhorn_cell_distance}[algorithm]
    result = compute_otscomics_c_index(file, num_features, num_subproblems, num_cells, args,
                                       metric=normalize_cell_cost_metric(cost),
                                       normalize_features=normalize_features,
                                       frequency=frequency, solver=solver, seed=seed,
                                       batch_size=batch_size, ctx=ctx)
    line = format_summary(result)
    print(line, flush=True)
    if output:
        with open(output, "w") as f:
            f.write(line + "\n")
    return result


def compute_otscomics_c_index(fpath: str, num_features: int, num_subproblems: int,
                              num_cells: int, args: EOTArgs, *, metric: str = "correlation",
                              normalize_features: bool = True, frequency: int = 100,
                              solver=extragradient_cell_distance, seed: int = 0,
                              batch_size: int = 512, ctx: Context | None = None) -> dict:
    """compute_otscomics_c_index (src/CellSimilarity.jl:257-274)."""
    cell_data = read_otscomics_cell_data(fpath, num_features, num_cells, seed)
    kernel = CellCostKernel.build(cell_data.data, metric, normalize_features)
    return streaming_c_index(cell_data, kernel, args, num_subproblems, frequency=frequency,
                             solver=solver, batch_size=batch_size, ctx=ctx)


def synthetic_cell_data(num_features: int, num_cells: int, num_clusters: int,
                        seed: int = 0, separation: float = 0.3) -> OTScOmicsCellData:
    """Clustered non-negative expression data (features x cells) with cell
    names ``Cell_<k>_<cluster>``: cluster profiles mix a shared profile and a
    cluster-specific one (weight ``separation``), cells are Poisson draws
    around their cluster's profile with a random library size (no network
    for the reference's OT-scOmics tables)."""
    rng = np.random.default_rng(seed)
    base = rng.gamma(0.6, 2.0, size=num_features)
    own = rng.gamma(0.6, 2.0, size=(num_features, num_clusters))
    centres = (1.0 - separation) * base[:, None] + separation * own
    labels = rng.integers(0, num_clusters, size=num_cells)
    lib = rng.uniform(0.5, 1.5, size=num_cells)
    data = rng.poisson(centres[:, labels] * lib).astype(np.float64)
    names = [f"Cell_{k + 1}_c{labels[k]}" for k in range(num_cells)]
    return OTScOmicsCellData(data, [f"f{i}" for i in range(num_features)], names,
                             infer_cell_clusters(names))


__all__ = ["OTScOmicsCellData", "CellCostKernel", "infer_cell_cluster", "infer_cell_clusters",
           "read_otscomics_cell_data", "fill_normalized_cell", "cell_marginal",
           "within_cluster_pairs", "extragradient_cell_distance", "sinkhorn_cell_distance",
           "annealed_sinkhorn_cell_distance", "pair_distances", "streaming_c_index",
           "compute_otscomics_c_index", "synthetic_cell_data", "format_summary", "run_cellsim"]
