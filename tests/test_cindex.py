"""Host logic of the cell-similarity C-index driver (culamp_amd.cindex,
src/CellSimilarity.jl:124-255, src/CellUtilities.jl:31-64, 175-186), checked
without a GPU: the pair distances come from a host stand-in solver."""
import numpy as np
import pytest

from culamp_amd import cindex


def _brute(values, keep):
    s = np.sort(np.asarray(values, dtype=np.float64))
    return s[:keep].sum(), s[len(s) - keep:].sum() if keep else 0.0


@pytest.mark.parametrize("keep", [0, 1, 3, 7, 40])
def test_extremes_match_sorted_sums(keep):
    rng = np.random.default_rng(keep)
    vals = np.concatenate([rng.uniform(0, 1, 25), rng.integers(0, 4, 10).astype(float)])
    rng.shuffle(vals)
    e = cindex._Extremes(keep)
    for v in vals:
        e.push(float(v))
    k = min(keep, len(vals))
    lo, hi = _brute(vals, k)
    assert abs(e.smin - lo) <= 1e-12 * max(1.0, abs(lo))
    assert abs(e.smax - hi) <= 1e-12 * max(1.0, abs(hi))


def test_extremes_integer_values_exact():
    # integers: every running sum is exact, so the incremental update must
    # land on the sorted sums exactly
    vals = [5, 1, 9, 3, 3, 7, 2, 8, 0, 6]
    e = cindex._Extremes(4)
    for v in vals:
        e.push(float(v))
    assert e.smin == 0 + 1 + 2 + 3 and e.smax == 9 + 8 + 7 + 6


def test_cluster_names_and_pairs():
    assert cindex.infer_cell_cluster("Cell_12_T_cell") == "T_cell"
    assert cindex.infer_cell_cluster("sample_B") == "B"
    assert cindex.infer_cell_cluster("plain") == "plain"
    assert cindex.within_cluster_pairs(["a", "b", "a", "a", "b", "c"]) == 3 + 1
    assert cindex._pair_order(4, 100) == [(1, 0), (2, 0), (2, 1), (3, 0), (3, 1), (3, 2)]
    # the reference checks the count after each solve, so 0 still solves one pair
    assert cindex._pair_order(4, 0) == [(1, 0)]
    assert cindex._pair_order(4, 4) == [(1, 0), (2, 0), (2, 1), (3, 0)]


def test_cell_marginals():
    data = np.array([[1.0, 0.0, 2.0], [3.0, 0.0, 2.0]])
    sums = data.sum(axis=0)
    assert np.array_equal(cindex.fill_normalized_cell(data, 0, sums), [0.25, 0.75])
    assert np.array_equal(cindex.fill_normalized_cell(data, 1, sums), [0.5, 0.5])
    m = cindex.cell_marginal(data, 0, sums)
    assert abs(m.sum() - 1.0) < 1e-15 and np.all(m > 0)
    np.testing.assert_allclose(m, (np.array([0.25, 0.75]) + 1e-7) / (1 + 2e-7), rtol=1e-15)


def _l1_solver(data, m1, m2, scale, metric, args, frequency, ctx):
    return float(np.abs(m1 - m2).sum())


def test_c_index_against_definition():
    cd = cindex.synthetic_cell_data(30, 11, 3, seed=4)
    kern = cindex.CellCostKernel.build(cd.data, "cosine")
    out = cindex.streaming_c_index(cd, kern, None, 10 ** 6, solver=_l1_solver)
    sums = cd.data.sum(axis=0)
    n = len(cd.cell_names)
    d, same = [], []
    for i in range(n):
        for j in range(i):
            d.append(np.abs(cindex.cell_marginal(cd.data, i, sums)
                            - cindex.cell_marginal(cd.data, j, sums)).sum())
            same.append(cd.clusters[i] == cd.clusters[j])
    d = np.array(d)
    nw = int(np.sum(same))
    lo, hi = _brute(d, nw)
    sw = d[np.array(same)].sum()
    assert out["Nw"] == nw and out["pairs"] == n * (n - 1) // 2
    assert abs(out["Sw"] - sw) < 1e-12 * sw
    assert abs(out["Smin"] - lo) < 1e-12 * lo and abs(out["Smax"] - hi) < 1e-12 * hi
    assert abs(out["c_index"] - (sw - lo) / (hi - lo)) < 1e-12


def test_c_index_separated_clusters_is_zero():
    # two clusters of point masses far apart: every within-cluster distance is
    # below every between-cluster one, so S_w = S_min and the index is 0
    data = np.zeros((6, 8))
    data[0:3, 0:4] = 1.0
    data[3:6, 4:8] = 1.0
    data[0, 0] += 0.1
    data[4, 5] += 0.2
    names = [f"Cell_{k}_{'a' if k < 4 else 'b'}" for k in range(8)]
    cd = cindex.OTScOmicsCellData(data, [f"f{i}" for i in range(6)], names,
                                  cindex.infer_cell_clusters(names))
    kern = cindex.CellCostKernel.build(data, "cosine")
    out = cindex.streaming_c_index(cd, kern, None, 10 ** 6, solver=_l1_solver)
    assert out["c_index"] == pytest.approx(0.0, abs=1e-15)


def test_c_index_denominator_rules(capsys):
    cd = cindex.OTScOmicsCellData(np.ones((4, 3)), ["a", "b", "c", "d"],
                                  ["Cell_1_x", "Cell_2_x", "Cell_3_y"], ["x", "x", "y"])
    kern = cindex.CellCostKernel.build(cd.data, "cosine")
    with pytest.raises(ValueError, match="Degenerate"):
        cindex.streaming_c_index(cd, kern, None, 3, solver=_l1_solver)
    out = cindex.streaming_c_index(cd, kern, None, 2, solver=_l1_solver)
    assert out["c_index"] == 0.0 and "Insufficient pairs" in capsys.readouterr().err


def test_read_otscomics_table(tmp_path):
    rng = np.random.default_rng(0)
    vals = rng.integers(0, 9, size=(7, 5)).astype(float)
    vals[2] = vals[5] = [0, 9, 0, 9, 0]  # equal spread: stable order keeps 2 before 5
    cells = [f"Cell_{k}_{'ab'[k % 2]}" for k in range(5)]
    path = tmp_path / "t.csv"
    with open(path, "w") as f:
        f.write("gene," + ",".join(cells) + "\n")
        for i in range(7):
            f.write(f"g{i}," + ",".join(str(v) for v in vals[i]) + "\n")
    cd = cindex.read_otscomics_cell_data(str(path), 3, 5, seed=1)
    pick = np.random.default_rng(1).choice(5, 5, replace=False)
    sub = vals[:, pick]
    order = np.argsort(-sub.std(axis=1, ddof=1), kind="stable")[:3]
    assert cd.feature_names == [f"g{i}" for i in order]
    assert np.array_equal(cd.data, sub[order])
    assert cd.cell_names == [cells[k] for k in pick]
    assert cd.clusters == [cells[k][-1] for k in pick]
