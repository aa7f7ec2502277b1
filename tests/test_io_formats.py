"""Host-side input and output formats (no GPU): the reference's text writers
(Julia number and matrix printing), the cellsim summary line, the DOTmark /
weights readers, the grid cost of the dense path and the --kernel inputs
(src/CmdLineUtils.jl:175-261, src/Utilities.jl:145-192,
src/solvers/MirrorProxKernel.jl:1143-1185)."""
import numpy as np
import pytest

from culamp_amd import cindex, dotmark
from culamp_amd.solvers import _write_assignments, julia_float_str, julia_matrix_str


def test_julia_matrix_strings():
    assert julia_matrix_str([[1.0, 2.0], [3.0, 4.0]]) == "[1.0 2.0; 3.0 4.0]"
    assert julia_matrix_str([[1.0], [2.5], [1e-5]]) == "[1.0; 2.5; 1.0e-5;;]"
    assert julia_matrix_str([[0.1, -0.0, 1e20]]) == "[0.1 -0.0 1.0e20]"


def test_assignment_file_layout(tmp_path):
    a = np.array([[0.5, 1.0, 0.0], [0.25, 2.0, 0.0]])  # two points, (N, 3)
    p = tmp_path / "out1"
    _write_assignments(str(p), a)
    lines = p.read_text().splitlines()
    assert len(lines) == 6
    assert lines[0] == "[0.5 0.25; 1.0 2.0; 0.0 0.0],1"
    assert lines[5] == "[0.5 0.25; 1.0 2.0; 0.0 0.0],6"


def test_cellsim_summary_line():
    r = {"c_index": 0.125, "Sw": 3.0, "Smin": 1e-5, "Smax": 2.5, "Nw": 7, "pairs": 45}
    assert cindex.format_summary(r) == "c_index=0.125,Sw=3.0,Smin=1.0e-5,Smax=2.5,Nw=7,pairs=45"
    assert julia_float_str(123456789.0) == "1.23456789e8"


def _write_csv(path, a):
    with open(path, "w") as f:
        for row in a:
            f.write(",".join(repr(float(v)) for v in row) + "\n")


def test_dotmark_reader_column_major(tmp_path):
    a = np.arange(1.0, 13.0).reshape(3, 4)
    p = tmp_path / "h.csv"
    _write_csv(p, a)
    m, h, w, n = dotmark.read_dotmark_data(str(p), (32, 32))
    assert (h, w, n) == (3, 4, 12)
    np.testing.assert_array_equal(m, a.reshape(12, order="F") / a.sum())
    assert abs(m.sum() - 1.0) < 1e-15
    # a smaller requested size resamples (bilinear at pixel centres)
    m2, h2, w2, n2 = dotmark.read_dotmark_data(str(p), (3, 2))
    assert (h2, w2, n2) == (3, 2, 6) and abs(m2.sum() - 1.0) < 1e-15
    # constant images stay constant under resampling
    _write_csv(p, np.full((6, 6), 2.0))
    m3, *_ = dotmark.read_dotmark_data(str(p), (3, 3))
    np.testing.assert_allclose(m3, np.full(9, 1 / 9), rtol=1e-15)
    np.testing.assert_array_equal(dotmark.read_weights(str(p)), np.full((6, 6), 2.0))


@pytest.mark.parametrize("h,w,p", [(4, 4, 2.0), (3, 5, 1.0), (5, 3, 3.0), (4, 4, 10.0), (3, 4, float("inf"))])
def test_grid_cost_matches_oracle(oracle, h, w, p):
    W = dotmark.get_euclidean_distance(h, w, p)
    np.testing.assert_array_equal(W, oracle.grid_cost_matrix(h, w, p))
    # the 64 x 64 closed forms: W_inf = 2 (g-1)^2 for p = 2, (g-1) for p >= 10
    assert dotmark.get_euclidean_distance(8, 8, 2.0).max() == 2 * 7 ** 2
    assert dotmark.get_euclidean_distance(8, 8, 12.0).max() == 7


def test_kernel_inputs(tmp_path):
    rng = np.random.default_rng(0)
    p1, p2 = tmp_path / "a.csv", tmp_path / "b.csv"
    _write_csv(p1, rng.uniform(size=(4, 6)))
    _write_csv(p2, rng.uniform(size=(4, 6)))
    x, y, r, c = dotmark.dot_kernel_inputs(str(p1), str(p2), (4, 6))
    assert x.shape == (24, 3) and np.array_equal(x, y)
    assert list(x[7]) == [1.0, 1.0, 0.0]  # k = 7: (7 div 6, 7 mod 6, 0)
    assert abs(r.sum() - 1) < 1e-15 and abs(c.sum() - 1) < 1e-15 and r.min() > 0
