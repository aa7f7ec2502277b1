"""Input side of the reference's ``run`` command (src/CmdLineUtils.jl:175-223,
src/Utilities.jl:145-192): DOTmark histograms, weight matrices, the grid cost
and the point layout of ``--kernel``.

``read_dotmark_data`` reads a headerless CSV histogram, resizes it to at most
the requested (height, width) and flattens it in column-major order (Julia's
``reshape``) divided by its sum.  When the requested size is at least the
file's, the reference's ``imresize`` keeps the matrix unchanged and so does
this reader; a smaller size is resampled bilinearly at pixel centres, a
restatement whose agreement with ImageTransformations' ``imresize`` is not
checked (parity unpinned for downsampling).
"""
from __future__ import annotations

import numpy as np


def _read_csv_matrix(fpath: str) -> np.ndarray:
    rows = []
    with open(fpath) as f:
        for line in f:
            line = line.strip()
            if line:
                rows.append([float(v) for v in line.split(",")])
    if not rows or any(len(r) != len(rows[0]) for r in rows):
        raise ValueError(f"{fpath}: expected a rectangular headerless CSV matrix")
    return np.array(rows, dtype=np.float64)


def _resize_bilinear(a: np.ndarray, h: int, w: int) -> np.ndarray:
    H, W = a.shape
    if (h, w) == (H, W):
        return a.copy()
    def axis(n_in, n_out):
        c = (np.arange(n_out) + 0.5) * (n_in / n_out) - 0.5
        c = np.clip(c, 0.0, n_in - 1.0)
        i0 = np.floor(c).astype(int)
        i1 = np.minimum(i0 + 1, n_in - 1)
        return i0, i1, c - i0
    r0, r1, fr = axis(H, h)
    c0, c1, fc = axis(W, w)
    top = a[r0][:, c0] * (1 - fc) + a[r0][:, c1] * fc
    bot = a[r1][:, c0] * (1 - fc) + a[r1][:, c1] * fc
    return top * (1 - fr)[:, None] + bot * fr[:, None]


def read_dotmark_data(fpath: str, sizes) -> tuple:
    """read_dotmark_data (src/Utilities.jl:166-176) -> (marginal, h, w, N)."""
    data = _read_csv_matrix(fpath)
    h = min(int(sizes[0]), data.shape[0])
    w = min(int(sizes[1]), data.shape[1])
    data = _resize_bilinear(data, h, w)
    n = h * w
    marginal = data.reshape(n, order="F") / data.sum()
    return marginal, h, w, n


def read_weights(fpath: str) -> np.ndarray:
    """read_weights (src/Utilities.jl:188-192): a headerless CSV matrix."""
    return _read_csv_matrix(fpath)


def get_euclidean_distance(height: int, width: int, p: float = 2.0) -> np.ndarray:
    """get_euclidean_distance (src/Utilities.jl:145-157): pixel k of an
    N = height * width image sits at (k div height, k mod height); the cost is
    sum |d|^p for p < 10 and the Chebyshev distance otherwise."""
    n = height * width
    k = np.arange(n)
    a, b = k // height, k % height
    da = np.abs(a[:, None] - a[None, :]).astype(np.float64)
    db = np.abs(b[:, None] - b[None, :]).astype(np.float64)
    if p < 10:
        return da ** p + db ** p
    return np.maximum(da, db)


def dot_kernel_inputs(file1: str, file2: str, sizes) -> tuple:
    """run_dot's --kernel inputs (src/CmdLineUtils.jl:183-213): both
    histograms, each + 1e-6 and L1-normalised, and the pixel locations
    ((k-1) div w, (k-1) mod w, 0) as an (N, 3) array.  -> (x, y, r, c)."""
    m1, h, w, n = read_dotmark_data(file1, sizes)
    m2, h2, w2, n2 = read_dotmark_data(file2, sizes)
    if (h, w, n) != (h2, w2, n2):
        raise ValueError("the two histograms resize to different shapes")
    r = m1 + 1e-6
    r = r / np.abs(r).sum()
    c = m2 + 1e-6
    c = c / np.abs(c).sum()
    k = np.arange(n)
    loc = np.stack([k // w, k % w, np.zeros(n)], axis=1).astype(np.float64)
    return loc, loc.copy(), r, c


__all__ = ["read_dotmark_data", "read_weights", "get_euclidean_distance", "dot_kernel_inputs"]
