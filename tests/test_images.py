"""Colour-transfer host side (no GPU): the kernel density marginal of
load_rgb (src/ImgUtilities.jl:30-52; KernelDensity.jl's kde / pdf, which the
reference does not vendor), the image <-> colour-cloud layout and the PNG
reader / writer.

The density estimate is pinned piece by piece against what the mathematics
fixes: Silverman's bandwidth on data whose sd and quartiles are known in closed
form; linear binning by its conserved mass and first moment and by hand-placed
points; the FFT convolution against the direct circular sum with the sampled
Gaussian; the whole estimate against the direct Gaussian kernel sum; the
quadratic B-spline interpolant against node values, exact reproduction of
linear data and scipy's de Boor evaluation of the same spline."""
import math

import numpy as np
import pytest

from culamp_amd import images as im


def test_default_bandwidth_closed_forms():
    n = 101
    x = np.arange(n, dtype=np.float64)
    sd = math.sqrt(n * (n + 1) / 12.0)           # sample sd of 0..n-1
    iqr = (n - 1) / 2.0                          # linear-interpolated quartiles
    assert im.default_bandwidth(x) == pytest.approx(0.9 * min(sd, iqr / 1.34) * n ** -0.2,
                                                    rel=1e-14)
    # heavy tail: the IQR term is the smaller; quartiles at 2.25 and 6.75
    y = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 1000.0])
    assert im.default_bandwidth(y) == pytest.approx(0.9 * (4.5 / 1.34) * 10 ** -0.2, rel=1e-14)
    # zero IQR falls back to the sd (sample variance 0.1), zero sd to 1
    z = np.array([0.0] * 9 + [1.0])
    assert im.default_bandwidth(z) == pytest.approx(0.9 * math.sqrt(0.1) * 10 ** -0.2, rel=1e-14)
    assert im.default_bandwidth(np.full(32, 0.25)) == pytest.approx(0.9 * 32 ** -0.2, rel=1e-14)
    assert im.default_bandwidth([0.7]) == 0.9


def test_tabulate_conserves_mass_and_mean():
    rng = np.random.default_rng(1)
    x = rng.beta(2.0, 5.0, 5000)
    bw = im.default_bandwidth(x)
    mid = im.kde_range(im.kde_boundary(x, bw), 512)
    s = mid[1] - mid[0]
    assert mid[0] == pytest.approx(x.min() - 4 * bw) and mid[-1] == pytest.approx(x.max() + 4 * bw)
    g = im.tabulate(x, mid).density
    assert (g >= 0).all()
    assert g.sum() * s == pytest.approx(1.0, abs=1e-12)
    assert (g * mid).sum() * s == pytest.approx(x.mean(), abs=1e-12)


def test_tabulate_hand_placed_points():
    mid = np.linspace(0.0, 1.0, 11)
    s = 0.1
    g = im.tabulate([mid[2] + 0.25 * s], mid).density
    assert g[2] == pytest.approx(0.75 / s) and g[3] == pytest.approx(0.25 / s)
    assert np.count_nonzero(g) == 2
    g = im.tabulate([mid[5]], mid).density
    assert g[5] == pytest.approx(1.0 / s) and np.count_nonzero(g) == 1
    # points outside the grid are dropped, as in the reference's binning
    assert not im.tabulate([-1.0, 2.0], mid).density.any()


def test_conv_equals_direct_circular_convolution():
    rng = np.random.default_rng(2)
    K = 256
    x = rng.normal(0.0, 1.0, 3000)
    mid = np.linspace(-5.0, 5.0, K)
    s = mid[1] - mid[0]
    bw = 4.0 * s  # aliases of the Gaussian's spectrum are below 1e-30
    grid = im.tabulate(x, mid)
    got = im.conv(grid, bw).density
    lag = np.arange(K)[:, None] - np.arange(K)[None, :]
    ker = sum(np.exp(-0.5 * (((lag + p * K) * s) / bw) ** 2) for p in range(-3, 4))
    ker *= s / (bw * math.sqrt(2.0 * math.pi))
    want = ker @ grid.density
    assert np.max(np.abs(got - want)) < 1e-12 * want.max()


def test_kde_matches_direct_gaussian_sum():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0.2, 0.05, 1200), rng.normal(0.7, 0.1, 800)])
    u = im.kde(x, npoints=4096)
    bw = im.default_bandwidth(x)
    d = (u.x[:, None] - x[None, :]) / bw
    want = np.exp(-0.5 * d * d).sum(1) / (x.size * bw * math.sqrt(2.0 * math.pi))
    assert np.max(np.abs(u.density - want)) < 1e-3 * want.max()
    assert u.density.sum() * u.step == pytest.approx(1.0, abs=1e-9)


def test_quadratic_bspline_interpolant():
    from scipy.interpolate import BSpline
    rng = np.random.default_rng(4)
    n = 50
    y = rng.uniform(0.0, 2.0, n)
    k = im.UnivariateKDE(np.linspace(1.0, 3.0, n), y)
    # interpolates the grid values
    assert np.max(np.abs(im.pdf(k, k.x) - y)) < 1e-12
    # zero outside the grid
    assert not im.pdf(k, np.array([0.999, 3.001])).any()
    # off the nodes: scipy's evaluation of the same spline (degree-2 basis
    # functions centred on the nodes, knots at the half-integers)
    c = im.quadratic_bspline_coefficients(y)
    t = np.arange(n + 5) - 2.5
    spl = BSpline(t, c, 2)
    u = rng.uniform(0.0, n - 1.0, 400)
    xs = k.x[0] + u * k.step
    assert np.max(np.abs(im.pdf(k, xs) - spl(u))) < 1e-12
    # the "Line" end conditions reproduce linear data exactly, ends included
    lin = im.UnivariateKDE(k.x, 3.0 + 0.5 * np.arange(n))
    u = np.concatenate([[0.2, n - 1.2], u])
    assert np.max(np.abs(im.pdf(lin, k.x[0] + u * k.step) - (3.0 + 0.5 * u))) < 1e-12
    assert np.max(np.abs(im.pdf(lin, k.x[[0, -1]]) - [3.0, 3.0 + 0.5 * (n - 1)])) < 1e-12


def test_kde_marginal_properties():
    rng = np.random.default_rng(5)
    img = rng.uniform(size=(300, 3))
    m = im.kde_marginal(img)
    assert m.shape == (300,) and (m > 0).all() and m.sum() == pytest.approx(1.0, abs=1e-14)
    perm = rng.permutation(300)
    assert np.max(np.abs(im.kde_marginal(img[perm]) - m[perm])) < 1e-15
    # every colour and its mirror image about 1/2: the per-channel densities
    # are symmetric, so a pixel and its mirror pixel get the same weight
    a = rng.uniform(size=(64, 3))
    ms = im.kde_marginal(np.concatenate([a, 1.0 - a]))
    assert np.max(np.abs(ms[:64] - ms[64:])) < 1e-9 * ms.max()
    assert np.ptp(ms) > 0.1 * ms.max()
    # the grid size is 2^ceil(log2 N): 300 pixels -> 512 points
    assert int(2 ** math.ceil(math.log2(300.0))) == 512
    with pytest.raises(ValueError):
        im.kde_marginal(img[:2])


def test_flatten_restore_layout():
    rng = np.random.default_rng(6)
    arr = rng.uniform(size=(3, 5, 7))
    flat, dims = im.flatten_image(arr)
    assert dims == (3, 5, 7) and flat.shape == (35, 3)
    for row in range(5):
        for col in range(7):
            assert (flat[row + 5 * col] == arr[:, row, col]).all()
    assert (im.restore_image(flat, dims) == arr).all()
    assert (im.apply_colormap(flat, [3, 1]) == flat[[2, 0]]).all()


def _write_png(path, px, mode="RGB"):
    from PIL import Image
    Image.fromarray(px, mode).save(path)


def test_load_rgb_and_save_image_round_trip(tmp_path):
    rng = np.random.default_rng(7)
    px = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    f = str(tmp_path / "a.png")
    _write_png(f, px)
    for res in [(), (5, 7)]:
        img, dims, marg = im.load_rgb(f, resolution=res)
        assert dims == (3, 5, 7) and img.shape == (35, 3)
        assert (img[2 + 5 * 4] == px[2, 4] / 255.0).all()
        assert marg.sum() == pytest.approx(1.0, abs=1e-14) and (marg > 0).all()
    img, dims, marg = im.load_rgb(f, uniform_marginal=True, dtype=np.float32)
    assert img.dtype == np.float32 and (marg == np.float32(1 / 35)).all()
    out = str(tmp_path / "b.png")
    im.save_image(out, im.load_rgb(f)[0], dims)
    assert (im.read_rgb_pixels(out) == px).all()


def test_to_rgb8_rounding_and_clamping():
    v = np.array([[0.26 / 255, 1.74 / 255, 254.6 / 255], [1.2, -0.1, np.nan]])
    got = im.to_rgb8(v, (3, 2, 1))
    assert got.shape == (2, 1, 3)
    assert got[0, 0].tolist() == [0, 2, 255] and got[1, 0].tolist() == [255, 0, 0]


def test_resample_corner_aligned(tmp_path):
    px = np.zeros((4, 4, 3), dtype=np.uint8)
    px[:, :, 0] = np.arange(16, dtype=np.uint8).reshape(4, 4) * 10
    f = str(tmp_path / "c.png")
    _write_png(f, px)
    img, dims, _ = im.load_rgb(f, resolution=(2, 2))
    blocks = px[:, :, 0].reshape(2, 2, 2, 2).mean(axis=(1, 3)) / 255.0
    assert dims == (3, 2, 2)
    assert np.allclose(img[:, 0], blocks.T.ravel(), atol=1e-15)


def test_image_modes(tmp_path):
    from PIL import Image
    pal = Image.new("P", (3, 2))
    pal.putpalette([10, 20, 30] + [0] * 765)
    f = str(tmp_path / "p.png")
    pal.save(f)
    assert (im.read_rgb_pixels(f) == [10, 20, 30]).all()
    g = str(tmp_path / "g.png")
    Image.new("L", (3, 2)).save(g)
    with pytest.raises(ValueError):
        im.read_rgb_pixels(g)


def test_run_ctransfer_rejects_unknown_algorithm(tmp_path):
    s = tmp_path / "s.json"
    s.write_text('{"eta_p": 0.0, "epsilon": 1e-4}')
    with pytest.raises(ValueError):
        im.run_ctransfer({"file1": "a", "file2": "b", "output1": "c", "output2": "d",
                          "settings": str(s), "algorithm": "greenkhorn"})
