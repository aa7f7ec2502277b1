"""Colour-transfer inputs and outputs (src/ImgUtilities.jl) and the
``ctransfer`` command (src/CmdLineUtils.jl:225-238).

An RGB image of height H and width W becomes a cloud of N = H * W colour points
in [0, 1]^3, pixel k = row + H * col (Julia's column-major ``reshape`` of the
3 x H x W ``channelview``); here the cloud is an (N, 3) C-contiguous array, the
same bytes as the reference's 3 x N matrix.  ``load_rgb`` weights each pixel by
the sum over channels of a kernel density estimate of that channel
(ImgUtilities.jl:34-44); ``save_image`` writes an (N, 3) cloud back as an 8-bit
RGB PNG.

The density estimate restates KernelDensity.jl's univariate ``kde`` and ``pdf``
(the reference pins no version: it has no Manifest.toml and no [compat] entry
for KernelDensity): Silverman's bandwidth 0.9 min(sd, IQR/1.34) n^-1/5
(``default_bandwidth``), a grid of ``npoints`` midpoints over the data range
widened by 4 bandwidths (``kde_boundary``, ``kde_range``), linear binning
(``tabulate``, Jones and Lotwick), circular convolution with the Gaussian
through the real FFT and its characteristic function (``conv``), negative
values clipped to zero; ``pdf`` evaluates the quadratic B-spline interpolant of
the grid density with linear ("Line") end conditions (Interpolations.jl's
``BSpline(Quadratic(Line(OnGrid())))``), zero outside the grid.

Image files go through Pillow.  Resizing to the requested resolution is the
identity when the sizes agree (then the pixels are exactly k/255); otherwise it
is a corner-aligned bilinear resample (ImageTransformations' ``imresize``
mapping), whose agreement with the reference's ``imresize`` is not checked
(parity unpinned for resampling).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .args import EOTArgs, read_args_json
from .dotmark import _resize_bilinear
from .solvers import _stdout_log, extragradient_color_transfer, sinkhorn_color_transfer


# --------------------------------------------------------------------------
# Kernel density estimate (KernelDensity.jl's univariate kde / pdf)
# --------------------------------------------------------------------------

@dataclasses.dataclass
class UnivariateKDE:
    """Grid midpoints ``x`` (uniform, ascending) and the density on them."""
    x: np.ndarray
    density: np.ndarray

    @property
    def step(self) -> float:
        return float(self.x[1] - self.x[0])


def default_bandwidth(data, alpha: float = 0.9) -> float:
    """Silverman's rule: alpha * min(sd, IQR / 1.34) * n^(-1/5); a zero IQR
    falls back to sd, zero sd to 1; one point gives alpha."""
    data = np.asarray(data, dtype=np.float64)
    n = data.size
    if n <= 1:
        return alpha
    var_width = float(np.std(data, ddof=1))
    q25, q75 = np.quantile(data, [0.25, 0.75])
    quantile_width = float(q75 - q25) / 1.34
    width = min(var_width, quantile_width)
    if width == 0.0:
        width = 1.0 if var_width == 0.0 else var_width
    return alpha * width * n ** -0.2


def kde_boundary(data, bandwidth: float) -> tuple:
    lo, hi = float(np.min(data)), float(np.max(data))
    return lo - 4.0 * bandwidth, hi + 4.0 * bandwidth


def kde_range(boundary, npoints: int) -> np.ndarray:
    lo, hi = boundary
    if not lo < hi:
        raise ValueError("boundary (a, b) must have a < b")
    return np.linspace(lo, hi, int(npoints))


def tabulate(data, midpoints: np.ndarray) -> UnivariateKDE:
    """Linear binning: each point splits 1 / (n s^2) between its two
    neighbouring midpoints in proportion to proximity."""
    data = np.asarray(data, dtype=np.float64)
    npoints = midpoints.size
    s = float(midpoints[1] - midpoints[0])
    ainc = 1.0 / (data.size * s * s)
    k = np.searchsorted(midpoints, data, side="left")
    j = k - 1
    ok = (j >= 0) & (j <= npoints - 2)
    k, j, xs = k[ok], j[ok], data[ok]
    grid = np.bincount(j, weights=(midpoints[k] - xs) * ainc, minlength=npoints)
    grid += np.bincount(k, weights=(xs - midpoints[j]) * ainc, minlength=npoints)
    return UnivariateKDE(midpoints, grid)


def conv(k: UnivariateKDE, bandwidth: float) -> UnivariateKDE:
    """Circular convolution of the binned grid with N(0, bandwidth^2): the real
    FFT of the grid times the Gaussian characteristic function
    exp(-bw^2 t^2 / 2) at t_j = 2 pi j / (s K), inverse FFT, clipped at 0."""
    K = k.density.size
    ft = np.fft.rfft(k.density)
    t = np.arange(ft.size) * (-2.0 * math.pi / (k.step * K))
    ft *= np.exp(-0.5 * (bandwidth * t) ** 2)
    dens = np.fft.irfft(ft, K)
    return UnivariateKDE(k.x, np.maximum(0.0, dens))


def kde(data, npoints: int = 2048, bandwidth: float | None = None) -> UnivariateKDE:
    """kde(data; npoints) with the Normal kernel and default boundary."""
    data = np.asarray(data, dtype=np.float64).ravel()
    if data.size == 0:
        raise ValueError("kde needs at least one data point")
    if npoints < 2:
        raise ValueError("kde needs at least two grid points")
    bw = default_bandwidth(data) if bandwidth is None else float(bandwidth)
    mid = kde_range(kde_boundary(data, bw), npoints)
    return conv(tabulate(data, mid), bw)


def quadratic_bspline_coefficients(y: np.ndarray) -> np.ndarray:
    """Coefficients c_0..c_{n+1} of the quadratic B-spline through y_0..y_{n-1}
    at the integer nodes: (c_{k} + 6 c_{k+1} + c_{k+2}) / 8 = y_k, with the
    "Line" end conditions c_0 - 2 c_1 + c_2 = 0 = c_{n-1} - 2 c_n + c_{n+1}
    (zero second difference of the coefficients at both ends)."""
    from scipy.linalg import solve_banded
    y = np.asarray(y, dtype=np.float64)
    n = y.size
    if n < 3:
        raise ValueError("a quadratic interpolant needs at least 3 grid values")
    # Substituting the end conditions into the first and last node equations
    # gives c_1 = y_0 and c_n = y_{n-1}; the rest is tridiagonal.
    c = np.empty(n + 2)
    c[1], c[n] = y[0], y[n - 1]
    m = n - 2
    if m > 0:
        rhs = y[1:n - 1].copy()
        rhs[0] -= c[1] / 8.0
        rhs[-1] -= c[n] / 8.0
        ab = np.empty((3, m))
        ab[0, :] = 1.0 / 8.0
        ab[1, :] = 3.0 / 4.0
        ab[2, :] = 1.0 / 8.0
        c[2:n] = solve_banded((1, 1), ab, rhs)
    c[0] = 2.0 * c[1] - c[2]
    c[n + 1] = 2.0 * c[n] - c[n - 1]
    return c


def pdf(k: UnivariateKDE, xs) -> np.ndarray:
    """pdf(::UnivariateKDE, xs): the quadratic B-spline interpolant of the grid
    density at xs, zero outside [x_0, x_{K-1}]."""
    xs = np.asarray(xs, dtype=np.float64)
    c = quadratic_bspline_coefficients(k.density)
    K = k.density.size
    t = (xs - k.x[0]) / k.step
    inside = (xs >= k.x[0]) & (xs <= k.x[-1])
    t = np.clip(t, 0.0, K - 1.0)
    ix = np.rint(t).astype(np.int64)
    d = t - ix
    v = (0.5 * (0.5 - d) ** 2 * c[ix] + (0.75 - d * d) * c[ix + 1]
         + 0.5 * (0.5 + d) ** 2 * c[ix + 2])
    return np.where(inside, v, 0.0)


def kde_marginal(img: np.ndarray) -> np.ndarray:
    """ImgUtilities.jl:35-44: per channel a kde on 2^ceil(log2 N) points,
    its density scaled to unit Euclidean norm (``normalize!``), the three pdfs
    at the pixels summed and scaled to unit sum."""
    img = np.asarray(img, dtype=np.float64)
    n = img.shape[0]
    if n < 3:
        raise ValueError("the kde marginal needs at least three pixels")
    npts = int(2 ** math.ceil(math.log2(float(n))))
    total = np.zeros(n)
    for ch in range(3):
        u = kde(img[:, ch], npoints=npts)
        u.density = u.density / np.linalg.norm(u.density)
        total += pdf(u, img[:, ch])
    return total / np.abs(total).sum()


# --------------------------------------------------------------------------
# Image layout and files
# --------------------------------------------------------------------------

def flatten_image(img: np.ndarray) -> tuple:
    """flatten_image (ImgUtilities.jl:6-19): a (channels, H, W) array, or
    (H, W) for one channel, to an (H * W, channels) cloud with pixel
    k = row + H * col, and the original shape."""
    img = np.asarray(img)
    if img.ndim == 3:
        ch, h, w = img.shape
    elif img.ndim == 2:
        (h, w), ch = img.shape, 1
        img = img[None]
    else:
        raise ValueError(f"Unsupported image shape {img.shape}")
    flat = np.ascontiguousarray(img.transpose(2, 1, 0).reshape(h * w, ch))
    return flat, (ch, h, w)


def restore_image(flat: np.ndarray, original_shape: tuple) -> np.ndarray:
    """restore_image (ImgUtilities.jl:25-28): (H * W, channels) back to
    (channels, H, W)."""
    ch, h, w = original_shape
    flat = np.asarray(flat)
    if flat.shape != (h * w, ch):
        raise ValueError(f"cloud of shape {flat.shape} does not fit image shape {original_shape}")
    return flat.reshape(w, h, ch).transpose(2, 1, 0)


def apply_colormap(img: np.ndarray, colormap) -> np.ndarray:
    """apply_colormap (ImgUtilities.jl:53-55): img[:, colormap] on the
    3 x N matrix, i.e. rows of the (N, 3) cloud; ``colormap`` is 1-based."""
    idx = np.asarray(colormap, dtype=np.int64) - 1
    return np.asarray(img)[idx]


def _pil():
    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - Pillow ships with this image
        raise ImportError("reading and writing PNG images needs Pillow (PIL)") from e
    return Image


def read_rgb_pixels(filepath: str) -> np.ndarray:
    """The file's pixels as (H, W, 3) uint8; palette images are expanded,
    other non-RGB modes are refused (the reference's colour clouds are RGB)."""
    Image = _pil()
    with Image.open(filepath) as im:
        if im.mode == "P":
            im = im.convert("RGB")
        if im.mode != "RGB":
            raise ValueError(f"{filepath}: expected an RGB image, got mode {im.mode}")
        return np.asarray(im, dtype=np.uint8).copy()


def load_rgb(filepath: str, resolution: tuple = (), uniform_marginal: bool = False,
             dtype=np.float64) -> tuple:
    """load_rgb (ImgUtilities.jl:30-52) -> (img (N, 3), original_dims, marginal).

    ``resolution`` is (height, width), () keeps the file's size;
    ``original_dims`` is (3, height, width)."""
    px = read_rgb_pixels(filepath).astype(np.float64) / 255.0
    if len(resolution):
        h, w = int(resolution[0]), int(resolution[1])
        if h < 1 or w < 1:
            raise ValueError("resolution must be positive")
        px = np.stack([_resize_bilinear(px[:, :, ch], h, w) for ch in range(3)], axis=2)
    chw = px.transpose(2, 0, 1)
    img, dims = flatten_image(chw)
    img = img.astype(dtype)
    if uniform_marginal:
        marginal = np.full(img.shape[0], 1.0 / img.shape[0], dtype=dtype)
    else:
        marginal = kde_marginal(img).astype(dtype)
    return img, dims, marginal


def to_rgb8(flat: np.ndarray, original_shape: tuple) -> np.ndarray:
    """An (N, 3) cloud as (H, W, 3) uint8: clamped to [0, 1] (NaN to 0) and
    rounded to the nearest of 0..255, ties to even."""
    chw = restore_image(np.asarray(flat, dtype=np.float64), original_shape)
    v = np.nan_to_num(chw, nan=0.0, posinf=1.0, neginf=0.0)
    v = np.rint(np.clip(v, 0.0, 1.0) * 255.0).astype(np.uint8)
    return np.ascontiguousarray(v.transpose(1, 2, 0))


def save_image(path: str, flattened_image: np.ndarray, original_shape: tuple) -> None:
    """save_image (ImgUtilities.jl:56-59): write the cloud as an RGB PNG."""
    if original_shape[0] != 3:
        raise ValueError("save_image writes RGB (3-channel) images")
    _pil().fromarray(to_rgb8(flattened_image, original_shape), "RGB").save(path)


# --------------------------------------------------------------------------
# Colour transfer between two image files
# --------------------------------------------------------------------------

def _compute_args(args: EOTArgs, dtype) -> EOTArgs:
    # Float32 inputs (the reference's --fp32) run the mixed-precision pairs:
    # FP32 cost and exponent arithmetic, FP64 potentials and sums.
    if np.dtype(dtype) == np.float32 and args.precision == "fp64":
        return args.copy(precision="mixed")
    return args


def extragradient_color_transfer_images(f1: str, f2: str, out_f1: str, out_f2: str,
                                        resolution: tuple, args: EOTArgs, frequency: int,
                                        p: float, *, dtype=np.float64, log=_stdout_log,
                                        ctx=None) -> tuple:
    """extragradient_color_transfer(f1::String, ...) (MirrorProxKernel.jl:1134-1140):
    LAMP between the two images' colour clouds; each image is written
    recoloured by its barycentric map.  -> (ν, φ, ψ, img1_new, img2_new)."""
    img1, dims1, m1 = load_rgb(f1, resolution=resolution, dtype=dtype)
    img2, dims2, m2 = load_rgb(f2, resolution=resolution, dtype=dtype)
    out = extragradient_color_transfer(img1, img2, m1, m2, _compute_args(args, dtype), frequency,
                                       p, log=log, ctx=ctx)
    save_image(out_f1, out[3], dims1)
    save_image(out_f2, out[4], dims2)
    return out


def sinkhorn_color_transfer_images(f1: str, f2: str, out_f1: str, out_f2: str,
                                   resolution: tuple, args: EOTArgs, frequency: int, p: float,
                                   *, dtype=np.float64, log=_stdout_log, ctx=None) -> tuple:
    """sinkhorn_color_transfer(f1::String, ...) (SinkhornKernel.jl:569-575)
    -> (φ, ψ, img1_new, img2_new)."""
    img1, dims1, m1 = load_rgb(f1, resolution=resolution, dtype=dtype)
    img2, dims2, m2 = load_rgb(f2, resolution=resolution, dtype=dtype)
    out = sinkhorn_color_transfer(img1, img2, m1, m2, _compute_args(args, dtype), frequency, p,
                                  log=log, ctx=ctx)
    save_image(out_f1, out[2], dims1)
    save_image(out_f2, out[3], dims2)
    return out


def run_ctransfer(parsed_args: dict, *, log=_stdout_log, ctx=None):
    """run_ctransfer (CmdLineUtils.jl:225-238), keys as the ``ctransfer``
    command's: file1, file2, output1, output2, algorithm ("lamp" or
    "sinkhorn", default "sinkhorn"), settings, frequency (100), p (2.0),
    height and width (128), fp32 (False)."""
    g = dict(algorithm="sinkhorn", settings="./test.json", frequency=100, p=2.0, height=128,
             width=128, fp32=False)
    g.update(parsed_args)
    dtype = np.float32 if g["fp32"] else np.float64
    args = read_args_json(g["settings"])
    size = (int(g["height"]), int(g["width"]))
    if g["algorithm"] == "lamp":
        fn = extragradient_color_transfer_images
    elif g["algorithm"] == "sinkhorn":
        fn = sinkhorn_color_transfer_images
    else:
        raise ValueError(f"algorithm must be 'lamp' or 'sinkhorn', got {g['algorithm']!r}")
    return fn(g["file1"], g["file2"], g["output1"], g["output2"], size, args, int(g["frequency"]),
              float(g["p"]), dtype=dtype, log=log, ctx=ctx)


__all__ = [
    "UnivariateKDE", "default_bandwidth", "kde_boundary", "kde_range", "tabulate", "conv", "kde",
    "quadratic_bspline_coefficients", "pdf", "kde_marginal", "flatten_image", "restore_image",
    "apply_colormap", "read_rgb_pixels", "load_rgb", "to_rgb8", "save_image",
    "extragradient_color_transfer_images", "sinkhorn_color_transfer_images", "run_ctransfer",
]
