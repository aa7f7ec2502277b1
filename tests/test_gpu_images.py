"""Colour transfer between two PNG files (the reference's ``ctransfer``
command: MirrorProxKernel.jl:1134-1140, SinkhornKernel.jl:569-575,
CmdLineUtils.jl:225-238): images read into colour clouds with kernel-density
marginals, solved on the GPU, each image written recoloured by its barycentric
map.  The maps are compared with the oracle's solve of the same clouds and the
written PNGs with the oracle maps rounded to 8 bits."""
import json
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pngs(tmp_path, h=12, w=10):
    """Two smooth synthetic RGB images (a warm and a cool gradient)."""
    from PIL import Image
    rng = np.random.default_rng(11)
    yy, xx = np.meshgrid(np.linspace(0, 1, h), np.linspace(0, 1, w), indexing="ij")
    a = np.stack([0.6 + 0.4 * xx, 0.3 + 0.3 * yy, 0.1 + 0.2 * xx * yy], axis=2)
    b = np.stack([0.1 + 0.2 * yy, 0.4 + 0.3 * xx, 0.5 + 0.5 * yy], axis=2)
    paths = []
    for k, im in enumerate((a, b)):
        px = np.clip(im + rng.normal(0, 0.03, im.shape), 0, 1)
        p = str(tmp_path / f"in{k}.png")
        Image.fromarray(np.rint(px * 255).astype(np.uint8), "RGB").save(p)
        paths.append(p)
    return paths


def _close_to_8bit(oracle_map, dims, written, tol=1e-6):
    """The written pixels equal the oracle map rounded to 8 bits, except where
    the oracle value lies within tol of a rounding boundary."""
    from culamp_amd import images as im
    want = im.to_rgb8(oracle_map, dims)
    v = im.restore_image(oracle_map, dims).transpose(1, 2, 0) * 255.0
    edge = np.abs(v - np.floor(v) - 0.5) < tol * 255
    assert ((written == want) | edge).all()


def test_lamp_color_transfer_files(gpu_ctx, oracle, tmp_path):
    import culamp_amd
    from culamp_amd import images as im
    f1, f2 = _pngs(tmp_path)
    o1, o2 = str(tmp_path / "out1.png"), str(tmp_path / "out2.png")
    kw = dict(eta_p=0.0, epsilon=1e-6, B=1.1, itermax=20000, verbose=True)
    say("ctransfer lamp 12x10")
    lines = []
    nu, phi, psi, a1, a2 = im.extragradient_color_transfer_images(
        f1, f2, o1, o2, (12, 10), culamp_amd.EOTArgs(**kw), 25, 2.0, log=lines.append,
        ctx=gpu_ctx)
    x, dims, r = im.load_rgb(f1, (12, 10))
    y, _, c = im.load_rgb(f2, (12, 10))
    o = oracle.lamp_solve(x, y, r, c, frequency=25, p=2.0, **kw)
    assert len(lines) == len(o["log"])
    assert np.max(np.abs(a1 - o["img1"])) < 1e-8 and np.max(np.abs(a2 - o["img2"])) < 1e-8
    assert np.max(np.abs(nu - o["nu"])) < 1e-8
    _close_to_8bit(o["img1"], dims, im.read_rgb_pixels(o1))
    _close_to_8bit(o["img2"], dims, im.read_rgb_pixels(o2))


def test_sinkhorn_color_transfer_via_run_ctransfer(gpu_ctx, oracle, tmp_path):
    from culamp_amd import images as im
    f1, f2 = _pngs(tmp_path)
    settings = tmp_path / "s.json"
    kw = dict(eta_p=0.05, epsilon=1e-8, itermax=5000, verbose=True)
    settings.write_text(json.dumps(kw))
    o1, o2 = str(tmp_path / "s1.png"), str(tmp_path / "s2.png")
    say("ctransfer sinkhorn 12x10 via run_ctransfer")
    phi, psi, a1, a2 = im.run_ctransfer(
        {"file1": f1, "file2": f2, "output1": o1, "output2": o2, "settings": str(settings),
         "algorithm": "sinkhorn", "frequency": 10, "p": 2.0, "height": 12, "width": 10},
        log=None, ctx=gpu_ctx)
    x, dims, r = im.load_rgb(f1, (12, 10))
    y, _, c = im.load_rgb(f2, (12, 10))
    o = oracle.sinkhorn_solve(x, y, r, c, frequency=10, p=2.0, **kw)
    assert np.max(np.abs(a1 - o["img1"])) < 1e-8 and np.max(np.abs(a2 - o["img2"])) < 1e-8
    assert np.max(np.abs(phi - o["phi"]) / np.maximum(np.abs(o["phi"]), 1.0)) < 1e-8
    _close_to_8bit(o["img1"], dims, im.read_rgb_pixels(o1))
    _close_to_8bit(o["img2"], dims, im.read_rgb_pixels(o2))


def test_fp32_color_transfer_runs_mixed(gpu_ctx, tmp_path):
    """--fp32: Float32 colours and marginals, mixed-precision pairs; the maps
    agree with the FP64 run to FP32 accuracy."""
    import culamp_amd
    from culamp_amd import images as im
    f1, f2 = _pngs(tmp_path, 16, 16)
    kw = dict(eta_p=0.02, epsilon=1e-6, B=1.1, itermax=20000, verbose=True)
    args = culamp_amd.EOTArgs(**kw)
    say("ctransfer lamp 16x16 fp64 and fp32")
    r64 = im.extragradient_color_transfer_images(f1, f2, str(tmp_path / "a.png"),
                                                 str(tmp_path / "b.png"), (16, 16), args, 50,
                                                 2.0, log=None, ctx=gpu_ctx)
    r32 = im.extragradient_color_transfer_images(f1, f2, str(tmp_path / "c.png"),
                                                 str(tmp_path / "d.png"), (16, 16), args, 50,
                                                 2.0, dtype=np.float32, log=None, ctx=gpu_ctx)
    err = max(np.max(np.abs(r64[3] - r32[3])), np.max(np.abs(r64[4] - r32[4])))
    say("fp32 vs fp64 map err", err)
    assert err < 1e-4
    d = np.abs(im.read_rgb_pixels(str(tmp_path / "a.png")).astype(int)
               - im.read_rgb_pixels(str(tmp_path / "c.png")).astype(int))
    assert d.max() <= 1
