"""Kernel-level entry points, one per reference device kernel (reference name
without the trailing ``!``).  Arguments are ``DeviceArray`` buffers (points as
3*n float64 in point-major order) plus the reference's scalar arguments; each
call is stream-ordered on the context and writes its output buffer(s).

    warp_logsumexp_spp_ct_fused / _opt  MirrorProxKernel.jl:833-938 / 702-829
    residual_spp_c                      MirrorProxKernel.jl:7-94
    max_logsumexp_spp_ct                MirrorProxKernel.jl:280-326
    warp_min_reduce                     MirrorProxKernel.jl:941-1016
    update_theta_residual               MirrorProx.jl:159-180
    warp_logsumexp_ct_fused / _opt      SinkhornKernel.jl:341-444 / 230-338
    residual_opt                        SinkhornKernel.jl:9-81
    naive_findmaxindex_ct               SinkhornKernel.jl:124-162
    naive_findmaxindex_spp_ct / _ct_t   MirrorProxKernel.jl:97-142 / 144-189

Feature (cell-similarity) kernels take a ``FeatureProblem`` (its packed rows
stay on the device) in place of the reference's data / row-statistics
arguments:

    warp_logsumexp_spp_sim_fused        MirrorProxCellSim.jl:259-342
    residual_cellsim_c                  MirrorProxCellSim.jl:190-257
    max_logsumexp_cellsim               MirrorProxCellSim.jl:1-41
    warp_min_reduce_cellsim             MirrorProxCellSim.jl:43-92
    warp_logsumexp_cellsim_fused        SinkhornCellSim.jl:153-238
    residual_cellsim_opt                SinkhornCellSim.jl:1-72
"""
from __future__ import annotations

from ._lib import FP64, Context, DeviceArray, check, dptr


def _n(a: DeviceArray) -> int:
    assert a.n % 3 == 0, "point buffers hold 3*n values"
    return a.n // 3


def max_logsumexp_spp_ct(ctx: Context, output: DeviceArray, img1: DeviceArray,
                         img2: DeviceArray, p: float) -> None:
    check(ctx.lib.culamp_max_cost(ctx.handle, img1.ptr, _n(img1), img2.ptr, _n(img2), float(p),
                                  output.ptr))


def warp_logsumexp_spp_ct_fused(ctx: Context, output: DeviceArray, img1: DeviceArray,
                                img2: DeviceArray, theta: DeviceArray, reg: float, st: float,
                                winf: float, p: float, precision: int = FP64) -> None:
    check(ctx.lib.culamp_lse_rows(ctx.handle, img1.ptr, _n(img1), img2.ptr, _n(img2), theta.ptr,
                                  float(reg), float(st), float(winf), float(p), precision,
                                  output.ptr))


warp_logsumexp_spp_ct_opt = warp_logsumexp_spp_ct_fused


def residual_spp_c(ctx: Context, output: DeviceArray, cost_output: DeviceArray | None,
                   img1: DeviceArray, img2: DeviceArray, marginal1: DeviceArray,
                   theta: DeviceArray, logZi: DeviceArray, reg: float, st: float, winf: float,
                   p: float, precision: int = FP64) -> None:
    check(ctx.lib.culamp_col_marginal_cost(ctx.handle, img1.ptr, _n(img1), img2.ptr, _n(img2),
                                           marginal1.ptr, theta.ptr, logZi.ptr, float(reg),
                                           float(st), float(winf), float(p), precision,
                                           output.ptr, dptr(cost_output)))


def warp_min_reduce(ctx: Context, output: DeviceArray, img1: DeviceArray, img2: DeviceArray,
                    theta: DeviceArray, winf: float, p: float) -> None:
    check(ctx.lib.culamp_minplus_rows(ctx.handle, img1.ptr, _n(img1), img2.ptr, _n(img2),
                                      theta.ptr, float(winf), float(p), output.ptr))


def update_theta_residual(ctx: Context, theta: DeviceArray, theta_0: DeviceArray,
                          residual: DeviceArray, c: DeviceArray, eta_mu_i: DeviceArray,
                          eta_mu: float, adjust: bool, minv: float, maxv: float,
                          winf: float) -> None:
    check(ctx.lib.culamp_lamp_theta_update(ctx.handle, theta.ptr, theta_0.ptr, residual.ptr,
                                           c.ptr, eta_mu_i.ptr, theta.n, float(eta_mu),
                                           1 if adjust else 0, float(minv), float(maxv),
                                           float(winf)))


def warp_logsumexp_ct_fused(ctx: Context, output: DeviceArray, img1: DeviceArray,
                            img2: DeviceArray, marginal: DeviceArray, theta: DeviceArray,
                            reg: float, winf: float, p: float, precision: int = FP64) -> None:
    check(ctx.lib.culamp_sinkhorn_half_step(ctx.handle, img1.ptr, _n(img1), img2.ptr, _n(img2),
                                            marginal.ptr, theta.ptr, float(reg), float(winf),
                                            float(p), precision, output.ptr))


def warp_logsumexp_ct_opt(ctx: Context, output: DeviceArray, img1: DeviceArray,
                          img2: DeviceArray, marginal: DeviceArray, psi: DeviceArray,
                          reg: float, winf: float, p: float, precision: int = FP64) -> None:
    """Two-pass Sinkhorn half-step; like the reference it evaluates p == 1 or
    squared L2 only (SinkhornKernel.jl:259-263)."""
    warp_logsumexp_ct_fused(ctx, output, img1, img2, marginal, psi, reg, winf,
                            1.0 if p == 1 else 2.0, precision)


def residual_opt(ctx: Context, output: DeviceArray, cost_output: DeviceArray | None,
                 img1: DeviceArray, img2: DeviceArray, marginal: DeviceArray, phi: DeviceArray,
                 psi: DeviceArray, reg: float, winf: float, p: float,
                 precision: int = FP64) -> None:
    """Row residual and cost; p == 1 or squared L2 as the reference (38-42)."""
    check(ctx.lib.culamp_row_residual_cost(ctx.handle, img1.ptr, _n(img1), img2.ptr, _n(img2),
                                           marginal.ptr, phi.ptr, psi.ptr, float(reg), float(winf),
                                           1.0 if p == 1 else 2.0, precision, output.ptr,
                                           dptr(cost_output)))


def naive_findmaxindex_ct(ctx: Context, output_img: DeviceArray, img1: DeviceArray,
                          img2: DeviceArray, phi: DeviceArray, psi: DeviceArray, reg: float,
                          winf: float, p: float) -> None:
    check(ctx.lib.culamp_barycentric_sinkhorn(ctx.handle, img1.ptr, _n(img1), img2.ptr, _n(img2),
                                              phi.ptr, psi.ptr, float(reg), float(winf),
                                              1.0 if p == 1 else 2.0, output_img.ptr))


def naive_findmaxindex_spp_ct(ctx: Context, output_img: DeviceArray, img1: DeviceArray,
                              img2: DeviceArray, mu: DeviceArray, marginal1: DeviceArray,
                              logZi: DeviceArray, reg: float, st: float, winf: float,
                              p: float) -> None:
    check(ctx.lib.culamp_barycentric_lamp_rows(ctx.handle, img1.ptr, _n(img1), img2.ptr,
                                               _n(img2), mu.ptr, marginal1.ptr, logZi.ptr,
                                               float(reg), float(winf), float(p), output_img.ptr))


def naive_findmaxindex_spp_ct_t(ctx: Context, output_img: DeviceArray, img1: DeviceArray,
                                img2: DeviceArray, marginal1: DeviceArray, mu: DeviceArray,
                                logZi: DeviceArray, reg: float, st: float, winf: float,
                                p: float) -> None:
    check(ctx.lib.culamp_barycentric_lamp_cols(ctx.handle, img1.ptr, _n(img1), img2.ptr,
                                               _n(img2), marginal1.ptr, mu.ptr, logZi.ptr,
                                               float(reg), float(winf), float(p), output_img.ptr))


def warp_logsumexp_spp_sim_fused(ctx: Context, output: DeviceArray, prob, theta: DeviceArray,
                                 reg: float, st: float, winf: float,
                                 precision: int = FP64) -> None:
    check(ctx.lib.culamp_feature_lse_rows(ctx.handle, prob.handle, theta.ptr, float(reg),
                                          float(st), float(winf), precision, output.ptr))


def residual_cellsim_c(ctx: Context, output: DeviceArray, cost_output: DeviceArray | None, prob,
                       marginal1: DeviceArray, theta: DeviceArray, logZi: DeviceArray,
                       reg: float, st: float, winf: float, precision: int = FP64) -> None:
    check(ctx.lib.culamp_feature_col_marginal_cost(ctx.handle, prob.handle, marginal1.ptr,
                                                   theta.ptr, logZi.ptr, float(reg), float(st),
                                                   float(winf), precision, output.ptr,
                                                   dptr(cost_output)))


def max_logsumexp_cellsim(ctx: Context, output: DeviceArray, prob) -> None:
    check(ctx.lib.culamp_feature_max_cost(ctx.handle, prob.handle, output.ptr))


def warp_min_reduce_cellsim(ctx: Context, output: DeviceArray, prob, theta: DeviceArray,
                            winf: float) -> None:
    check(ctx.lib.culamp_feature_minplus_rows(ctx.handle, prob.handle, theta.ptr, float(winf),
                                              output.ptr))


def warp_logsumexp_cellsim_fused(ctx: Context, output: DeviceArray, prob, rows: bool,
                                 marginal: DeviceArray, theta: DeviceArray, reg: float,
                                 winf: float, precision: int = FP64) -> None:
    check(ctx.lib.culamp_feature_sinkhorn_half_step(ctx.handle, prob.handle, 1 if rows else 0,
                                                    marginal.ptr, theta.ptr, float(reg),
                                                    float(winf), precision, output.ptr))


def residual_cellsim_opt(ctx: Context, output: DeviceArray, cost_output: DeviceArray | None,
                         prob, marginal: DeviceArray, phi: DeviceArray, psi: DeviceArray,
                         reg: float, winf: float, precision: int = FP64) -> None:
    check(ctx.lib.culamp_feature_row_residual_cost(ctx.handle, prob.handle, marginal.ptr,
                                                   phi.ptr, psi.ptr, float(reg), float(winf),
                                                   precision, output.ptr, dptr(cost_output)))
