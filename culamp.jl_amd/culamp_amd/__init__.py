"""culamp_amd -- host-side mirror of CuLAMP.jl's on-the-fly solvers over the
libculamp C ABI (MI355X / gfx950 HIP kernels).

Solver entry points keep the reference's names, argument order and return
values (src/solvers/MirrorProxKernel.jl, SinkhornKernel.jl,
AnnealedSinkhorn.jl); kernel-level functions keep the reference kernel names
(without the trailing ``!``).  Point clouds are ``(n, 3)`` C-contiguous float64
arrays, i.e. the same bytes as the reference's ``3 x n`` column-major
matrices.
"""
from ._lib import (BF16, BF16X3, FP64, MIXED, CulampError, Context, DeviceArray, device_count,
                   load)
from .args import EOTArgs, args_from_dict, read_args_json
from .solvers import (annealed_sinkhorn_euclidean, default_context,
                      extragradient_color_transfer, extragradient_euclidean,
                      sinkhorn_color_transfer, sinkhorn_euclidean, Problem)
from . import kernels
from .batch import Batch
from .diff import divergence_opts, sinkhorn_divergence
from . import diff
from .cellsim import (FeatureProblem, annealed_sinkhorn_cellsim, cost_scale,
                      extragradient_cellsim, normalize_cell_cost_metric, sinkhorn_cellsim)
from . import cindex
from .dense import DenseProblem
from . import dotmark
from . import images
from .images import load_rgb, save_image
from .cindex import (CellCostKernel, OTScOmicsCellData, compute_otscomics_c_index,
                     streaming_c_index)

__all__ = [
    "BF16", "BF16X3", "FP64", "MIXED", "CulampError", "Context", "DeviceArray", "device_count", "load",
    "EOTArgs", "args_from_dict", "read_args_json", "annealed_sinkhorn_euclidean",
    "default_context", "extragradient_color_transfer", "extragradient_euclidean",
    "sinkhorn_color_transfer", "sinkhorn_euclidean", "Problem", "kernels", "FeatureProblem",
    "annealed_sinkhorn_cellsim", "cost_scale", "extragradient_cellsim",
    "normalize_cell_cost_metric", "sinkhorn_cellsim", "Batch", "divergence_opts",
    "sinkhorn_divergence", "diff", "cindex", "CellCostKernel", "OTScOmicsCellData",
    "compute_otscomics_c_index", "streaming_c_index", "DenseProblem", "dotmark", "images",
    "load_rgb", "save_image",
]
