"""Settings: EOTArgs and read_args_json (src/Utilities.jl:18-32, 88-92).

Same JSON keys and defaults as the reference's ``EOTArgs`` (missing keys keep
their defaults, unknown keys are ignored as JSON3 does for a mutable struct).
Two optional keys are new, both defaulting to the reference's behaviour:
``"precision"`` in {"fp64", "mixed", "bf16", "bf16x3"} (default "fp64";
the bf16 modes apply to the cosine / correlation feature costs) and ``"stop"`` in
{"reference", "infeas"} (default "reference"; "infeas" stops LAMP on
infeasibility < epsilon/6 alone).
"""
from __future__ import annotations

import ctypes as C
import json
import math
from dataclasses import dataclass, fields

from ._lib import BF16, BF16X3, FP64, MIXED, culamp_args

_PRECISIONS = {"fp64": FP64, "mixed": MIXED, "bf16": BF16, "bf16x3": BF16X3}
_STOPS = {"reference": 0, "infeas": 1}


@dataclass
class EOTArgs:
    eta_p: float = 0.0
    eta_mu: float = 0.0
    tau_p: float = 1.0
    tau_mu: float = 1.0
    alpha: float = 0.01
    B: float = 1.0
    epsilon: float = 1e-4
    anneal_mult: float = 0.95
    itermax: int = 10_000
    inner_iter: int = 10
    tmax: float = math.inf
    verbose: bool = True
    precision: str = "fp64"
    stop: str = "reference"

    def to_c(self) -> culamp_args:
        if self.precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {self.precision!r}")
        a = culamp_args()
        for f in ("eta_p", "eta_mu", "tau_p", "tau_mu", "alpha", "B", "epsilon", "anneal_mult", "tmax"):
            setattr(a, f, float(getattr(self, f)))
        a.itermax = int(self.itermax)
        a.inner_iter = int(self.inner_iter)
        a.verbose = 1 if self.verbose else 0
        a.precision = _PRECISIONS[self.precision]
        if self.stop not in _STOPS:
            raise ValueError(f"stop must be one of {sorted(_STOPS)}, got {self.stop!r}")
        a.stop = _STOPS[self.stop]
        return a

    def copy(self, **kw) -> "EOTArgs":
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d.update(kw)
        return EOTArgs(**d)


_INT_FIELDS = {"itermax", "inner_iter"}
_BOOL_FIELDS = {"verbose"}


def args_from_dict(d: dict) -> EOTArgs:
    a = EOTArgs()
    names = {f.name for f in fields(EOTArgs)}
    for k, v in d.items():
        if k not in names:
            continue
        if k in _INT_FIELDS:
            if isinstance(v, bool) or (isinstance(v, float) and not v.is_integer()):
                raise ValueError(f"{k} must be an integer, got {v!r}")
            v = int(v)
        elif k in _BOOL_FIELDS:
            if not isinstance(v, bool):
                raise ValueError(f"{k} must be a boolean, got {v!r}")
        elif k == "precision":
            if v not in _PRECISIONS:
                raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {v!r}")
        elif k == "stop":
            if v not in _STOPS:
                raise ValueError(f"stop must be one of {sorted(_STOPS)}, got {v!r}")
        else:
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise ValueError(f"{k} must be a number, got {v!r}")
            v = float(v)
        setattr(a, k, v)
    return a


def read_args_json(fpath: str) -> EOTArgs:
    """read_args_json (src/Utilities.jl:88-92)."""
    with open(fpath) as f:
        return args_from_dict(json.load(f))
