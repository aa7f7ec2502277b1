"""Build libculamp.so (gfx950) in-tree with hipcc.

    python culamp.jl_amd/build.py            # compile + link
    python culamp.jl_amd/build.py --clean

Each HIP translation unit is compiled in parallel, then linked into
``culamp.jl_amd/libculamp.so``.  Every HIP compile carries
``-Xclang -target-feature -Xclang -packed-fp32-ops`` (packed FP32 VALU ops are
kept out of the generated code).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libculamp.so")
SOURCES = ["culamp_api.hip", "culamp_batch.hip", "culamp_dense.hip", "culamp_diff.hip", "culamp_noalloc.hip", "culamp_pairs.hip",
           "culamp_pass_fp32.hip", "culamp_pass_fp64.hip", "culamp_probe.hip"]
ARCH = os.environ.get("CULAMP_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

# Per-file additions: the MFMA feature kernels (instantiated in the FP32 pass
# unit) write their accumulators to VGPRs, which their exp-sum reads
# directly, instead of AGPRs that would need a v_accvgpr_read per value.
EXTRA = {"culamp_pass_fp32.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}

_NOISE = ("packed-fp32-ops' is not a recognized feature", "packed-fp32-ops\" is not a recognized")


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if not any(n in l for n in _NOISE))
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    _run([HIPCC, *CFLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp,
          *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        for f in [LIB] + [os.path.join(OBJ, s.replace(".hip", ".o")) for s in SOURCES]:
            if os.path.exists(f):
                os.remove(f)
        return
    build()


if __name__ == "__main__":
    main()
