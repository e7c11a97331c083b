"""Build the HIP library ``biotfd/libbiotfd.so`` for gfx950, and beside it the
diagnostics library ``biotfd/libbiotfd_probes.so`` (bandwidth probes).

Flags that matter for correctness:
  -ffp-contract=off   every a*b+c rounds twice, as the reference's source order
                      does (the kernels are bit-exact against it)
  -fno-gpu-flush-denormals-to-zero   binary32 subnormals kept, as on the host
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BIOTFD_ARCH", "gfx950")
OUT = os.path.join(HERE, "biotfd", "libbiotfd.so")
SOURCES = [os.path.join(HERE, "csrc", f)
           for f in ("bq_device.hip", "bq_batch.hip", "bq_misfit.hip", "bq_learned.hip",
                     "bq_host.cpp", "bq_output.cpp")]
PROBES_OUT = os.path.join(HERE, "biotfd", "libbiotfd_probes.so")
PROBES_SOURCES = [os.path.join(HERE, "csrc", "bq_probes.hip")]

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"), "-pthread",
]


def _link(sources, out, verbose):
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *sources, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    _link(SOURCES, OUT, verbose)
    _link(PROBES_SOURCES, PROBES_OUT, verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
