"""Build of the native libraries (run by ``__graft_entry__.build()``).

* ``openvort_amd/libopenvort_gpu.so`` — the HIP library behind
  ``include/openvort_gpu.h``, compiled for gfx950 with hipcc.
* ``oracle/liboracle.so`` and, where the reference tree is present,
  ``oracle/_ref/libovref.so`` — the test-only checkers (``oracle/Makefile``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libopenvort_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are switched off for every device pass: the
# compiler's own v_pk_* forms have given sporadically wrong results on these
# machines, so all FP32 arithmetic is issued as scalar VALU operations.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-value",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_gpu(extra=None) -> str:
    srcs = [os.path.join(CSRC, "ov_gpu.hip")]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, *(extra or []), "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    if os.path.isdir("/root/reference/src"):
        _run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all() -> None:
    build_gpu()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "gpu" in what:
            build_gpu()
        if "oracle" in what:
            build_oracle()
