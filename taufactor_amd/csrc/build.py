"""Build recipe for libtauhip.so (gfx950) from taufactor_amd/csrc/tauhip.hip.

Run as ``python -m taufactor_amd.csrc.build`` or through ``__graft_entry__.build()``.
The library is written in-tree next to the package so that it travels with
the source tree and is found by ``taufactor_amd._lib``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
SRC = os.path.join(HERE, "tauhip.hip")
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(PKG, "libtauhip.so")

ARCH = os.environ.get("TAUHIP_ARCH", "gfx950")

# -ffp-contract=off: the update must round each float32 operation on its own
#   (no fused multiply-add), as the reference's eager torch ops do.
# -fno-gpu-flush-denormals-to-zero: keep float32 subnormals, as the CPU does.
# -packed-fp32-ops disabled: no v_pk_{add,mul,fma}_f32 in generated code.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc_path() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found; ROCm is required to build libtauhip.so")
    return p


def build_library(verbose: bool = False) -> str:
    """Compile libtauhip.so; returns its path. Raises on compiler failure."""
    tmp = OUT + ".tmp"
    cmd = [hipcc_path(), *HIPCC_FLAGS, f"-I{INCLUDE}", SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        # host pass warns about the packed-fp32 feature flag; harmless
        print(res.stderr.strip()[-2000:], flush=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    path = build_library(verbose="-q" not in sys.argv)
    print(path)
