"""Build recipe for libtauhip.so (gfx950) from the HIP sources in taufactor_amd/csrc.

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
# the C-ABI and non-sweep kernels, then one unit per sweep kind (compiled in
# parallel: the sweep templates dominate the build)
SOURCES = ["tauhip.hip", "sweep_solver.hip", "sweep_multiphase.hip", "sweep_mpfaces.hip", "sweep_field.hip",
           "sweep_anisotropic.hip", "sweep_electrode.hip", "sweep_batch.hip"]
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(PKG, "libtauhip.so")
OBJ_DIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("TAUHIP_ARCH", "gfx950")

# -ffp-contract=off: the update must round each float32 operation on its own,
#   (no fused multiply-add), as the reference's eager torch ops do.
# -fno-gpu-flush-denormals-to-zero: keep float32 subnormals, as the CPU does.
# -packed-fp32-ops disabled: no v_pk_{add,mul,fma}_f32 in generated code.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
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


def _run(cmd):
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def build_library(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile libtauhip.so (every unit, in parallel, then one link); returns
    its path.  Raises on compiler failure."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    hipcc = hipcc_path()
    jobs = jobs or max(1, min(len(SOURCES), os.cpu_count() or 1))
    pending = [(src, os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")) for src in SOURCES]
    running, objs, errors = [], [], []
    while pending or running:
        while pending and len(running) < jobs:
            src, obj = pending.pop(0)
            cmd = [hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", "-c", os.path.join(HERE, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            running.append((src, obj, _run(cmd)))
        src, obj, proc = running.pop(0)
        out, err = proc.communicate()
        if proc.returncode != 0:
            errors.append(f"{src}: hipcc failed ({proc.returncode}):\n{out}\n{err}")
        objs.append(obj)
    if errors:
        raise RuntimeError("\n".join(errors))
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    path = build_library(verbose="-q" not in sys.argv)
    print(path)
