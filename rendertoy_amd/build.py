"""Builds librtx_hip.so (the gfx950 kernels + C ABI) in-tree with hipcc.

Usage: python -m rendertoy_amd.build   (or rendertoy_amd.build.build_hip() from Python)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librtx_hip.so")
SOURCES = ["rt_api.hip", "rt_build.hip", "rt_trace.hip", "rt_paths.hip", "rt_mesh.hip",
           "rt_image.hip"]
ARCH = "gfx950"

# -ffp-contract=off + correctly rounded div/sqrt: every FP32 result of the kernels is a fixed
# function of its inputs (the explicit fmaf calls are the only fused operations).
# -target-feature -packed-fp32-ops: no packed FP32 VALU instructions are emitted.
HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc_path():
    p = shutil.which("hipcc")
    if p:
        return p
    p = "/opt/rocm/bin/hipcc"
    if os.path.exists(p):
        return p
    raise RuntimeError("hipcc not found: librtx_hip.so cannot be built")


def build_hip(verbose=False):
    """Compiles every source under csrc/ into rendertoy_amd/librtx_hip.so (always rebuilds)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc_path(), *HIPCC_FLAGS, "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_hip(verbose=True))
