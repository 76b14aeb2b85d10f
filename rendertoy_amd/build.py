"""Builds librtx_hip.so (the gfx950 kernels + C ABI) in-tree with hipcc.

Every source is compiled to its own object in parallel, then the objects are linked. The library
carries a build id, the SHA-256 of its sources and header (`source_id()`), which the loader
compares with the sources next to it, so a library older than its sources is never loaded.

Usage: python -m rendertoy_amd.build   (or rendertoy_amd.build.build_hip() from Python)
"""
import concurrent.futures
import hashlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "librtx_hip.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["rt_api.hip", "rt_build.hip", "rt_trace.hip", "rt_paths.hip", "rt_mesh.hip",
           "rt_image.hip"]
ARCH = "gfx950"

# -ffp-contract=off + correctly rounded div/sqrt: every FP32 result of the kernels is a fixed
# function of its inputs (the explicit fmaf calls are the only fused operations).
# -target-feature -packed-fp32-ops: no packed FP32 VALU instructions are emitted.
HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def source_files():
    """Every file the library is built from: csrc/*.hip, csrc/*.h and include/*.h, sorted."""
    files = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".h"))]
    files += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h")]
    return sorted(files)


def source_id():
    """SHA-256 (hex, 16 digits) over the names and bytes of source_files()."""
    h = hashlib.sha256()
    for f in source_files():
        h.update(os.path.basename(f).encode() + b"\0")
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def hipcc_path():
    p = shutil.which("hipcc")
    if p:
        return p
    p = "/opt/rocm/bin/hipcc"
    if os.path.exists(p):
        return p
    raise RuntimeError("hipcc not found: librtx_hip.so cannot be built")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    return r.returncode, err


def build_hip(verbose=False, jobs=None):
    """Compiles every source under csrc/ into rendertoy_amd/librtx_hip.so (always rebuilds)."""
    os.makedirs(OBJDIR, exist_ok=True)
    hipcc = hipcc_path()
    sid = source_id()
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        extra = [f'-DRT_BUILD_ID="{sid}"'] if src == "rt_api.hip" else []
        cmds.append([hipcc, *HIPCC_FLAGS, *extra, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_run, cmds))
    for cmd, (rc, err) in zip(cmds, results):
        if verbose:
            print(" ".join(cmd), flush=True)
        if rc != 0:
            raise RuntimeError(f"hipcc failed ({rc}) on {cmd[-3]}:\n{err}")
        if verbose and err.strip():
            print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    rc, err = _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC",
                    *[c[-1] for c in cmds], "-o", tmp])
    if rc != 0:
        raise RuntimeError(f"link failed ({rc}):\n{err}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_hip(verbose=True))
