"""Build the in-tree native libraries.

libplens_gpu.so  -- the product: HIP kernels for gfx950 + C ABI + host mesh/metric layer.
Built with hipcc directly (no JIT cache), so the shared object sits next to this file and travels
with the repository snapshot.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libplens_gpu.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libplens_gpu.so")
    return exe


def build(verbose=False):
    """Compile libplens_gpu.so; raises on failure."""
    obj_host = os.path.join(CSRC, "host_mesh.o")
    obj_out = os.path.join(CSRC, "host_output.o")
    obj_expr = os.path.join(CSRC, "expr.o")
    obj_dev = os.path.join(CSRC, "plens_gpu.o")
    obj_mlp = os.path.join(CSRC, "mlp.o")
    cmds = [
        ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}", "-c",
         os.path.join(CSRC, "host_mesh.cc"), "-o", obj_host],
        ["g++", "-O2", "-std=c++17", "-fPIC", f"-I{INCLUDE}", "-c", os.path.join(CSRC, "host_output.cc"), "-o", obj_out],
        ["g++", "-O2", "-std=c++17", "-fPIC", f"-I{INCLUDE}", "-c", os.path.join(CSRC, "expr.cc"), "-o", obj_expr],
        [hipcc(), *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, "plens_gpu.hip"),
         "-o", obj_dev],
        [hipcc(), *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, "mlp.hip"), "-o", obj_mlp],
        [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", obj_dev, obj_mlp, obj_host, obj_out, obj_expr, "-o", LIB + ".tmp"],
    ]
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(c)}\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
    os.replace(LIB + ".tmp", LIB)
    for o in (obj_host, obj_out, obj_expr, obj_dev, obj_mlp):
        if os.path.exists(o):
            os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
