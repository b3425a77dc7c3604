"""Build the CPU oracle (test infrastructure only) into oracle/liboracle.so with g++.

No fused multiply-add contraction, so the oracle evaluates the reference's FP64 expressions as
written. x86-64-v3 keeps the shared object runnable on any AVX2 host. Also builds
oracle/liboracle_flops.so, the flop-counting form of the same source.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the same source with Real = the flop-counting double of flopcount.h, single-threaded: the
# algorithmic flop counts of the measurement
LIB_FLOPS = os.path.join(HERE, "liboracle_flops.so")


def _compile(extra, out, verbose):
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", *extra, "-std=c++17", "-fPIC",
           "-shared", os.path.join(HERE, "oracle.cc"), "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)


def build(verbose=False):
    _compile(["-fopenmp"], LIB, verbose)
    _compile(["-DORACLE_COUNT_FLOPS"], LIB_FLOPS, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
