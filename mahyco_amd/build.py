"""Builds the HIP library (gfx950) in-tree: mahyco_amd/_build/libmahyco_hip.so."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmahyco_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # every multiply and add rounded separately, as in the CPU reference
    "-ffp-contract=off",
    # no packed FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP library)")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    src = os.path.join(CSRC, "mahyco.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


RUNNER = os.path.join(OUT_DIR, "mahyco_run")


def build_host(verbose: bool = False) -> str:
    """C++ host interface (MahycoModule / IRemap / IEquationOfState) and its driver."""
    host = os.path.join(HERE, "host")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(host, "mahyco_module.cpp"), os.path.join(host, "main.cpp"),
           "-L", OUT_DIR, "-lmahyco_hip", "-Wl,-rpath,$ORIGIN", "-o", RUNNER]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return RUNNER


if __name__ == "__main__":
    print(build(verbose=True))
