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


# translation units: (source, extra flags, object); the sweep unit is compiled once per
# environment count so that the units build in parallel
UNITS = [("mahyco.hip", [], "mahyco.o"), ("mh_lag.hip", [], "mh_lag.o"), ("mh_noalloc.hip", [], "mh_noalloc.o")] + [
    ("mh_sweep.hip", [f"-DMH_TU_M={m}"], f"mh_sweep_m{m}.o") for m in (1, 2, 3, 4)]


def build(verbose: bool = False, jobs: int = 0) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cc = hipcc()
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    cmds = []
    for src, extra, obj in UNITS:
        cmds.append([cc] + compile_flags + extra + ["-c", "-o", os.path.join(OUT_DIR, obj),
                                                    os.path.join(CSRC, src)])
    jobs = jobs or min(len(cmds), os.cpu_count() or 1)
    procs, failed = [], []
    pending = list(cmds)
    while pending or procs:
        while pending and len(procs) < jobs:
            cmd = pending.pop(0)
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd)))
        cmd, p = procs.pop(0)
        if p.wait() != 0:
            failed.append(cmd)
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB + ".tmp"
    link = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [os.path.join(OUT_DIR, o) for _, _, o in UNITS]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
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
