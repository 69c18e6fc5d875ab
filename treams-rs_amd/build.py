"""Build libtreams_gpu.so (gfx950) in-tree from csrc/ with hipcc.

    python treams-rs_amd/build.py [-v]
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtreams_gpu.so")
DRIVER = os.path.join(HERE, "cluster_driver")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    host_sources = sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    jobs = []
    for src in sources + host_sources:
        obj = os.path.join(build_dir, os.path.basename(src) + ".o")
        if src.endswith(".cpp"):
            cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-c", src, "-o", obj]
        else:
            cmd = ["hipcc"] + HIP_FLAGS + ["-c", src, "-o", obj]
        objs.append(obj)
        jobs.append(cmd)
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    failed = [cmd for cmd, p in zip(jobs, procs) if p.wait() != 0]
    if failed:
        raise RuntimeError("compilation failed: " + " ".join(failed[0]))
    link = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT] + objs
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.check_call(link)
    # C++ host driver over the C ABI (the call sequence of a Rust host)
    drv = ["hipcc", "-O2", "-std=c++17", os.path.join(HERE, "host", "cluster_driver.cpp"), "-o", DRIVER,
           "-L" + HERE, "-ltreams_gpu", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(drv), flush=True)
    subprocess.check_call(drv)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
