"""Builds the native parts of the package in-tree.

* libcrsolve.so  -- HIP kernels for gfx950 + the C ABI (include/crsolve.h),
                    compiled with hipcc.
* bin/k_cops_visibility, bin/k_cops_visibility_v2
                 -- the reference-compatible CLI (app/main.cpp behaviour) with the
                    six-step plug-in for each reference solver TU, linked against
                    libcrsolve.so.

Usage: python -m copsandrobbers_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("CR_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-value",
    "-Wno-unused-result",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops",)
    for line in (proc.stdout + proc.stderr).splitlines():
        if not any(n in line for n in noise):
            print(line)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


HIP_SOURCES = ("crsolve.hip", "macro.hip", "batch.hip", "columns.hip", "game.hip", "cr_host.cpp")


def build_library():
    """Compiles every source to an object (in parallel; each holds the kernels it
    launches) and links them into libcrsolve.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = max(1, min(len(HIP_SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))

    def compile_one(name):
        obj = os.path.join(objdir, os.path.splitext(name)[0] + ".o")
        run([HIPCC, *HIP_FLAGS, "-c", f"-I{INCLUDE}", f"-I{CSRC}", os.path.join(CSRC, name), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(compile_one, HIP_SOURCES))
    out = os.path.join(HERE, "libcrsolve.so")
    tmp = out + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli():
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    outs = []
    for variant, name in ((1, "k_cops_visibility"), (2, "k_cops_visibility_v2")):
        out = os.path.join(bindir, name)
        run([CXX, "-O2", "-std=c++17", "-Wall", f"-DCR_PLUGIN_VARIANT={variant}", f"-I{INCLUDE}",
             os.path.join(CSRC, "cli_main.cpp"), os.path.join(CSRC, "plugin_solver.cpp"),
             f"-L{HERE}", "-lcrsolve", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def main():
    build_library()
    build_cli()


if __name__ == "__main__":
    sys.exit(main())
