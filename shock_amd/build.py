"""Build libshockx.so (gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "shockx.hip")
OUT_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(OUT_DIR, "libshockx.so")

# Packed FP32 VALU forms are disabled for all device code on this target.  Multiply-adds are
# fused only within one source expression (-ffp-contract=on), so a kernel body gives the same
# rounding in every kernel it is inlined into (single-case and batched launches agree bit for bit).
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


DRIVER_SRC = os.path.join(HERE, "host", "shockx_drive.c")
DRIVER = os.path.join(HERE, "host", "shockx_drive")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_driver(verbose=False):
    """The C host loop (plain gcc against include/shockx.h and libshockx.so)."""
    cmd = ["gcc", "-O2", "-Wall", "-I", INCLUDE, DRIVER_SRC, "-L", OUT_DIR, "-lshockx",
           "-Wl,-rpath,$ORIGIN/../lib", "-lm", "-o", DRIVER]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return DRIVER


COMPAT_DIR = os.path.join(HERE, "compat")
COMPAT_LIB = os.path.join(OUT_DIR, "libshock_compat.so")
COMPAT_RUN = os.path.join(COMPAT_DIR, "compat_run")


def build_compat(verbose=False):
    """The reference-named entry points (shock_compat.h) over libshockx.so, and their step-loop
    driver (plain gcc)."""
    cmds = [
        ["gcc", "-O2", "-Wall", "-fPIC", "-shared", "-I", INCLUDE, "-I", COMPAT_DIR,
         os.path.join(COMPAT_DIR, "shock_compat.c"), os.path.join(COMPAT_DIR, "cgns_mesh.c"),
         "-L", OUT_DIR, "-lshockx", "-Wl,-rpath,$ORIGIN", "-pthread",
         "-o", COMPAT_LIB],
        ["gcc", "-O2", "-Wall", "-I", COMPAT_DIR, os.path.join(COMPAT_DIR, "compat_run.c"), "-L", OUT_DIR,
         "-lshock_compat", "-lshockx", "-Wl,-rpath,$ORIGIN/../lib", "-o", COMPAT_RUN],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return COMPAT_LIB


TESTCASE_SRC = os.path.join(HERE, "host", "testcase_run.c")
TESTCASE = os.path.join(HERE, "host", "testcase_run")


def build_testcase(verbose=False):
    """The C program that runs a reference testcase from its mesh and division files (plain gcc
    against both libraries)."""
    cmd = ["gcc", "-O2", "-Wall", "-I", INCLUDE, "-I", COMPAT_DIR, TESTCASE_SRC, "-L", OUT_DIR, "-lshock_compat",
           "-lshockx", "-Wl,-rpath,$ORIGIN/../lib", "-lm", "-o", TESTCASE]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return TESTCASE


def build(verbose=False, extra=()):
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + HIPCC_FLAGS + list(extra) + ["-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libshockx.so")
    os.replace(OUT + ".tmp", OUT)
    build_driver(verbose)
    build_compat(verbose)
    build_testcase(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
