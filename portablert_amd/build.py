"""Builds the native libraries of the MI355X nearest-hit backend in-tree.

    lib/libprt_mi355x.so   HIP kernels (gfx950 only) + the C ABI of include/prt_mi355x.h
    lib/libprt_scenes.so   deterministic workload generators (host C++)

Every HIP translation unit is compiled with -ffp-contract=off (the exact leaf predicates must
round every operation on its own) and without packed fp32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["api.hip", "build.hip", "trace.hip", "mlp.hip", "learned.hip", "batch.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r.stdout


def build_scenes():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libprt_scenes.so")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          os.path.join(CSRC, "scenes.cpp"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip(jobs=4):
    """Compiles each HIP source to an object in parallel, then links the shared library."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("HIP compilation failed")
    out = os.path.join(LIBDIR, "libprt_mi355x.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cpp_test():
    """The C++ API test program (tests/cpp/api_test.cpp) against the built library; its rpath is
    relative ($ORIGIN), so the copy of the tree on another machine finds the library."""
    out = os.path.join(LIBDIR, "api_test")
    _run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "cpp", "api_test.cpp"), "-L", LIBDIR, "-lprt_mi355x",
          "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_scenes()
    build_hip()
    build_cpp_test()


if __name__ == "__main__":
    build_all()
    print("built", LIBDIR)
