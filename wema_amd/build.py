"""Builds the gfx950 pipeline library wema_amd/libwema_dev.so with hipcc, and the C11 host
library wema_amd/libwema.so (the include/wema.h API over it), its benchmark driver
wema_amd/wema_h_bench and the command-line front end wema_amd/wema with gcc.

Usage: python wema_amd/build.py   (a standalone script: importing the package needs the
library it builds)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwema_dev.so")
HOST_LIB = os.path.join(HERE, "libwema.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["wema_kernels.hip", "wema_dev.hip", "wema_synth.hip", "wema_pool.hip", "wema_serve.hip",
           "wema_tune.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every fused multiply-add is written out explicitly, in the reference's shape
    "-ffp-contract=off",
    # the reference runs with FTZ/DAZ (-Ofast links crtfastmath)
    "-fgpu-flush-denormals-to-zero",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp%d" % os.getpid()
    # (rocprofiler-sdk-roctx: named host ranges of calls and stage launches for tracers)
    cmd = [hipcc()] + HIPCC_FLAGS + srcs + ["-lrocprofiler-sdk-roctx", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libwema_dev.so")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose=False):
    """The C11 host: plain gcc, linked against libwema_dev.so found next to it."""
    tmp = HOST_LIB + ".tmp%d" % os.getpid()
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Wpedantic", "-fPIC", "-shared",
           "-I" + INCLUDE, os.path.join(HERE, "host", "wema_host.c"), "-L" + HERE, "-lwema_dev",
           "-Wl,-rpath,$ORIGIN", "-Wl,-Bsymbolic", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gcc failed building libwema.so")
    os.replace(tmp, HOST_LIB)
    # the command-line front end (the reference CLI's option set, FFmpeg pipes)
    cli = os.path.join(HERE, "wema")
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-I" + INCLUDE, os.path.join(HERE, "host", "wema_cli.c"),
           "-L" + HERE, "-lwema_dev", "-Wl,-rpath,$ORIGIN", "-o", cli + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gcc failed building the wema CLI")
    os.replace(cli + ".tmp", cli)
    # the drop-in benchmark (bench.py runs it): the wema.h path driven from C
    exe = os.path.join(HERE, "wema_h_bench")
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-I" + INCLUDE,
           os.path.join(HERE, "host", "wema_h_bench.c"), "-L" + HERE, "-lwema", "-lwema_dev",
           "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gcc failed building wema_h_bench")
    os.replace(exe + ".tmp", exe)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
