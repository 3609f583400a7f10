"""Builds the native parts of voce_amd in-tree for gfx950.

  voce_amd/libvoce.so   HIP kernels + host context + file formats (the C ABI
                        of include/voce.h)
  voce_amd/bin/voce     the command-line host driver
  voce_amd/bin/hbm_copy a plain copy kernel of known size (traffic-counter
                        calibration)

Run ``python -m voce_amd.build``.  hipcc cross-compiles without a GPU.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("VOCE_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 arithmetic is disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose_resources=False):
    obj = os.path.join(HERE, "_obj")
    os.makedirs(obj, exist_ok=True)
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources else []
    objs = []
    jobs = []
    for src, flags in (("voce_kernels.hip", extra), ("voce_ctx.cpp", []), ("voce_tune.cpp", []),
                       ("voce_batch.cpp", []), ("voce_serve.cpp", []), ("voce_gen.cpp", []),
                       ("voce_noalloc.hip", [])):
        o = os.path.join(obj, os.path.splitext(src)[0] + ".o")
        objs.append(o)
        jobs.append([HIPCC] + HIP_FLAGS + flags + ["-c", os.path.join(CSRC, src), "-o", o])
    io_o = os.path.join(obj, "voce_io.o")
    objs.append(io_o)
    jobs.append(["gcc", "-O2", "-fPIC", "-std=gnu99", "-Wall", "-c", os.path.join(CSRC, "voce_io.c"), "-o", io_o])
    # compile the translation units side by side
    procs = []
    for cmd in jobs:
        print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    bad = [cmd for cmd, p in procs if p.wait() != 0]
    if bad:
        raise subprocess.CalledProcessError(1, bad[0])
    lib = os.path.join(HERE, "libvoce.so")
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-lpthread", "-o", lib])
    run(["gcc", "-O2", "-std=gnu99", "-Wall", os.path.join(CSRC, "voce_main.c"), "-o",
         os.path.join(HERE, "bin", "voce"), "-L" + HERE, "-lvoce", "-Wl,-rpath,$ORIGIN/.."])
    # calibration kernel for the HBM traffic counters
    run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "hbm_copy.hip"), "-o", os.path.join(HERE, "bin", "hbm_copy")])
    return lib


if __name__ == "__main__":
    build(verbose_resources="-v" in sys.argv)
