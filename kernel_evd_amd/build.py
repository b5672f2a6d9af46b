"""Build libkevd.so (HIP, gfx950) and the C++ drivers in-tree.

Every HIP translation unit is compiled by hipcc for gfx950 only; host-only
C++ (the KS generator, .npy I/O, drivers) by g++.  Outputs:
  kernel_evd_amd/libkevd.so
  kernel_evd_amd/bin/kevd-{gen,arpc,evd,bandwidth}
Object files go to kernel_evd_amd/_build/.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
DRV = os.path.join(PKG, "drivers")
OBJ = os.path.join(PKG, "_build")
BIN = os.path.join(PKG, "bin")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libkevd.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["ctx.hip", "arpc.hip", "factor.hip", "evd.hip", "bandwidth.hip", "ensemble.hip", "peaks.hip"]
CPP_SOURCES = ["ksgen.cpp"]
DRIVERS = ["kevd_gen.cpp", "kevd_arpc.cpp", "kevd_evd.cpp", "kevd_bandwidth.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops stay disabled in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC, "-I" + CSRC,
]
# per-file device flags: the bandwidth sweep feeds MFMA results straight to
# v_exp_f32, so its MFMAs write ordinary VGPRs (no accumulator-register reads)
FILE_FLAGS = {"bandwidth.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v3", "-I" + INC, "-I" + CSRC,
             "-Wall", "-Wextra"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  %s\n%s" % (" ".join(cmd), out))
    return out


def _compile(src):
    name = os.path.basename(src)
    obj = os.path.join(OBJ, name + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + FILE_FLAGS.get(name, []) + ["-c", src, "-o", obj]
    else:
        cmd = ["g++"] + CXX_FLAGS + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build_library(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + CPP_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas", "-lrocprofiler-sdk-roctx",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    return LIB


def find_hdf5():
    """(include dir, lib dir) of an HDF5 C library (the reference's file
    format), or None: the drivers then read and write .npy only."""
    for prefix in (os.environ.get("HDF5_DIR", ""), "/opt/conda", "/usr", "/usr/local"):
        if not prefix:
            continue
        for inc in (os.path.join(prefix, "include"), os.path.join(prefix, "include", "hdf5", "serial")):
            if not os.path.exists(os.path.join(inc, "hdf5.h")):
                continue
            for lib in (os.path.join(prefix, "lib"), os.path.join(prefix, "lib", "x86_64-linux-gnu"),
                        os.path.join(prefix, "lib", "x86_64-linux-gnu", "hdf5", "serial")):
                if os.path.exists(os.path.join(lib, "libhdf5.so")):
                    return inc, lib
    return None


def build_drivers(jobs=8):
    os.makedirs(BIN, exist_ok=True)
    common = [os.path.join(DRV, "npy.cpp"), os.path.join(DRV, "arrays.cpp")]
    h5 = find_hdf5()
    h5flags = ([], [])
    if h5:
        # Link the library by path and resolve it at run time through a
        # private directory holding only its soname: a search path to the
        # whole prefix (conda's) would also shadow the system libstdc++
        # that the ROCm libraries need.
        import glob
        import re
        sonames = [f for f in glob.glob(os.path.join(h5[1], "libhdf5.so.*"))
                   if re.fullmatch(r"libhdf5\.so\.\d+", os.path.basename(f))]
        if sonames:
            priv = os.path.join(BIN, "hdf5")
            os.makedirs(priv, exist_ok=True)
            link = os.path.join(priv, os.path.basename(sonames[0]))
            if os.path.lexists(link):
                os.remove(link)
            os.symlink(os.path.realpath(sonames[0]), link)
            h5flags = (["-DKEVD_HAVE_HDF5", "-isystem", h5[0]], [sonames[0], "-Wl,-rpath,$ORIGIN/hdf5"])

    def one(d):
        exe = os.path.join(BIN, os.path.splitext(d)[0].replace("_", "-"))
        _run(["g++"] + CXX_FLAGS + h5flags[0] + ["-o", exe, os.path.join(DRV, d)] + common +
             ["-L" + PKG, "-lkevd", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + os.path.join(ROCM, "lib")] +
             h5flags[1])
        return exe

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(one, DRIVERS))


def build_asan():
    """The drivers' host code (KS generation, delay embedding, .npy and HDF5
    I/O, point loading) with the address and undefined-behaviour sanitizers:
    _build/asan/asan_check, run by the CPU tests.  g++ only (no device code is
    instrumented); the library supplies the remaining symbols."""
    out_dir = os.path.join(OBJ, "asan")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "asan_check")
    h5 = find_hdf5()
    h5flags = ([], [])
    priv = os.path.join(BIN, "hdf5")
    if h5 and os.path.isdir(priv):
        import glob
        libs = glob.glob(os.path.join(priv, "libhdf5.so.*"))
        if libs:
            h5flags = (["-DKEVD_HAVE_HDF5", "-isystem", h5[0]], [libs[0], "-Wl,-rpath," + priv])
    srcs = [os.path.join(DRV, "asan_check.cpp"), os.path.join(DRV, "npy.cpp"), os.path.join(DRV, "arrays.cpp"),
            os.path.join(CSRC, "ksgen.cpp")]
    _run(["g++", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-std=c++17",
          "-I" + INC, "-I" + CSRC, "-I" + DRV] + h5flags[0] + ["-o", exe] + srcs +
         ["-L" + PKG, "-lkevd", "-Wl,-rpath," + PKG, "-Wl,-rpath," + os.path.join(ROCM, "lib")] + h5flags[1])
    return exe


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all():
    lib = build_library()
    drivers = build_drivers()
    orc = build_oracle()
    return lib, drivers, orc


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
