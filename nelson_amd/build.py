"""Builds libnelson_amd.so (HIP kernels for gfx950 + the C ABI) in-tree.

Usage: python -m nelson_amd.build   (or nelson_amd.build.build() from Python)
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnelson_amd.so")

ARCH = os.environ.get("NSN_OFFLOAD_ARCH", "gfx950")
HIP_SOURCES = ["ba_kernels.hip", "chol_kernels.hip", "api.hip", "batch_kernels.hip", "batch_api.hip", "calib.hip",
               "structure_dev.hip"]
CPP_SOURCES = ["generate.cpp", "util.cpp", "structure.cpp", "io.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
# packed FP32 instructions are disabled for all device code
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


# a workgroup may not rely on more than 64 KiB of LDS (the kernels are sized
# to that budget; the resource report below enforces it)
LDS_LIMIT = 65536


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _check_resources(src, out):
    """Parses -Rpass-analysis=kernel-resource-usage remarks: LDS within
    LDS_LIMIT and no scratch (register spills) in any kernel of this library.
    rocPRIM's kernels (the radix sorts and scans of the structure stage) are
    the library's own code, with their own stack frames: not checked."""
    fn = None
    for line in out.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            fn = m.group(1)
        if fn and "rocprim" in fn:
            continue
        m = re.search(r"LDS Size \[bytes[^\]]*\]: (\d+)", line)
        if m and int(m.group(1)) > LDS_LIMIT:
            raise RuntimeError(f"{src}: kernel {fn} uses {m.group(1)} bytes of LDS (> {LDS_LIMIT})")
        m = re.search(r"ScratchSize \[bytes[^\]]*\]: (\d+)", line)
        if m and int(m.group(1)) > 0:
            raise RuntimeError(f"{src}: kernel {fn} spills ({m.group(1)} bytes/lane of scratch)")


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        cmds.append(["hipcc", "-c"] + COMMON + HIP_FLAGS + ["-Rpass-analysis=kernel-resource-usage",
                                                            os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CPP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        cmds.append(["g++", "-c"] + COMMON + ["-Wall", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    for s, out in zip(HIP_SOURCES, outs):
        _check_resources(s, out)
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
    tmp = LIB + ".tmp"
    # roctx ranges around the C ABI calls (rocprofiler-sdk-roctx)
    _run(["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"])
    os.replace(tmp, LIB)
    build_cpp_tests()
    return LIB


CPP_TESTS = os.path.join(ROOT, "tests", "cpp")


def build_cpp_tests():
    """The C++ programs under tests/cpp, linked against the library (rpath to
    the in-tree copy)."""
    out = []
    for f in sorted(os.listdir(CPP_TESTS)):
        if f.endswith(".cpp"):
            exe = os.path.join(CPP_TESTS, f[:-4])
            _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(CPP_TESTS, f),
                  "-L", HERE, "-lnelson_amd", "-Wl,-rpath,$ORIGIN/../../nelson_amd", "-o", exe])
            out.append(exe)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
