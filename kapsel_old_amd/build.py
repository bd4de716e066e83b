"""Builds kapsel_old_amd/libspm.so for gfx950 with hipcc, in tree.

Usage: python -m kapsel_old_amd.build [--jobs N] [--define NAME=VALUE ...] [--out PATH]

--define overrides a launch-shape constant of csrc/fft_core.h
(SPM_LINE_POINTS_F64, SPM_LINE_RADIX, SPM_Z_POINTS_F64, SPM_FUSED_RADIX,
SPM_FUSED_X_POINTS, SPM_ZADV_POINTS_F64, ...) and
--out writes that build somewhere other than the in-tree library.
"""
import hashlib
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspm.so")
EXE = os.path.join(HERE, "bin", "spm")
SOURCES = ["fft.hip", "fused.hip", "spectral.hip", "particles.hip", "electrolyte.hip", "shear.hip", "noalloc.hip", "closure.hip",
           "spm_api.cpp", "ensemble.cpp", "checkpoint.cpp", "udf.cpp"]
# plain host C++ (no device code), built with the host compiler
HOST_SOURCES = {"udf.cpp"}
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src, objdir=OBJ, defines=()):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    if src in HOST_SOURCES:
        cmd = [CXX, "-O2", "-fPIC", "-std=c++17", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed on %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(jobs=4, verbose=False, defines=(), out=LIB):
    defines = tuple(defines)
    objdir = OBJ if not defines else os.path.join(OBJ, hashlib.sha1(" ".join(defines).encode()).hexdigest()[:10])
    os.makedirs(objdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, objdir, defines), SOURCES))
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    if out == LIB:
        exe = build_program()
        if verbose:
            print("built", exe)
    return out


def build_program():
    """kapsel_old_amd/bin/spm: the C++ command-line program (csrc/spm_main.cpp),
    host code over the library's C ABI, linked to the in-tree libspm.so."""
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    tmp = EXE + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "spm_main.cpp"), "-o", tmp,
           "-L" + HERE, "-l:libspm.so", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("building the spm program failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, EXE)
    return EXE


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--define", action="append", default=[])
    ap.add_argument("--out", default=LIB)
    a = ap.parse_args()
    build(jobs=a.jobs, verbose=True, defines=a.define, out=os.path.abspath(a.out))
