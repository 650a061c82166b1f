"""Build the gfx950 shared library src_amd/libopalsc.so with hipcc.

Every translation unit is compiled with -ffp-contract=off (no FMA
contraction: the per-particle arithmetic must round like the reference's
scalar C++) and without packed-FP32 instructions.  Objects go to
src_amd/_build/, the library next to this file.
"""
import concurrent.futures as cf
import json
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libopalsc.so")

TUNING = os.path.join(HERE, "tuning.json")  # written by src_amd/autotune.py on an MI355X
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OPALSC_ARCH", "gfx950")
FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", BUILD,
]

DEFAULT_LARGE_BOX_MAX = 8 << 20  # particles; the deposit's larger LDS box up to this size


def write_tuning_header():
    """_build/tuning.hpp from the autotuner's table: FFT tile entries
    {L, fp64, E} and the deposit's large-box threshold (defaults without a table)."""
    tiles, large, drop = [], DEFAULT_LARGE_BOX_MAX, 0
    if os.path.exists(TUNING):
        with open(TUNING) as f:
            t = json.load(f)
        for e in t.get("fft_tiles", []):
            tiles.append("{%d, %d, %d}, " % (int(e["L"]), 1 if e["precision"] == "fp64" else 0, int(e["E"])))
        large = int(t.get("deposit", {}).get("large_box_max_particles", large))
        drop = int(t.get("sort", {}).get("drop_bits", drop))
    text = ("// generated by src_amd/build.py from src_amd/tuning.json; do not edit\n"
            "#pragma once\n"
            "#define OPALSC_TUNED_FFT_TILES %s\n"
            "#define OPALSC_TUNED_LARGE_BOX_MAX ((int64_t)%dLL)\n"
            "#define OPALSC_TUNED_SORT_DROP_BITS %d\n" % ("".join(tiles), large, drop))
    path = os.path.join(BUILD, "tuning.hpp")
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
SOURCES = ["api.hip", "bins.hip", "particles.hip", "poisson.hip", "sort.hip", "noalloc.hip", "io.cpp"]


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature; drop that noise
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, out))
    if out.strip() and verbose:
        print(out)
    return obj


H5LIB = os.path.join(HERE, "libopalsc_h5.so")
HDF5_PREFIXES = [os.environ.get("OPALSC_HDF5", ""), "/opt/conda", "/usr", "/usr/local"]


def hdf5_prefix():
    for p in HDF5_PREFIXES:
        if p and os.path.exists(os.path.join(p, "include", "hdf5.h")):
            return p
    return None


def build_h5(verbose=False):
    """libopalsc_h5.so (H5Part files, host code) against the HDF5 C library;
    skipped with a message where no HDF5 is installed."""
    prefix = hdf5_prefix()
    if prefix is None:
        print("build: no HDF5 found (OPALSC_HDF5); libopalsc_h5.so not built", file=sys.stderr)
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(prefix, "include"),
           os.path.join(CSRC, "h5part.cpp"), "-o", H5LIB + ".tmp", "-L", os.path.join(prefix, "lib"),
           "-Wl,-rpath," + os.path.join(prefix, "lib"), "-lhdf5"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for h5part.cpp:\n" + r.stdout)
    os.replace(H5LIB + ".tmp", H5LIB)
    return H5LIB


def build(verbose=False, jobs=4):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(BUILD, exist_ok=True)
    write_tuning_header()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
