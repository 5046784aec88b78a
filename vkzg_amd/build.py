#!/usr/bin/env python3
"""Builds vkzg_amd/libvkzg.so for gfx950 with hipcc (no cmake needed).

Each HIP translation unit compiles in parallel; the mainnet setup is linked
in through embed_setup.S.  Run:  python3 vkzg_amd/build.py [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvkzg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VKZG_AMD_ARCH", "gfx950")

SOURCES = ["setup.hip", "setup_fft.hip", "setup_tables.hip", "fk20_scalars.hip", "phase_a.hip",
           "ladder.hip", "phase_b.hip", "reduce.hip", "recover.hip", "msm.hip", "pairing.hip",
           "verify.hip", "prover.cpp", "noalloc.cpp", "server.cpp", "verify.cpp", "tables_v1.cpp", "sha256.cc"]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall"]  # plain host C++ (.cc): g++
HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    # Packed FP32 instructions are kept out of every device code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"),
]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in line and "is not a recognized feature" not in line)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def compile_one(src):
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    if src.endswith(".cc"):
        run(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj
    cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    run(cmd)
    return obj


def build(jobs=8, verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    setup_file = os.path.join(HERE, "data", "g1_monomial_mainnet.bin")
    embed_obj = os.path.join(OUT_DIR, "embed_setup.o")
    run(["gcc", "-c", os.path.join(CSRC, "embed_setup.S"), "-o", embed_obj,
         '-DVKZG_SETUP_FILE="%s"' % setup_file])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    # Export only the C ABI (vkzg_* and vkzg_amd_*): a version script, so the
    # extern "C" functions keep default visibility in the objects.
    run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [embed_obj] +
        ["-Wl,--version-script=" + os.path.join(CSRC, "exports.map")])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    try:
        build(a.j)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
