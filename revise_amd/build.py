"""Build librevise_hip.so in-tree for gfx950.

HIP sources (*.hip) are compiled by hipcc for gfx950 only; host sources
(*.cpp) by g++.  All floating-point code is compiled with -ffp-contract=off so
every fp32 operation is rounded on its own (the ray-cast results are required
to match the CPU restatement bit for bit), and HIP code additionally with
packed-fp32 instruction selection disabled.

Usage: python -m revise_amd.build   (or revise_amd.build.build())
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librevise_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = [
    "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-Wall", "-Wextra",
    "-Wno-unused-parameter",
]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _compile(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return cmd, p.returncode, p.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    hip, cpp = _sources()
    jobs = []
    objs = []
    for f in hip:
        o = os.path.join(OBJ, f + ".o")
        objs.append(o)
        jobs.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, f), "-o", o])
    for f in cpp:
        o = os.path.join(OBJ, f + ".o")
        objs.append(o)
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, f), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        results = list(ex.map(_compile, jobs))
    failed = False
    for cmd, rc, out in results:
        if verbose or rc:
            sys.stderr.write(" ".join(cmd) + "\n" + out)
        failed |= rc != 0
    if failed:
        raise RuntimeError("librevise_hip build failed")
    tmp = LIB + ".tmp"
    link = [hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp, *objs,
            "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lpthread"]
    cmd, rc, out = _compile(link)
    if rc:
        sys.stderr.write(" ".join(cmd) + "\n" + out)
        raise RuntimeError("librevise_hip link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
