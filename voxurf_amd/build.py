"""Ahead-of-time build of libvoxhip.so (all HIP sources under csrc/) for gfx950.

Usage: python -m voxurf_amd.build   (or __graft_entry__.build()).  Importing the
package never compiles anything; the loader only opens the built library.
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(OUT_DIR, "libvoxhip.so")
OBJ_DIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("VOXHIP_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 instructions off: only scalar FP32 VALU ops are emitted.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE single-precision division and square root (also hipcc's default).
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


HEADER = os.path.join(os.path.dirname(HERE), "include", "voxhip.h")


def header_hash(path=HEADER):
    """sha256 of the C-ABI header; compiled into the library and checked at load time."""
    import hashlib
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def _compile(src, obj):
    cmd = [_hipcc(), *HIP_FLAGS, f'-DVX_ABI_HASH="{header_hash()}"', "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{msgs}")
    return msgs


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(OBJ_DIR, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, msgs in zip(srcs, ex.map(_compile, srcs, objs)):
            if verbose and msgs.strip():
                print(f"[{os.path.basename(src)}]\n{msgs}")
    tmp = LIB_PATH + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link of libvoxhip.so failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB_PATH)
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
