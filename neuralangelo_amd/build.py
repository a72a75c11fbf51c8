"""Build recipe for the gfx950 HIP library ``libneuralangelo_amd.so``.

Every ``csrc/*.hip`` file is compiled with hipcc for gfx950 into
``_build/`` and linked into ``neuralangelo_amd/libneuralangelo_amd.so``
(in-tree, so the built library travels with the source tree).

Usage:  python -m neuralangelo_amd.build [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.path.join(REPO_DIR, "_build", "hip")
LIB_PATH = os.path.join(PKG_DIR, "libneuralangelo_amd.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(REPO_DIR, "include"),
    # Packed-FP32 VALU forms are kept out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, proc.returncode, proc.stdout + proc.stderr, cmd


def build(jobs=None, verbose=False):
    """Compile all HIP sources and link the shared library.  Raises on error."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)), 16)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, out, cmd in ex.map(_compile, srcs):
            if rc != 0:
                raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{out}")
            if verbose and out.strip():
                print(out, file=sys.stderr)
            objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
