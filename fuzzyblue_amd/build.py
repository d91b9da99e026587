"""Build libfuzzyblue_hip.so (HIP kernels + C-ABI runtime) for gfx950.

Usage: ``python -m fuzzyblue_amd.build``.  Compiles every source under
csrc/ with hipcc into objects under csrc/_build/ and links the shared
library next to this file, in-tree, so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libfuzzyblue_hip.so")
# Micro-kernels measuring the VALU and LDS ceilings (loaded by bench.py only).
CEILINGS_LIB = os.path.join(HERE, "libfb_ceilings.so")
CEILINGS_SOURCES = ["ceilings.hip"]

ARCH = os.environ.get("FB_OFFLOAD_ARCH", "gfx950")

SOURCES = ["precompute.hip", "render.hip", "runtime.hip", "batch.hip", "server.hip",
           "tune.hip", "noalloc.hip", "lutio.hip", "interop.hip"]

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Every written multiply and add rounds separately (the reference's
    # binary32 semantics); kernels fuse explicitly with fmaf where intended.
    "-ffp-contract=off",
    # Packed FP32 VALU forms stay off (compiler-generated v_pk_* f32 ops).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [hipcc(), *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature.
    err = "\n".join(
        l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()
    )
    if r.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{err}")
    if err:
        sys.stderr.write(err + "\n")
    return obj


def _link(objs, out):
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of {os.path.basename(out)} failed:\n{r.stderr}")
    os.replace(tmp, out)


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES + CEILINGS_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    _link([objs[s] for s in SOURCES], LIB)
    _link([objs[s] for s in CEILINGS_SOURCES], CEILINGS_LIB)
    if verbose:
        print(f"built {LIB} and {CEILINGS_LIB}")
    return LIB


if __name__ == "__main__":
    build()
