"""Builds libhexa.so (gfx950) in-tree with hipcc.

Every source is compiled for gfx950 only, with packed FP32 instructions
disabled, then linked into ``libhexa.so`` beside this file.  Compilation of
the translation units runs in parallel.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhexa.so")

SOURCES = ["capi.cpp", "capi_topopt.cpp", "capi_noalloc.hip", "server.cpp", "io.cpp", "ke.cpp", "k_elastic.hip", "k_elastic_fact.hip", "k_elastic_rw.hip", "k_elastic_mfma.hip", "k_mg.hip", "k_vector.hip", "k_pcg.hip", "k_filter.hip",
           "k_topopt.hip", "k_probe.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


# Per-file flags.  The row-wave operator writes its fused multiply-adds out
# and is compiled without automatic contraction: each copy of its unrolled
# layer loop must round the same way, so that results do not depend on the
# z-chunk length.
FILE_FLAGS = {"k_elastic_rw.hip": ["-ffp-contract=off"]}


def _compile(src: str, build_dir: str = BUILD, extra=()) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = ([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + FLAGS + FILE_FLAGS.get(src, [])
           + list(extra))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False, lib: str = LIB, build_dir: str = BUILD, defines=()) -> str:
    """Compile and link; ``defines`` (e.g. ["-DHEXA_FACT_PF=2"]) and another
    ``lib``/``build_dir`` build tuning variants beside the default library."""
    os.makedirs(build_dir, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda f: _compile(f, build_dir, defines), SOURCES))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
