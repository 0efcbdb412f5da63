"""Build recipe for libgraphit_amd.so (gfx950 only).

Compiles every csrc/*.hip translation unit with hipcc in parallel and links
them into easy_graphit_amd/libgraphit_amd.so, in-tree. Run as
``python -m easy_graphit_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgraphit_amd.so")
ARCH = os.environ.get("GRAPHIT_AMD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 instructions stay off: ordinary float code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    workers = min(len(sources), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ apps over the mirror API (include/graphit_amd.hpp), linked to the library
    include = os.path.join(os.path.dirname(HERE), "include")
    for app in sorted(glob.glob(os.path.join(HERE, "apps", "*.cpp"))):
        exe = app[:-4] + "_app"
        cmd = ["g++", "-std=c++17", "-O2", f"-I{include}", app, f"-L{HERE}", "-lgraphit_amd",
               "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"app build failed for {app}:\n{r.stderr}")
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    # the reference's apps through the staged front end (when its tree is here)
    from easy_graphit_amd.frontend import build_apps
    build_apps.build(verbose=verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
