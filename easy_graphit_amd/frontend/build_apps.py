"""Build the GraphIt apps through the staged front end.

For each app of the reference (apps/{bfs,sssp,cc,pr,bc}.cpp, compiled
unchanged from the reference tree with this front end's headers in place of
BuildIt's and GraphIt's): g++ builds the generator, running it prints the
app's HIP translation unit (as the reference's generator prints CUDA), and
hipcc compiles that for gfx950 into a shared library over libgraphit_amd.so.
Outputs go to easy_graphit_amd/frontend/_gen/ (not tracked):

    <app>_gen            the generator
    <app>[_power].hip    its output (the app's "power" mode: hybrid schedule, unfused)
    lib<app>[_power].so  the compiled program; its entry point is the staged
                         function with C linkage (BFS, SSSP, CC, PR, BC)

Run as ``python -m easy_graphit_amd.frontend.build_apps`` or through the
package build (when the reference tree is present).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
GEN = os.path.join(HERE, "_gen")
INCLUDE = os.path.join(HERE, "include")
SOURCES = [os.path.join(HERE, "src", "staging.cpp"), os.path.join(HERE, "src", "graphit.cpp")]
CSRC = os.path.join(PKG, "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("GRAPHIT_AMD_ARCH", "gfx950")

# app -> (modes); "power" is the app's hybrid/unfused mode (argv[3])
APPS = {"bfs": ("", "power"), "sssp": ("", "power"), "cc": ("",), "pr": ("",), "bc": ("", "power")}
MAX_CTA, CTA_SIZE = "0", "256"  # argv[1], argv[2]: 0 = a persistent grid sized by the runtime


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd[:3])}... failed:\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}")
    return r


def _generator(app: str, apps_dir: str) -> str:
    exe = os.path.join(GEN, f"{app}_gen")
    _run(["g++", "-std=c++17", "-O0", f"-I{INCLUDE}", os.path.join(apps_dir, f"{app}.cpp"), *SOURCES, "-o", exe])
    return exe


def _program(app: str, mode: str, exe: str) -> str:
    tag = app + (f"_{mode}" if mode else "")
    args = [exe, MAX_CTA, CTA_SIZE] + ([mode] if mode else [])
    hip = os.path.join(GEN, f"{tag}.hip")
    with open(hip, "w") as f:
        r = subprocess.run(args, stdout=f, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"generator {app} {mode} failed:\n{r.stderr[-4000:]}")
    lib = os.path.join(GEN, f"lib{tag}.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-shared",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-variable",
          "-Wno-unused-command-line-argument", f"-I{CSRC}", hip, "-o", lib + ".tmp",
          f"-L{PKG}", "-lgraphit_amd", "-Wl,-rpath,$ORIGIN/../.."])
    os.replace(lib + ".tmp", lib)
    return lib


def build(reference_root: str = "/root/reference", verbose: bool = False) -> list[str]:
    apps_dir = os.path.join(reference_root, "apps")
    if not os.path.isdir(apps_dir):
        if verbose:
            print(f"front end: {apps_dir} not present; staged apps not rebuilt")
        return []
    os.makedirs(GEN, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        gens = dict(zip(APPS, ex.map(lambda a: _generator(a, apps_dir), APPS)))
        jobs = [(a, m, gens[a]) for a, modes in APPS.items() for m in modes]
        libs = list(ex.map(lambda j: _program(*j), jobs))
    if verbose:
        print(f"front end: built {len(libs)} staged programs in {GEN}")
    return libs


if __name__ == "__main__":
    build(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", verbose=True)
