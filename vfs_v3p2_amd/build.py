"""Build the HIP engine (libvfs_pp.so) in-tree with hipcc for gfx950.

Each .hip translation unit is compiled in parallel to an object under
vfs_v3p2_amd/_build/, then linked into vfs_v3p2_amd/libvfs_pp.so.  Every HIP
unit is compiled with packed FP32 instructions disabled.  The C++ host side
(host/vfs_userctx.cpp) is linked against it into host/libvfs_host.so, and the
one-block pressure-step driver host/pressure_step against both.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libvfs_pp.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["pp_core.hip", "pp_mg.hip", "pp_setup.hip", "pp_engine.hip", "pp_batch.hip", "pp_metrics.hip", "pp_grid.hip", "pp_ibsearch.hip", "pp_gradient.hip", "pp_noalloc.hip", "pp_momentum.hip"]

ARCH = os.environ.get("VFS_PP_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-parameter", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{SRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [hipcc()] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn:
        print(warn, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_host()
    if verbose:
        print(f"built {OUT}")
    return OUT


HOST = os.path.join(HERE, "host")


def build_host() -> None:
    """libvfs_host.so (the C++ UserCtx and per-step calls over the C ABI) and the
    pressure_step driver; host code only, linked against libvfs_pp.so."""
    base = [hipcc(), "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-Wall", f"-I{os.path.join(ROOT, 'include')}", f"-I{HOST}"]
    lib = os.path.join(HOST, "libvfs_host.so")
    steps = [
        base + ["-fPIC", "-shared", os.path.join(HOST, "vfs_userctx.cpp"), f"-L{HERE}", "-lvfs_pp",
                "-Wl,-rpath,$ORIGIN/..", "-o", lib + ".tmp"],
        base + [os.path.join(HOST, "pressure_step.cpp"), f"-L{HOST}", "-lvfs_host", f"-L{HERE}", "-lvfs_pp",
                "-Wl,-rpath,$ORIGIN:$ORIGIN/..", "-o", os.path.join(HOST, "pressure_step")],
    ]
    for i, cmd in enumerate(steps):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed:\n{r.stdout}\n{r.stderr}")
        if i == 0:
            os.replace(lib + ".tmp", lib)


if __name__ == "__main__":
    build()
