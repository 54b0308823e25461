"""In-tree build of libfemgpu.so (hipcc, gfx950) and of the FEMproject-compatible
host executable ``fem_main``.

The HIP code is compiled for gfx950 only, with packed-FP32 instructions
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libfemgpu.so")

ARCH = os.environ.get("FEMGPU_ARCH", "gfx950")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]

LIB_SOURCES = ["kernels.hip", "cg1.hip", "batch.hip", "batch_ls.hip", "noalloc.hip", "femgpu_api.cpp",
               "batch.cpp", "mesh_plan.cpp", "segy.cpp"]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_library(out: str = LIB) -> str:
    """Compile libfemgpu.so (always recompiles)."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = _hipcc()
    objs = [os.path.join(CSRC, src.replace(".", "_") + ".o") for src in LIB_SOURCES]
    cmds = [[hipcc, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, src), "-o", o]
            for src, o in zip(LIB_SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=min(len(cmds), 4)) as ex:
        list(ex.map(_run, cmds))
    tmp = out + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


FEM_MAIN = os.path.join(PKG, "fem_main")
FEM_MAIN_STAMP = FEM_MAIN + ".srchash"


def _fem_main_sources() -> list[str]:
    src_dir = os.path.join(PKG, "femproject")
    if not os.path.isdir(src_dir):
        return []
    return sorted(os.path.join(src_dir, f) for f in os.listdir(src_dir)
                  if f.endswith((".cpp", ".h")))


def _fem_main_hash() -> str:
    """sha256 over the driver's sources and the C-ABI header it compiles against."""
    import hashlib
    h = hashlib.sha256()
    for p in _fem_main_sources() + [os.path.join(INCLUDE, "femgpu.h")]:
        h.update(os.path.basename(p).encode())
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def fem_main_current() -> bool:
    """fem_main exists and was built from the sources as they are now."""
    if not (os.path.exists(FEM_MAIN) and os.path.exists(FEM_MAIN_STAMP)):
        return False
    with open(FEM_MAIN_STAMP) as f:
        return f.read().strip() == _fem_main_hash()


def build_fem_main() -> str | None:
    """Compile the FEMproject-compatible driver executable (femproject/),
    linked against libfemgpu.so with an rpath to this directory, and record
    the hash of its sources beside it."""
    src_dir = os.path.join(PKG, "femproject")
    srcs = [p for p in _fem_main_sources() if p.endswith(".cpp")]
    if not srcs:
        return None
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{src_dir}", "-o", FEM_MAIN, *srcs,
          f"-L{PKG}", "-lfemgpu", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"])
    with open(FEM_MAIN_STAMP, "w") as f:
        f.write(_fem_main_hash() + "\n")
    return FEM_MAIN


def ensure_fem_main() -> str:
    """fem_main as built from the current sources (rebuilt with g++ when it is
    missing or stale; needs libfemgpu.so)."""
    if not fem_main_current():
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run __graft_entry__.build()")
        build_fem_main()
    return FEM_MAIN


PROBE_LIB = os.path.join(PKG, "libfemgpu_probe.so")


def build_probe(out: str = PROBE_LIB) -> str:
    """Compile libfemgpu_probe.so: the measurement kernels of the
    stacked-load-case element contraction (on the fly against K_e with the
    matrix instruction), not part of the solver library."""
    hipcc = _hipcc()
    tmp = out + ".tmp"
    _run([hipcc, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-shared", "-o", tmp,
          os.path.join(CSRC, "contraction_probe.hip")])
    os.replace(tmp, out)
    return out


def build_all():
    build_library()
    build_probe()
    build_fem_main()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
