"""Builds libaerolbm.so (the HIP library behind include/aerolbm.h) in-tree for gfx950.

Usage: python -m aerosim_amd.build   (or aerosim_amd.build.build() from Python)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaerolbm.so")

ARCH = os.environ.get("AERO_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# (source, extra flags)
SOURCES = [
    # contraction decided per source expression, so every cells-per-thread variant rounds identically
    ("lbm_step.hip", ["-ffp-contract=on"]),
    ("voxelize.hip", ["-ffp-contract=off"]),
    ("aerolbm_api.cpp", []),
    ("aerolbm_noalloc.hip", []),
    ("aerolbm_batch.cpp", []),
    ("mlp_train.hip", []),
    ("geometry.cpp", ["-ffp-contract=off"]),
    ("stl.cpp", ["-ffp-contract=off"]),
    # readouts round where the reference's double arithmetic does
    ("fields.hip", ["-ffp-contract=off"]),
    ("aerolbm_fields.cpp", ["-ffp-contract=off"]),
]


def _hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src, extra):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{p.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_node_addon(verbose)
    return LIB


NODE_DIR = os.path.join(HERE, "node")
NODE_ADDON = os.path.join(NODE_DIR, "aerolbm.node")


def build_node_addon(verbose=False):
    """N-API addon (aerolbm.node) for the JavaScript host; skipped when the node headers are absent."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        if verbose:
            print("node headers not found: N-API addon not built")
        return None
    tmp = NODE_ADDON + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{inc}", os.path.join(NODE_DIR, "aerolbm_napi.cpp"),
           f"-L{HERE}", "-laerolbm", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"N-API addon build failed: {' '.join(cmd)}\n{p.stderr}")
    os.replace(tmp, NODE_ADDON)
    if verbose:
        print(f"built {NODE_ADDON}")
    return NODE_ADDON


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
