"""Builds libpacking.so in-tree with hipcc for gfx950 (no torch, no cmake)."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libpacking.so")
OBJ = os.path.join(HERE, "_obj")

SOURCES = [
    "pk_overlap.hip",
    "pk_lut.hip",
    "pk_relax.hip",
    "pk_feasible.hip",
    "pk_lap.hip",
    "pk_lbfgs.hip",
    "pk_sep.hip",
]

ARCH = os.environ.get("PK_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 is disabled for all device code: hipcc -O3 would otherwise form
# v_pk_{add,mul,fma}_f32 from scalar/float2 code.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-I" + INCLUDE, "-I" + CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-ffp-contract=on",
    # atomics to one address from a whole wave: reduce with DPP first (the default
    # lane-by-lane loop costs ~64 iterations per atomic)
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=DPP",
]


# Per-file additions.  The separation builder decides containment by the signs of exact
# cross products (pieces of a decomposition share edges and vertices, so many are exactly
# zero); fused multiply-adds would turn those zeros into rounding noise.
EXTRA = {"pk_sep.hip": ["-ffp-contract=off"]}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{p.stderr}")
    return obj


def build(verbose: bool = False, only=None) -> str:
    """Compile every source (or, with only, just those, reusing the other objects) and
    link libpacking.so."""
    os.makedirs(OBJ, exist_ok=True)
    todo = SOURCES if only is None else [s for s in SOURCES if s in only]
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True, only=sys.argv[1:] or None)
