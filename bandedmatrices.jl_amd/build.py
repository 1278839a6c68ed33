"""Build libbandedhip.so (gfx950 only) from csrc/*.hip with hipcc.

Usage: python bandedmatrices.jl_amd/build.py [-j N]
The library is written next to this file; it is the only native artefact the
package loads.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libbandedhip.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: only explicit fma() calls fuse, so the rounding recipe of
#   every kernel is what the source says.
# -packed-fp32-ops off: no v_pk_*_f32 instructions are generated.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


_HEAVY = ("gbtrs_ring_", "gbtrs_part_up", "gbtrs_part_lo", "lu_lane_", "gbmv_slab_", "tbsv_", "lu_small",
          "gbtrf_mfma", "gbqr", "gbtrf_reg")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def _stale(src: str, headers_mtime: float) -> bool:
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return os.path.getmtime(src) > t or headers_mtime > t


def build(jobs: int = 8, verbose: bool = False, incremental: bool = False) -> str:
    """Compile every csrc/*.hip and link the library.  With incremental=True
    only sources newer than their object (or than any header) are compiled."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    hdrs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(HERE, "..", "include", "*.h"))
    hm = max((os.path.getmtime(h) for h in hdrs), default=0.0)
    todo = [s for s in srcs if _stale(s, hm)] if incremental else srcs
    # the longest compiles start first (the units that instantiate many kernel
    # variants), so the parallel build ends with the short ones
    todo = sorted(todo, key=lambda s: not any(k in os.path.basename(s) for k in _HEAVY))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ, os.path.basename(s).replace(".hip", ".o")) for s in srcs]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources ({len(todo)} compiled)")
    return OUT


if __name__ == "__main__":
    j = 8
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    build(j, verbose=True, incremental="--incremental" in sys.argv)
