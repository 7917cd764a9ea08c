"""Builds libvoxelhip.so (all HIP kernels + the C-ABI) for gfx950 with hipcc, in-tree.

    python matrix-free-voxel-elasticity_amd/build.py [--verbose]

Output: matrix-free-voxel-elasticity_amd/voxelhip/libvoxelhip.so
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "voxelhip" / "libvoxelhip.so"
# matvec.hip is compiled through matvec_{f64,f32}.hip and matvec_batch_{f64,f32}.hip (one
# translation unit each, built in parallel)
SOURCES = ["capi.hip", "matvec_f64.hip", "matvec_f32.hip", "matvec_batch_f64.hip", "matvec_batch_f32.hip",
           "grid_ops.hip", "blas.hip", "simp.hip", "batch.hip", "stencil.hip", "lanes.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VOXELHIP_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for every kernel (scalar and float2 code only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd[-3:])}")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)


def build(verbose=False, extra=(), jobs=None):
    """Compiles every translation unit in parallel (one hipcc per source) and links the .so."""
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory(prefix="voxelhip-build-") as d:
        objs = [str(Path(d) / (Path(s).stem + ".o")) for s in SOURCES]
        cmds = [[HIPCC, *FLAGS, *extra, "-c", str(CSRC / s), "-o", o] for s, o in zip(SOURCES, objs)]
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
