"""Build the HIP library (liblesgpu.so) in-tree for gfx950 with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "liblesgpu.so")
SOURCES = [
    "k_lines.hip", "k_lines_rg.hip", "k_tdma.hip", "k_point.hip", "k_stats.hip", "k_lin.hip", "les_api.hip",
    "les_lin.hip", "les_files.hip", "les_prof.hip",
]
ARCH = os.environ.get("LES_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # every product rounded, as the reference's arithmetic (no fused multiply-adds)
    "-ffp-contract=off",
    "-Wno-unused-result",
]


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every .hip source and link liblesgpu.so (always recompiles)."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0), procs)
    while procs:
        _wait(procs.pop(0), procs)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, LIB)
    return LIB


def _wait(item, others=()):
    """Wait for one compile; on failure stop the other compiles still
    running (so none is left writing its object) and raise."""
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    if p.returncode != 0:
        for _, q in others:
            q.kill()
            q.communicate()
        raise RuntimeError(f"hipcc failed on {src}:\n{text}")
    msgs = [l for l in text.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if msgs:
        sys.stderr.write("\n".join(msgs[:20]) + "\n")


if __name__ == "__main__":
    print(build(verbose=True))
