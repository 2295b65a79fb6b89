"""Build libth2d.so in-tree for gfx950 with hipcc (no JIT, no cache dir),
and the two C libraries with the reference's function signatures on top
of it (libthirring_hmc.so, libthirring_bag.so).

    python -m thirring2d_amd.build
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libth2d.so")
COMPAT = os.path.join(HERE, "compat")
COMPAT_LIBS = {"libthirring_hmc.so": "thirring_hmc.c", "libthirring_bag.so": "thirring_bag.c"}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for every HIP source.
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def c_sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.c")))


def build(verbose: bool = False) -> str:
    """Compile every HIP/C source and link libth2d.so; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in c_sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = ["gcc", "-O2", "-std=c99", "-fPIC", "-Wall", "-Wextra", "-c", src, "-o", obj]
        _run(cmd, verbose)
        objs.append(obj)
    # One hipcc process per HIP source, run side by side, then one link.
    hip_objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in sources()]
    cmds = [[HIPCC] + HIP_FLAGS + ["-c", s, "-o", o] for s, o in zip(sources(), hip_objs)]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + hip_objs + objs
    _run(cmd, verbose)
    os.replace(tmp, OUT)
    for lib, src in COMPAT_LIBS.items():
        out = os.path.join(HERE, lib)
        cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-Wextra",
               "-o", out + ".tmp", os.path.join(COMPAT, src), "-L" + HERE, "-lth2d",
               "-Wl,-rpath,$ORIGIN", "-lm"]
        _run(cmd, verbose)
        os.replace(out + ".tmp", out)
    return OUT


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
