"""Builds libwignerhip.so in-tree with hipcc for gfx950.

Usage: python -m pywignercuda_amd._build [-v]
Objects go to pywignercuda_amd/_build/, the library to pywignercuda_amd/libwignerhip.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libwignerhip.so"
SOURCES = ["wh_sweeps_f64.hip", "wh_sweeps_f32.hip", "wh_sweeps_f32w.hip", "wh_misc.hip", "wh_capi.cpp", "wh_adjoint.hip", "wh_dirac.hip",
           "wh_odm.hip", "wh_bloch.hip", "wh_noalloc.hip", "wh_expr.hip"]
ARCH = os.environ.get("WH_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        # packed FP32 instructions are disabled for all HIP code of this library
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
    ]


def build(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    cc = hipcc()
    procs, objs = [], []
    for src in SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = OUT / (src + ".o")
        objs.append(obj)
        cmd = [cc, *flags(), "-c", str(path), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        msg = "\n".join(l for l in out.splitlines() if "unknown target feature" not in l
                        and "packed-fp32-ops" not in l and l.strip())
        if p.returncode != 0:
            failed.append((src, msg))
        elif verbose and msg:
            print(msg)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(f"[{s}]\n{m}" for s, m in failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
