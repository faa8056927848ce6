"""Builds libsaruman_hip.so (hipcc, gfx950) in-tree.

The library is compiled from saruman_amd/csrc with hipcc directly; objects go to
saruman_amd/_build/ and the shared object to saruman_amd/lib/.  Compile steps run
in parallel, and only for sources newer than their object unless force=True.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libsaruman_hip.so"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["nw_kernels.hip", "band_bits.hip", "nw_runtime.hip", "records.cpp", "filter.cpp", "filter_gpu.hip",
               "records_gpu.hip", "synth.cpp", "serve.cpp", "ingest.cpp", "static_align.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # packed FP32 instructions stay off (they are not needed and are avoided here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_lib(force: bool = False, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    hipcc = _hipcc()
    headers = [p.stat().st_mtime for p in list(CSRC.glob("*.h")) + list(INCLUDE.glob("*.h"))]
    newest_header = max(headers) if headers else 0.0
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        sp = CSRC / src
        if not sp.exists():
            continue
        op = BUILD / (sp.stem + ".o")
        objs.append(op)
        stale = force or not op.exists() or op.stat().st_mtime < max(sp.stat().st_mtime,
                                                                     newest_header)
        if stale:
            lang = ["-x", "hip"] if sp.suffix == ".hip" else []
            cmd = [hipcc, *HIPCC_FLAGS, "-I", str(INCLUDE), "-I", str(CSRC), *lang, "-c",
                   str(sp), "-o", str(op)]
            jobs.append(cmd)
    workers = min(len(jobs), os.cpu_count() or 4) or 1
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out, file=sys.stderr)
    if force or jobs or not LIB.exists():
        tmp = LIB.with_suffix(".so.tmp")
        _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
              *[str(o) for o in objs], "-lpthread"])
        os.replace(tmp, LIB)
    return LIB


def build_all(force: bool = False) -> None:
    build_lib(force=force)


if __name__ == "__main__":
    build_all(force="--force" in sys.argv)
    print(LIB)
