"""The unit-cost band kernels (saruman_amd/csrc/band_bits.hip: bits_band_kernel and
bits_trace_kernel), compiled unmodified for the CPU through a stand-in HIP header
(tests/emu/hip/hip_runtime.h) and run lane by lane over random candidates of every
band class, against the oracle (oracle_fill and oracle_traceback, which restate
cuda-functions.cu:413-497 and :108-315): score, transcript and end cell of every
candidate.

Cases: k 0-14, flags 0-8, reads 1-900 bp (both read-code paths: staged in LDS up
to 250/600 codes, from global memory beyond), random reads, substituted copies,
copies with 1-bp indels and shifted copies, Q1 on one candidate in five.  A
mistake planted in the checker (the oracle's gap rows read the wrong way round)
must be reported.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EMU = ROOT / "tests" / "emu"


def _clang():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and Path(c).exists():
            return c
    return None


@pytest.fixture(scope="module")
def emu_bin(tmp_path_factory):
    clang = _clang()
    if clang is None:
        pytest.skip("clang++ not found")
    d = tmp_path_factory.mktemp("emu")
    obj = d / "oracle.o"
    subprocess.run(["gcc", "-O2", "-c", str(ROOT / "oracle" / "nw_oracle.c"), "-o", str(obj)],
                   check=True)
    exe = d / "band_bits_emu"
    subprocess.run([clang, "-O2", "-std=c++17", "-I", str(EMU), "-I", str(ROOT / "include"),
                    "-I", str(ROOT / "saruman_amd" / "csrc"), str(EMU / "band_bits_emu.cpp"),
                    str(obj), "-o", str(exe)], check=True)
    return exe


def _run(exe, n, seed, plant=False):
    args = [str(exe), str(n), str(seed)] + (["plant"] if plant else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    m = re.search(r"cases (\d+) hits (\d+) bad (\d+) diag (\d+) top (\d+) left (\d+)", r.stdout)
    assert m, r.stdout + r.stderr
    return r.returncode, dict(zip(("cases", "hits", "bad", "diag", "top", "left"),
                                  map(int, m.groups())))


@pytest.mark.parametrize("seed", [1, 0x5A2F])
def test_bits_kernels_match_oracle(emu_bin, seed):
    rc, t = _run(emu_bin, 20000, seed)
    assert rc == 0 and t["bad"] == 0, t
    # the cases reach every kind of step
    assert t["hits"] > 5000 and t["top"] > 1000 and t["left"] > 1000, t


def test_planted_mistake_is_reported(emu_bin):
    rc, t = _run(emu_bin, 4000, 3, plant=True)
    assert rc == 1 and t["bad"] > 0, t
