"""The unit-cost band kernels (saruman_amd/csrc/band_bits.hip: bits_band_kernel and
bits_trace_kernel, in both scratch layouts: stored D bits, and the checkpoints
from which the walk recomputes each 32-column segment), compiled unmodified for
the CPU through a stand-in HIP header
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
    """Runs the emulation; returns (exit status, {layout: tallies}) for the two
    scratch layouts (stored D bits, checkpoints the walk recomputes from)."""
    args = [str(exe), str(n), str(seed)] + (["plant"] if plant else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    out = {}
    for m in re.finditer(r"(stored|checkpoint): cases (\d+) hits (\d+) bad (\d+) diag (\d+) "
                         r"top (\d+) left (\d+)", r.stdout):
        out[m.group(1)] = dict(zip(("cases", "hits", "bad", "diag", "top", "left"),
                                   map(int, m.groups()[1:])))
    assert set(out) == {"stored", "checkpoint"}, r.stdout + r.stderr
    return r.returncode, out


@pytest.mark.parametrize("seed", [1, 0x5A2F])
def test_bits_kernels_match_oracle(emu_bin, seed):
    rc, out = _run(emu_bin, 20000, seed)
    assert rc == 0, out
    for layout, t in out.items():
        assert t["bad"] == 0, (layout, t)
        # the cases reach every kind of step
        assert t["hits"] > 5000 and t["top"] > 1000 and t["left"] > 1000, (layout, t)


def test_planted_mistake_is_reported(emu_bin):
    rc, out = _run(emu_bin, 4000, 3, plant=True)
    assert rc == 1 and out["stored"]["bad"] > 0 and out["checkpoint"]["bad"] > 0, out
