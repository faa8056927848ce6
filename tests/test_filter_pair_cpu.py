"""CPU build of the device filter's per-read code (saruman_amd/csrc/filter_pair.h)
under AddressSanitizer and UBSan: the count / scan / fill / pairing steps of the
kernels in filter_gpu.hip, run on the host with exactly-sized buffers, give the
host filter's candidates, batches and perfect hits (tests/cpu_build/filter_pair_cpu.cpp).
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpu_build" / "filter_pair_cpu.cpp"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    from saruman_amd import _build
    if not _build.LIB.exists():
        pytest.skip("libsaruman_hip.so is not built")
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("cpu_build") / "filter_pair_cpu"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I", str(ROOT / "include"), "-I", str(ROOT / "saruman_amd" / "csrc"), str(SRC),
           "-L", str(_build.LIBDIR), "-lsaruman_hip", f"-Wl,-rpath,{_build.LIBDIR}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("args", [
    # seed, genome, reads, Lmin, Lmax, chunk, repeats, non-ACGT bytes
    ("11", "30000", "400", "36", "36", "60", "0", "0"),       # uniform, small chunks
    ("12", "40000", "3000", "40", "40", "7", "1", "0"),       # repeats: HBM-slot passes, spills
    ("13", "40000", "3000", "40", "40", "100000", "1", "0"),  # one chunk
    ("14", "60000", "3000", "36", "120", "50", "0", "1"),     # N runs, other letters
    ("15", "200000", "20000", "50", "250", "5000", "1", "1"),  # mixed lengths, everything
], ids=["uniform", "repeats_tiny_chunks", "repeats_one_chunk", "non_acgt", "mixed"])
def test_device_filter_code_on_cpu(exe, args):
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK"), r.stdout
    if args[6] == "1":
        assert "overflow_reads=0" not in r.stdout    # the HBM-slot passes ran
