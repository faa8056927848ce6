"""CPU build of the record writer's word-level transcript scans
(saruman_amd/csrc/rec_ops.h: leading and closing gaps, the indel scan and its
end-of-row count) against the reference's op-by-op loops
(src/cuda-functions.cu:866-980) restated one op at a time in
tests/cpu_build/rec_measure_cpu.cpp, under AddressSanitizer and UBSan."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpu_build" / "rec_measure_cpu.cpp"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("rec_measure") / "rec_measure_cpu"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", str(ROOT / "saruman_amd" / "csrc"), str(SRC),
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_word_scans_equal_op_scans(exe, seed):
    r = subprocess.run([str(exe), str(seed), "100000"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
