"""The C-ABI library loads and exports every function and global declared in
include/saruman_hip.h (no GPU needed: nothing here launches work)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "saruman_hip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    funcs = set()
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M):
        name = m.group(1)
        if name in ("if", "return", "sizeof") or name.startswith("SRM_"):
            continue
        funcs.add(name)
    funcs.discard("srm_ops_words")   # static inline in the header
    globs = set(re.findall(r"^extern\s+[\w \*]+?\b(\w+)(?:\[\d+\])?\s*;", text, flags=re.M))
    return funcs, globs


@pytest.fixture(scope="module")
def lib():
    from saruman_amd import _lib
    return _lib.load()


def test_header_parse():
    funcs, globs = declared()
    for f in ("cuda_needleman_wunsch", "cuda_allocate_memory", "cuda_free_memory",
              "get_cuda_infos", "srm_job_upload", "srm_format_record"):
        assert f in funcs
    for g in ("read_names", "forward_output_control", "alignment_time", "main_alloc_ptr_rev"):
        assert g in globs


def test_exports(lib):
    funcs, globs = declared()
    nm = subprocess.run(["nm", "-D", "--defined-only", str(lib._name)], capture_output=True,
                        text=True, check=True).stdout
    syms = {line.split()[-1] for line in nm.splitlines() if line.strip()}
    missing = sorted((funcs | globs) - syms)
    assert not missing, missing
    for f in funcs:
        assert getattr(lib, f) is not None


def test_python_bindings_cover_header():
    from saruman_amd import _lib
    funcs, globs = declared()
    assert funcs <= set(_lib.EXPORTED), sorted(funcs - set(_lib.EXPORTED))
    assert globs <= set(_lib.GLOBALS)


def test_version(lib):
    assert b"gfx950" in lib.srm_version()


def test_globals_defaults(lib):
    # util.c:192-194: use_gpu_value defaults to 5000 (automatic choice)
    assert C.c_ushort.in_dll(lib, "use_gpu_value").value == 5000
    assert C.c_uint.in_dll(lib, "forward_alignments_done").value == 0


def test_product_has_no_oracle_dependency():
    """The product path must never import, load or call the CPU oracle."""
    pat = re.compile(r"import\s+oracle|from\s+oracle|liboracle|oracle_\w+\s*\(|oracle\.")
    for p in (ROOT / "saruman_amd").rglob("*"):
        if p.suffix in (".py", ".hip", ".cpp", ".h", ".c"):
            assert not pat.search(p.read_text(errors="replace")), p
