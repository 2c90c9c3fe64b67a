"""The C-ABI library loads without a GPU and exports every symbol include/zkgpu.h declares;
the reference hot-path symbols keep the reference's names.  No compute calls are made here."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / "include" / "zkgpu.h"
LIB = REPO / "zikkurat-algebra_amd" / "libzkgpu.so"

HOT_PATH_HEADERS = [
    "lib/cbits/curves/g1/proj/bn128_G1_proj.h", "lib/cbits/curves/g1/proj/bls12_381_G1_proj.h",
    "lib/cbits/curves/g1/jac/bn128_G1_jac.h", "lib/cbits/curves/g1/jac/bls12_381_G1_jac.h",
    "lib/cbits/curves/poly/mont/bn128_poly_mont.h", "lib/cbits/curves/poly/mont/bls12_381_poly_mont.h",
]
# hot-path entry points of those headers (MSM with std/mont scalars, NTT family)
HOT_RE = re.compile(r"\b((?:bn128|bls12_381)_(?:G1_(?:proj|jac)_MSM_(?:std|mont)_coeff_\w+_out"
                    r"|poly_mont_ntt_\w+))\s*\(")


def declared(text):
    body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:zkg|bn128|bls12_381)_\w+)\s*\(", body)))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


@pytest.fixture(scope="module")
def built():
    if not LIB.exists():
        subprocess.run(["python3", str(REPO / "zikkurat-algebra_amd" / "build.py")], check=True)
    return LIB


def test_library_loads_without_gpu(built):
    ctypes.CDLL(str(built))


def test_every_declared_symbol_is_exported(built):
    decl = declared(HEADER.read_text())
    assert len(decl) > 40
    missing = [s for s in decl if s not in exported()]
    assert not missing, missing


def test_reference_hot_path_symbols_present(built):
    ref = Path("/root/reference")
    if not ref.is_dir():
        pytest.skip("reference sources not mounted")
    names = set()
    for h in HOT_PATH_HEADERS:
        names |= set(HOT_RE.findall((ref / h).read_text()))
    assert len(names) == 2 * (8 + 5), sorted(names)
    ours = exported()
    assert not sorted(names - ours)
