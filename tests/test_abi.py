"""The C-ABI: the built library loads on a machine without a GPU and exports every function
that include/*.h declares (the drop-in API of d3qd2_src/bem3_emf_qd2.h:17-76, the QPGF and
Faddeeva entry points of qgf2_src/d3_qpgf_d2.h:101-126 and qgf2_src/Faddeeva.h:39, and the
batched entry points).  No compute call is made here."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def _declared(path):
    txt = Path(path).read_text()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    txt = re.sub(r"#[^\n]*", " ", txt)
    txt = re.sub(r"typedef\s+struct\s*\w*\s*\{.*?\}\s*\w+\s*;", " ", txt, flags=re.S)
    names = []
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(([^;{}]*)\)\s*;", txt):
        if m.group(2) not in ("if", "while", "for", "return", "sizeof"):
            names.append(m.group(2))
    return names


def _exported():
    from bem3_emf_qd2_amd import LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True, check=True)
    return {ln.split()[-1] for ln in out.stdout.splitlines() if " T " in ln}


def test_headers_declare_functions():
    for h in HEADERS:
        assert len(_declared(h)) > 5, h


@pytest.mark.parametrize("header", [h.name for h in HEADERS])
def test_library_exports_every_declared_function(header):
    from bem3_emf_qd2_amd import lib
    L = lib()  # loads without a GPU
    exported = _exported()
    missing = [n for n in _declared(ROOT / "include" / header) if n not in exported]
    assert not missing, missing
    for n in _declared(ROOT / "include" / header):
        assert isinstance(getattr(L, n), ctypes._CFuncPtr)
