"""Extract the known-answer table of w(z) that the reference's Faddeeva package checks in its
own test program (qgf2_src/Faddeeva.cc:1916-2060, values from WolframAlpha/Maple) into
faddeeva_kat.json.  Only the finite entries are kept.  Also writes faddeeva_grid.npz:
w(z) on a grid covering the arguments the Ewald sums use, evaluated with mpmath at 40 digits
(ground truth from the definition w = exp(-z^2) erfc(-iz)).

Run in a container that holds /root/reference:  python tests/golden/make_faddeeva_kat.py
"""
import json
import math
import re
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = Path("/root/reference/qgf2_src/Faddeeva.cc")


def _num(s: str) -> float:
    s = s.strip()
    return {"Inf": math.inf, "-Inf": -math.inf, "NaN": math.nan}.get(s, None) if s in ("Inf", "-Inf", "NaN") else float(s)


def _pairs(block: str):
    body = re.sub(r"/\*.*?\*/", "", block, flags=re.S)
    return [(_num(a), _num(b)) for a, b in re.findall(r"C\(\s*([^,()]+?)\s*,\s*([^()]+?)\s*\)", body)]


def kat():
    text = SRC.read_text()
    start = text.index("printf(\"############# w(z) tests #############\\n\");")
    zb = text[text.index("cmplx z[NTST] = {", start):]
    zb = zb[: zb.index("};")]
    wb = text[text.index("cmplx w[NTST] = {", start):]
    wb = wb[: wb.index("};")]
    z, w = _pairs(zb), _pairs(wb)
    assert len(z) == len(w) == 57, (len(z), len(w))
    rows = [dict(z=list(a), w=list(b)) for a, b in zip(z, w)
            if all(map(math.isfinite, a + b))]
    (HERE / "faddeeva_kat.json").write_text(json.dumps(rows, indent=1))
    return len(rows)


def grid():
    import mpmath as mp
    mp.mp.dps = 40
    xs = np.r_[np.linspace(0.0, 8.0, 33)[1:], [0.01, 1.7320508, 2.95, 6.5, 12.0]]
    ys = np.r_[np.linspace(-7.5, 9.0, 34), [0.0, 1e-9, 7.2, 11.0]]
    pts = [(x, y) for x in xs for y in ys] + [(-x, y) for x in xs[::3] for y in ys[::3]]
    z = np.array(pts)
    w = np.array([complex(mp.exp(-mp.mpc(a, b) ** 2) * mp.erfc(-1j * mp.mpc(a, b))) for a, b in pts])
    np.savez_compressed(HERE / "faddeeva_grid.npz", z=z, w_re=w.real, w_im=w.imag)
    return len(pts)


if __name__ == "__main__":
    print("kat rows", kat())
    print("grid points", grid())
