"""Field evaluation at arbitrary points (EH_t, d3qd2_src/d3qd2_field.c:203-297) on the GPU
against the oracle's restatement, on the reference's own ex2.dat solution (553 bilinear quads
+ 28 linear triangles), for all four rule types: the element nodes (a triangle's three nodes
and centroid), the 9-/7-point rules, and the 12- and 64-point Gauss rules (n^2 points on a
quad, 3 n^2 on a triangle).  Points lie in the open periodic region (one of them in a
neighbouring lattice cell, so the Bloch phase of the folding is exercised) and inside the
cone (free-space kernel with complex k).

Tolerance 1e-11 relative: both sides are FP64 with the same rules; the difference is
summation order and lattice-sum truncation."""
from pathlib import Path

import numpy as np
import pytest

from datfile import read_dat

G = Path(__file__).resolve().parent / "golden"
TOL = 1e-11


@pytest.mark.gpu
def test_gpu_fields_vs_oracle_all_rules(gpu_lib):
    from bem3_emf_qd2_amd.solver import Model
    from oracle_pipeline import OracleModel
    dat = G / "cone_m1" / "ex2.dat"
    d = read_dat(dat)
    o = OracleModel.from_dat(d)
    sol = {k: [np.ascontiguousarray(dd[k]) for dd in d["dom"]] for k in ("E", "dE", "H", "dH")}
    inside = d["rn"][1:].mean(0)  # centroid of the cone's nodes
    pts0 = [[0.3, 0.0, 1.1], [0.1, 0.2, -0.9], [0.3 + 0.75, 0.0, 1.1]]
    pts1 = [list(inside)]
    m = Model.from_dat(dat)
    try:
        for t in (0, 1, 2, 3):
            pts = pts0 + pts1 if t < 3 else pts0[:1] + pts1
            E, H, did = m.fields(pts, t)
            for i, p in enumerate(pts):
                odid, oE, oH = o.EH_t(sol, p, t)
                assert did[i] == odid, (t, p, did[i], odid)
                eE = np.abs(E[i] - oE).max() / np.abs(oE).max()
                eH = np.abs(H[i] - oH).max() / np.abs(oH).max()
                print(f"type {t} point {p} domain {odid}: E rel {eE:.2e} H rel {eH:.2e}")
                assert eE < TOL and eH < TOL
        assert 1 in did
    finally:
        m.close()
