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


def _read_map(path):
    """A gnuplot block file of example2: header line, then 'a b value' rows, blank line per row."""
    lines = Path(path).read_text().splitlines()
    rows = [list(map(float, l.split())) for l in lines[1:] if l.strip()]
    return lines[0], np.array(rows)


@pytest.mark.gpu
def test_gpu_example2_field_maps(tmp_path, gpu_lib):
    """The example2 tool (bem3_emf_qd2_amd/bin/example2, the reference's example2.c:11-105)
    writes the intensity maps of the three planes; each value is |E|^2 or |H|^2 of EH_t at
    that point (type 1), as the batched evaluation gives it point by point."""
    import subprocess
    from bem3_emf_qd2_amd.solver import Model
    exe = Path(__file__).resolve().parent.parent / "bem3_emf_qd2_amd" / "bin" / "example2"
    dat = Path(__file__).resolve().parent / "golden" / "sample3" / "ex3.dat"
    mx = 9
    p = subprocess.run([str(exe), str(dat), str(mx)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "Intensity plot is finished" in p.stdout
    m = Model.from_dat(dat)
    lam = read_dat(dat)["lambda_0"]
    for plane, (fixed, a, b) in {"yz": (0, 1, 2), "xz": (1, 0, 2), "xy": (2, 0, 1)}.items():
        he, Ie = _read_map(tmp_path / f"Ie_{plane}.txt")
        hh, Ih = _read_map(tmp_path / f"Ih_{plane}.txt")
        assert he == f"# {plane[0]} {plane[1]} electric_field_intensity"
        assert hh == f"# {plane[0]} {plane[1]} magnetic_field_intensity"
        assert Ie.shape == (mx * mx, 3) and Ih.shape == (mx * mx, 3)
        g = -lam + np.arange(mx) * (2 * lam / (mx - 1))
        r = np.zeros((mx * mx, 3))
        r[:, a] = np.repeat(g, mx)
        r[:, b] = np.tile(g, mx)
        assert np.allclose(Ie[:, 0], r[:, a], atol=1e-5) and np.allclose(Ie[:, 1], r[:, b], atol=1e-5)
        E, H, _ = m.fields(r, 1)
        ie = (np.abs(E) ** 2).sum(1)
        ih = (np.abs(H) ** 2).sum(1)
        assert np.abs(Ie[:, 2] - ie).max() <= 1e-13 * ie.max()
        assert np.abs(Ih[:, 2] - ih).max() <= 1e-13 * ih.max()
    m.close()
