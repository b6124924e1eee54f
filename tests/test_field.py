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


def _read_png(path):
    """Minimal PNG reader for the example3 images (8-bit RGB, filter 0 rows): checks every
    chunk's CRC and returns the pixel array [h][w][3]."""
    import struct
    import zlib
    b = Path(path).read_bytes()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    o, idat, w, h = 8, b"", None, None
    while o < len(b):
        n = struct.unpack(">I", b[o:o + 4])[0]
        typ, data = b[o + 4:o + 8], b[o + 8:o + 8 + n]
        assert struct.unpack(">I", b[o + 8 + n:o + 12 + n])[0] == zlib.crc32(typ + data)
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", data[:10])
            assert depth == 8 and ctype == 2
        elif typ == b"IDAT":
            idat += data
        o += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 3 * w + 1)
    assert (raw[:, 0] == 0).all()
    return raw[:, 1:].reshape(h, w, 3)


_CT1 = np.array([[0x00, 0x00, 0x90], [0x00, 0x0f, 0xff], [0x00, 0x90, 0xff], [0x0f, 0xff, 0xee], [0xff, 0xff, 0xff],
                 [0xff, 0xee, 0x00], [0xff, 0x70, 0x00], [0xee, 0x00, 0x00], [0x7f, 0x00, 0x00]])


def _color(x):
    """The reference's colour scale (example3.c:403-439): 2^23 steps over [-1, 1]."""
    if not -1.0 <= x <= 1.0:
        return np.zeros(3, int)
    n = int(np.floor(2.0 ** 23 * (x + 1.0)))
    nc = 2 ** 21
    i = n // nc + 1
    if i > 8:
        return _CT1[8]
    nd = n - nc * (i - 1)
    return np.array([int(np.floor(_CT1[i - 1][c] + (_CT1[i][c] - _CT1[i - 1][c]) * (1.0 / nc) * nd))
                     for c in range(3)])


@pytest.mark.gpu
def test_gpu_example3_field_images(tmp_path, gpu_lib):
    """The example3 tool (the reference's example3.c) writes, per plane and phase, the real
    part of each E and H component times exp(-i omega t) as PNG images in the reference's
    colour scale, normalised by the component's largest modulus, plus the range files and
    the colour bar.  The images are decoded here and compared pixel by pixel with the scale
    applied to EH_t at the same points."""
    import shutil
    import subprocess
    from bem3_emf_qd2_amd.solver import Model
    exe = Path(__file__).resolve().parent.parent / "bem3_emf_qd2_amd" / "bin" / "example3"
    dat = tmp_path / "ex3.dat"
    shutil.copy(Path(__file__).resolve().parent / "golden" / "sample3" / "ex3.dat", dat)
    mx, ts = 7, 3
    p = subprocess.run([str(exe), str(dat), str(mx), str(ts)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = tmp_path / "ex3_images"
    assert len(list(out.glob("*.png"))) == 3 * ts * 6 + 1
    d = read_dat(dat)
    lam, omega = d["lambda_0"], d["omega"]
    m = Model.from_dat(dat)
    g = np.arange(mx) * (2 * lam / (mx - 1))
    for pl, (fixed, row, col) in {"yz": (0, 2, 1), "xz": (1, 2, 0), "xy": (2, 1, 0)}.items():
        r = np.zeros((mx * mx, 3))
        r[:, row] = np.repeat(lam - g, mx)
        r[:, col] = np.tile(-lam + g, mx)
        E, H, _ = m.fields(r, 1)
        info = (out / f"{pl}_info.txt").read_text().splitlines()
        for f, V in (("E", E), ("H", H)):
            mxv = np.abs(V).max(0)
            for dd, cn in enumerate("xyz"):
                line = next(l for l in info if l.startswith(f"{f}{cn} is"))
                assert abs(float(line.split()[-1]) - mxv[dd]) <= 1e-6 * mxv[dd]
                for nt in (0, ts - 1):
                    img = _read_png(out / f"{pl}_{f}{cn}_{nt:03d}.png")
                    assert img.shape == (2 * mx, 2 * mx, 3)
                    cet = np.exp(-1j * omega * (lam / ts) * nt)
                    want = np.array([_color((cet * V[k, dd]).real / mxv[dd]) for k in range(mx * mx)])
                    got = img[::2, ::2].reshape(-1, 3)
                    # a value within rounding of a colour step may land on either side of it
                    assert np.abs(got.astype(int) - want).max() <= 1
                    assert (img[::2, ::2] == img[1::2, 1::2]).all()
    bar = _read_png(out / "color_bar.png")
    assert bar.shape == (2 * mx, max(2 * mx // 16, 1), 3)
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sample3/ex3.dat", "sphere_m2/ex.dat", "cone_m1/ex2.dat"])
def test_gpu_domain_ids_match_host(gpu_lib, case):
    """bq_domain_id_batch (one device launch, signed solid angles summed per domain by a
    workgroup) against the host's q0_domain_id (the reference's fid_calc_solid_angle walk,
    d3qd2_src/d3qd2_setup.c:157-245,1351-1422) for random points in the cell and points just
    off the surface on both sides of element nodes: identical ids."""
    import ctypes
    from datfile import read_dat
    from bem3_emf_qd2_amd.solver import Model, L as SL
    G = Path(__file__).resolve().parent / "golden"
    d = read_dat(G / case)
    m = Model.from_dat(G / case)
    lb = SL()
    D = ctypes.POINTER(ctypes.c_double)
    lb.q0_domain_id.argtypes = [D, ctypes.c_void_p]
    lb.bq_domain_id_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, D, ctypes.POINTER(ctypes.c_int)]
    rng = np.random.default_rng(3)
    qd = d["qd"]
    d1, d2 = qd[4:6], qd[6:8]
    a = rng.uniform(-0.5, 0.5, (3000, 2))
    pts = np.c_[a[:, :1] * d1 + a[:, 1:] * d2, rng.uniform(-0.6, 0.6, 3000)]
    ren, wen = d["ren"][1:].reshape(-1, 3), d["wen"][1:].reshape(-1, 3)
    live = np.linalg.norm(wen, axis=1) > 0
    nrm = wen[live] / np.linalg.norm(wen[live], axis=1, keepdims=True)
    sel = rng.choice(len(nrm), 500, replace=False)
    near = np.r_[ren[live][sel] + 2e-3 * nrm[sel], ren[live][sel] - 2e-3 * nrm[sel]]
    pts = np.ascontiguousarray(np.r_[pts, near])
    got = np.zeros(len(pts), np.int32)
    assert lb.bq_domain_id_batch(m.ptr, len(pts), pts.ctypes.data_as(D), got.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
    host = np.array([lb.q0_domain_id(np.ascontiguousarray(p).ctypes.data_as(D), m.ptr) for p in pts])
    m.close()
    print(case, "ids", np.bincount(host + 1), "mismatches", int(np.sum(got != host)))
    assert np.array_equal(got, host)
