"""The oracle's element integrals against the reference's own solved models.

The result files the reference ships hold its surface solution (U, dU, E, dE, H, dH per
domain).  That solution satisfies, to the accuracy of the reference's sparse solve, every
row of the systems the reference assembled:

  * H_d U_d - G_d dU_d = 0 for each domain d (the BIE rows of create_matrix_csr_dac,
    d3qd2_src/d3qd2_solve_bieq.c:284-421, rewritten per domain after tmatrix_bd_store
    :456-544; domain 0 carries the scattered part);
  * G_d dE_d - H_d E_d = 0 (solve_deh_bv :662-791), the same for H;
  * the boundary E and H, which q1_solve_eh_bv (:931-1030) builds from U, dU and the
    tangential-derivative rows tdG, tdH, tdF.

So the oracle's rows for a few collocation nodes, applied to the reference's solution, must
give residuals at rounding level, and its derivative rows must rebuild the stored boundary
E and H.  Each row touches every source element of its domain, so a wrong factor, sign,
node or rule in any element's coefficients (regular, near-singular GH, singular polar or
principal value) leaves a residual of the size of that element's term: a 1e-3 error in
the triangle Jacobian alone gives residuals of 1e-4 here, against 3e-15 when right.

Cases: ex3.dat (384 bilinear quads, hexagonal 0.75 um), sphere_m2 ex.dat (two-layer
sphere, three domains) and cone_m1 ex2.dat (553 quads + 28 linear triangles, the only
shipped model with triangles; elements 1-28 are the triangles).  The triangle field rules
(4- and 7-point) are checked against the shipped example1 output for ex2.dat.
"""
from pathlib import Path

import numpy as np
import pytest

from datfile import read_dat
from oracle_lib import D, ptr
from oracle_pipeline import OracleModel

G = Path(__file__).resolve().parent / "golden"
ROW_TOL = 2e-14   # residual / sum of |terms|; measured <= 2.5e-15
EH_TOL = 1e-12    # rebuilt boundary E, H vs stored; measured <= 1.3e-13

# (dat file, domain, first element, last element) -- rows assembled in parallel per block
ROW_CASES = [
    ("sample3/ex3.dat", 0, 1, 2),
    ("sample3/ex3.dat", 1, 1, 8),
    ("sphere_m2/ex.dat", 1, 300, 303),
    ("sphere_m2/ex.dat", 2, 1, 4),
    ("cone_m1/ex2.dat", 0, 27, 30),
    ("cone_m1/ex2.dat", 1, 1, 32),
]

_models = {}


def _model(case):
    if case not in _models:
        d = read_dat(G / case)
        _models[case] = (d, OracleModel.from_dat(d))
    return _models[case]


def test_triangle_node_data_matches_reference():
    """Collocation nodes, normals and incident field at the triangle nodes and centroid
    (init_boundary_data d3qd2_src/d3qd2_setup.c:1309-1349) vs the arrays stored in ex2.dat."""
    d, o = _model("cone_m1/ex2.dat")
    ex = o.export()
    tri = d["ed"][1:, 3] == 0
    assert tri.sum() == 28
    assert np.abs(ex["ren"] - d["ren"]).max() < 1e-15
    assert np.abs(ex["wen"] - d["wen"]).max() < 1e-15
    assert np.abs(ex["Ui"] - d["Ui"]).max() < 1e-14 * np.abs(d["Ui"]).max()
    assert np.abs(ex["dUi"] - d["dUi"]).max() < 1e-14 * np.abs(d["dUi"]).max()
    assert o.NN == 4 * 553 + 3 * 28


@pytest.mark.parametrize("case,did,t0,t1", ROW_CASES)
def test_oracle_rows_annihilate_reference_solution(case, did, t0, t1):
    d, o = _model(case)
    Gm, Hm, *_ = o.assemble(did, t0, t1, deriv=False)
    dom = d["dom"][did]
    U, dU = dom["U"][1:].reshape(-1, 4), dom["dU"][1:].reshape(-1, 4)
    E, dE = dom["E"][1:].reshape(-1, 3), dom["dE"][1:].reshape(-1, 3)
    Hf, dHf = dom["H"][1:].reshape(-1, 3), dom["dH"][1:].reshape(-1, 3)
    live = o.live_rows(did)
    worst = 0.0
    for r in range(4 * (t0 - 1), 4 * t1):
        if not live[r]:
            assert not np.any(Hm[r]) and not np.any(Gm[r])
            continue
        g, h = Gm[r], Hm[r]
        for a, b, sgn in ((U, dU, 1), (E, dE, -1), (Hf, dHf, -1)):
            res = sgn * (h @ a - g @ b)
            scale = (np.abs(h)[:, None] * np.abs(a) + np.abs(g)[:, None] * np.abs(b)).sum(0)
            worst = max(worst, (np.abs(res) / scale).max())
    print(case, did, t0, t1, f"worst residual {worst:.2e}")
    assert worst < ROW_TOL


def test_oracle_derivative_rows_rebuild_boundary_EH():
    """Free-space rows with tangential derivatives (the principal-value self terms of
    lit_dcoef_bd_tz0_PV / bil_dcoef_bd_tz0_PV among them) for triangles 1-2 and quads
    29-30 of ex2.dat rebuild E and H on both sides of those elements."""
    d, o = _model("cone_m1/ex2.dat")
    MN = d["MN"]
    rows = [1, 2, 29, 30]
    mats = []
    for did in range(MN + 1):
        N = 4 * o.dNe[did]
        dG = np.zeros((N, 2, N), complex); dH = np.zeros((N, 2, N), complex); dF = np.zeros((N, 3))
        if did >= 1:
            for t0, t1 in ((1, 2), (29, 30)):
                _, _, a, b, c = o.assemble(did, t0, t1)
                sl = slice(4 * (t0 - 1), 4 * t1)
                dG[sl], dH[sl], dF[sl] = a[sl], b[sl], c[sl]
        mats.append((dG, dH, dF))
    U = [np.ascontiguousarray(dd["U"]) for dd in d["dom"]]
    dU = [np.ascontiguousarray(dd["dU"]) for dd in d["dom"]]
    E = [np.zeros_like(dd["E"]) for dd in d["dom"]]
    H = [np.zeros_like(dd["H"]) for dd in d["dom"]]
    P = lambda L: (D * len(L))(*[ptr(x.view(float)) for x in L])
    with np.errstate(all="ignore"):  # rows not assembled divide by zero; they are not read
        o.o.or_boundary_EH(o.h, P(U), P(dU), P([m[0] for m in mats]), P([m[1] for m in mats]),
                           (D * (MN + 1))(*[ptr(m[2]) for m in mats]), P(E), P(H))
    sid0 = list(np.abs(o.sid(0)))
    sid1 = o.sid(1)
    worst = 0.0
    for t in rows:
        s = abs(sid1[t])
        nl = 3 if d["ed"][s, 3] == 0 else 4
        for key, arr in (("E", E), ("H", H)):
            for dd, idx in ((1, t), (0, sid0.index(s))):
                ref = d["dom"][dd][key][idx][:nl]
                worst = max(worst, np.abs(arr[dd][idx][:nl] - ref).max() / np.abs(ref).max())
    print(f"worst boundary E/H {worst:.2e}")
    assert worst < EH_TOL


def test_oracle_triangle_field_rules_match_example1():
    """EH_t at (0.3, 0, 1.1) with the 4- and 7-point triangle rules (q0_lit_coef_4p/_7p)
    on the reference's ex2.dat solution vs the example1 output shipped with it."""
    from test_end_to_end import _example1
    d, o = _model("cone_m1/ex2.dat")
    exp = _example1(G / "cone_m1" / "example1_out.txt")
    sol = {k: [np.ascontiguousarray(dd[k]) for dd in d["dom"]] for k in ("E", "dE", "H", "dH")}
    for t in (0, 1):
        did, E, H = o.EH_t(sol, [0.3, 0.0, 1.1], t)
        assert did == 0
        assert np.abs(E - exp[t][:3]).max() < 1e-13 * np.abs(exp[t][:3]).max()
        assert np.abs(H - exp[t][3:]).max() < 1e-13 * np.abs(exp[t][3:]).max()
