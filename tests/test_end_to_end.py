"""End-to-end parity with the reference's own results.

tests/golden/cone_m1/ex2.dat (analysis_sample2/) is the reference's only shipped result with
linear triangles: a cone of 553 quads and 28 triangles (n = 0.06 + 4i) on a hexagonal
0.75 um lattice, rotated by pi about x.
tests/golden/sample3/ex3.dat and tests/golden/verification/veri.dat are result files the
reference ships (analysis_sample3/, verification/): a 384-quad cube-sphere (r = 0.125 um,
n = 1.6) in water (n0 = 1.33) on hexagonal lattices of 0.75 and 2.0 um, lambda0 = 1 um,
theta = 0.2.  tests/golden/sphere_m2/ex.dat is the reference's top-level result: a two-layer
sphere (1086 quads; shell n = 1.6, metallic core n = 0.06 + 4i, so MN = 2 and three domains)
on a 0.75 um square lattice.  They hold the model and the reference's surface solution
U, dU, E, dE, H, dH.
The GPU solve of the same model must reproduce them to 1e-10 (max |diff| / max |ref| per
array); the oracle reproduces ex3.dat to 4e-12 (tests/oracle_pipeline.py, slow test below).
example1_out.txt (the output example1 printed, as shipped in example1.log) holds EH_t at r = (0.3, 0, 1.1) for the 4- and 9-point rules.
"""
import re
from pathlib import Path

import numpy as np
import pytest

from datfile import read_dat

G = Path(__file__).resolve().parent / "golden"
TOL = 1e-10
# cone_m1: the cone's flat base was rotated by pi about x, which left its node z at
# 0.125 +- 1e-11, and a few Gauss points of base pairs land at |R_z| = 2^-56 (one ulp of
# 0.125).  There the reference's d3hm_qpgf_d2_v1 stops its spectral sum on a comparison of
# erfc(b v + a) and erfc(b v - a) that is decided by libm rounding (its value is then neither
# the truncated r_z == 0 one nor the converged one; tests/golden/qpgf_ref.npz and
# oracle/_ref show the band 2^-58 < |R_z| < 2^-52).  The library takes the converged value
# there.  The reference's own algorithm restated in C (oracle/) differs from ex2.dat by up to
# 6e-10 for the same reason; the GPU solution differs by up to 9.6e-10 (H), 1.5e-10 (U).
TOL_CASE = {"cone_m1/ex2.dat": 2e-9}
KEYS = ["U", "dU", "E", "dE", "H", "dH"]


def _example1(path):
    txt = Path(path).read_text()
    out = {}
    for block in re.split(r"type=", txt)[1:3]:
        t = int(block[0])
        vals = re.findall(r"= ([ -]?[0-9.]+e[+-]\d+) ([+-][0-9.]+e[+-]\d+) I", block)
        out[t] = np.array([complex(float(a), float(b)) for a, b in vals[:6]])
    return out


def _compare(sol, ref, tol):
    worst = {}
    for k in KEYS:
        for d in range(len(ref["dom"])):
            a = sol["dom"][d][k][1:]
            b = ref["dom"][d][k][1:]
            worst[f"{k}{d}"] = np.abs(a - b).max() / np.abs(b).max()
    print({k: f"{v:.2e}" for k, v in worst.items()})
    bad = {k: v for k, v in worst.items() if not v < tol}
    assert not bad, bad


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sample3/ex3.dat", "verification/veri.dat", "sphere_m2/ex.dat", "cone_m1/ex2.dat"])
def test_gpu_solve_matches_reference_dat(tmp_path, gpu_lib, case):
    from bem3_emf_qd2_amd.solver import Model
    ref = read_dat(G / case)
    m = Model.from_dat(G / case)
    st = m.solve()
    m.write(tmp_path / "gpu.dat")
    m.close()
    print(case, f"assembly {st.t_assemble_ms:.1f} ms solve {st.t_solve_ms:.1f} ms gh pairs {st.n_gh_pairs}")
    sol = read_dat(tmp_path / "gpu.dat")
    _compare(sol, ref, TOL_CASE.get(case, TOL))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sample3/ex3.dat", "sphere_m2/ex.dat", "cone_m1/ex2.dat"])
def test_gpu_mixed_precision_solve_matches_reference_dat(tmp_path, gpu_lib, case):
    """The global system solved with FP32 LU factors and FP64 iterative refinement reaches
    the same 1e-10 agreement with the reference's results as the FP64 LU, and agrees with
    the FP64 LU solution itself to 1e-10.  (The refined solution is the closer one to the
    reference, whose PARDISO solve also refines: on sphere_m2 dE1 differs from it by 3.9e-12
    against the plain FP64 LU's 1.5e-11.)"""
    from bem3_emf_qd2_amd.solver import Model
    ref = read_dat(G / case)
    sols = {}
    for mode in ("fp64", "mixed"):
        gpu_lib.solve_precision(mode)
        try:
            m = Model.from_dat(G / case)
            st = m.solve()
            m.write(tmp_path / f"{mode}.dat")
            m.close()
        finally:
            gpu_lib.solve_precision("fp64")
        print(case, mode, f"solve {st.t_solve_ms:.1f} ms mixed {st.mixed} it {st.ir_iterations} "
              f"corr {st.ir_correction:.2e}")
        assert st.mixed == (mode == "mixed")
        sols[mode] = read_dat(tmp_path / f"{mode}.dat")
    assert gpu_lib.solve_precision() == "fp64"
    _compare(sols["mixed"], ref, TOL_CASE.get(case, TOL))
    _compare(sols["mixed"], sols["fp64"], TOL)


@pytest.mark.gpu
def test_gpu_auto_precision_keeps_small_systems_in_fp64(gpu_lib):
    """'auto' takes the FP64 LU for a block system of order 2 NN = 3072 (below 16384)."""
    from bem3_emf_qd2_amd.solver import Model
    gpu_lib.solve_precision("auto")
    try:
        m = Model.from_dat(G / "sample3" / "ex3.dat")
        st = m.solve()
        mixed = st.mixed
        m.close()
    finally:
        gpu_lib.solve_precision("fp64")
    assert mixed == 0


@pytest.mark.gpu
def test_gpu_fields_follow_a_new_solve(gpu_lib):
    """Fields evaluated after solve_bieq use the new solution, not the device copy of the
    solution the model held before (the arrays are rewritten in place): a model read with the
    reference's solution, evaluated, then solved again must give exactly the fields of a
    freshly read and solved model."""
    from bem3_emf_qd2_amd.solver import Model
    pts = [[0.3, 0.0, 1.1], [0.1, 0.2, -0.9]]
    m = Model.from_dat(G / "sample3" / "ex3.dat")
    E_ref, H_ref, _ = m.fields(pts, 1)
    m.solve()
    E_re, H_re, _ = m.fields(pts, 1)
    m.close()
    m2 = Model.from_dat(G / "sample3" / "ex3.dat")
    m2.solve()
    E_new, H_new, _ = m2.fields(pts, 1)
    m2.close()
    assert not np.array_equal(E_ref, E_new)  # the two solutions differ in the last digits
    assert np.array_equal(E_re, E_new) and np.array_equal(H_re, H_new)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sample3", "verification", "sphere_m2", "cone_m1"])
def test_gpu_fields_on_reference_solution(gpu_lib, case):
    """EH_t on the reference's own solution vs the output of the reference example1 (15 digits)."""
    from bem3_emf_qd2_amd.solver import Model
    dat = G / case / {"sample3": "ex3.dat", "verification": "veri.dat", "sphere_m2": "ex.dat",
                      "cone_m1": "ex2.dat"}[case]
    exp = _example1(G / case / "example1_out.txt")
    m = Model.from_dat(dat)
    for t in (0, 1):
        E, H, did = m.fields([[0.3, 0.0, 1.1]], t)
        got = np.r_[E[0], H[0]]
        scale = max(np.abs(exp[t][:3]).max(), 1e-300)
        eE = np.abs(got[:3] - exp[t][:3]).max() / np.abs(exp[t][:3]).max()
        eH = np.abs(got[3:] - exp[t][3:]).max() / np.abs(exp[t][3:]).max()
        print(case, "type", t, "E rel", eE, "H rel", eH)
        assert did[0] == 0
        assert eE < 1e-12 and eH < 1e-12
    m.close()


@pytest.mark.slow
def test_oracle_pipeline_matches_ex3():
    """The CPU restatement end to end against ex3.dat (~7 min on 8 cores; BQ_SLOW=1)."""
    from oracle_pipeline import OracleModel
    ref = read_dat(G / "sample3" / "ex3.dat")
    o = OracleModel.from_dat(ref)
    sol = o.solve()
    fake = {"dom": [{k: sol[k][d] for k in KEYS} for d in range(ref["MN"] + 1)]}
    _compare(fake, ref, 1e-11)


@pytest.mark.gpu
def test_gpu_mini_sphere_cone_vs_oracle(tmp_path, gpu_lib):
    """The benchmark scene in miniature: two-layer cube-sphere (2x2 quads per face) and a
    coarse linear-triangle cone with jittered vertices in one hexagonal cell, oblique TM
    incidence; four domains (open, shell, core, cone), quads and triangles in one open
    region.  GPU solve vs the oracle's complete solve of the same stored model."""
    from bem3_emf_qd2_amd import meshgen
    from bem3_emf_qd2_amd.solver import Model
    from oracle_pipeline import OracleModel
    meshgen.sphere_cone_scene(tmp_path, n=2, cone_h=0.2)
    m = Model.from_dir(tmp_path)
    m.write(tmp_path / "setup.dat")
    m.solve()
    m.write(tmp_path / "gpu.dat")
    m.close()
    d = read_dat(tmp_path / "setup.dat")
    assert d["MN"] == 3 and np.sum(d["ed"][1:, 3] == 0) > 40
    o = OracleModel.from_dat(d, stored_nodes=True)
    sol = o.solve(log=lambda *a: None)
    fake = {"dom": [{k: sol[k][dd] for k in KEYS} for dd in range(d["MN"] + 1)]}
    _compare(read_dat(tmp_path / "gpu.dat"), fake, TOL)


# ---- non-periodic mode (ps = 0): the open region with the free-space Green's function ----

def _nonperiodic_sphere(tmp_path, n):
    from bem3_emf_qd2_amd import meshgen
    from bem3_emf_qd2_amd.solver import Model
    meshgen.sphere_scene(tmp_path, n, 0.4, 1.5, 0.6, ps=0)
    m = Model.from_dir(tmp_path)
    m.write(tmp_path / "setup.dat")
    return m


MIE_POINTS = [[0.0, 0.0, 1.0], [0.5, 0.3, -0.8], [-0.6, 0.2, 0.3], [0.0, 0.0, -0.7]]


def _mie_error(fields):
    """Largest relative difference of total E and Z0 H from the Mie series over MIE_POINTS."""
    import mie
    eE = eH = 0.0
    for r, (E, H) in zip(MIE_POINTS, fields):
        Es, Hs = mie.scattered(r, 0.4, 1.5, 1.0, 0.6)
        Ei, Hi = mie.incident(np.array(r), 1.0, 0.6)
        eE = max(eE, np.abs(E - Es - Ei).max() / np.abs(Es + Ei).max())
        eH = max(eH, np.abs(H - Hs - Hi).max() / np.abs(Hs + Hi).max())
    return eE, eH


def test_oracle_nonperiodic_sphere_converges_to_mie(tmp_path):
    """A lone dielectric sphere (r 0.4 um, n 1.5, lambda0 0.6 um, ps = 0): the oracle's BIE
    solution approaches the Mie series at second order in the element size (measured E
    0.191, 0.086, 0.048 and H 0.101, 0.048, 0.028 with 4, 6, 8 quads per cube-face edge), an
    independent check of the whole formulation in the non-periodic mode."""
    from oracle_pipeline import OracleModel
    errs = []
    for n in (4, 6):
        m = _nonperiodic_sphere(tmp_path / f"n{n}", n)
        m.close()
        d = read_dat(tmp_path / f"n{n}" / "setup.dat")
        assert d["ps"] == 0
        o = OracleModel.from_dat(d, stored_nodes=True)
        sol = o.solve(log=lambda *a: None)
        s4 = {k: sol[k] for k in ("E", "dE", "H", "dH")}
        errs.append(_mie_error([o.EH_t(s4, r, 1)[1:] for r in MIE_POINTS]))
    print(errs)
    assert errs[1][0] < 0.1 and errs[1][1] < 0.06
    assert 1.8 < errs[0][0] / errs[1][0] < 2.8  # (6/4)^2 = 2.25


@pytest.mark.gpu
def test_gpu_nonperiodic_sphere_vs_oracle_and_mie(tmp_path, gpu_lib):
    """ps = 0 end to end on the GPU: the solution matches the oracle's complete solve of the
    same stored model (1e-10) and the fields approach the Mie series as the oracle's do."""
    from oracle_pipeline import OracleModel
    m = _nonperiodic_sphere(tmp_path, 6)
    m.solve()
    m.write(tmp_path / "gpu.dat")
    E, H, did = m.fields(MIE_POINTS, 1)
    m.close()
    d = read_dat(tmp_path / "setup.dat")
    o = OracleModel.from_dat(d, stored_nodes=True)
    sol = o.solve(log=lambda *a: None)
    fake = {"dom": [{k: sol[k][dd] for k in KEYS} for dd in range(d["MN"] + 1)]}
    _compare(read_dat(tmp_path / "gpu.dat"), fake, TOL)
    assert list(did) == [0, 0, 0, 0]
    eE, eH = _mie_error(list(zip(E, H)))
    print("GPU vs Mie (6 quads per edge): E", eE, "H", eH)
    assert eE < 0.1 and eH < 0.06
