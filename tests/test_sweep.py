"""Single-object reduction, resident sessions and incidence sweeps.

A model with one object (every element between the open region and domain 1) is solved by
the reduction to systems of order NN (csrc/dev/bq_reduced.hip): LUs of G0, G1, B = P + M and
the Schur complement S.  Its FP64 result is checked against the block elimination of order
2 NN that the multi-domain models take (BQ_BLOCK_ELIM=1 forces it), against the oracle's
complete solve (a restatement of the reference's sparse system, d3qd2_src/d3qd2_solve_bieq.c:
284-454), and the mixed-precision result (complex FP32 factors as the preconditioner of FP64
GMRES on the unreduced system) against the FP64 one.  The reference's own single-object
result files (ex3.dat, veri.dat) are solved through the same path by test_end_to_end.py.

bq_set_incidence replaces the wave of a set-up model; the resident session then reassembles
for the new wave, or reuses matrices and factors when only the polarisation changed.  Either
way the result must equal that of a model set up from input files for the new wave.
"""
import os
from pathlib import Path

import numpy as np
import pytest

from datfile import read_dat

KEYS = ["U", "dU", "E", "dE", "H", "dH"]
PROBES = [[0.1, 0.2, 0.9], [-0.3, 0.25, -0.8], [0.45, -0.1, 0.6]]


def _rel(a, b):
    worst = 0.0
    for k in KEYS:
        for d in range(len(b["dom"])):
            x, y = a["dom"][d][k][1:], b["dom"][d][k][1:]
            worst = max(worst, float(np.abs(x - y).max() / np.abs(y).max()))
    return worst


def _scene(d, n=3, lam=0.55, theta=0.3, phi=0.4, pol="TM", radius=0.45, n_sphere=1.5 + 0.001j, core=None):
    from bem3_emf_qd2_amd import meshgen
    meshgen.sphere_scene(d, n, radius, n_sphere, lam, theta=theta, phi=phi, pol=pol, core=core)
    return d


def _solve(d, tmp, name, precision="fp64", env=None, gpu_lib=None):
    from bem3_emf_qd2_amd.solver import Model
    old = {}
    for k, v in (env or {}).items():
        old[k] = os.environ.get(k)
        os.environ[k] = v
    gpu_lib.solve_precision(precision)
    try:
        m = Model.from_dir(d)
        st = m.solve()
        m.write(tmp / f"{name}.dat")
        E, H, did = m.fields(PROBES, 1)
        m.close()
    finally:
        gpu_lib.solve_precision("fp64")
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return st, read_dat(tmp / f"{name}.dat"), np.r_[E.ravel(), H.ravel()]


@pytest.mark.gpu
def test_gpu_reduced_equals_block_elimination(tmp_path, gpu_lib):
    """The reduction of order NN and the block elimination of order 2 NN solve the same
    linear system: their surface solutions agree to 1e-10 (rounding of two different
    eliminations)."""
    d = _scene(tmp_path / "s")
    st_r, red, fr = _solve(d, tmp_path, "red", gpu_lib=gpu_lib)
    st_b, blk, fb = _solve(d, tmp_path, "blk", env={"BQ_BLOCK_ELIM": "1"}, gpu_lib=gpu_lib)
    e = _rel(red, blk)
    ef = float(np.abs(fr - fb).max() / np.abs(fb).max())
    print(f"reduced vs block elimination: surface {e:.2e}, probes {ef:.2e}, backward error "
          f"{st_r.backward_error:.2e}, refinement steps {st_r.ir_iterations}")
    assert st_r.path == 1 and st_b.path == 0
    assert e < 1e-10 and ef < 1e-10
    assert st_r.backward_error < 1e-14


@pytest.mark.gpu
def test_gpu_reduced_vs_oracle(tmp_path, gpu_lib):
    """The 6144-quad sweep's object (sphere r 0.45 um, n 1.5 + 0.001i, square lattice 1.0
    um) at 54 quads, oblique TM incidence at 0.55 um: the FP64 reduction against the oracle's
    complete solve of the same stored model (1e-10), and the mixed-precision solve against
    the FP64 one (surface solution and probe fields 1e-8, the metric's bound; the backward
    error of the unreduced system 1e-13)."""
    from bem3_emf_qd2_amd.solver import Model
    from oracle_pipeline import OracleModel
    d = _scene(tmp_path / "s")
    m = Model.from_dir(d)
    m.write(tmp_path / "setup.dat")
    m.close()
    st, gpu, f64 = _solve(d, tmp_path, "fp64", gpu_lib=gpu_lib)
    stm, mix, fmx = _solve(d, tmp_path, "mixed", precision="mixed", gpu_lib=gpu_lib)
    dd = read_dat(tmp_path / "setup.dat")
    o = OracleModel.from_dat(dd, stored_nodes=True)
    sol = o.solve(log=lambda *a: None)
    fake = {"dom": [{k: sol[k][q] for k in KEYS} for q in range(dd["MN"] + 1)]}
    e_or = _rel(gpu, fake)
    e_mx = _rel(mix, gpu)
    e_fx = float(np.abs(fmx - f64).max() / np.abs(f64).max())
    print(f"FP64 vs oracle {e_or:.2e}; mixed vs FP64: surface {e_mx:.2e} probes {e_fx:.2e}, "
          f"GMRES steps {stm.gmres_iterations}, backward error {stm.backward_error:.2e}")
    assert st.path == 1 and stm.path == 1 and stm.mixed == 1
    assert e_or < 1e-10
    assert e_mx < 1e-8 and e_fx < 1e-8
    assert stm.backward_error < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("layers", [1, 2])
def test_gpu_set_incidence_equals_fresh_model(tmp_path, gpu_lib, layers):
    """A sweep on one resident model: after a solve at one wavelength and direction,
    bq_set_incidence moves it to another; the solution and probe fields equal those of a model
    set up from input files written for the second wave (the same arithmetic: 1e-13).  The
    single sphere takes the reduction, the two-layer sphere of the batched spectral sweep (shell
    n 1.45, core r 0.3 um n 1.6 + 0.01i) the block elimination."""
    from bem3_emf_qd2_amd.solver import Model
    kw = dict(n_sphere=1.45, core=(0.3, 1.6 + 0.01j)) if layers == 2 else {}
    d1 = _scene(tmp_path / "a", lam=0.55, theta=0.3, phi=0.4, **kw)
    d2 = _scene(tmp_path / "b", lam=0.62, theta=0.5, phi=1.1, **kw)
    m = Model.from_dir(d1)
    m.solve()
    m.set_incidence(0.62, 0.5, 1.1)
    st = m.solve()
    m.write(tmp_path / "swept.dat")
    E, H, _ = m.fields(PROBES, 1)
    nbytes = m.device_bytes()
    m.close()
    st2, fresh, f2 = _solve(d2, tmp_path, "fresh", gpu_lib=gpu_lib)
    e = _rel(read_dat(tmp_path / "swept.dat"), fresh)
    ef = float(np.abs(np.r_[E.ravel(), H.ravel()] - f2).max() / np.abs(f2).max())
    print(f"swept vs fresh: surface {e:.2e} probes {ef:.2e}; session {nbytes / 2**20:.1f} MiB")
    assert st.reused == 0 and nbytes > 0
    assert e < 1e-13 and ef < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_polarisation_change_reuses_factors(tmp_path, gpu_lib, precision):
    """Same wavelength and direction, TE after TM: the resident matrices and factors are
    reused (no assembly, no factorisation) and the result equals a fresh TE model's."""
    from bem3_emf_qd2_amd.solver import Model
    d1 = _scene(tmp_path / "tm", pol="TM")
    d2 = _scene(tmp_path / "te", pol="TE")
    gpu_lib.solve_precision(precision)
    try:
        m = Model.from_dir(d1)
        m.solve()
        m.set_incidence(0.55, 0.3, 0.4, e0x=0.0, e0y=1.0)
        st = m.solve()
        m.write(tmp_path / "reuse.dat")
        m.close()
    finally:
        gpu_lib.solve_precision("fp64")
    st2, fresh, _ = _solve(d2, tmp_path, "fresh", precision=precision, gpu_lib=gpu_lib)
    e = _rel(read_dat(tmp_path / "reuse.dat"), fresh)
    print(precision, f"reused {st.reused} factor {st.t_factor_ms:.2f} ms; vs fresh {e:.2e}")
    assert st.reused == 1 and st.t_assemble_ms == 0.0 and st.t_factor_ms == 0.0
    assert e < (1e-13 if precision == "fp64" else 1e-10)


@pytest.mark.gpu
def test_gpu_set_incidence_rejects_invalid(tmp_path, gpu_lib):
    """Normal incidence on a lattice (cos theta = 1) and a zero polarisation are rejected
    with -1, as the reference's set-up rejects them, and leave the model usable."""
    from bem3_emf_qd2_amd.solver import Model
    m = Model.from_dir(_scene(tmp_path / "s", n=2))
    with pytest.raises(ValueError):
        m.set_incidence(0.55, 0.0, 0.0)
    with pytest.raises(ValueError):
        m.set_incidence(0.55, 0.3, 0.0, e0x=0.0, e0y=0.0)
    m.set_incidence(0.6, 0.2, 0.1)
    m.solve()
    m.close()


@pytest.mark.gpu
def test_gpu_assembly_tolerance_keeps_the_metric_bound(tmp_path, gpu_lib):
    """The tuned assembly (bq_assembly_tolerance(1e-11): Gauss orders bounded at 1e-11, the
    lattice-sum cut chosen per incidence from sampled element offsets) with the mixed-precision
    solve, against the plain FP64 path at the reference's accuracy: surface solution and probe
    fields within 1e-8 (the metric's bound), the cut below the default 40, and tolerance 0
    restoring the default cut."""
    d = _scene(tmp_path / "s", n=4)
    st64, ref, f64 = _solve(d, tmp_path, "fp64", gpu_lib=gpu_lib)
    gpu_lib.assembly_tolerance(1e-11)
    try:
        stt, tun, ft = _solve(d, tmp_path, "tuned", precision="mixed", gpu_lib=gpu_lib)
    finally:
        gpu_lib.assembly_tolerance(0.0)
    e = _rel(tun, ref)
    ef = float(np.abs(ft - f64).max() / np.abs(f64).max())
    print(f"tuned cut {stt.asm_cut} (tol {stt.asm_tol}), QPGF evaluations {stt.n_qpgf} vs {st64.n_qpgf}; "
          f"surface {e:.2e}, probes {ef:.2e}")
    assert st64.asm_cut == 40.0 and st64.asm_tol == 0.0
    assert stt.asm_tol == 1e-11 and stt.asm_cut < 40.0
    assert stt.n_qpgf < st64.n_qpgf
    assert e < 1e-8 and ef < 1e-8
    assert gpu_lib.assembly_tolerance() == 0.0
