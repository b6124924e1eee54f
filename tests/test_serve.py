"""Request loop (csrc/host/bq_serve.cpp): continuous batching over resident models.

Each request is the reference's one-wave run (d3qd2_src/d3qd2_bv_solver.c:3-21: wave set-up,
solve_bieq, then EH_t at probe points as example1 does) on a model kept on the GPU.  The
checks: a request's fields are the same whether it is served alone or in a mix of concurrent
requests on several slots (1e-10 of the largest value: the mixed-precision solve is refined to
an FP64 backward error near 1e-16, so only rounding can differ); a polarisation-only repeat on
the slot that holds its wave reuses matrices and factors; a memory limit that admits one
session at a time serialises the requests and changes no result.
"""
from pathlib import Path

import numpy as np
import pytest

PROBES = [[0.1, 0.2, 0.9], [-0.3, 0.25, -0.8], [0.45, -0.1, 0.6], [0.0, 0.0, 1.5]]


def test_serving_stream_is_seeded_and_shaped():
    """The serving workload's stream: same seed, same requests; Poisson gaps with the stated
    mean; meshes with the stated probabilities; 30 % polarisation-flipped repeats of an earlier
    request's mesh and direction."""
    from bem3_emf_qd2_amd.serve import serving_requests
    a = serving_requests(2000, 0.5, seed=5)
    b = serving_requests(2000, 0.5, seed=5)
    assert a == b
    t = np.array([q["arrival_s"] for q in a])
    gaps = np.diff(np.r_[0.0, t])
    assert np.all(gaps > 0) and abs(gaps.mean() - 2.0) < 0.15
    rep = [q for q in a if "repeat_of" in q]
    assert abs(len(rep) / len(a) - 0.3) < 0.04
    for q in rep:
        o = a[q["repeat_of"]]
        assert (q["mesh"], q["lambda0"], q["theta"], q["phi"]) == (o["mesh"], o["lambda0"], o["theta"], o["phi"])
        assert q["pol"] != o["pol"]
    fresh = [q for q in a if "repeat_of" not in q]
    frac = np.bincount([q["mesh"] for q in fresh], minlength=4) / len(fresh)
    assert np.abs(frac - [0.4, 0.3, 0.2, 0.1]).max() < 0.04


def _meshes(tmp):
    from bem3_emf_qd2_amd import meshgen
    from bem3_emf_qd2_amd.solver import Model
    files = []
    for n in (2, 3):
        d = meshgen.sphere_scene(tmp / f"n{n}", n, 0.45, 1.5 + 0.001j, 0.55, theta=0.3, phi=0.4)
        m = Model.from_dir(d)
        m.write(d / "setup.dat")
        m.close()
        files.append(d / "setup.dat")
    return files


def _requests():
    rng = np.random.default_rng(21)
    out = []
    for i in range(8):
        out.append(dict(mesh=i % 2, lambda0=float(rng.uniform(0.4, 0.8)), theta=float(rng.uniform(0.05, 0.6)),
                        phi=float(rng.uniform(0, 1.5)), pol="TM" if i % 3 else "TE", arrival_s=0.02 * i))
    for j in (1, 4):  # polarisation repeats
        q = dict(out[j])
        q["pol"] = "TE" if q["pol"] == "TM" else "TM"
        q["arrival_s"] = 0.2 + 0.01 * j
        out.append(q)
    return out


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.gpu
def test_gpu_serve_alone_equals_in_the_mix(tmp_path, gpu_lib):
    from bem3_emf_qd2_amd.serve import Engine
    files = _meshes(tmp_path)
    reqs = _requests()
    eng = Engine(files, nslots=3, mem_limit_gb=64.0, precision="mixed")
    E, H, outs, rc = eng.run(reqs, PROBES)
    eng.close()
    assert rc == 0 and all(o["status"] == 0 for o in outs)
    slots = {o["slot"] for o in outs}
    spans = sorted((o["t_start_s"], o["t_done_s"], o["slot"]) for o in outs)
    overlap = any(spans[i + 1][0] < spans[i][1] and spans[i + 1][2] != spans[i][2] for i in range(len(spans) - 1))
    print("slots used", slots, "overlapping requests", overlap, "reused", sum(o["reused"] for o in outs))
    assert len(slots) > 1
    worst = 0.0
    for i, q in enumerate(reqs):
        one = Engine(files, nslots=1, mem_limit_gb=64.0, precision="mixed")
        E1, H1, o1, rc1 = one.run([dict(q, arrival_s=0.0)], PROBES)
        one.close()
        assert rc1 == 0 and o1[0]["status"] == 0
        worst = max(worst, _rel(E[i], E1[0]), _rel(H[i], H1[0]))
    print("alone vs in the mix: max rel diff", worst)
    assert worst < 1e-10


@pytest.mark.gpu
def test_gpu_serve_polarisation_repeat_reuses_factors(tmp_path, gpu_lib):
    from bem3_emf_qd2_amd.serve import Engine
    files = _meshes(tmp_path)
    a = dict(mesh=1, lambda0=0.61, theta=0.35, phi=0.7, pol="TM", arrival_s=0.0)
    b = dict(a, pol="TE", arrival_s=0.0)
    eng = Engine(files, nslots=1, mem_limit_gb=64.0, precision="mixed")
    E, H, outs, rc = eng.run([a, b], PROBES)
    eng.close()
    assert rc == 0 and [o["reused"] for o in outs] == [0, 1]
    one = Engine(files, nslots=1, mem_limit_gb=64.0, precision="mixed")
    E1, H1, _, _ = one.run([b], PROBES)
    one.close()
    print("reused vs fresh TE:", _rel(E[1], E1[0]), _rel(H[1], H1[0]))
    assert _rel(E[1], E1[0]) < 1e-10 and _rel(H[1], H1[0]) < 1e-10


@pytest.mark.gpu
def test_gpu_serve_memory_limit_serialises(tmp_path, gpu_lib):
    """With a limit that holds one session of the mesh but not two, requests run one at a time (idle sessions of other slots are released to admit the next) and the
    results equal those of the unconstrained run."""
    from bem3_emf_qd2_amd.serve import Engine
    from bem3_emf_qd2_amd.solver import Model
    files = _meshes(tmp_path)
    reqs = [dict(q, mesh=1) for q in _requests()[:6]]
    gpu_lib.solve_precision("mixed")
    try:
        m = Model.from_dat(files[1])
        m.solve()
        one = m.device_bytes()
        m.close()
    finally:
        gpu_lib.solve_precision("fp64")
    eng = Engine(files, nslots=3, mem_limit_gb=1.4 * one / 2**30, precision="mixed")
    E, H, outs, rc = eng.run(reqs, PROBES)
    eng.close()
    assert rc == 0
    spans = sorted((o["t_start_s"], o["t_done_s"]) for o in outs)
    assert all(spans[i + 1][0] >= spans[i][1] for i in range(len(spans) - 1))
    free = Engine(files, nslots=3, mem_limit_gb=64.0, precision="mixed")
    E2, H2, _, _ = free.run(reqs, PROBES)
    free.close()
    assert _rel(E, E2) < 1e-10 and _rel(H, H2) < 1e-10
