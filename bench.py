"""Benchmark: solved incidences per second on one MI355X, with p50/p99 latency and the error of
every result against the plain FP64 GPU path.

One step = one solved incidence, the work the reference's d3qd2_bv_solver + example1 do for
one plane wave (d3qd2_src/d3qd2_solve_bieq.c:36-84, d3qd2_src/d3qd2_field.c:271): assembly
of every domain's coefficient matrices, the global solve, boundary E/H and dE/dH, then E and
H at 100 probe points (EH_t, type 1).  The mesh stays resident; each step recomputes
everything above for its incident wave.

Default workload (c4, the sweep the metric is quoted on): a cube-sphere of 6 x 32 x 32 = 6144
bilinear quads (radius 0.45 um, n 1.5 + 0.001i) in vacuum on a 1.0 um square lattice; 16
incidence points from default_rng(4): lambda0 uniform in [0.4, 0.8] um, theta in [0, 40] deg,
phi in [0, 90] deg, TE or TM; 100 probes uniform in the cell with z in [-1, 1] um.  The model
is set up once; each point moves it to a new wave (bq_set_incidence) and solves on the
resident device session.  The same points are solved first on the plain FP64 path (timed:
fp64_value) and then on the mixed-precision path (timed: value): complex FP32 factors of the
single-object reduction as the preconditioner of FP64 GMRES on the unreduced system.  Each
mixed result is compared with the FP64 result of the same point: surface solution (U, dU, E,
dE, H, dH) and probe E, H, max |diff| / max |FP64| per array.
Workload c4s: the same at 6 x 8 x 8 quads (per-evaluation counters); c1: the 384-quad C1
sphere, one incidence (plain FP64); c2: the C2 two-layer sphere + cone scene (plain FP64).

Prints ONE JSON line.  `roofline` describes the kernel with the largest share of device time
(HIP events on the session stream around every launch); `cpu_baseline` times the CPU
restatement of the same path (oracle/) on a bounded sample of rows, scaled to one incidence.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6

# FP64 flops per QPGF evaluation, 64 x (ADD + MUL + TRANS + 2 FMA) instructions of one launch
# (rocprofv3 PMC) over its device-counted evaluations; C2's figures are read on the same scene
# at reduced resolution (workload c2s: same lattice, wavelength, materials and incidence).
# profiles/k_q0_gl_c1_pmc.md and profiles/k_q0_gl_c2s_pmc.md hold the counter values.
FLOPS_PER_EVAL = {"c1": {"k_q0_gl": 42868, "k_q0_gh": 36660, "k_q0_sing": 36531},
                  "c2": {"k_q0_gl": 35489, "k_q0_gh": 36690, "k_q0_sing": 36984}}
FLOPS_PER_EVAL["c2s"] = FLOPS_PER_EVAL["c2"]
# the 6144-quad sweep, read on its 384-quad variant (c4s) at the same six timed points
# (profiles/k_q0_gl_c4s_pmc.md): the reference's accuracy (cut 40) and the mixed path's tuned
# assembly (tolerance 1e-12: cut 26 at every point)
FLOPS_PER_EVAL["c4"] = {"k_q0_gl": 38705}
FLOPS_PER_EVAL["c4_tuned"] = {"k_q0_gl": 29427}
FLOPS_PER_EVAL["c4s"], FLOPS_PER_EVAL["c4s_tuned"] = FLOPS_PER_EVAL["c4"], FLOPS_PER_EVAL["c4_tuned"]
# the 1536-quad two-layer sweep has the open region, lattice, wavelength and direction ranges of
# the 6144-quad sweep, which fix the per-evaluation cost at the reference's cut
FLOPS_PER_EVAL["c3"] = FLOPS_PER_EVAL["c4"]
# HBM bytes per evaluation of k_q0_gl: (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00) KiB x 1024 of one
# launch over its evaluations (factors from a 1 GiB copy read with the same counters)
TRAFFIC_PER_EVAL = {"c1": {"k_q0_gl": (6234276.5 * 2.0 + 7736990.2) * 1024 / 41990176},
                    "c2": {"k_q0_gl": (47751495.6 * 2.0 + 50115568.2) * 1024 / 350386455}}
TRAFFIC_PER_EVAL["c2s"] = TRAFFIC_PER_EVAL["c2"]
# FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 per evaluation on c4s at the reference's cut and at the
# tuned cut (profiles/k_q0_gl_c4s_pmc.md): nearly all of it the register spills of the kernel's
# 128-register budget; the 1536-quad two-layer sweep shares the open region's physics
TRAFFIC_PER_EVAL["c4"] = TRAFFIC_PER_EVAL["c3"] = TRAFFIC_PER_EVAL["c4s"] = {"k_q0_gl": 1241.2}
TRAFFIC_PER_EVAL["c4_tuned"] = TRAFFIC_PER_EVAL["c4s_tuned"] = {"k_q0_gl": 1277.3}
# k_q0_gl's issue profile on c4s (same counters, six points): active lanes per VALU instruction
# (SQ_THREAD_CYCLES_VALU / 64 SQ_ACTIVE_INST_VALU; the tail lanes of a pair's last sweep count as
# active, they evaluate a point whose result is discarded) and the FP64 share of the VALU
# instructions (ADD + MUL + FMA + TRANS F64 over SQ_INSTS_VALU)
VALU_PROFILE = {"c4": {"k_q0_gl": {"valu_lane_util": 0.951, "fp64_share_of_valu": 0.779}},
                "c4_tuned": {"k_q0_gl": {"valu_lane_util": 0.952, "fp64_share_of_valu": 0.765}}}
VALU_PROFILE["c4s"], VALU_PROFILE["c4s_tuned"] = VALU_PROFILE["c4"], VALU_PROFILE["c4_tuned"]
VALU_PROFILE["c3"] = VALU_PROFILE["c4"]

WORKLOADS = {
    "c1": dict(kind="sphere", n=8, radius=0.4, n_sphere=1.5, lam=0.6, n0=1.0, theta=1e-6, phi=0.0, pol="TM",
               d1=(1.0, 0.0), d2=(0.0, 1.0), probes=100, seed=1, steps=10, warmup=2,
               label="C1: 384-quad cube-sphere r=0.4um n=1.5 in n0=1.0, square lattice 1.0um, "
                     "lambda0=0.6um, x-pol, normal incidence, 100 probes (EH_t type 1)"),
    "c2": dict(kind="sphere_cone", n=16, lam=0.55, n0=1.0, theta=np.radians(30), phi=np.radians(15), pol="TM",
               d1=(0.9, 0.0), d2=(0.45, 0.779), probes=100, seed=1, steps=2, warmup=1,
               label="C2: two-layer cube-sphere (core r=0.3um n=1.6+0.01i, shell r=0.40um n=1.45, 1536 quads "
                     "per surface) + cone (base r=0.3um, h=0.6um, n=2.0, 2012 linear triangles, 1% edge jitter, "
                     "seed 2) in one hexagonal cell a1=(0.9,0) a2=(0.45,0.779)um, lambda0=0.55um, theta=30deg, "
                     "phi=15deg, TM, 100 probes (EH_t type 1); shell 0.40 instead of 0.45um so that it does "
                     "not overlap its periodic image (|a2| = 0.8996um)"),
    # the C2 scene at a quarter of the element count (384 quads per sphere surface, ~500
    # cone triangles), same lattice, wavelength, materials and incidence: used to read the
    # QPGF kernels' per-evaluation counters, which depend on the lattice and the spread of
    # source-target offsets, not on the element count
    # the batched spectral sweep: two-layer 1536-quad sphere, 64 seeded points, FP64 assembly
    # and solves at the reference's accuracy (a lattice-sum cut tuned at tolerance 1e-14 moves
    # the probe fields by up to 1.9e-9 on this model, beyond the FP64 path's 1e-10 parity)
    "c3": dict(kind="sweep", n=16, radius=0.45, n_sphere=1.45, core=(0.3, 1.6 + 0.01j), n0=1.0, d1=(1.0, 0.0),
               d2=(0.0, 1.0), points=64, seed=3, probes=100, steps=12, warmup=1, precision="fp64", asm_tol=0.0,
               label="C3: two-layer cube-sphere (shell r=0.45um n=1.45, core r=0.3um n=1.6+0.01i, 6x16x16 = "
                     "1536 quads per surface) in n0=1.0, square lattice 1.0um; sweep of 64 points from "
                     "default_rng(3): lambda0 U[0.4,0.8]um, theta U[0,40]deg, phi U[0,90]deg, TE/TM; FP64 "
                     "assembly and solve; 100 probes (EH_t type 1)"),
    # the 6144-quad sweep: one resident model, 16 seeded incidence points, FP64 and mixed
    "c4": dict(kind="sweep", n=32, radius=0.45, n_sphere=1.5 + 0.001j, n0=1.0, d1=(1.0, 0.0), d2=(0.0, 1.0),
               points=16, seed=4, probes=100, steps=6, warmup=1,
               label="C4: 6x32x32 = 6144-quad cube-sphere r=0.45um n=1.5+0.001i in n0=1.0, square lattice "
                     "1.0um; sweep of 16 points from default_rng(4): lambda0 U[0.4,0.8]um, theta U[0,40]deg, "
                     "phi U[0,90]deg, TE/TM; 100 probes (EH_t type 1)"),
    # serving: Poisson request stream over four cube-sphere sizes (the C4 object and lattice)
    "c5": dict(kind="serve", sizes=(8, 16, 24, 32), radius=0.45, n_sphere=1.5 + 0.001j, n0=1.0, d1=(1.0, 0.0),
               d2=(0.0, 1.0), stream=2000, rate=0.5, seed=5, probes=100, steps=40, warmup=4,
               label="C5: serving 2000 Poisson requests (seed 5, 0.5 requests/s) over cube-spheres of 384/1536/3456/"
                     "6144 quads (p 0.4/0.3/0.2/0.1; r=0.45um n=1.5+0.001i, square lattice 1.0um), lambda0 "
                     "U[0.4,0.8]um, theta U[0,40]deg, phi U[0,90]deg, TE/TM, 30% polarisation repeats; mixed "
                     "precision, 100 probes (EH_t type 1)"),
    # the same sweep at 6x8x8 quads (per-evaluation counters, quick checks)
    "c4s": dict(kind="sweep", n=8, radius=0.45, n_sphere=1.5 + 0.001j, n0=1.0, d1=(1.0, 0.0), d2=(0.0, 1.0),
                points=16, seed=4, probes=100, steps=6, warmup=1,
                label="C4 sweep physics on a 6x8x8 = 384-quad cube-sphere (counters and checks)"),
    "c2s": dict(kind="sphere_cone", n=8, cone_h=0.06, lam=0.55, n0=1.0, theta=np.radians(30), phi=np.radians(15),
                pol="TM", d1=(0.9, 0.0), d2=(0.45, 0.779), probes=100, seed=1, steps=3, warmup=1,
                label="C2 geometry at reduced resolution (384 quads per sphere surface, cone edge 0.06um), "
                      "for per-evaluation kernel counters"),
}


def _probes(w, n, seed):
    rng = np.random.default_rng(seed)
    d1, d2 = np.array(w["d1"]), np.array(w["d2"])
    a = rng.uniform(-0.5, 0.5, (n, 2))
    xy = a[:, :1] * d1 + a[:, 1:] * d2
    z = rng.uniform(-1.0, 1.0, (n, 1))
    return np.ascontiguousarray(np.c_[xy, z])


def sweep_points(w):
    """The seeded (lambda0, theta, phi, polarisation) points of a sweep workload."""
    rng = np.random.default_rng(w["seed"])
    pts = []
    for _ in range(w["points"]):
        lam = float(rng.uniform(0.4, 0.8))
        th = float(np.radians(rng.uniform(0.0, 40.0)))
        ph = float(np.radians(rng.uniform(0.0, 90.0)))
        pol = "TE" if rng.integers(2) == 0 else "TM"
        pts.append((lam, th, ph, pol))
    return pts


def _set_point(m, p):
    lam, th, ph, pol = p
    ex, ey = (1.0, 0.0) if pol == "TM" else (0.0, 1.0)
    m.set_incidence(lam, th, ph, e0x=ex, e0y=ey)


def _surface(path):
    from datfile import read_dat
    d = read_dat(path)
    return [{k: d["dom"][q][k][1:].copy() for k in ("U", "dU", "E", "dE", "H", "dH")} for q in range(d["MN"] + 1)]


def _surface_err(a, b):
    e = 0.0
    for da, db in zip(a, b):
        for k in da:
            e = max(e, float(np.abs(da[k] - db[k]).max() / np.abs(db[k]).max()))
    return e


def _scene(w, d):
    from bem3_emf_qd2_amd import meshgen
    if w["kind"] == "sweep":
        lam, th, ph, pol = sweep_points(w)[0]
        meshgen.sphere_scene(d, w["n"], w["radius"], w["n_sphere"], lam, n0=w["n0"], theta=th, phi=ph, pol=pol,
                             d1=w["d1"], d2=w["d2"], core=w.get("core"))
    elif w["kind"] == "sphere_cone":
        meshgen.sphere_cone_scene(d, n=w["n"], cone_h=w.get("cone_h", 0.03), lam=w["lam"], n0=w["n0"],
                                  theta=w["theta"], phi=w["phi"], pol=w["pol"], d1=w["d1"], d2=w["d2"])
    else:
        meshgen.sphere_scene(d, w["n"], w["radius"], w["n_sphere"], w["lam"], n0=w["n0"], theta=w["theta"],
                             phi=w["phi"], pol=w["pol"], d1=w["d1"], d2=w["d2"])


class _QuietStdout:
    """Send the library's progress prints (solve_bieq's phase lines) to /dev/null while the
    steps run, so that stdout carries only the JSON line."""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.libc.fflush(None)
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *exc):
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)
        return False


def _kernel_times(L):
    import ctypes
    names = ctypes.create_string_buffer(48 * 64)
    ms = (ctypes.c_double * 64)()
    cnt = (ctypes.c_longlong * 64)()
    n = L.bq_kernel_times(64, names, ms, cnt)
    out = {}
    for i in range(n):
        nm = names.raw[48 * i:48 * (i + 1)].split(b"\0")[0].decode()
        out[nm] = (ms[i], cnt[i])
    return out


def _roofline(ktimes, stats_per_step, steps, workload):
    """Roofline object for the hand-written kernel (k_*) with the largest total device time; the
    timed ranges of library calls (rocBLAS updates inside getrf/getrs) are not candidates."""
    own = {k: v for k, v in ktimes.items() if k.startswith("k_")} or ktimes
    name = max(own, key=lambda k: own[k][0])
    tot_ms, cnt = ktimes[name]
    avg_s = tot_ms / cnt / 1e3
    r = {"kernel": name, "bound": "mfma", "pipe": "fp64-valu",
         "bound_note": "compute bound on the FP64 vector pipe (no matrix instructions); MI355X's FP64 vector "
                       "and FP64 matrix peaks are both 78.6 TFLOP/s",
         "unit": "TFLOP/s", "peak": FP64_PEAK_TFS,
         "avg_launch_ms": avg_s * 1e3, "launches_per_step": cnt / steps, "traffic": None, "achieved": None,
         "frac": None}
    if name == "zgetrf_update":
        r["note"] = "rocBLAS zgemm trailing updates of the blocked LU"
        return r
    # QPGF evaluations per launch, as the device counted them from the pair classification
    evals = {"k_q0_gl": stats_per_step.get("n_qpgf_gl", 0), "k_q0_gh": stats_per_step.get("n_qpgf_gh", 0),
             "k_q0_sing": stats_per_step.get("n_qpgf_sing", 0)}
    if name in evals:
        r["evals_per_launch"] = evals[name]
        fpe = FLOPS_PER_EVAL.get(workload, {}).get(name)
        tpe = TRAFFIC_PER_EVAL.get(workload, {}).get(name)
        r["traffic"] = tpe * evals[name] if tpe else None
        if fpe:
            ach = evals[name] * fpe / avg_s / 1e12
            r["achieved"] = ach
            r["frac"] = ach / FP64_PEAK_TFS
            r["flops_per_eval"] = fpe
        vp = VALU_PROFILE.get(workload, {}).get(name)
        if vp:
            r.update(vp)
    return r


def gemm_probe():
    """k_cgemm_split (C -= A B, complex FP32 operands, bf16x3 products) against rocBLAS cgemm at the
    update shapes of the 6144-quad factorisation: kernel time from the library's HIP events."""
    import ctypes
    sys.path.insert(0, str(ROOT / "tests"))
    import bem3_emf_qd2_amd as bq
    from hipmem import DeviceArray, _rt
    L = bq.lib()
    if not bq.device_ok():
        raise SystemExit(f"bench.py: no usable gfx950 device: {L.bq_last_error().decode()}")
    V = ctypes.c_void_p
    L.bq_cgemm_split_dev.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, ctypes.c_long, V, ctypes.c_long,
                                     V, ctypes.c_long]
    L.bq_cgemm_op_dev.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, V, V,
                                  ctypes.c_long, V, ctypes.c_long, V, V, ctypes.c_long]
    rt = _rt()
    al = (ctypes.c_float * 2)(-1.0, 0.0)
    be = (ctypes.c_float * 2)(1.0, 0.0)
    for (m, n, k) in [(24576, 24576, 512), (16384, 16384, 512), (8192, 8192, 512), (24576, 24576, 256)]:
        rng = np.random.default_rng(1)
        A = DeviceArray((rng.standard_normal((k, m)) + 1j * rng.standard_normal((k, m))).astype(np.complex64))
        B = DeviceArray((rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))).astype(np.complex64))
        C = DeviceArray(np.zeros((n, m), np.complex64))
        L.bq_cgemm_split_dev(None, m, n, k, A.ptr, m, B.ptr, k, C.ptr, m)
        rt.hipDeviceSynchronize()
        reps = 5
        L.bq_kernel_timing(1)
        t = time.perf_counter()
        for _ in range(reps):
            if L.bq_cgemm_split_dev(None, m, n, k, A.ptr, m, B.ptr, k, C.ptr, m):
                raise SystemExit(L.bq_last_error().decode())
        rt.hipDeviceSynchronize()
        wall = (time.perf_counter() - t) / reps
        kt = _kernel_times(L)
        L.bq_kernel_timing(0)
        kms = kt["k_cgemm_split"][0] / kt["k_cgemm_split"][1]
        L.bq_cgemm_op_dev(None, 0, 0, m, n, k, al, A.ptr, m, B.ptr, k, be, C.ptr, m)
        rt.hipDeviceSynchronize()
        t = time.perf_counter()
        for _ in range(reps):
            L.bq_cgemm_op_dev(None, 0, 0, m, n, k, al, A.ptr, m, B.ptr, k, be, C.ptr, m)
        rt.hipDeviceSynchronize()
        rb = (time.perf_counter() - t) / reps
        fl = 8.0 * m * n * k
        for d in (A, B, C):
            d.free()
        import subprocess
        lib_run = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--library-bf16", str(m), str(n), str(k)],
                                 capture_output=True, text=True, check=True)
        lmm, lup = json.loads(lib_run.stdout.strip().splitlines()[-1])
        print(json.dumps(dict(m=m, n=n, k=k, split_kernel_ms=kms, split_wall_ms=wall * 1e3, split_tfs=fl / kms / 1e9,
                              split_bf16_frac=3 * fl / kms / 1e9 / 2500.0, rocblas_ms=rb * 1e3,
                              rocblas_tfs=fl / rb / 1e12, library_bf16_product_ms=lmm,
                              library_bf16_product_frac=3 * fl / lmm / 1e9 / 2500.0, library_bf16_update_ms=lup,
                              library_bf16_update_frac=3 * fl / lup / 1e9 / 2500.0)), flush=True)


def _library_bf16_ms(m, n, k):
    """The same bf16 matrix work as one real GEMM in the vendor library (torch -> hipBLASLt):
    C' (2m x n, FP32) -= A' B' with A' = 2m x 6k, B' = 6k x n in bf16 (the complex product as a
    2 x 2 real block, each operand's hi, lo, hi split stacked along k), i.e. the library's rate for
    the matrix-core work the split kernel does.  Returns the product alone (C' = A' B', C not
    read) and the update (addmm, beta = 1, alpha = -1), each the mean of five calls after one
    (torch's events)."""
    import torch
    a = torch.randn(2 * m, 6 * k, device="cuda", dtype=torch.bfloat16)
    b = torch.randn(6 * k, n, device="cuda", dtype=torch.bfloat16)
    c = torch.zeros(2 * m, n, device="cuda", dtype=torch.float32)
    out = []
    for f in (lambda: torch.mm(a, b, out_dtype=torch.float32),
              lambda: torch.addmm(c, a, b, out_dtype=torch.float32, beta=1.0, alpha=-1.0)):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            f()
        e1.record()
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1) / 5)
    del a, b, c
    torch.cuda.empty_cache()
    return out


def _cpu_baseline(model_dat, w, probes, gpu_solve_dims):
    """Oracle (CPU restatement, OpenMP) on a bounded sample, scaled to one incidence."""
    sys.path.insert(0, str(ROOT / "tests"))
    from datfile import read_dat
    from oracle_pipeline import OracleModel
    import scipy.linalg as sla

    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    d = read_dat(model_dat)
    o = OracleModel.from_dat(d)
    rng = np.random.default_rng(7)
    total = 0.0
    measured = 0.0
    parts = []
    for dom in range(o.MN + 1):
        ne = o.dNe[dom]
        rows = min(ne, 2 * cores)
        t0 = int(rng.integers(1, ne - rows + 2))
        t = time.perf_counter()
        o.assemble(dom, t0, t0 + rows - 1, deriv=(dom > 0))
        dt = time.perf_counter() - t
        measured += dt
        total += dt * ne / rows
        parts.append(f"domain {dom}: {rows}/{ne} element rows in {dt:.1f}s")
    # dense LUs (the global system of order 8 NN, and per domain the dE/dH system of order
    # N_d): timed at order min(n, 3072) and scaled by (n / n_s)^3
    def lu_time(n):
        n_s = min(n, 3072)
        A = rng.standard_normal((n_s, n_s)) + 1j * rng.standard_normal((n_s, n_s))
        t = time.perf_counter()
        sla.lu_factor(A, check_finite=False)
        return time.perf_counter() - t, (n / n_s) ** 3, n_s

    na = 8 * o.NN
    dt, sc, n_s = lu_time(na)
    measured += dt
    total += dt * sc
    parts.append(f"global LU {na} timed at {n_s} in {dt:.1f}s x{sc:.0f}")
    for dom in range(o.MN + 1):
        n = int(o.live_rows(dom).sum())
        dt, sc, n_s = lu_time(n)
        measured += dt
        total += dt * sc
    parts.append("per-domain dE/dH LUs timed at <= 3072 and scaled")
    return {"value": 1.0 / total, "unit": "incidences/s", "cores": cores, "kind": "port",
            "seconds_per_incidence": total,
            "sample": "oracle/ (C restatement, OpenMP) on the same model: " + "; ".join(parts) +
                      f"; element rows scaled by row count, probes not included; {measured:.0f}s measured"}


def run_sweep(args, w, bq, L):
    """Sweep workload: one resident model, each step a new incidence point.  The FP64 path
    runs the points first (warm-up, then timed), then the mixed path the same points; every
    mixed result is compared with the FP64 result of its point."""
    from bem3_emf_qd2_amd.solver import Model
    sys.path.insert(0, str(ROOT / "tests"))
    os.environ.setdefault("BQ_SERVE_LOG", "1")  # the request loop's per-request progress lines
    pts = sweep_points(w)
    if args.warmup + args.steps > len(pts):
        raise SystemExit(f"bench.py: {w['points']} sweep points; --warmup + --steps must not exceed them")
    warm, timed = pts[:args.warmup], pts[args.warmup:args.warmup + args.steps]
    probes = _probes(w, w["probes"], w["seed"])
    res = {"fp64": [], "mixed": []}
    with tempfile.TemporaryDirectory() as td:
        _scene(w, td)
        m = Model.from_dir(td)
        m.write(Path(td) / "setup.dat")

        def step(p):
            _set_point(m, p)
            st = m.solve()
            E, H, did = m.fields(probes, 1)
            return st, np.r_[E.ravel(), H.ravel()]

        kt = kt_serial = {}
        if args.passes:
            modes = [m_ for m_ in args.passes.split(",") if m_ in ("fp64", "mixed")]
            if not modes:
                raise SystemExit("bench.py: --passes needs a serial pass (fp64 or mixed) besides batched")
        elif args.precision == "mixed":
            # the plain FP64 path, then the batched mixed-precision path (below); the serial
            # mixed pass (surface-solution errors) runs when asked for with --passes fp64,mixed
            modes = ["fp64"] if args.slots > 0 else ["fp64", "mixed"]
        else:
            modes = [args.precision]
        with _QuietStdout():
            for mode in modes:
                # each pass starts from a fresh session, so that no timed point reuses the
                # matrices the other pass assembled for the same wave (allocation happens in
                # the untimed warm-up)
                m.release_device()
                bq.solve_precision(mode)
                bq.assembly_tolerance(args.asm_tol if mode == "mixed" else 0.0)
                for i, p in enumerate(warm):
                    step(p)
                    print(f"bench.py: {mode} warm-up point {i + 1}/{len(warm)} done", file=sys.stderr, flush=True)
                if mode == modes[-1]:
                    L.bq_kernel_timing(1)
                for i, p in enumerate(timed):
                    t = time.perf_counter()
                    st, f = step(p)
                    dt = time.perf_counter() - t
                    m.write(Path(td) / "pt.dat")
                    if args.save_dat and i == 0:
                        # the first timed point's solved model, for checks on the host
                        Path(args.save_dat).mkdir(parents=True, exist_ok=True)
                        shutil.copy(Path(td) / "pt.dat", Path(args.save_dat) / f"{args.workload}_{mode}_point1.dat")
                    res[mode].append(dict(lat=dt, st=st.as_dict(), probes=f, surf=_surface(Path(td) / "pt.dat")))
                    print(f"bench.py: {mode} point {i + 1}/{len(timed)} lambda0={p[0]:.4f} theta="
                          f"{np.degrees(p[1]):.2f} phi={np.degrees(p[2]):.2f} {p[3]}: {dt:.2f} s "
                          f"(assembly {st.t_assemble_ms:.0f} ms, solve {st.t_solve_ms:.0f} ms, "
                          f"GMRES {st.gmres_iterations}, eta {st.backward_error:.1e})", file=sys.stderr, flush=True)
                if mode == modes[-1]:
                    kt = kt_serial = _kernel_times(L)
                    L.bq_kernel_timing(0)
            bq.solve_precision("fp64")
            bq.assembly_tolerance(0.0)
        batched = None
        if args.slots > 0 and args.precision in ("fp64", "mixed") and (not args.passes or "batched" in args.passes):
            # the batched path: the request loop with args.slots worker slots, every timed point
            # queued at once; one warm-up run first so that each slot holds its session
            from bem3_emf_qd2_amd.serve import Engine
            m.release_device()
            bq.assembly_tolerance(args.asm_tol)
            eng = Engine([Path(td) / "setup.dat"], nslots=args.slots, mem_limit_gb=args.mem_gb,
                         precision=args.precision)
            as_req = lambda ps: [dict(mesh=0, lambda0=p[0], theta=p[1], phi=p[2], pol=p[3], arrival_s=0.0) for p in ps]
            wpts = (warm + pts[args.warmup + args.steps:])[:args.slots] or warm
            _, _, _, rcw = eng.run(as_req(wpts), probes)
            print(f"bench.py: batched warm-up ({len(wpts)} points) done", file=sys.stderr, flush=True)
            L.bq_kernel_timing(1)
            t = time.perf_counter()
            Eb, Hb, outs, rcb = eng.run(as_req(timed), probes)
            wall = time.perf_counter() - t
            kt = _kernel_times(L)
            L.bq_kernel_timing(0)
            eng.close()
            bq.assembly_tolerance(0.0)
            if rcw or rcb or any(o["status"] for o in outs):
                raise SystemExit(f"bench.py: batched run failed ({rcw}, {rcb}, {[o['status'] for o in outs]})")
            makespan = max(o["t_done_s"] for o in outs)
            batched = dict(wall=wall, makespan=makespan, outs=outs,
                           probes=[np.r_[Eb[i].ravel(), Hb[i].ravel()] for i in range(len(timed))])
            for i, o in enumerate(outs):
                print(f"bench.py: batched point {i + 1}/{len(timed)} slot {o['slot']}: start {o['t_start_s']:.2f} s, "
                      f"done {o['t_done_s']:.2f} s", file=sys.stderr, flush=True)
        sys.path.insert(0, str(ROOT / "tests"))
        from datfile import read_dat
        dd = read_dat(Path(td) / "setup.dat")
        tris = int(np.sum(dd["ed"][1:, 3] == 0))
        quads, NN = int(dd["Ne"]) - tris, int(dd["NN"])
        serial = modes[-1]                     # the last serial pass
        head = "batched" if batched is not None else serial
        rows = []
        for i, p in enumerate(timed):
            r = {"lambda0": p[0], "theta_deg": float(np.degrees(p[1])), "phi_deg": float(np.degrees(p[2])),
                 "pol": p[3]}
            for mode in modes:
                q = res[mode][i]
                r[f"ms_{mode}"] = q["lat"] * 1e3
                r[f"assemble_ms_{mode}"] = q["st"]["t_assemble_ms"]
                r[f"solve_ms_{mode}"] = q["st"]["t_solve_ms"]
                r[f"backward_error_{mode}"] = q["st"]["backward_error"]
                r[f"n_qpgf_gl_{mode}"] = q["st"]["n_qpgf_gl"]
                r[f"asm_cut_{mode}"] = q["st"]["asm_cut"]
            if res["fp64"] and res["mixed"]:
                a, b = res["mixed"][i], res["fp64"][i]
                n3 = len(b["probes"]) // 2
                r["surface_rel_err_vs_fp64"] = _surface_err(a["surf"], b["surf"])
                r["probe_E_rel_err_vs_fp64"] = float(np.abs(a["probes"][:n3] - b["probes"][:n3]).max() /
                                                     np.abs(b["probes"][:n3]).max())
                r["probe_H_rel_err_vs_fp64"] = float(np.abs(a["probes"][n3:] - b["probes"][n3:]).max() /
                                                     np.abs(b["probes"][n3:]).max())
                r["gmres_iterations"] = a["st"]["gmres_iterations"]
            if batched is not None:
                o = batched["outs"][i]
                r["batched_slot"] = o["slot"]
                r["batched_start_ms"] = o["t_start_s"] * 1e3
                r["batched_done_ms"] = o["t_done_s"] * 1e3
                r["batched_gmres_iterations"] = o["gmres_iterations"]
                r["batched_backward_error"] = o["backward_error"]
                if res["fp64"]:
                    pb, pf = batched["probes"][i], res["fp64"][i]["probes"]
                    n3 = len(pf) // 2
                    r["batched_probe_E_rel_err_vs_fp64"] = float(np.abs(pb[:n3] - pf[:n3]).max() / np.abs(pf[:n3]).max())
                    r["batched_probe_H_rel_err_vs_fp64"] = float(np.abs(pb[n3:] - pf[n3:]).max() / np.abs(pf[n3:]).max())
            rows.append(r)
        # QPGF evaluations per launch of the pass the kernel times come from
        st_ref = res[serial][-1]["st"]
        evals = {k: float(np.mean([q["st"][k] for q in res[serial]])) for k in ("n_qpgf_gl", "n_qpgf_gh", "n_qpgf_sing")}
        # roofline from the last serial pass: its launches do not share the GPU with another
        # session's, so a launch's HIP-event duration is the kernel's own
        fkey = args.workload + "_tuned" if serial == "mixed" and args.asm_tol > 0 else args.workload
        roof = _roofline(kt_serial, evals, args.steps, fkey) if kt_serial else None
        if roof is not None:
            roof["pass"] = f"serial {serial}"
        cpu = None
        if not args.no_cpu_baseline:
            _set_point(m, timed[0])
            m.write(Path(td) / "point.dat")
            cpu = _cpu_baseline_pairs(Path(td) / "point.dat")
        nbytes = m.device_bytes()
        m.close()
    prec_head = args.precision if batched is not None else serial
    if batched is not None:
        total = batched["makespan"]
        lat = np.array([o["t_done_s"] for o in batched["outs"]]) * 1e3
        service = np.array([o["t_done_s"] - o["t_start_s"] for o in batched["outs"]]) * 1e3
    else:
        lat = np.array([q["lat"] for q in res[serial]]) * 1e3
        total = float(lat.sum() / 1e3)
    errkeys = [k for k in ("surface_rel_err_vs_fp64", "probe_E_rel_err_vs_fp64", "probe_H_rel_err_vs_fp64",
                           "batched_probe_E_rel_err_vs_fp64", "batched_probe_H_rel_err_vs_fp64") if k in rows[0]]
    out = {
        "metric": "solved incidences/sec + p50/p99 latency at <=1e-8 rel err vs FP64 plain path",
        "value": args.steps / total,
        "unit": "incidences/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 factors + f64 (assembly, GMRES residual)" if prec_head == "mixed" else "f64",
        "data": "synthetic (generated cube-sphere mesh and seeded plane-wave sweep)",
        "config": {"workload": w["label"], "quads": quads, "triangles": tris, "element_nodes": NN,
                   "na": 8 * NN, "probes": w["probes"], "sweep_points_timed": args.steps,
                   "parallelism": "single GPU",
                   "solve_precision": prec_head,
                   "assembly_tolerance": args.asm_tol if head != "fp64" else 0.0,
                   "path": (f"batched: request loop, {args.slots} slots, phase-gated (one assembly and one "
                            f"factorisation at a time), all points queued at t=0; latency = time to completion"
                            if batched is not None else f"serial {serial}")},
        "max_rel_err_vs_fp64": {k: max(r[k] for r in rows) for k in errkeys} or None,
        "points": rows,
        "phase_ms_last_fp64_point": st_ref,
        "device_bytes": nbytes,
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in sorted(kt.items(), key=lambda x: -x[1][0])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if batched is not None and kt_serial:
        out["kernels_ms_per_step_serial"] = {k: v[0] / args.steps
                                             for k, v in sorted(kt_serial.items(), key=lambda x: -x[1][0])}
    if batched is not None:
        out["batched_service_ms"] = {"p50": float(np.percentile(service, 50)), "p99": float(np.percentile(service, 99))}
        out["batched_wall_s"] = batched["wall"]
    if res["mixed"]:
        ml = np.array([q["lat"] for q in res["mixed"]])
        out["serial_mixed_value"] = args.steps / float(ml.sum())
    if res["fp64"] and head != "fp64":
        fl = np.array([q["lat"] for q in res["fp64"]])
        out["fp64_value"] = args.steps / float(fl.sum())
        out["fp64_latency_ms"] = {"p50": float(np.percentile(fl * 1e3, 50)),
                                  "p99": float(np.percentile(fl * 1e3, 99))}
        out["speedup_vs_fp64"] = out["value"] / out["fp64_value"]
    print(json.dumps(out), flush=True)


def _cpu_baseline_pairs(model_dat, budget_s=8.0):
    """Oracle (CPU restatement of the reference's element-pair routines) on the box's host
    cores: random (collocation node, element) pairs of every domain for about budget_s each,
    scaled to all pairs; plus the reference's dense-equivalent solve of the 8 NN system (LU
    timed at order 3072 and scaled by n^3).  One element row of the 6144-quad model alone
    takes the oracle ~45 s per core, so pairs, not rows, are the sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    from concurrent.futures import ThreadPoolExecutor
    from datfile import read_dat
    from oracle_pipeline import OracleModel
    import scipy.linalg as sla

    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    d = read_dat(model_dat)
    o = OracleModel.from_dat(d)
    rng = np.random.default_rng(7)
    total, measured, parts = 0.0, 0.0, []
    with ThreadPoolExecutor(cores) as ex:
        for dom in range(o.MN + 1):
            ne = o.dNe[dom]
            sid = o.sid(dom)
            live = o.live_rows(dom)
            done, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < budget_s:
                batch = []
                for _ in range(4 * cores):
                    while True:
                        row = int(rng.integers(0, 4 * ne))
                        if live[row]:
                            break
                    s = int(rng.integers(1, ne + 1))
                    batch.append((sid[row // 4 + 1], row % 4, sid[s]))
                list(ex.map(lambda q: o.pair(dom, q[0], q[1], q[2]), batch))
                done += len(batch)
            dt = time.perf_counter() - t0
            npairs = int(live.sum()) * ne
            measured += dt
            total += dt * npairs / done
            parts.append(f"domain {dom}: {done} of {npairs} pairs in {dt:.1f}s")
    n_s = 3072
    A = rng.standard_normal((n_s, n_s)) + 1j * rng.standard_normal((n_s, n_s))
    t = time.perf_counter()
    sla.lu_factor(A, check_finite=False)
    dt = time.perf_counter() - t
    na = 8 * o.NN
    measured += dt
    lu = dt * (na / n_s) ** 3
    asm = total
    total += lu
    parts.append(f"dense LU of the 8NN = {na} system timed at {n_s} in {dt:.1f}s, scaled x{(na / n_s) ** 3:.0f} "
                 f"= {lu:.0f}s (the reference's PARDISO solve of this dense-block system; its factors would "
                 f"need {na * na * 16 / 2**30:.0f} GiB)")
    return {"value": 1.0 / total, "unit": "incidences/s", "cores": cores, "kind": "port",
            "seconds_per_incidence": total, "assembly_seconds": asm, "solve_seconds": lu,
            "measured_seconds": measured,
            "sample": "oracle/ (C restatement of the reference's pair routines, 12-point rule everywhere) on "
                      "the same model at the first timed point's wave: " + "; ".join(parts) +
                      "; pairs scaled by count; probes not included"}


def run_serve(args, w, bq, L):
    """Serving workload: the first args.steps requests of the seeded stream replayed in real
    time (their Poisson arrival times) through the request loop; throughput over the span from
    the first arrival to the last completion, latency = completion - arrival.  A sample (the
    first request of each mesh) is solved again alone on the plain FP64 path for the error."""
    from bem3_emf_qd2_amd import meshgen
    from bem3_emf_qd2_amd.serve import Engine, serving_requests
    from bem3_emf_qd2_amd.solver import Model
    os.environ.setdefault("BQ_SERVE_LOG", "1")
    reqs = serving_requests(w["stream"], w["rate"], seed=w["seed"])[:args.steps]
    probes = _probes(w, w["probes"], w["seed"])
    with tempfile.TemporaryDirectory() as td:
        files = []
        for n in w["sizes"]:
            d = Path(td) / f"n{n}"
            meshgen.sphere_scene(d, n, w["radius"], w["n_sphere"], 0.55, n0=w["n0"], theta=0.3, phi=0.4,
                                 d1=w["d1"], d2=w["d2"])
            m = Model.from_dir(d)
            m.write(d / "setup.dat")
            m.close()
            files.append(d / "setup.dat")
        bq.assembly_tolerance(args.asm_tol)
        eng = Engine(files, nslots=args.slots or 2, mem_limit_gb=args.mem_gb, precision="mixed")
        warm = [dict(mesh=i, lambda0=0.5, theta=0.2, phi=0.3, pol="TM", arrival_s=0.0) for i in range(len(files))]
        _, _, _, rcw = eng.run(warm[:max(args.warmup, 0)], probes)
        print(f"bench.py: serving warm-up done ({rcw})", file=sys.stderr, flush=True)
        t = time.perf_counter()
        E, H, outs, rc = eng.run(reqs, probes)
        wall = time.perf_counter() - t
        eng.close()
        bq.assembly_tolerance(0.0)
        bad = [i for i, o in enumerate(outs) if o["status"]]
        if rc or bad:
            raise SystemExit(f"bench.py: serving run failed ({rc}); requests {bad[:10]}")
        # accuracy sample: the first request of each mesh alone on the plain FP64 path
        errs = []
        ref = Engine(files, nslots=1, mem_limit_gb=args.mem_gb, precision="fp64")
        for mi in range(len(files)):
            idx = next((i for i, q in enumerate(reqs) if q["mesh"] == mi), None)
            if idx is None:
                continue
            E1, H1, o1, rc1 = ref.run([dict(reqs[idx], arrival_s=0.0)], probes)
            a, b = np.r_[E[idx].ravel(), H[idx].ravel()], np.r_[E1[0].ravel(), H1[0].ravel()]
            errs.append(dict(request=idx, mesh=mi, probe_rel_err_vs_fp64=float(np.abs(a - b).max() / np.abs(b).max())))
        ref.close()
    t_first = min(q["arrival_s"] for q in reqs)
    span = max(o["t_done_s"] for o in outs) - t_first
    lat = np.array([o["t_done_s"] - o["t_arrival_s"] for o in outs]) * 1e3
    service = np.array([o["t_done_s"] - o["t_start_s"] for o in outs]) * 1e3
    quads = [6 * n * n for n in w["sizes"]]
    per_mesh = {}
    for q, o in zip(reqs, outs):
        per_mesh.setdefault(quads[q["mesh"]], []).append((o["t_done_s"] - o["t_start_s"]) * 1e3)
    out = {
        "metric": "solved incidences/sec + p50/p99 latency at <=1e-8 rel err vs FP64 plain path",
        "value": len(reqs) / span,
        "unit": "incidences/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": args.warmup,
        "ms_per_step": span / len(reqs) * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "service_ms": {"p50": float(np.percentile(service, 50)), "p99": float(np.percentile(service, 99))},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 factors + f64 (assembly, GMRES residual)",
        "data": "synthetic (generated cube-sphere meshes and seeded request stream)",
        "config": {"workload": w["label"], "requests_served": len(reqs),
                   "requests_note": f"the first {len(reqs)} of the {w['stream']}-request stream, arrivals in real time",
                   "slots": args.slots or 2, "mem_limit_gb": args.mem_gb, "parallelism": "single GPU",
                   "solve_precision": "mixed", "assembly_tolerance": args.asm_tol},
        "reused_factors": int(sum(o["reused"] for o in outs)),
        "repeats_in_sample": int(sum("repeat_of" in q for q in reqs)),
        "service_ms_by_quads": {str(k): float(np.mean(v)) for k, v in sorted(per_mesh.items())},
        "max_backward_error": float(max(o["backward_error"] for o in outs)),
        "accuracy_sample": errs,
        "wall_s": wall,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default per workload: c2 2, c1 10)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default per workload: c2 1, c1 2)")
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--save-dat", default=None, metavar="DIR",
                    help="sweep workloads: write the first timed point's solved model (ex.dat layout) of each "
                         "serial pass to DIR")
    ap.add_argument("--slots", type=int, default=2,
                    help="sweep workloads: worker slots of the batched mixed-precision pass (request loop); "
                         "0 skips it and reports the serial mixed pass")
    ap.add_argument("--mem-gb", type=float, default=270.0, help="device memory limit of the request loop (GiB)")
    ap.add_argument("--asm-tol", type=float, default=None,
                    help="assembly tolerance of the mixed-precision and batched passes (bq_assembly_tolerance: "
                         "Gauss-order bound and per-incidence lattice-sum cut; default 1e-12, c3 1e-14); 0 keeps "
                         "the reference's accuracy.  The serial FP64 pass always runs at the reference's accuracy")
    ap.add_argument("--passes", default=None,
                    help="sweep workloads: comma-separated solve precisions to run in order (default "
                         "'fp64,mixed': the FP64 pass gives fp64_value and the reference results); "
                         "'mixed' alone for profiling runs")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed", "auto"],
                    help="global solve: the plain complex FP64 LU ('fp64', default: the C1 and C2 "
                         "configurations are the plain FP64 path), complex FP32 LU + FP64 iterative "
                         "refinement ('mixed'; one untimed FP64 step first gives rel_err_vs_fp64), or "
                         "'auto': mixed for block systems of order >= 16384")
    ap.add_argument("--gemm-probe", action="store_true",
                    help="only time the FP32 factorisation's update kernel (bf16x3 matrix cores) against rocBLAS "
                         "cgemm at the C4 update shapes, one JSON line per shape")
    ap.add_argument("--library-bf16", type=int, nargs=3, default=None, metavar=("M", "N", "K"),
                    help="only time the vendor library's bf16 GEMM for the split kernel's matrix work at complex "
                         "shape M x N x K (a separate process of --gemm-probe: torch and the library each bring "
                         "their own HIP runtime)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(args.copy_probe * 262144, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4}))
        return
    if args.library_bf16:
        print(json.dumps(_library_bf16_ms(*args.library_bf16)))
        return
    if args.gemm_probe:
        gemm_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent incidences would be spread one per GPU with no exchange).", file=sys.stderr)
        sys.exit(2)

    import bem3_emf_qd2_amd as bq
    from bem3_emf_qd2_amd.solver import Model
    L = bq.lib()
    if not bq.device_ok():
        raise SystemExit(f"bench.py: no usable gfx950 device: {L.bq_last_error().decode()}")
    w = WORKLOADS[args.workload]
    if args.steps is None:
        args.steps = w["steps"]
    if args.warmup is None:
        args.warmup = w["warmup"]
    if args.precision is None:
        args.precision = w.get("precision", "mixed" if w["kind"] == "sweep" else "fp64")
    if args.asm_tol is None:
        args.asm_tol = w.get("asm_tol", 1e-12)
    if w["kind"] == "sweep":
        run_sweep(args, w, bq, L)
        return
    if w["kind"] == "serve":
        run_serve(args, w, bq, L)
        return
    probes = _probes(w, w["probes"], w["seed"])
    with tempfile.TemporaryDirectory() as td:
        _scene(w, td)
        m = Model.from_dir(td)
        m.write(Path(td) / "setup.dat")

        def step():
            st = m.solve()
            E, H, did = m.fields(probes, 1)
            return st, E, H

        # the scene repeats one wave, whose resident matrices and factors a solve would reuse:
        # each step starts from a released device session (released outside the timed region),
        # so that it assembles, factorises and solves the incidence anew

        with _QuietStdout():
            fp64_ref = None
            if args.precision != "fp64":
                bq.solve_precision("fp64")
                m.release_device()
                _, Er, Hr = step()
                fp64_ref = np.r_[Er.ravel(), Hr.ravel()]
                print("bench.py: FP64 reference step done", file=sys.stderr, flush=True)
            bq.solve_precision(args.precision)
            for i in range(args.warmup):
                m.release_device()
                st, E0, H0 = step()
                print(f"bench.py: warm-up step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
            L.bq_kernel_timing(1)
            lat = []
            for i in range(args.steps):
                m.release_device()
                t = time.perf_counter()
                st, E, H = step()
                lat.append(time.perf_counter() - t)
                print(f"bench.py: timed step {i + 1}/{args.steps}: {lat[-1]:.2f} s", file=sys.stderr, flush=True)
            elapsed = float(np.sum(lat))
        kt = _kernel_times(L)
        L.bq_kernel_timing(0)
        bq.solve_precision("fp64")
        stats = {"t_assemble_ms": st.t_assemble_ms, "t_solve_ms": st.t_solve_ms, "t_fields_ms": st.t_fields_ms,
                 "n_gh_pairs": st.n_gh_pairs, "n_qpgf": st.n_qpgf, "n_qpgf_gl": st.n_qpgf_gl,
                 "n_qpgf_gh": st.n_qpgf_gh, "n_qpgf_sing": st.n_qpgf_sing}
        mixed_used = bool(st.mixed)
        if mixed_used:
            stats.update({"ir_iterations": st.ir_iterations, "ir_correction": st.ir_correction})
        rel_fp64 = 0.0
        if fp64_ref is not None:
            got = np.r_[E.ravel(), H.ravel()]
            rel_fp64 = float(np.abs(got - fp64_ref).max() / np.abs(fp64_ref).max())
        ref = np.r_[E0.ravel(), H0.ravel()] if args.warmup else None
        rep = None
        if ref is not None:
            got = np.r_[E.ravel(), H.ravel()]
            rep = float(np.abs(got - ref).max() / np.abs(ref).max())
        sys.path.insert(0, str(ROOT / "tests"))
        from datfile import read_dat  # the ex.dat reader (layout only)
        dd = read_dat(Path(td) / "setup.dat")
        tris = int(np.sum(dd["ed"][1:, 3] == 0))
        quads, NN = int(dd["Ne"]) - tris, int(dd["NN"])
        roof = _roofline(kt, stats, args.steps, args.workload)
        cpu = None if args.no_cpu_baseline else _cpu_baseline(Path(td) / "setup.dat", w, probes, None)
        m.close()

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "solved incidences/sec + p50/p99 latency at <=1e-8 rel err vs FP64 plain path",
        "value": args.steps / elapsed,
        "unit": "incidences/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated cube-sphere mesh and plane wave)",
        "config": {"workload": w["label"], "quads": quads, "triangles": tris, "element_nodes": NN, "na": 8 * NN,
                   "probes": w["probes"], "parallelism": "single GPU",
                   "solve_precision": "mixed" if mixed_used else "fp64"},
        "rel_err_vs_fp64": rel_fp64,
        "repeat_rel_diff": rep,
        "phase_ms": stats,
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in sorted(kt.items(), key=lambda x: -x[1][0])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
