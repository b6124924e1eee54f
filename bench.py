"""Benchmark: solved incidences per second of the plain FP64 GPU path.

One step = one solved incidence, the work the reference's d3qd2_bv_solver + example1 do for
one plane wave (d3qd2_src/d3qd2_solve_bieq.c:36-84, d3qd2_src/d3qd2_field.c:271): assembly
of every domain's coefficient matrices, the global solve, boundary E/H and dE/dH, then E and
H at 100 probe points (EH_t, type 1).  The mesh and incidence stay resident; each step
recomputes everything above from them.

Default workload (c2, the scene the metric is quoted on): in one hexagonal lattice cell
(a1 = (0.9, 0), a2 = (0.45, 0.779) um) a two-layer cube-sphere (core r 0.3 um, n 1.6 + 0.01i;
shell n 1.45; 6x16x16 = 1536 bilinear quads per surface) and a right circular cone (base
radius 0.3 um, height 0.6 um, n 2.0) of 2012 linear triangles with a seeded 1 %-of-edge jitter
of its non-feature vertices; oblique incidence theta = 30 deg, phi = 15 deg, TM, lambda0 =
0.55 um; 100 probes uniform in the cell with z in [-1, 1] um from default_rng(1).  The shell
radius is 0.40 um, not 0.45: on this lattice |a2| = 0.8996 um, so a 0.45 um shell would
overlap its own periodic image; the cone stands on z = 0.5 um above the sphere.
Workload c1: the 384-quad cube-sphere of radius 0.4 um, index 1.5, in vacuum on a 1.0 um
square lattice, lambda0 = 0.6 um, x-polarised, normal incidence (theta = 1e-6 rad, the
reference rejects theta = 0 on a lattice), same probes.

Prints ONE JSON line.  `roofline` describes the kernel with the largest share of device time
(HIP events on the library stream around every launch); `cpu_baseline` times the CPU
restatement of the same path (oracle/) on a bounded sample of rows, scaled to one incidence.
"""
from __future__ import annotations

import argparse
import json
import os
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
FLOPS_PER_EVAL = {"c1": {"k_q0_gl": 55999, "k_q0_gh": 51625, "k_q0_sing": 50274},
                  "c2": {"k_q0_gl": 51253, "k_q0_gh": 50007, "k_q0_sing": 50175}}
FLOPS_PER_EVAL["c2s"] = FLOPS_PER_EVAL["c2"]
# HBM bytes per evaluation of k_q0_gl: (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00) KiB x 1024 of one
# launch over its evaluations (factors from a 1 GiB copy read with the same counters)
TRAFFIC_PER_EVAL = {"c1": {"k_q0_gl": (1163230.7 * 2.0 + 9545073.7) * 1024 / 46912896},
                    "c2": {"k_q0_gl": (12430181.2 * 2.0 + 72190715.0) * 1024 / 385592752}}
TRAFFIC_PER_EVAL["c2s"] = TRAFFIC_PER_EVAL["c2"]

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


def _scene(w, d):
    from bem3_emf_qd2_amd import meshgen
    if w["kind"] == "sphere_cone":
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
    """Roofline object for the kernel with the largest total device time."""
    name = max(ktimes, key=lambda k: ktimes[k][0])
    tot_ms, cnt = ktimes[name]
    avg_s = tot_ms / cnt / 1e3
    r = {"kernel": name, "bound": "mfma", "pipe": "fp64-valu", "unit": "TFLOP/s", "peak": FP64_PEAK_TFS,
         "avg_launch_ms": avg_s * 1e3, "launches_per_step": cnt / steps, "traffic": None, "achieved": None,
         "frac": None}
    if name == "zgetrf_update":
        r["note"] = "rocBLAS zgemm trailing updates of the blocked LU"
        return r
    # QPGF evaluations per launch, as the device counted them from the pair classification
    evals = {"k_q0_gl": stats_per_step["n_qpgf_gl"], "k_q0_gh": stats_per_step["n_qpgf_gh"],
             "k_q0_sing": stats_per_step["n_qpgf_sing"]}
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
    return r


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default per workload: c2 2, c1 10)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default per workload: c2 1, c1 2)")
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "auto"],
                    help="global solve: the plain complex FP64 LU ('fp64', default: the C1 and C2 "
                         "configurations are the plain FP64 path), complex FP32 LU + FP64 iterative "
                         "refinement ('mixed'; one untimed FP64 step first gives rel_err_vs_fp64), or "
                         "'auto': mixed for block systems of order >= 16384")
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
    probes = _probes(w, w["probes"], w["seed"])
    with tempfile.TemporaryDirectory() as td:
        _scene(w, td)
        m = Model.from_dir(td)
        m.write(Path(td) / "setup.dat")

        def step():
            st = m.solve()
            E, H, did = m.fields(probes, 1)
            return st, E, H

        with _QuietStdout():
            fp64_ref = None
            if args.precision != "fp64":
                bq.solve_precision("fp64")
                _, Er, Hr = step()
                fp64_ref = np.r_[Er.ravel(), Hr.ravel()]
                print("bench.py: FP64 reference step done", file=sys.stderr, flush=True)
            bq.solve_precision(args.precision)
            for i in range(args.warmup):
                st, E0, H0 = step()
                print(f"bench.py: warm-up step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
            L.bq_kernel_timing(1)
            lat = []
            t_all = time.perf_counter()
            for i in range(args.steps):
                t = time.perf_counter()
                st, E, H = step()
                lat.append(time.perf_counter() - t)
                print(f"bench.py: timed step {i + 1}/{args.steps}: {lat[-1]:.2f} s", file=sys.stderr, flush=True)
            elapsed = time.perf_counter() - t_all
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
