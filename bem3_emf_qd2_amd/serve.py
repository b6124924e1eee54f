"""ctypes driver for the request loop (bq_serve_* in include/bq_amd.h) and the seeded request
streams of the serving and sweep workloads.  The scheduling, admission and solves are the
library's (csrc/host/bq_serve.cpp); this module only builds request arrays and reads results.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _lib

_D = ctypes.POINTER(ctypes.c_double)


class Req(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_int), ("lambda0", ctypes.c_double), ("theta", ctypes.c_double),
                ("phi", ctypes.c_double), ("pol", ctypes.c_int), ("arrival_s", ctypes.c_double)]


class Out(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("slot", ctypes.c_int), ("reused", ctypes.c_int),
                ("gmres_iterations", ctypes.c_int), ("t_arrival_s", ctypes.c_double),
                ("t_start_s", ctypes.c_double), ("t_done_s", ctypes.c_double), ("backward_error", ctypes.c_double),
                ("n_qpgf_gl", ctypes.c_longlong), ("asm_cut", ctypes.c_double)]


def _L():
    L = _lib()
    if not getattr(L, "_serve_declared", False):
        L.bq_serve_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int]
        L.bq_serve_create.restype = ctypes.c_void_p
        L.bq_serve_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Req), ctypes.c_int, _D, _D, _D,
                                   ctypes.POINTER(Out)]
        L.bq_serve_destroy.argtypes = [ctypes.c_void_p]
        L._serve_declared = True
    return L


PRECISION = {"fp64": 0, "mixed": 1, "auto": 2}


class Engine:
    """Resident meshes (set-up .dat files) served by nslots worker slots."""

    def __init__(self, setup_files, nslots=2, mem_limit_gb=270.0, precision="mixed"):
        files = [str(f).encode() for f in setup_files]
        arr = (ctypes.c_char_p * len(files))(*files)
        self.h = _L().bq_serve_create(len(files), arr, nslots, mem_limit_gb, PRECISION[precision])
        if not self.h:
            raise ValueError("bq_serve_create: bad arguments")

    def run(self, requests, probes):
        """requests: list of dicts (mesh, lambda0, theta, phi, pol 'TE'/'TM', arrival_s).
        Returns (E, H [nreq][nprobe][3] complex, list of per-request dicts, rc)."""
        n = len(requests)
        R = (Req * max(n, 1))()
        for i, q in enumerate(requests):
            R[i] = Req(q["mesh"], q["lambda0"], q["theta"], q["phi"], 1 if q["pol"] == "TM" else 0,
                       q.get("arrival_s", 0.0))
        P = np.ascontiguousarray(np.asarray(probes, float).reshape(-1, 3))
        E = np.zeros((n, len(P), 3), complex)
        H = np.zeros((n, len(P), 3), complex)
        O = (Out * max(n, 1))()
        rc = _L().bq_serve_run(self.h, n, R, len(P), P.ctypes.data_as(_D), E.ctypes.data_as(_D),
                               H.ctypes.data_as(_D), O)
        outs = [{k: getattr(O[i], k) for k, _ in Out._fields_} for i in range(n)]
        return E, H, outs, rc

    def close(self):
        if self.h:
            _L().bq_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serving_requests(n, rate, seed=5, mesh_p=(0.4, 0.3, 0.2, 0.1), repeat=0.3):
    """The serving workload's request stream: Poisson arrivals at `rate` per second; mesh index
    drawn with probabilities mesh_p; lambda0 ~ U[0.4, 0.8] um, theta ~ U[0, 40] deg, phi ~
    U[0, 90] deg, TE or TM; a fraction `repeat` of the requests (from the second on) repeats an
    earlier request's mesh and direction with the other polarisation."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        if i > 0 and rng.random() < repeat:
            j = int(rng.integers(0, i))
            q = dict(out[j])
            q["pol"] = "TE" if q["pol"] == "TM" else "TM"
            q["repeat_of"] = j
        else:
            q = dict(mesh=int(rng.choice(len(mesh_p), p=mesh_p)), lambda0=float(rng.uniform(0.4, 0.8)),
                     theta=float(np.radians(rng.uniform(0.0, 40.0))), phi=float(np.radians(rng.uniform(0.0, 90.0))),
                     pol="TE" if rng.integers(2) == 0 else "TM")
        q["arrival_s"] = t
        out.append(q)
    return out
