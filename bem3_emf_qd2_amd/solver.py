"""ctypes driver for the drop-in C API (read/solve/write and field evaluation).

Thin plumbing over libbem3qd2.so for tests and benchmarks; the computation is the
library's.  DQD2 is held as an opaque buffer of the size the library reports.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from . import lib as _lib

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class Stats(ctypes.Structure):
    _fields_ = [("t_assemble_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double),
                ("t_fields_ms", ctypes.c_double), ("n_gh_pairs", ctypes.c_int), ("n_qpgf", ctypes.c_longlong),
                ("n_qpgf_gl", ctypes.c_longlong), ("n_qpgf_gh", ctypes.c_longlong), ("n_qpgf_sing", ctypes.c_longlong),
                ("mixed", ctypes.c_int), ("ir_iterations", ctypes.c_int), ("ir_correction", ctypes.c_double),
                ("path", ctypes.c_int), ("reused", ctypes.c_int), ("backward_error", ctypes.c_double),
                ("t_factor_ms", ctypes.c_double), ("gmres_iterations", ctypes.c_int), ("qpgf_error", ctypes.c_int),
                ("asm_tol", ctypes.c_double), ("asm_cut", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def L():
    lb = _lib()
    if not getattr(lb, "_solver_declared", False):
        lb.bq_sizeof_dqd2.restype = ctypes.c_int
        for f in ("dat_read", "dat_write"):
            getattr(lb, f).argtypes = [ctypes.c_char_p, ctypes.c_void_p]
        for f in ("initialize_dqd2", "finalize_dqd2", "solve_bieq", "print_dqd2"):
            getattr(lb, f).argtypes = [ctypes.c_void_p]
        lb.read_dqd2.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_void_p]
        lb.output_node_particles.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
        lb.bq_EH_t_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, _D, _D, _I]
        lb.bq_EH_t_bd_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _I, _I, _D, _D, ctypes.c_int, _D, _D]
        lb.bq_EH_mEMP_t_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_int, _D, _D, _I]
        for f in ("EH_mEMP_s", "EH_mEMP_t", "EH_t", "EH_s"):
            getattr(lb, f).argtypes = [_D, _D, _D, ctypes.c_int, ctypes.c_void_p]
        lb.force_FN.argtypes = [_D, _D, _D, ctypes.c_int, ctypes.c_void_p]
        lb.mEMP_s.argtypes = [_D, _D, ctypes.c_int, ctypes.c_void_p]
        lb.mEMP_t.argtypes = [_D, _D, ctypes.c_int, ctypes.c_void_p]
        lb.EH_s_bd.argtypes = [_D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.c_void_p]
        lb.bq_assemble_domain_host.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, _D]
        lb.bq_solve_last_stats.restype = ctypes.POINTER(Stats)
        lb.bq_set_incidence.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 8
        lb.bq_release_solver.argtypes = [ctypes.c_void_p]
        lb.bq_solver_bytes.argtypes = [ctypes.c_void_p]
        lb.bq_solver_bytes.restype = ctypes.c_longlong
        lb.bq_solve_stats.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lb._solver_declared = True
    return lb


class Model:
    """A DQD2 owned by Python; released with close()."""

    def __init__(self):
        self.buf = ctypes.create_string_buffer(L().bq_sizeof_dqd2())
        self.ptr = ctypes.cast(self.buf, ctypes.c_void_p)
        self.live = False

    @classmethod
    def from_dat(cls, path) -> "Model":
        m = cls()
        L().dat_read(str(path).encode(), m.ptr)
        m.live = True
        return m

    @classmethod
    def from_inputs(cls, ipw, per, med, msh, out="out.dat") -> "Model":
        m = cls()
        args = [b"bq", *(str(x).encode() for x in (ipw, per, med, msh, out))]
        argv = (ctypes.c_char_p * len(args))(*args)
        L().read_dqd2(len(args), argv, m.ptr)
        L().initialize_dqd2(m.ptr)
        m.live = True
        return m

    @classmethod
    def from_dir(cls, d, out="out.dat") -> "Model":
        d = Path(d)
        return cls.from_inputs(d / "ipw.txt", d / "periodicity_data.txt", d / "medium_data.txt", d / "mesh.msh",
                               d / out)

    def solve(self) -> Stats:
        L().solve_bieq(self.ptr)
        st = Stats()
        if L().bq_solve_stats(self.ptr, ctypes.byref(st)) != 0:
            raise RuntimeError("solve_bieq left no statistics")
        return st

    def set_incidence(self, lambda0, theta, phi, e0x=1.0, e0y=0.0, power=None):
        """Replace the incident wave (bq_set_incidence); power None means amplitude 1."""
        if power is None:
            power = 0.0
        e0x, e0y = complex(e0x), complex(e0y)
        rc = L().bq_set_incidence(self.ptr, lambda0, theta, phi, e0x.real, e0x.imag, e0y.real, e0y.imag, power)
        if rc:
            raise ValueError(f"bq_set_incidence rejected lambda0={lambda0} theta={theta} phi={phi}")

    def release_device(self) -> None:
        """Free the resident solver session (bq_release_solver)."""
        L().bq_release_solver(self.ptr)

    def device_bytes(self) -> int:
        return int(L().bq_solver_bytes(self.ptr))

    def write(self, path) -> None:
        L().dat_write(str(path).encode(), self.ptr)

    def fields(self, r, type_=1):
        r = np.ascontiguousarray(np.asarray(r, float).reshape(-1, 3))
        n = len(r)
        E = np.zeros((n, 3), complex)
        H = np.zeros((n, 3), complex)
        did = np.zeros(n, np.int32)
        L().bq_EH_t_batch(self.ptr, n, r.ctypes.data_as(_D), type_, E.ctypes.data_as(_D), H.ctypes.data_as(_D),
                          did.ctypes.data_as(_I))
        return E, H, did

    def fields_memp(self, r, type_=1):
        """Total E, H from the modified potential (EH_mEMP_t) at points r[n][3]; ret[n] = domain
        id, or -did (-99 for the open region) for a point too close to the boundary."""
        r = np.ascontiguousarray(np.asarray(r, float).reshape(-1, 3))
        n = len(r)
        E = np.zeros((n, 3), complex)
        H = np.zeros((n, 3), complex)
        ret = np.zeros(n, np.int32)
        rc = L().bq_EH_mEMP_t_batch(self.ptr, n, r.ctypes.data_as(_D), type_, E.ctypes.data_as(_D),
                                    H.ctypes.data_as(_D), ret.ctypes.data_as(_I))
        if rc:
            raise ValueError("bq_EH_mEMP_t_batch: bad arguments")
        return E, H, ret

    def fields_memp_s(self, r, type_=1):
        """Scattered E, H from the modified potential at one point through EH_mEMP_s."""
        r = np.ascontiguousarray(np.asarray(r, float).reshape(3))
        E = np.zeros(3, complex)
        H = np.zeros(3, complex)
        ret = L().EH_mEMP_s(E.ctypes.data_as(_D), H.ctypes.data_as(_D), r.ctypes.data_as(_D), type_, self.ptr)
        return E, H, ret

    def fields_bd(self, did, t, zt, et, type_=1):
        """Total E, H at boundary points (zt[p], et[p]) of the t[p]-th element of domain did[p]."""
        did = np.ascontiguousarray(np.atleast_1d(did), np.int32)
        t = np.ascontiguousarray(np.atleast_1d(t), np.int32)
        zt = np.ascontiguousarray(np.atleast_1d(zt), float)
        et = np.ascontiguousarray(np.atleast_1d(et), float)
        n = len(t)
        E = np.zeros((n, 3), complex)
        H = np.zeros((n, 3), complex)
        rc = L().bq_EH_t_bd_batch(self.ptr, n, did.ctypes.data_as(_I), t.ctypes.data_as(_I), zt.ctypes.data_as(_D),
                                  et.ctypes.data_as(_D), type_, E.ctypes.data_as(_D), H.ctypes.data_as(_D))
        if rc:
            raise ValueError("invalid boundary query")
        return E, H

    def fields_s_bd(self, did, t, zt, et, type_=1):
        """Scattered (or interior) E, H at one boundary point through EH_s_bd."""
        E = np.zeros(3, complex)
        H = np.zeros(3, complex)
        L().EH_s_bd(E.ctypes.data_as(_D), H.ctypes.data_as(_D), did, t, zt, et, type_, self.ptr)
        return E, H

    def potential(self, r, type_=1, total=True):
        """Modified electromagnetic potential U = (A, phi) at one point (mEMP_t / mEMP_s)."""
        r = np.ascontiguousarray(np.asarray(r, float).reshape(3))
        U = np.zeros(4, complex)
        f = L().mEMP_t if total else L().mEMP_s
        did = f(U.ctypes.data_as(_D), r.ctypes.data_as(_D), type_, self.ptr)
        return U, did

    def force(self, rc=(0.0, 0.0, 0.0), type_=0):
        """Force and torque about rc on the objects (force_FN)."""
        F = np.zeros(3)
        N = np.zeros(3)
        c = np.ascontiguousarray(rc, float)
        rcode = L().force_FN(F.ctypes.data_as(_D), N.ctypes.data_as(_D), c.ctypes.data_as(_D), type_, self.ptr)
        if rcode:
            raise RuntimeError("force_FN: domain 0 has no element")
        return F, N

    def assemble(self, d, dNe, deriv=True):
        N = 4 * dNe
        G = np.zeros((N, N), complex)
        H = np.zeros((N, N), complex)
        dG = np.zeros((2 * N, N), complex) if deriv else None
        dH = np.zeros((2 * N, N), complex) if deriv else None
        tdF = np.zeros((N, 3)) if deriv else None
        nul = ctypes.cast(None, _D)
        p = lambda a: a.ctypes.data_as(_D) if a is not None else nul
        rc = L().bq_assemble_domain_host(self.ptr, d, p(G), p(H), p(dG), p(dH), p(tdF))
        if rc:
            raise RuntimeError(f"device assembly failed ({rc}): {_lib().bq_last_error().decode()}")
        return G, H, dG, dH, tdF

    def close(self):
        if self.live:
            L().finalize_dqd2(self.ptr)
            self.live = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def smoke_solve():
    """Small end-to-end solve (24-quad sphere on a square lattice) checked for sanity."""
    import tempfile
    from . import meshgen
    with tempfile.TemporaryDirectory() as td:
        meshgen.sphere_scene(td, 2, 0.3, 1.5, 0.6, theta=0.2)
        m = Model.from_dir(td)
        st = m.solve()
        E, H, did = m.fields([[0.1, 0.2, 0.9]], 1)
        m.close()
        print(f"smoke: 24-quad solve assembly {st.t_assemble_ms:.2f} ms, solve {st.t_solve_ms:.2f} ms, "
              f"E(0.1,0.2,0.9) = {E[0]}")
        if not np.all(np.isfinite(E)):
            raise RuntimeError("smoke: non-finite field")
