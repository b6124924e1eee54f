"""ctypes handle on oracle/liboracle.so -- the CPU restatement used as the checker.

Test infrastructure only: the product never imports this module.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle" / "liboracle.so"
REF = ROOT / "oracle" / "_ref" / "libqpgf_ref.so"
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)

_o = None


def oracle() -> ctypes.CDLL:
    global _o
    if _o is None:
        if not ORACLE.exists():
            raise RuntimeError("oracle/liboracle.so missing: run make -C oracle")
        o = ctypes.CDLL(str(ORACLE))
        o.or_faddeeva_w.argtypes = [D, ctypes.c_double, D]
        o.or_qpgf_batch.argtypes = [ctypes.c_int, ctypes.c_int, D, ctypes.c_double, D, D]
        o.or_qpgf_batch.restype = ctypes.c_int
        o.or_model_new.restype = ctypes.c_void_p
        o.or_model_new.argtypes = [ctypes.c_int, D, ctypes.c_int, D, D, ctypes.c_int, D, ctypes.c_int, D,
                                   ctypes.c_int, I, I, I]
        o.or_model_free.argtypes = [ctypes.c_void_p]
        o.or_model_export.argtypes = [ctypes.c_void_p, D, D, D, D, D, D]
        o.or_model_dNe.argtypes = [ctypes.c_void_p, ctypes.c_int]
        o.or_model_sid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        o.or_model_NN.argtypes = [ctypes.c_void_p]
        o.or_model_set_nodes.argtypes = [ctypes.c_void_p, D, D, D, D]
        o.or_node_tangents.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, D, D]
        o.or_pair.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, D, D, D]
        o.or_coef_rt.argtypes = [ctypes.c_void_p, ctypes.c_int, D, ctypes.c_int, ctypes.c_int, D]
        o.or_assemble_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D, D, D, D]
        o.or_global_system.argtypes = [ctypes.c_void_p, ctypes.POINTER(D), ctypes.POINTER(D), D, D]
        o.or_store_solution.argtypes = [ctypes.c_void_p, D, ctypes.POINTER(D), ctypes.POINTER(D)]
        o.or_boundary_EH.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(D)] * 7
        o.or_deh_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, D, D, D, D]
        o.or_domain_id.argtypes = [ctypes.c_void_p, D]
        o.or_EH_t.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(D)] * 4 + [D, ctypes.c_int, D, D]
        o.or_EH_bd.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(D)] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, D, D]
        o.or_EH_mEMP.argtypes = [ctypes.c_void_p, ctypes.POINTER(D), ctypes.POINTER(D), D, ctypes.c_int, ctypes.c_int,
                                 D, D]
        o.or_coef_rt_grad.argtypes = [ctypes.c_void_p, ctypes.c_int, D, ctypes.c_int, ctypes.c_int, D, D]
        o.or_mEMP.argtypes = [ctypes.c_void_p, ctypes.POINTER(D), ctypes.POINTER(D), D, ctypes.c_int, ctypes.c_int, D]
        o.or_force_FN.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(D)] * 4 + [D, ctypes.c_int, D, D]
        o.or_coef_bd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, D]
        _o = o
    return _o


def ptr(a: np.ndarray):
    return a.ctypes.data_as(D)


def faddeeva_w(z, relerr=1e-15) -> np.ndarray:
    o = oracle()
    z = np.asarray(z, complex).ravel()
    out = np.empty(len(z), complex)
    buf = np.zeros(2)
    for i, v in enumerate(z):
        zz = np.array([v.real, v.imag])
        o.or_faddeeva_w(ptr(zz), relerr, ptr(buf))
        out[i] = complex(buf[0], buf[1])
    return out


def qpgf(qd8, r, mode=1, eps=1e-15):
    """(n, 10) complex: qG, dqG[3], d2qG[6] (mode 1) from the oracle."""
    o = oracle()
    q = np.ascontiguousarray(qd8, float)
    rr = np.ascontiguousarray(np.asarray(r, float).reshape(-1, 3))
    out = np.zeros((len(rr), 20))
    rc = o.or_qpgf_batch(mode, len(rr), ptr(rr), eps, ptr(q), ptr(out))
    return out[:, 0::2] + 1j * out[:, 1::2], rc
