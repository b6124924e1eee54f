"""Python handle on libbem3qd2.so, the MI355X (gfx950) implementation of the
bem3_emf_qd2 boundary-integral path.

The product is the C-ABI library (include/bem3_emf_qd2.h, include/bq_amd.h).  This module
only loads it with ctypes and wraps the batch entry points with numpy conversions; it
computes nothing itself and has no CPU fallback: every call fails loudly (RuntimeError)
when the library is not built or no gfx950 device is usable.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libbem3qd2.so"

_D = ctypes.POINTER(ctypes.c_double)
_lib = None


def lib() -> ctypes.CDLL:
    """Load the built library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.bq_last_error.restype = ctypes.c_char_p
        _declare(_lib)
    return _lib


def _declare(L: ctypes.CDLL) -> None:
    L.bq_device_ok.restype = ctypes.c_int
    L.bq_faddeeva_w_batch.argtypes = [ctypes.c_int, _D, _D]
    L.bq_qpgf_v1_batch.argtypes = [_D, ctypes.c_int, _D, _D]
    L.bq_qpgf_batch.argtypes = [_D, ctypes.c_int, ctypes.c_int, _D, _D]
    L.bq_zgesv_host.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D]
    L.bq_cgesv_host.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float)]
    L.bq_solve_precision.argtypes = [ctypes.c_int]
    L.bq_solve_precision.restype = ctypes.c_int


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib().bq_last_error().decode()}")


def device_ok() -> bool:
    return lib().bq_device_ok() == 0


def _as(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def faddeeva_w(z) -> np.ndarray:
    """w(z) on the GPU for an array of complex z."""
    z = np.asarray(z, complex).ravel()
    zz = _as(np.stack([z.real, z.imag], 1))
    w = np.empty_like(zz)
    _check(lib().bq_faddeeva_w_batch(len(z), zz.ctypes.data_as(_D), w.ctypes.data_as(_D)), "faddeeva_w")
    return w[:, 0] + 1j * w[:, 1]


def qpgf_v1(qd8, r, converged=False) -> np.ndarray:
    """Quasi-periodic Green's function and gradient, shape (n, 4) complex.

    converged=False reproduces d3hm_qpgf_d2_v1 (whose spectral sum stops early at r_z == 0);
    converged=True gives the converged sum there, as d3hm_qpgf_d2_qG does for the value."""
    q = _as(qd8)
    assert q.shape == (8,)
    rr = _as(np.asarray(r).reshape(-1, 3))
    out = np.empty((len(rr), 8))
    _check(lib().bq_qpgf_batch(q.ctypes.data_as(_D), int(bool(converged)), len(rr), rr.ctypes.data_as(_D),
                               out.ctypes.data_as(_D)), "qpgf")
    return out[:, 0::2] + 1j * out[:, 1::2]


def qpgf_v2(qd8, r) -> np.ndarray:
    """d3hm_qpgf_d2_v2 on the GPU: shape (n, 10) complex = qG, dqG[3], d2qG[6] (xx, yy, zz, xy,
    yz, zx)."""
    q = _as(qd8)
    assert q.shape == (8,)
    rr = _as(np.asarray(r).reshape(-1, 3))
    out = np.empty((len(rr), 20))
    _check(lib().bq_qpgf_batch(q.ctypes.data_as(_D), 2, len(rr), rr.ctypes.data_as(_D), out.ctypes.data_as(_D)),
           "qpgf_v2")
    return out[:, 0::2] + 1j * out[:, 1::2]


def zgesv(A, B):
    """Solve A X = B on the GPU with the library's LU (complex double); returns X."""
    A = np.asarray(A, complex)
    B = np.asarray(B, complex)
    vec = B.ndim == 1
    B2 = B.reshape(len(B), -1)
    n, nrhs = A.shape[0], B2.shape[1]
    Af = np.asfortranarray(A).copy(order="F")
    Bf = np.asfortranarray(B2).copy(order="F")
    rc = lib().bq_zgesv_host(n, nrhs, Af.ctypes.data_as(_D), Bf.ctypes.data_as(_D))
    if rc != 0:
        raise RuntimeError(f"zgesv failed ({rc}): {lib().bq_last_error().decode()}")
    return Bf[:, 0] if vec else Bf


def cgesv(A, B):
    """Solve A X = B on the GPU with the library's LU in complex single precision (the
    correction solver of the mixed-precision solve); returns X as complex64."""
    A = np.asarray(A, np.complex64)
    B = np.asarray(B, np.complex64)
    vec = B.ndim == 1
    B2 = B.reshape(len(B), -1)
    n, nrhs = A.shape[0], B2.shape[1]
    Af = np.asfortranarray(A).copy(order="F")
    Bf = np.asfortranarray(B2).copy(order="F")
    _F = ctypes.POINTER(ctypes.c_float)
    rc = lib().bq_cgesv_host(n, nrhs, Af.ctypes.data_as(_F), Bf.ctypes.data_as(_F))
    if rc != 0:
        raise RuntimeError(f"cgesv failed ({rc}): {lib().bq_last_error().decode()}")
    return Bf[:, 0] if vec else Bf


def solve_precision(mode=None):
    """Query (mode None) or set the precision of the global solve: "fp64" (complex FP64 LU,
    the default), "mixed" (complex FP32 LU + FP64 iterative refinement) or "auto" (mixed for
    block systems of order 2 NN >= 16384, FP64 below)."""
    m = {None: -1, "fp64": 0, "mixed": 1, "auto": 2}[mode]
    return ["fp64", "mixed", "auto"][lib().bq_solve_precision(m)]


def assembly_tolerance(tol=None) -> float:
    """Query (None) or set the assembly tolerance (bq_assembly_tolerance): 0 = the reference's
    accuracy; 0 < tol < 1e-3 = matrices to about tol of their largest entries, with the
    lattice-sum cut tuned per incidence."""
    L = lib()
    L.bq_assembly_tolerance.argtypes = [ctypes.c_double]
    L.bq_assembly_tolerance.restype = ctypes.c_double
    return float(L.bq_assembly_tolerance(-1.0 if tol is None else float(tol)))


def lu_update_mode(mode=None) -> str:
    """Query (None) or set the update arithmetic of the complex FP32 factorisation:
    "bf16x3" (split bf16 products on the matrix cores, the default) or "fp32" (rocBLAS
    cgemm).  Returns the mode in force before the call."""
    m = {None: -1, "fp32": 0, "bf16x3": 1}[mode]
    return ["fp32", "bf16x3"][lib().bq_lu_update_mode(m)]


def lu_lookahead(on=None) -> bool:
    """Query (None) or switch (True / False) the look-ahead of the dense factorisations
    (bq_lu_lookahead).  Returns the setting in force before the call."""
    return bool(lib().bq_lu_lookahead(-1 if on is None else int(bool(on))))
