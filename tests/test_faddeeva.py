"""Faddeeva w(z): oracle pinned against the reference package's own known-answer table
(qgf2_src/Faddeeva.cc:1916-2060) and 40-digit mpmath values; the GPU kernel against both.

Tolerance: |w - w_ref| <= 1e-13 |w_ref| (the Ewald sums need ~1e-15 of |w|; the gap covers
the inherent loss of the y < 0 reflection w(z) = 2 exp(-z^2) - w(-z))."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib

G = Path(__file__).resolve().parent / "golden"
TOL = 1e-13


def _kat():
    rows = json.loads((G / "faddeeva_kat.json").read_text())
    z = np.array([complex(*r["z"]) for r in rows])
    w = np.array([complex(*r["w"]) for r in rows])
    return z, w


def _grid():
    d = np.load(G / "faddeeva_grid.npz")
    return d["z"][:, 0] + 1j * d["z"][:, 1], d["w_re"] + 1j * d["w_im"]


def _relerr(a, b):
    return np.abs(a - b) / np.abs(b)


def test_oracle_kat():
    z, w = _kat()
    # the table spans |z| up to 1e160; keep the finite-precision-meaningful ones
    keep = np.abs(z) < 1e7
    e = _relerr(oracle_lib.faddeeva_w(z[keep]), w[keep])
    assert e.max() < TOL, (z[keep][e.argmax()], e.max())


def test_oracle_grid_mpmath():
    z, w = _grid()
    e = _relerr(oracle_lib.faddeeva_w(z), w)
    assert e.max() < TOL, (z[e.argmax()], e.max())


def test_oracle_symmetry():
    # w(-conj z) = conj w(z); planted-mistake check: a sign slip in Im w breaks this
    rng = np.random.default_rng(3)
    z = rng.uniform(-6, 6, 200) + 1j * rng.uniform(0, 6, 200)
    a = oracle_lib.faddeeva_w(z)
    b = oracle_lib.faddeeva_w(-np.conj(z))
    assert np.max(np.abs(b - np.conj(a)) / np.abs(a)) < 1e-14


@pytest.mark.gpu
def test_gpu_kat(gpu_lib):
    z, w = _kat()
    keep = np.abs(z) < 1e7
    e = _relerr(gpu_lib.faddeeva_w(z[keep]), w[keep])
    print("faddeeva gpu kat max rel", e.max())
    assert e.max() < TOL


@pytest.mark.gpu
def test_gpu_grid_mpmath(gpu_lib):
    z, w = _grid()
    e = _relerr(gpu_lib.faddeeva_w(z), w)
    print("faddeeva gpu grid max rel", e.max())
    assert e.max() < TOL, (z[e.argmax()], e.max())


@pytest.mark.gpu
def test_gpu_matches_oracle_dense(gpu_lib):
    rng = np.random.default_rng(11)
    z = rng.uniform(0, 8, 4000) + 1j * rng.uniform(-7, 9, 4000)
    a = gpu_lib.faddeeva_w(z)
    b = oracle_lib.faddeeva_w(z)
    assert np.max(_relerr(a, b)) < TOL


# ---- erfcx by table (the assembly's lattice sums) ------------------------------------------
DEV = Path(__file__).resolve().parent.parent / "bem3_emf_qd2_amd" / "csrc" / "dev"


def _erfcx_points():
    t0 = 4.0 / 54.0
    tk = t0 + (1.0 - t0) * np.arange(65) / 64.0  # interval ends in t = 4/(4+x)
    edges = 4.0 / tk - 4.0
    x = np.concatenate([np.linspace(0.0, 12.0, 1201), np.linspace(12.0, 60.0, 241), edges,
                        edges * (1 + 1e-12), [50.0, 50.000001, 1e3, 1e8]])
    return np.concatenate([x, -np.linspace(0.01, 5.9, 60)])


def _erfcx_mp(x):
    import mpmath as mp
    mp.mp.dps = 40
    return np.array([float(mp.exp(mp.mpf(v) ** 2) * mp.erfc(mp.mpf(v))) for v in x])


def _erfcx_table_np(x):
    """the device evaluator (bq_faddeeva.hpp erfcx_tab / erfcx_any) restated in numpy on the
    committed coefficients"""
    rows = [r for r in (DEV / "bq_erfcx_table.inc").read_text().splitlines() if not r.startswith("//")]
    c = np.array([[float(v) for v in r.strip().rstrip(",").split(",")] for r in rows if r.strip()])
    assert c.shape == (64, 8)

    def pos(a):
        if a > 50.0:
            r = 1.0 / a
            r2 = 0.5 * r * r
            return r / np.sqrt(np.pi) * (1 - r2 * (1 - 3 * r2 * (1 - 5 * r2 * (1 - 7 * r2 * (1 - 9 * r2)))))
        t0 = 4.0 / 54.0
        s = (4.0 / (4.0 + a) - t0) * (64.0 / (1.0 - t0))
        k = max(0, min(63, int(s)))
        u = 2.0 * (s - k) - 1.0
        h = c[k, 7]
        for p in range(6, -1, -1):
            h = h * u + c[k, p]
        return h
    return np.array([pos(a) if a >= 0 else 2.0 * np.exp(a * a) - pos(-a) for a in x])


def test_erfcx_table_against_mpmath():
    """The committed coefficient table evaluated as the device evaluates it: within 2e-15 of
    40-digit erfcx on [0, 60] (interval ends and both sides of x = 50 included) and 1e-14 for
    x < 0 (the reflection 2 exp(x^2) - erfcx(-x))."""
    x = _erfcx_points()
    e = np.abs(_erfcx_table_np(x) - _erfcx_mp(x)) / _erfcx_mp(x)
    print("erfcx table max rel err", e[x >= 0].max(), e[x < 0].max())
    assert e[x >= 0].max() < 2e-15
    assert e[x < 0].max() < 1e-14


def test_erfcx_generator_reproduces_table(capsys):
    """gen_erfcx_table.py writes exactly the committed table."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_erfcx_table", DEV / "gen_erfcx_table.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    capsys.readouterr()
    mod.main()
    assert capsys.readouterr().out == (DEV / "bq_erfcx_table.inc").read_text()


@pytest.mark.gpu
def test_gpu_erfcx_table(gpu_lib):
    """bq_erfcx_batch (the LDS-table erfcx of the lattice sums) against 40-digit values and
    against the numpy restatement of the same table: 3e-15 (x >= 0), 1e-14 (x < 0)."""
    import ctypes
    L = gpu_lib.lib()
    x = _erfcx_points()
    out = np.zeros_like(x)
    P = ctypes.POINTER(ctypes.c_double)
    L.bq_erfcx_batch.argtypes = [ctypes.c_int, P, P]
    assert L.bq_erfcx_batch(len(x), x.ctypes.data_as(P), out.ctypes.data_as(P)) == 0
    ref = _erfcx_mp(x)
    e = np.abs(out - ref) / ref
    d = np.abs(out - _erfcx_table_np(x)) / ref
    print("gpu erfcx max rel err", e[x >= 0].max(), e[x < 0].max(), "vs numpy restatement", d.max())
    assert e[x >= 0].max() < 3e-15
    assert e[x < 0].max() < 1e-14
    assert d.max() < 3e-15
