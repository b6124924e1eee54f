"""Golden vectors for the quasi-periodic Green's function from the reference's own
qgf2_src/d3_qpgf_d2.c, compiled in place by oracle/Makefile into oracle/_ref/libqpgf_ref.so.

Writes qpgf_ref.npz: for each lattice/incidence case, seeded sample points r and the
reference's d3hm_qpgf_d2_v2 output (qG, dqG[3], d2qG[6]) with eps = 1e-15 and its return
code, plus the _v1 output and the _qG / _dqG outputs at the same points.  The _v1 and _dqG
variants stop their spectral sums early when r_z == 0 exactly (their stop test compares a
zero z-derivative with a zero sum), so at the z = 0 sample points they differ from _v2.  Run where oracle/_ref has been built:
    python tests/golden/make_qpgf_golden.py
"""
import ctypes
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = HERE.parent.parent / "oracle" / "_ref" / "libqpgf_ref.so"


class QPD2(ctypes.Structure):
    _fields_ = [("k", ctypes.c_double), ("vk", ctypes.c_double * 3),
                ("vd1", ctypes.c_double * 2), ("vd2", ctypes.c_double * 2)]


def qpd2(lam, n0, th, ph, d1, d2):
    k = 2 * math.pi * n0 / lam
    vk = [k * math.sin(th) * math.cos(ph), k * math.sin(th) * math.sin(ph), k * math.cos(th)]
    vk = [v if v != 0.0 else k * 1e-15 for v in vk[:2]] + [vk[2]]
    return [k, *vk, *d1, *d2]


CASES = {
    # square 1.0 um, lambda 0.6, n0 1, theta ~ 0 (the 384-quad sphere case)
    "sq1_l06": qpd2(0.6, 1.0, 1e-6, 0.0, (1.0, 0.0), (0.0, 1.0)),
    # hexagonal 0.75 um, lambda 1.0 in n0 = 1.33, theta 0.2 (shipped sample 3)
    "hex075": qpd2(1.0, 1.33, 0.2, 0.0, (0.75, 0.0), (0.375, 0.6495190528383290)),
    # hexagonal 2.0 um, many propagating orders (shipped verification case)
    "hex2": qpd2(1.0, 1.33, 0.2, 0.0, (2.0, 0.0), (1.0, 1.7320508075688772)),
    # hexagonal 0.9 um, oblique TM incidence, lambda 0.55
    "hex09_obl": qpd2(0.55, 1.0, math.radians(30), math.radians(15), (0.9, 0.0), (0.45, 0.779)),
    # square 1.0 um, lambda 0.8, theta 40 deg, phi 70 deg (sweep corner)
    "sq1_sweep": qpd2(0.8, 1.0, math.radians(40), math.radians(70), (1.0, 0.0), (0.0, 1.0)),
}


def main():
    lib = ctypes.CDLL(str(REF))
    D = ctypes.POINTER(ctypes.c_double)
    lib.d3hm_qpgf_d2_v2.argtypes = [D, D, D, D, ctypes.c_double, ctypes.POINTER(QPD2)]
    lib.d3hm_qpgf_d2_v2.restype = ctypes.c_int
    lib.d3hm_qpgf_d2_v1.argtypes = [D, D, D, ctypes.c_double, ctypes.POINTER(QPD2)]
    lib.d3hm_qpgf_d2_qG.argtypes = [D, D, ctypes.c_double, ctypes.POINTER(QPD2)]
    lib.d3hm_qpgf_d2_dqG.argtypes = [D, D, ctypes.c_double, ctypes.POINTER(QPD2)]
    rng = np.random.default_rng(20261203)
    out = {}
    for name, q in CASES.items():
        qd = QPD2(q[0], (ctypes.c_double * 3)(*q[1:4]), (ctypes.c_double * 2)(*q[4:6]), (ctypes.c_double * 2)(*q[6:8]))
        R = np.concatenate([
            rng.uniform(-1.0, 1.0, (160, 3)),                 # generic
            rng.normal(0.0, 0.03, (24, 3)),                   # near-singular
            np.c_[rng.uniform(-1, 1, (16, 2)), np.zeros(16)],  # z = 0 plane
            np.c_[rng.uniform(-2, 2, (16, 2)), rng.choice([-1, 1], 16) * rng.uniform(1.5, 3.0, 16)],  # far in z
        ])
        vals = np.zeros((len(R), 10), complex)
        v1 = np.zeros((len(R), 4), complex)      # _v1: qG, dqG
        vq = np.zeros((len(R), 4), complex)      # _qG value, _dqG gradient
        codes = np.zeros(len(R), np.int32)
        for i, p in enumerate(R):
            pr = (ctypes.c_double * 3)(*p)
            g = (ctypes.c_double * 2)(); d = (ctypes.c_double * 6)(); h = (ctypes.c_double * 12)()
            codes[i] = lib.d3hm_qpgf_d2_v2(g, d, h, pr, 1e-15, ctypes.byref(qd))
            vals[i, 0] = complex(g[0], g[1])
            vals[i, 1:4] = np.array(d[0::2]) + 1j * np.array(d[1::2])
            vals[i, 4:10] = np.array(h[0::2]) + 1j * np.array(h[1::2])
            lib.d3hm_qpgf_d2_v1(g, d, pr, 1e-15, ctypes.byref(qd))
            v1[i, 0] = complex(g[0], g[1])
            v1[i, 1:4] = np.array(d[0::2]) + 1j * np.array(d[1::2])
            lib.d3hm_qpgf_d2_qG(g, pr, 1e-15, ctypes.byref(qd))
            lib.d3hm_qpgf_d2_dqG(d, pr, 1e-15, ctypes.byref(qd))
            vq[i, 0] = complex(g[0], g[1])
            vq[i, 1:4] = np.array(d[0::2]) + 1j * np.array(d[1::2])
        out[name + "_qd"] = np.array(q)
        out[name + "_r"] = R
        out[name + "_val"] = vals
        out[name + "_code"] = codes
        out[name + "_v1"] = v1
        out[name + "_vq"] = vq
    np.savez_compressed(HERE / "qpgf_ref.npz", names=np.array(list(CASES)), **out)
    print("cases", list(CASES), "points per case", len(R))


if __name__ == "__main__":
    main()
