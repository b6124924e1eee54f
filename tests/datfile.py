"""Reader for the reference's ex.dat result files (d3qd2_src/d3qd2_setup.c:247-298
dat_write), x86-64 layout.  Test infrastructure: used to pull mesh, inputs and the
reference's surface solution out of the shipped .dat files and to read the files the
drop-in library writes.
"""
from __future__ import annotations

import numpy as np

C16 = np.complex128


class _R:
    def __init__(self, b: bytes):
        self.b, self.o = b, 0

    def take(self, dtype, count=1):
        a = np.frombuffer(self.b, dtype=dtype, count=count, offset=self.o)
        self.o += a.nbytes
        return a

    def i(self):
        return int(self.take(np.int32)[0])


IPW_FIELDS = ["lambda0", "ni", "power", "e0x", "e0y", "fx", "fy", "fz", "theta", "phi"]


def read_dat(path) -> dict:
    b = open(path, "rb").read()
    r = _R(b)
    d = {}
    d["ped_fn"], d["med_fn"], d["msh_fn"] = [r.take(np.uint8, 128).tobytes().split(b"\0")[0].decode() for _ in range(3)]
    d["rv"] = r.take(np.float64, 3).copy()
    d["th"] = float(r.take(np.float64)[0])
    d["tv"] = r.take(np.float64, 3).copy()
    MN = d["MN"] = r.i()
    d["n"] = r.take(C16, MN + 1).copy()
    d["kn"] = r.take(C16, MN + 1).copy()
    mipw = r.take(np.uint8, 104).tobytes()
    d["fn_ipw"] = mipw[:64].split(b"\0")[0].decode()
    n_ipw = d["n_ipw"] = int(np.frombuffer(mipw, np.int32, 1, 64)[0])
    d["n_0"], d["lambda_0"], d["omega"] = np.frombuffer(mipw, np.float64, 3, 72)
    ipws = []
    for _ in range(n_ipw):
        raw = r.take(np.uint8, 240).tobytes()
        f = np.frombuffer(raw, np.float64)
        ipws.append(dict(lambda0=f[0], ni=f[1], power=f[2], e0x=complex(f[3], f[4]), e0y=complex(f[5], f[6]),
                         fx=f[7], fy=f[8], fz=f[9], theta=f[10], phi=f[11], raw=raw))
    d["ipw"] = ipws
    d["ps"] = r.i()
    d["qd"] = r.take(np.float64, 8).copy()
    Nn, Ne = d["Nn"], d["Ne"] = r.i(), r.i()
    d["rn"] = r.take(np.float64, 3 * (Nn + 1)).reshape(Nn + 1, 3).copy()
    d["ed"] = r.take(np.int32, 4 * (Ne + 1)).reshape(Ne + 1, 4).copy()
    d["NN"] = r.i()
    d["eni"] = r.take(np.int32, 4 * (Ne + 1)).reshape(Ne + 1, 4).copy()
    d["md"] = r.take(np.int32, Ne + 1).copy()
    d["sd"] = r.take(np.int32, Ne + 1).copy()
    d["gd"] = r.take(np.int32, Ne + 1).copy()
    d["ren"] = r.take(np.float64, 12 * (Ne + 1)).reshape(Ne + 1, 4, 3).copy()
    d["wen"] = r.take(np.float64, 12 * (Ne + 1)).reshape(Ne + 1, 4, 3).copy()
    d["Ui"] = r.take(C16, 16 * (Ne + 1)).reshape(Ne + 1, 4, 4).copy()
    d["dUi"] = r.take(C16, 16 * (Ne + 1)).reshape(Ne + 1, 4, 4).copy()
    dom = []
    for _ in range(MN + 1):
        ne = r.i()
        s = dict(Ne=ne, sid=r.take(np.int32, ne + 1).copy())
        for key, c in (("U", 4), ("dU", 4), ("E", 3), ("dE", 3), ("H", 3), ("dH", 3)):
            s[key] = r.take(C16, c * 4 * (ne + 1)).reshape(ne + 1, 4, c).copy()
        dom.append(s)
    d["dom"] = dom
    assert r.o == len(b), (r.o, len(b))
    return d


def ipw_params(d: dict) -> np.ndarray:
    """12 doubles per wave, as the oracle's or_model_new expects."""
    rows = []
    for w in d["ipw"]:
        rows.append([w["lambda0"], w["ni"], w["power"], w["e0x"].real, w["e0x"].imag, w["e0y"].real,
                     w["e0y"].imag, w["fx"], w["fy"], w["fz"], w["theta"], w["phi"]])
    return np.array(rows, float).ravel()
