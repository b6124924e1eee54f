"""BIE-row residuals of a solved model at full size: the oracle's rows against the GPU's solution.

    python tests/check_rows.py MODEL.dat [--block 8] [--seed 7]

MODEL.dat is a solved model in the ex.dat layout (e.g. `bench.py --save-dat DIR` writes the
first timed point of a sweep).  For every domain a seeded block of `block` consecutive
source-side elements is chosen; the oracle (tests/oracle_pipeline.py, a CPU restatement of
the reference's element routines with the reference's 12-point rule on every regular pair)
assembles their 4 * block collocation rows of G_d and H_d, and each row is applied to the
stored solution:

    H_d U_d - G_d dU_d,   G_d dE_d - H_d E_d,   G_d dH_d - H_d H_d

(the rows of create_matrix_csr_dac / solve_deh_bv, d3qd2_src/d3qd2_solve_bieq.c:284-421,
:662-791).  The residual of each row relative to the sum of the magnitudes of its terms is
reported; the GPU's FP64 path (adaptive Gauss orders bounded at 1e-15, the reference's
lattice-sum cut) gives a few 1e-15, a wrong coefficient anywhere in a row gives the size of
that element's term.  Only the sampled rows are held in memory (a full matrix of the
6144-quad model would take 9.7 GB).  Test infrastructure: nothing in the library imports it.
"""
import argparse
import ctypes
import json
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent))
from datfile import read_dat              # noqa: E402
from oracle_lib import D                  # noqa: E402
from oracle_pipeline import OracleModel   # noqa: E402


def rows(o, did, t0, t1):
    """G, H rows 4 (t0 - 1) .. 4 t1 - 1 of domain did (complex, rows x N)."""
    N = 4 * o.dNe[did]
    r0, nr = 4 * (t0 - 1), 4 * (t1 - t0 + 1)
    G = np.zeros((nr, N), complex)
    H = np.zeros((nr, N), complex)
    # the oracle writes row r at offset r * N of the matrix it is given: hand it base
    # pointers shifted back by r0 rows so that only this block's rows are stored
    sh = lambda a: ctypes.cast(ctypes.c_void_p(a.ctypes.data - r0 * N * 16), D)
    nul = ctypes.cast(None, D)
    o.o.or_assemble_rows(o.h, did, t0, t1, sh(G), sh(H), nul, nul, nul)
    return G, H


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dat")
    ap.add_argument("--block", type=int, default=8)
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args()
    d = read_dat(a.dat)
    o = OracleModel.from_dat(d)
    rng = np.random.default_rng(a.seed)
    out = {"model": str(a.dat), "Ne": int(d["Ne"]), "NN": int(o.NN), "domains": []}
    for did in range(o.MN + 1):
        ne = o.dNe[did]
        t0 = int(rng.integers(1, max(ne - a.block + 1, 1) + 1))
        t1 = min(ne, t0 + a.block - 1)
        t = time.time()
        G, H = rows(o, did, t0, t1)
        dt = time.time() - t
        dom = d["dom"][did]
        U, dU = dom["U"][1:].reshape(-1, 4), dom["dU"][1:].reshape(-1, 4)
        E, dE = dom["E"][1:].reshape(-1, 3), dom["dE"][1:].reshape(-1, 3)
        Hf, dHf = dom["H"][1:].reshape(-1, 3), dom["dH"][1:].reshape(-1, 3)
        live = o.live_rows(did)
        worst = {"U": 0.0, "E": 0.0, "H": 0.0}
        nrows = 0
        for i, r in enumerate(range(4 * (t0 - 1), 4 * t1)):
            if not live[r]:
                continue
            nrows += 1
            g, h = G[i], H[i]
            for key, x, dx, sgn in (("U", U, dU, 1), ("E", E, dE, -1), ("H", Hf, dHf, -1)):
                res = sgn * (h @ x - g @ dx)
                scale = (np.abs(h)[:, None] * np.abs(x) + np.abs(g)[:, None] * np.abs(dx)).sum(0)
                worst[key] = max(worst[key], float((np.abs(res) / scale).max()))
        rec = {"domain": did, "elements": [t0, t1], "rows": nrows, "oracle_s": round(dt, 1),
               "worst_rel_residual": worst}
        print(json.dumps(rec), flush=True)
        out["domains"].append(rec)
    out["worst"] = max(max(r["worst_rel_residual"].values()) for r in out["domains"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
