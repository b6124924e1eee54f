"""End-to-end solve with the oracle: the reference's solve_bieq pipeline
(d3qd2_src/d3qd2_solve_bieq.c:36-84) on the CPU, dense LAPACK standing in for PARDISO.

Test infrastructure only.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import scipy.linalg as sla

from oracle_lib import D, oracle, ptr
from datfile import ipw_params


class OracleModel:
    def __init__(self, MN, n_sub, ps, vd1, vd2, ipw, rn, ed, md, sd):
        o = oracle()
        self.o = o
        n_ri = np.ascontiguousarray(np.c_[np.real(n_sub), np.imag(n_sub)].ravel(), float)
        self.keep = [np.ascontiguousarray(x) for x in (n_ri, np.asarray(vd1, float), np.asarray(vd2, float),
                                                        np.asarray(ipw, float), np.asarray(rn, float))]
        self.ed = np.ascontiguousarray(ed, np.int32)
        self.md = np.ascontiguousarray(md, np.int32)
        self.sd = np.ascontiguousarray(sd, np.int32)
        Ip = ctypes.POINTER(ctypes.c_int)
        self.h = o.or_model_new(MN, ptr(self.keep[0]), ps, ptr(self.keep[1]), ptr(self.keep[2]),
                                len(self.keep[3]) // 12, ptr(self.keep[3]), len(rn) - 1, ptr(self.keep[4]),
                                len(ed) - 1, self.ed.ctypes.data_as(Ip), self.md.ctypes.data_as(Ip),
                                self.sd.ctypes.data_as(Ip))
        self.MN, self.Ne = MN, len(ed) - 1
        self.dNe = [o.or_model_dNe(self.h, d) for d in range(MN + 1)]
        self.NN = o.or_model_NN(self.h)
        self.tri = self.ed[:, 3] == 0

    @classmethod
    def from_dat(cls, d, stored_nodes=False):
        """Model of a result file.  stored_nodes: take the element node points, normals and
        incident field from the file (as the library's dat_read does) instead of recomputing
        them from the mesh; they can differ in the last bit from this restatement's own."""
        q = d["qd"]
        m = cls(d["MN"], d["n"][1:], d["ps"], q[4:6], q[6:8], ipw_params(d), d["rn"], d["ed"], d["md"], d["sd"])
        if stored_nodes:
            keep = [np.ascontiguousarray(d[k]) for k in ("ren", "wen", "Ui", "dUi")]
            m.o.or_model_set_nodes(m.h, ptr(keep[0]), ptr(keep[1]), ptr(keep[2].view(float)), ptr(keep[3].view(float)))
        return m

    def __del__(self):
        try:
            self.o.or_model_free(self.h)
        except Exception:
            pass

    def export(self):
        Ne = self.Ne
        ren = np.zeros((Ne + 1, 4, 3)); wen = np.zeros((Ne + 1, 4, 3))
        Ui = np.zeros((Ne + 1, 4, 4), complex); dUi = np.zeros((Ne + 1, 4, 4), complex)
        qd = np.zeros(8); kn = np.zeros(self.MN + 1, complex)
        self.o.or_model_export(self.h, ptr(ren), ptr(wen), ptr(Ui.view(float)), ptr(dUi.view(float)), ptr(qd),
                               ptr(kn.view(float)))
        return dict(ren=ren, wen=wen, Ui=Ui, dUi=dUi, qd=qd, kn=kn)

    def pair(self, did, t, tn, s, type_=2):
        CC = np.zeros(9, complex); dz = np.zeros(9, complex); de = np.zeros(9, complex)
        rc = self.o.or_pair(self.h, did, t, tn, s, type_, ptr(CC.view(float)), ptr(dz.view(float)), ptr(de.view(float)))
        return CC, dz, de, rc

    def sid(self, d):
        return np.array([0] + [self.o.or_model_sid(self.h, d, i) for i in range(1, self.dNe[d] + 1)], np.int32)

    def live_rows(self, did):
        """Rows/columns of the 4-per-element domain layout that hold a node (a triangle's
        fourth slot is empty; solve_deh_bv d3qd2_src/d3qd2_solve_bieq.c:680-760 drops it)."""
        sid = np.abs(self.sid(did)[1:])
        keep = np.ones((len(sid), 4), bool)
        keep[self.tri[sid], 3] = False
        return keep.ravel()

    def assemble(self, did, t0=1, t1=None, deriv=True):
        N = 4 * self.dNe[did]
        t1 = self.dNe[did] if t1 is None else t1
        G = np.zeros((N, N), complex); H = np.zeros((N, N), complex)
        dG = np.zeros((N, 2, N), complex) if deriv else None
        dH = np.zeros((N, 2, N), complex) if deriv else None
        dF = np.zeros((N, 3)) if deriv else None
        nul = ctypes.cast(None, D)
        self.o.or_assemble_rows(self.h, did, t0, t1, ptr(G.view(float)), ptr(H.view(float)),
                                ptr(dG.view(float)) if deriv else nul, ptr(dH.view(float)) if deriv else nul,
                                ptr(dF) if deriv else nul)
        return G, H, dG, dH, dF

    def solve(self, log=print):
        MN, NN = self.MN, self.NN
        mats = []
        for d in range(MN + 1):
            t = time.time()
            mats.append(self.assemble(d))
            log(f"oracle: domain {d} assembled in {time.time() - t:.1f}s")
        na = 8 * NN
        A = np.zeros((na, na), complex); B = np.zeros(na, complex)
        Gp = (D * (MN + 1))(*[ptr(m[0].view(float)) for m in mats])
        Hp = (D * (MN + 1))(*[ptr(m[1].view(float)) for m in mats])
        self.o.or_global_system(self.h, Gp, Hp, ptr(A.view(float)), ptr(B.view(float)))
        t = time.time()
        X = sla.solve(A, B, check_finite=False)
        log(f"oracle: global system {na} solved in {time.time() - t:.1f}s")
        del A
        U = [np.zeros((self.dNe[d] + 1, 4, 4), complex) for d in range(MN + 1)]
        dU = [np.zeros_like(u) for u in U]
        E = [np.zeros((self.dNe[d] + 1, 4, 3), complex) for d in range(MN + 1)]
        H = [np.zeros_like(e) for e in E]
        dE = [np.zeros_like(e) for e in E]
        dH = [np.zeros_like(e) for e in E]
        P = lambda L: (D * len(L))(*[ptr(x.view(float)) for x in L])
        self.o.or_store_solution(self.h, ptr(X.view(float)), P(U), P(dU))
        self.o.or_boundary_EH(self.h, P(U), P(dU), P([m[2] for m in mats]), P([m[3] for m in mats]),
                              (D * (MN + 1))(*[ptr(m[4]) for m in mats]), P(E), P(H))
        for d in range(MN + 1):
            N = 4 * self.dNe[d]
            Bd = np.zeros((N, 6), complex)
            self.o.or_deh_rhs(self.h, d, ptr(mats[d][1].view(float)), ptr(E[d].view(float)), ptr(H[d].view(float)),
                              ptr(Bd.view(float)))
            keep = self.live_rows(d)
            Xd = np.zeros((N, 6), complex)
            Xd[keep] = sla.solve(mats[d][0][np.ix_(keep, keep)], Bd[keep], check_finite=False)
            dE[d][1:] = Xd[:, :3].reshape(-1, 4, 3)
            dH[d][1:] = Xd[:, 3:].reshape(-1, 4, 3)
        return dict(U=U, dU=dU, E=E, H=H, dE=dE, dH=dH, X=X, mats=mats)

    def EH_t(self, sol, r, type_=1):
        MN = self.MN
        P = lambda L: (D * len(L))(*[ptr(x.view(float)) for x in L])
        E = np.zeros(3, complex); H = np.zeros(3, complex)
        rr = np.ascontiguousarray(r, float)
        did = self.o.or_EH_t(self.h, P(sol["E"]), P(sol["dE"]), P(sol["H"]), P(sol["dH"]), ptr(rr), type_,
                             ptr(E.view(float)), ptr(H.view(float)))
        return did, E, H

    def mEMP(self, sol, r, type_=1, total=True):
        """(A, phi) at r from sol["U"], sol["dU"] ([Ne+1][4][4] complex per domain)."""
        P = lambda L: (D * len(L))(*[ptr(x.view(float)) for x in L])
        U = np.zeros(4, complex)
        rr = np.ascontiguousarray(r, float)
        did = self.o.or_mEMP(self.h, P(sol["U"]), P(sol["dU"]), ptr(rr), type_, int(total), ptr(U.view(float)))
        return did, U

    def EH_mEMP(self, sol, r, type_=1, total=True):
        """E, H at r from (A, phi) and its gradient; returns (did or the near-boundary code, E, H)."""
        P = lambda L: (D * len(L))(*[ptr(x.view(float)) for x in L])
        E = np.zeros(3, complex); H = np.zeros(3, complex)
        rr = np.ascontiguousarray(r, float)
        did = self.o.or_EH_mEMP(self.h, P(sol["U"]), P(sol["dU"]), ptr(rr), type_, int(total), ptr(E.view(float)),
                                ptr(H.view(float)))
        return did, E, H

    def coef_rt(self, did, r, s, type_=1, grad=False):
        """CC[9] (and dC[3][9] = d CC / d r) of signed element s at field point r."""
        CC = np.zeros(9, complex)
        rr = np.ascontiguousarray(r, float)
        if not grad:
            self.o.or_coef_rt(self.h, did, ptr(rr), int(s), type_, ptr(CC.view(float)))
            return CC
        dC = np.zeros((3, 9), complex)
        self.o.or_coef_rt_grad(self.h, did, ptr(rr), int(s), type_, ptr(CC.view(float)), ptr(dC.view(float)))
        return CC, dC

    def EH_bd(self, sol, did, t, zt, et, type_=0, total=True):
        P = lambda L: (D * len(L))(*[ptr(x.view(float)) for x in L])
        E = np.zeros(3, complex); H = np.zeros(3, complex)
        self.o.or_EH_bd(self.h, P(sol["E"]), P(sol["dE"]), P(sol["H"]), P(sol["dH"]), did, t, zt, et, type_,
                        int(total), ptr(E.view(float)), ptr(H.view(float)))
        return E, H

    def force_FN(self, sol, rc=(0.0, 0.0, 0.0), type_=0):
        P = lambda L: (D * len(L))(*[ptr(x.view(float)) for x in L])
        F = np.zeros(3); N = np.zeros(3); r = np.ascontiguousarray(rc, float)
        rc_ = self.o.or_force_FN(self.h, P(sol["E"]), P(sol["dE"]), P(sol["H"]), P(sol["dH"]), ptr(r), type_, ptr(F), ptr(N))
        return rc_, F, N
