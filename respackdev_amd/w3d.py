"""Python host of the calc_w3d path (calc_intW.F) over librespack_gpu's C ABI:
Wannier-basis bare and screened Coulomb integrals V_ij(R), W_ij(R, w).

Mirrors the reference program's inputs: dir-wfn (k list, G lists, lattice,
symmetry), dir-wan/dat.wan (Wannier functions per full-mesh k), dir-eps
(dat.chi_cutoff, dat.wgrid, and per irreducible q dat.sq and dat.epsqw.600,
or 600/601/602 averaged at Gamma).
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from ._lib import check, ptr
from .chiqw import WfnData, _records

AU = 27.21151  # calc_int/config_w3d.h:2


def read_dir_eps_grid(root: str):
    """dat.chi_cutoff (Ry) and dat.wgrid (nen complex, read as '(2f15.10)')."""
    d = os.path.join(root, "dir-eps")
    with open(os.path.join(d, "dat.chi_cutoff")) as f:
        ecut = float(f.read().split()[0].replace("D", "E"))
    with open(os.path.join(d, "dat.wgrid")) as f:
        lines = f.read().split("\n")
    nen = int(lines[0].split()[0])
    em = np.array([complex(float(l[:15]), float(l[15:30])) for l in lines[1:1 + nen]])
    return ecut, em


def read_epsqw(path: str, ng: int, nen: int) -> np.ndarray:
    """One unformatted record eps(ng,ng,nen) complex(8) -> array [ie][jg][ig]."""
    (rec,) = list(_records(path))
    return np.frombuffer(rec, np.complex128).reshape(nen, ng, ng).copy()


def read_dat_wan(root: str, ng0) -> tuple[np.ndarray, int]:
    """dir-wan/dat.wan -> C0 as [k][w][G] (padded to max NG0)."""
    recs = _records(os.path.join(root, "dir-wan", "dat.wan"))
    nwf = int(np.frombuffer(next(recs)[:4], np.int32)[0])
    ntg = int(max(ng0))
    wan = np.zeros((len(ng0), nwf, ntg), np.complex128)
    for k, n in enumerate(ng0):
        c = np.frombuffer(next(recs), np.complex128).reshape(nwf, n)
        wan[k, :, :n] = c
    return wan, nwf


class W3d:
    """One calc_w3d system on one GPU (one handle, one HIP stream)."""

    def __init__(self, device: int = 0):
        self._h = _lib.P()
        check(_lib.lib().rp_init(device, _lib.C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.lib().rp_finalize(self._h)
            self._h = None

    def open(self, w: WfnData, ecut_for_eps: float, em):
        nk = len(w.ski)
        ntg = w.kgi.shape[1]
        lat = np.ascontiguousarray(w.lattice, np.float64)
        rg = np.ascontiguousarray(w.rg, np.int32)
        pg = np.ascontiguousarray(w.pg, np.int32)
        ski = np.ascontiguousarray(w.ski, np.float64)
        ngi = np.ascontiguousarray(w.ngi, np.int32)
        kgi = np.ascontiguousarray(w.kgi, np.int32)
        em = np.ascontiguousarray(em, np.complex128)
        check(_lib.lib().rp_w3d_open(self._h, ptr(lat), len(rg), int(w.nnp), ptr(rg), ptr(pg), nk,
                                     ptr(ski), ptr(ngi), ntg, ptr(kgi), float(ecut_for_eps),
                                     len(em), ptr(em)))
        return self.info()

    def info(self) -> dict:
        iout = np.zeros(16, np.int32)
        dout = np.zeros(8)
        check(_lib.lib().rp_w3d_info(self._h, ptr(iout), ptr(dout)))
        keys = ["ntk", "nkb1", "nkb2", "nkb3", "na1", "na2", "na3", "nwx", "nwy", "nwz", "nen",
                "ntg", "nwf", "nq_irr", "ntq", "nog0"]
        d = {k: int(v) for k, v in zip(keys, iout)}
        d.update(ecut_psi=dout[0], ecut_eps=dout[1], volume=dout[2], chead=dout[3])
        return d

    def kmesh(self):
        inf = self.info()
        sk0 = np.zeros((inf["ntk"], 3))
        ng0 = np.zeros(inf["ntk"], np.int32)
        check(_lib.lib().rp_w3d_kmesh(self._h, ptr(sk0), ptr(ng0)))
        return sk0, ng0

    def set_wannier(self, wan):
        """wan: [k][w][G] complex (G in the full-mesh G-list order)."""
        inf = self.info()
        ntg = inf["ntg"]
        nk, nwf, n = wan.shape
        a = np.zeros((nk, nwf, ntg), np.complex128)
        a[:, :, :n] = wan
        check(_lib.lib().rp_w3d_set_wannier(self._h, nwf, ptr(a)))

    def set_q(self, sqi) -> np.ndarray:
        sqi = np.ascontiguousarray(sqi, np.float64)
        ng = np.zeros(len(sqi), np.int32)
        check(_lib.lib().rp_w3d_set_q(self._h, len(sqi), ptr(sqi), ptr(ng)))
        return ng

    def set_epsinv(self, iq: int, eps):
        """eps: [ie][jg][ig] complex (Fortran (ig,jg,ie))."""
        e = np.ascontiguousarray(eps, np.complex128)
        check(_lib.lib().rp_w3d_set_epsinv(self._h, iq, ptr(e)))

    def funcs(self):
        inf = self.info()
        nwf, nen, ntq = inf["nwf"], inf["nen"], inf["ntq"]
        func = np.zeros((ntq, nwf, nwf), np.complex128)       # [q][j][i]
        funcw = np.zeros((ntq, nen, nwf, nwf), np.complex128)  # [q][ie][j][i]
        check(_lib.lib().rp_w3d_funcs(self._h, ptr(func), ptr(funcw)))
        return func, funcw

    def vw(self):
        inf = self.info()
        nwf, nen = inf["nwf"], inf["nen"]
        n1, n2, n3 = 2 * inf["na1"] + 1, 2 * inf["na2"] + 1, 2 * inf["na3"] + 1
        v = np.zeros((n3, n2, n1, nwf, nwf), np.complex128)        # Fortran V(i,j,a1,a2,a3)
        wm = np.zeros((nen, n3, n2, n1, nwf, nwf), np.complex128)  # W(i,j,a1,a2,a3,ie)
        heads = np.zeros(2 + 2 * nen)
        check(_lib.lib().rp_w3d_vw(self._h, ptr(v), ptr(wm), ptr(heads)))
        return v, wm, heads[0], heads[2::2] + 1j * heads[3::2]

    def stage_times(self) -> dict:
        """Device seconds of the last funcs() (rho, func, W contraction, total)
        and j3d() (rho of every R, func and sums, J contraction, total)."""
        t = np.zeros(8)
        check(_lib.lib().rp_w3d_stage_times(self._h, ptr(t)))
        return {"w_rho": t[0], "w_func": t[1], "w_contract": t[2], "w_total": t[3],
                "j_rho": t[4], "j_func_sum": t[5], "j_contract": t[6], "j_total": t[7]}

    def j3d(self, rlist):
        """calc_j3d exchange X_ij(R), J_ij(R, w) for the lattice vectors rlist.
        Returns X [r][j][i], J [ie][r][j][i], chead, cheadw."""
        inf = self.info()
        nwf, nen = inf["nwf"], inf["nen"]
        r = np.ascontiguousarray(rlist, np.int32).reshape(-1, 3)
        x = np.zeros((len(r), nwf, nwf), np.complex128)
        j = np.zeros((nen, len(r), nwf, nwf), np.complex128)
        heads = np.zeros(2 + 2 * nen)
        check(_lib.lib().rp_j3d_run(self._h, len(r), ptr(r), ptr(x), ptr(j), ptr(heads)))
        return x, j, heads[0], heads[2::2] + 1j * heads[3::2]
