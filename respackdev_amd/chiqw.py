"""Python host of the chiqw hot path (mirrors PROGRAM DIELECTRIC, chiqw.F).

``WfnData`` holds the contents of ``dir-wfn`` (read by :func:`read_dir_wfn`, as
chiqw.F:88-329 reads them, or taken from a generated :class:`gen.System`).
:class:`Chiqw` opens it on the GPU and runs the per-q stages through the C ABI:
pair densities, tetrahedron weights, the chi contraction and the Dyson step.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr

AU_EV = 27.21151


@dataclass
class WfnData:
    lattice: np.ndarray   # (3,3) rows a1, a2, a3
    rg: np.ndarray        # (nsym, 9) int32, Fortran order rg(i,j) (i fastest)
    pg: np.ndarray        # (nsym, 3) int32
    nnp: int
    ski: np.ndarray       # (nk, 3)
    ngi: np.ndarray       # (nk,) int32
    kgi: np.ndarray       # (nk, ntg, 3) int32
    eig: np.ndarray       # (nk, ntb) Ha
    wfn: np.ndarray       # (nk, ntb, ntg) complex128
    ef: float             # Ha
    ecut: float = 0.0     # Ry (dat.bandcalc line 1; the program re-derives it)

    @property
    def ntg(self) -> int:
        return self.kgi.shape[1]

    @property
    def ntb(self) -> int:
        return self.eig.shape[1]


def from_system(s) -> WfnData:
    nk = len(s.ski)
    ngi = np.array([len(k) for k in s.kg], np.int32)
    ntg = int(ngi.max())
    kgi = np.zeros((nk, ntg, 3), np.int32)
    wfn = np.zeros((nk, s.ntb, ntg), np.complex128)
    for ik in range(nk):
        kgi[ik, : ngi[ik]] = s.kg[ik]
        wfn[ik, :, : ngi[ik]] = s.coef[ik]
    return WfnData(lattice=np.ascontiguousarray(s.lattice, np.float64),
                   rg=np.eye(3, dtype=np.int32).reshape(1, 9).copy(),
                   pg=np.zeros((1, 3), np.int32), nnp=1, ski=np.ascontiguousarray(s.ski),
                   ngi=ngi, kgi=kgi, eig=np.ascontiguousarray(s.eig), wfn=wfn, ef=s.ef,
                   ecut=s.ecut_psi)


def _numbers(path: str) -> list[str]:
    with open(path) as f:
        return f.read().replace(",", " ").split()


def _records(path: str):
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        yield data[pos + 4: pos + 4 + n]
        pos += 8 + n


def read_dir_wfn(root: str) -> WfnData:
    """Reads ./dir-wfn as chiqw.F:88-329 and :1100-1160 do."""
    d = os.path.join(root, "dir-wfn")
    bc = _numbers(os.path.join(d, "dat.bandcalc"))
    ecut, ef = float(bc[0].replace("D", "E")), float(bc[1].replace("D", "E"))
    lat = np.array([float(x.replace("D", "E")) for x in _numbers(os.path.join(d, "dat.lattice"))[:9]]).reshape(3, 3)
    sy = [int(x) for x in _numbers(os.path.join(d, "dat.symmetry"))]
    nsym, nnp = sy[0], sy[1]
    rg = np.zeros((nsym, 9), np.int32)
    pg = np.zeros((nsym, 3), np.int32)
    for i in range(nsym):
        blk = sy[2 + 12 * i: 14 + 12 * i]
        rg[i] = blk[:9]
        pg[i] = blk[9:12]
    sk = _numbers(os.path.join(d, "dat.sample-k"))
    nk = int(sk[0])
    ski = np.array([float(x.replace("D", "E")) for x in sk[1: 1 + 3 * nk]]).reshape(nk, 3)
    ngi = np.array([int(x) for x in _numbers(os.path.join(d, "dat.nkm"))[:nk]], np.int32)
    ntg = int(ngi.max())
    kgt = [int(x) for x in _numbers(os.path.join(d, "dat.kg"))]
    kgi = np.zeros((nk, ntg, 3), np.int32)
    pos = 0
    for ik in range(nk):
        n = kgt[pos]
        if n != ngi[ik]:
            raise ValueError("ERROR;STOP;NG_for_psi should be NGI(ik)")
        kgi[ik, :n] = np.array(kgt[pos + 1: pos + 1 + 3 * n]).reshape(n, 3)
        pos += 1 + 3 * n
    ev = _numbers(os.path.join(d, "dat.eigenvalue"))
    ntb = int(ev[0])
    eig = np.array([float(x.replace("D", "E")) for x in ev[1: 1 + nk * ntb]]).reshape(nk, ntb)
    recs = _records(os.path.join(d, "dat.wfn"))
    (ncomp,) = struct.unpack("<i", next(recs)[:4])
    if ncomp != 1:
        raise ValueError("This program not suport ncomp/=1")
    wfn = np.zeros((nk, ntb, ntg), np.complex128)
    for ik in range(nk):
        for ib in range(ntb):
            r = np.frombuffer(next(recs), np.complex128)
            wfn[ik, ib, : ngi[ik]] = r[: ngi[ik]]
    return WfnData(lattice=lat, rg=rg, pg=pg, nnp=nnp, ski=ski, ngi=ngi, kgi=kgi, eig=eig,
                   wfn=wfn, ef=ef, ecut=ecut)


def default_params(**kw) -> dict:
    """Namelist &param_chiqw defaults (m_rdinput.f90:52-72)."""
    p = dict(n_calc_band=0, num_freq_grid=70, flg_crpa=0, precision=0, ecut_for_eps=0.0,
             shift_ef=0.0, max_excitation_energy=200.0, green_func_delt=0.1,
             ttrhdrn_dmna=0.001, ttrhdrn_dmnr=0.001)
    p.update(kw)
    return p


class Chiqw:
    """One chiqw system on one GPU (one handle, one HIP stream)."""

    def __init__(self, device: int = 0):
        self._h = _lib.P()
        check(_lib.lib().rp_init(device, _lib.C.byref(self._h)))
        self.info = None

    def close(self) -> None:
        if self._h:
            check(_lib.lib().rp_finalize(self._h))
            self._h = _lib.P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def open(self, w: WfnData, **params) -> dict:
        p = default_params(**params)
        prm = _lib.ChiqwParams(**p)
        lat = np.ascontiguousarray(w.lattice, np.float64)
        rg = np.ascontiguousarray(w.rg, np.int32)
        pg = np.ascontiguousarray(w.pg, np.int32)
        ski = np.ascontiguousarray(w.ski, np.float64)
        ngi = np.ascontiguousarray(w.ngi, np.int32)
        kgi = np.ascontiguousarray(w.kgi, np.int32)
        eig = np.ascontiguousarray(w.eig, np.float64)
        wfn = np.ascontiguousarray(w.wfn, np.complex128)
        check(_lib.lib().rp_chiqw_open(self._h, ptr(lat), len(rg), w.nnp, ptr(rg), ptr(pg),
                                       len(ski), ptr(ski), ptr(ngi), w.ntg, ptr(kgi), w.ntb,
                                       ptr(eig), ptr(wfn), w.ef, _lib.C.byref(prm)))
        iout = np.zeros(16, np.int32)
        dout = np.zeros(8)
        check(_lib.lib().rp_chiqw_info(self._h, ptr(iout), ptr(dout), None))
        nen = int(iout[7])
        em = np.zeros(nen, np.complex128)
        check(_lib.lib().rp_chiqw_info(self._h, None, None, ptr(em)))
        self.info = dict(ntk=int(iout[0]), nkb=tuple(int(x) for x in iout[1:4]),
                         nw=tuple(int(x) for x in iout[4:7]), nen=nen, n_occ=int(iout[8]),
                         n_vir=int(iout[9]), n_docc=int(iout[10]), ntb=int(iout[11]),
                         nk_irr=int(iout[12]), ecut_psi=float(dout[0]), ecut_eps=float(dout[1]),
                         volume=float(dout[2]), emax=float(dout[3]), ef=float(dout[4]),
                         delt=float(dout[5]), em=em)
        return self.info

    def set_crpa(self, ns, nb, umat) -> None:
        ntk = self.info["ntk"]
        ns = np.ascontiguousarray(ns, np.int32)
        nb = np.ascontiguousarray(nb, np.int32)
        umat = np.ascontiguousarray(umat, np.complex128)  # (ntk, nwf, mb)
        check(_lib.lib().rp_chiqw_set_crpa(self._h, ntk, ptr(ns), ptr(nb), umat.shape[2],
                                           umat.shape[1], ptr(umat)))

    def kmesh(self):
        ntk = self.info["ntk"]
        sk0 = np.zeros((ntk, 3))
        numirr = np.zeros(ntk, np.int32)
        trs = np.zeros(ntk, np.int32)
        check(_lib.lib().rp_chiqw_kmesh(self._h, ptr(sk0), ptr(numirr), ptr(trs)))
        return sk0, numirr, trs

    def prepare_q(self, q) -> np.ndarray:
        q = np.ascontiguousarray(q, np.float64)
        ng = np.zeros(1, np.int32)
        check(_lib.lib().rp_chiqw_prepare_q(self._h, ptr(q), ptr(ng), None, 0))
        lg = np.zeros((int(ng[0]), 3), np.int32)
        check(_lib.lib().rp_chiqw_prepare_q(self._h, ptr(q), ptr(ng), ptr(lg), len(lg)))
        self.ng = int(ng[0])
        self.q0 = bool(np.all(q == 0.0))
        n = np.zeros(1, np.int32)
        check(_lib.lib().rp_chiqw_npair(self._h, ptr(n)))
        self.npair = int(n[0])
        return lg

    @property
    def ndir(self) -> int:
        return 3 if self.q0 else 1

    def pair_densities(self, copy: bool = True):
        K = self.npair * self.info["ntk"]
        out = np.zeros((self.ndir, K, self.ng), np.complex64) if copy else None
        check(_lib.lib().rp_pairdens_batch(self._h, ptr(out)))
        return out

    def weights(self, copy: bool = True):
        K = self.npair * self.info["ntk"]
        out = np.zeros((self.info["nen"], K), np.complex128) if copy else None
        check(_lib.lib().rp_tetra_weights_batch(self._h, ptr(out)))
        return out

    def accumulate(self, precision: str = "fp64", copy: bool = True):
        """chi as [dir][ie][jg][ig] (Fortran chiqw(ig,jg,ie) per direction)."""
        nen, ng = self.info["nen"], self.ng
        out = np.zeros((self.ndir, nen, ng, ng), np.complex128) if copy else None
        check(_lib.lib().rp_chi_accumulate(self._h, _lib.PREC[precision], ptr(out)))
        return out

    def dyson_from(self, chi) -> np.ndarray:
        """eps^-1 of the prepared q from a given chi ([dir][ie][jg][ig]; e.g.
        with the band-tail closure applied) through the GPU Dyson step."""
        chi = np.ascontiguousarray(chi, np.complex128)
        assert chi.shape == (self.ndir, self.info["nen"], self.ng, self.ng)
        check(_lib.lib().rp_chi_set(self._h, ptr(chi)))
        return self.dyson()

    def dyson(self, copy: bool = True):
        nen, ng = self.info["nen"], self.ng
        out = np.zeros((self.ndir, nen, ng, ng), np.complex128) if copy else None
        check(_lib.lib().rp_dyson(self._h, ptr(out)))
        return out

    def autotune(self, enable: bool = True) -> None:
        check(_lib.lib().rp_chiqw_autotune(self._h, 1 if enable else 0))

    def tune_table(self) -> list[dict]:
        n = np.zeros(1, np.int32)
        check(_lib.lib().rp_chiqw_tune_count(self._h, ptr(n)))
        out = []
        for i in range(int(n[0])):
            iv = np.zeros(10, np.int32)
            dv = np.zeros(2)
            check(_lib.lib().rp_chiqw_tune_info(self._h, i, ptr(iv), ptr(dv)))
            out.append(dict(precision=int(iv[0]), ng=int(iv[1]), nen=int(iv[2]), ndir=int(iv[3]),
                            kpad=int(iv[4]), candidates=int(iv[5]), nsplit=int(iv[6]),
                            nwg=int(iv[7]), heuristic_nsplit=int(iv[8]),
                            heuristic_nwg=int(iv[9]), ms=float(dv[0]),
                            heuristic_ms=float(dv[1])))
        return out

    def batch_plan(self, qs, ecut_for_eps=None, n_vir=None, precision: str = "bf16x3") -> None:
        """Prepares every q (per-q windows optional), checks each stage once
        and captures the launch sequence of all q into one HIP graph."""
        qq = np.ascontiguousarray(qs, np.float64).reshape(-1, 3)
        ec = None if ecut_for_eps is None else np.ascontiguousarray(ecut_for_eps, np.float64)
        nv = None if n_vir is None else np.ascontiguousarray(n_vir, np.int32)
        self._batch_n = len(qq)
        check(_lib.lib().rp_chiqw_batch_plan(self._h, len(qq), ptr(qq),
                                             None if ec is None else ptr(ec),
                                             None if nv is None else ptr(nv),
                                             _lib.PREC[precision]))

    def batch_run(self) -> float:
        t = np.zeros(1)
        check(_lib.lib().rp_chiqw_batch_run(self._h, ptr(t)))
        return float(t[0])

    def batch_result(self, iq: int, want_eps: bool = True):
        ng = np.zeros(1, np.int32)
        nd = np.zeros(1, np.int32)
        check(_lib.lib().rp_chiqw_batch_result(self._h, iq, ptr(ng), ptr(nd), None, 0, None))
        ng, nd = int(ng[0]), int(nd[0])
        lg = np.zeros((ng, 3), np.int32)
        eps = np.zeros((nd, self.info["nen"], ng, ng), np.complex128) if want_eps else None
        check(_lib.lib().rp_chiqw_batch_result(self._h, iq, None, None, ptr(lg), ng,
                                               None if eps is None else ptr(eps)))
        return dict(lg=lg, epsinv=eps)

    def batch_free(self) -> None:
        check(_lib.lib().rp_chiqw_batch_free(self._h))

    def set_span(self, span: int = 0) -> None:
        """FP32 accumulation span (K) of the split-BF16 contraction (0: automatic)."""
        check(_lib.lib().rp_chiqw_set_span(self._h, int(span)))

    def set_window(self, ecut_for_eps: float = 0.0, n_vir: int = 0):
        """Per-q chi cutoff (Ry) and empty-band window for the q points prepared
        next (0: the values of open)."""
        check(_lib.lib().rp_chiqw_set_window(self._h, float(ecut_for_eps), int(n_vir)))

    def run_q(self, q, precision: str = "fp64", want_chi: bool = False, want_eps: bool = True):
        """One q through all four stages.  want_eps / want_chi copy eps^-1 / chi
        to the host; without them the results stay in HBM."""
        lg = self.prepare_q(q)
        nen, ng = self.info["nen"], self.ng
        eps = np.zeros((self.ndir, nen, ng, ng), np.complex128) if want_eps else None
        chi = np.zeros((self.ndir, nen, ng, ng), np.complex128) if want_chi else None
        t = np.zeros(8)
        qq = np.ascontiguousarray(q, np.float64)
        check(_lib.lib().rp_chiqw_q(self._h, ptr(qq), _lib.PREC[precision], ptr(eps), ptr(chi), ptr(t)))
        return dict(lg=lg, epsinv=eps, chi=chi, timings=dict(
            pairdens=t[0], weights=t[1], contraction=t[2], dyson=t[3], device_total=t[4],
            host_prepare=t[5], gemm_kernel=t[6], gemm_launches=int(t[7])))
