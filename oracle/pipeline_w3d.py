"""The oracle's calc_w3d run (TEST INFRASTRUCTURE, the checker).

Restates the setup of calc_intW.F (k and q unfolding, make_LG0, make_eps
packing) with the oracle's C routines and Python loops, builds rho_w(q+G) with
the oracle's pair density (ro_pairdens, the twin of calc_int_sub.F:102-220), and
evaluates func / funcw / V / W with ro_w3d_func, ro_w3d_funcw and ro_w3d_vw in
the reference's loop order (calc_intW.F:647-985).
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import lib, ptr
from .pipeline import setup as chiqw_setup

AU = 27.21151


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def setup_w3d(w, wan, sqi, epsinv, em, ecut_eps):
    """The setup shared by calc_w3d and calc_j3d (calc_intW.F:8-630).
    w: dir-wfn contents (WfnData); wan: [k][w][G] over the full mesh in the
    reference's G-list order; sqi: [nq_irr][3]; epsinv: list per irreducible q of
    [ie][jg][ig] complex (already averaged at Gamma); em: nen complex."""
    L = lib()
    # k mesh, G lists, FFT grid: identical to chiqw's setup (calc_intW.F:124-330)
    # (calc_int reads no band wavefunctions: one band keeps the shared setup cheap)
    s = chiqw_setup(w, ecut_for_eps=ecut_eps, num_freq_grid=1, n_calc_band=1)
    b, vol = s["b"], s["volume"]
    ntk, ntg = s["ntk"], s["ntg"]
    nkb = [int(x) for x in s["nkb"]]
    na = [n // 2 for n in nkb]
    ecut_psi = s["ecut_psi"]
    nwf = wan.shape[1]
    nen = len(em)
    rg = _i32(w.rg)
    nsym = len(rg)
    nq_irr = len(sqi)
    # q unfolding (calc_intW.F:420-482)
    sq = [np.asarray(q, np.float64) for q in sqi]
    numirrq = list(range(nq_irr)); numrotq = [0] * nq_irr; trsq = [1] * nq_irr
    rwq = [np.zeros(3, np.int32) for _ in range(nq_irr)]
    for iq in range(nq_irr):
        for op in range(nsym):
            r = rg[op].reshape(3, 3).T  # r[i, j] = rg(i,j)
            for trs in (1, -1):
                k = np.ascontiguousarray(r @ sqi[iq], np.float64)
                rw = np.zeros(3, np.int32)
                (L.ro_kcheck if trs == 1 else L.ro_kcheck_trs)(ptr(k), ptr(rw))
                kk = k * trs
                if any(np.all(np.abs(x - kk) < 1e-4) for x in sq):
                    continue
                sq.append(kk.copy()); numirrq.append(iq); numrotq.append(op); trsq.append(trs)
                rwq.append(rw)
    ntq = len(sq)
    assert ntq == ntk, "NTQ should be NTK"
    # G lists per q (calc_intW.F:500-545)
    def one_q(iq):
        # the C calls release the GIL: the q points run in threads
        tmp = np.zeros((ntg, 3), np.int32)
        if iq < nq_irr:
            q = np.asarray(sqi[iq], np.float64)
        else:
            r = rg[numrotq[iq]].reshape(3, 3).T
            q = r @ np.asarray(sqi[numirrq[iq]], np.float64) + rwq[iq]
        ne = np.zeros(1, np.int32)
        n = L.ro_make_lg0(ntg, ptr(b[0]), ptr(b[1]), ptr(b[2]), ecut_eps, ecut_psi, q[0], q[1], q[2],
                          ptr(tmp), ptr(ne))
        assert n <= ntg
        return np.ascontiguousarray(tmp[:n] * trsq[iq]), int(ne[0]), n

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        res = list(ex.map(one_q, range(ntq)))
    lg0 = [r_[0] for r_ in res]
    nge = [r_[1] for r_ in res]
    ngp = [r_[2] for r_ in res]
    ntgq = max(nge)
    # packing of the irreducible eps spheres, epsirr in single precision
    lq = [int(max(np.abs(l[:, d]).max() for l in lg0)) + 1 for d in range(3)]
    # epsinv: a list, or a function of the irreducible index (then each epsirr
    # is made when first needed, for workloads too large to hold them all)
    epsirr = _LazyEps(epsinv, nge, nen, ntgq) if callable(epsinv) else []
    packs = []
    for iq in range(nq_irr):
        if not callable(epsinv):
            epsirr.append(_epsirr_one(epsinv[iq], nge[iq], nen, ntgq))
        n = nge[iq]
        pk = np.zeros((2 * lq[2] + 1, 2 * lq[1] + 1, 2 * lq[0] + 1), np.int32)
        g = lg0[iq][:n]
        pk[g[:, 2] + lq[2], g[:, 1] + lq[1], g[:, 0] + lq[0]] = np.arange(1, n + 1)
        packs.append(pk)
    rginv = [np.ascontiguousarray(np.linalg.inv(rg[op].reshape(3, 3).T).T.reshape(-1))
             for op in range(nsym)]
    return dict(L=L, s=s, b=b, vol=vol, ntk=ntk, ntg=ntg, na=na, nwf=nwf, nen=nen, nq_irr=nq_irr,
                sq=sq, numirrq=numirrq, numrotq=numrotq, trsq=trsq, rwq=rwq, ntq=ntq, lg0=lg0,
                nge=nge, ngp=ngp, ntgq=ntgq, lq=lq, epsirr=epsirr, packs=packs, rginv=rginv)


def _epsirr_one(e_in, n, nen, ntgq):
    e = np.zeros((nen, ntgq, ntgq), np.complex64)  # Fortran (ig,jg,ie) with ld ntgq
    e[:, :n, :n] = np.asarray(e_in).astype(np.complex64)
    return np.ascontiguousarray(e)


class _LazyEps:
    """epsirr made on first use from fn(iq) (the last few kept)."""

    def __init__(self, fn, nge, nen, ntgq, keep=4):
        self.fn, self.nge, self.nen, self.ntgq, self.keep = fn, nge, nen, ntgq, keep
        self.cache = {}

    def __getitem__(self, iq):
        if iq not in self.cache:
            if len(self.cache) >= self.keep:
                self.cache.pop(next(iter(self.cache)))
            self.cache[iq] = _epsirr_one(self.fn(iq), self.nge[iq], self.nen, self.ntgq)
        return self.cache[iq]


def _epsmk(c, w, iq, out):
    """make_eps for full-mesh q iq into out (single precision)."""
    ir, op = c["numirrq"][iq], c["numrotq"][iq]
    c["L"].ro_make_eps(c["ntgq"], c["nen"], c["trsq"][iq], c["nge"][iq], ptr(c["lg0"][iq]),
                       ptr(_i32(c["rwq"][iq])), ptr(c["rginv"][op]), ptr(_i32(w.pg[op])), int(w.nnp),
                       c["lq"][0], c["lq"][1], c["lq"][2], ptr(c["packs"][ir]), ptr(c["epsirr"][ir]),
                       ptr(out))


def _heads(c, w):
    """Hybertsen-Louie heads (calc_intW.F:782-795): Gamma of the k list, G = 0."""
    iqgm = [i for i in range(c["nq_irr"]) if np.all(np.abs(w.ski[i]) < 1e-5)][-1]
    g0 = int(np.where((c["lg0"][iqgm] == 0).all(1))[0][0])
    qsz = (6.0 * np.pi ** 2 / c["ntq"] / c["vol"]) ** (1.0 / 3.0)
    chead = (2.0 / np.pi) * qsz
    cheadw = np.array([(2.0 / np.pi) * complex(c["epsirr"][iqgm][ie, g0, g0]) * qsz
                       for ie in range(c["nen"])])
    return chead, cheadw


def _rho_q(c, wan, iq):
    """rho_w(q+G) of full-mesh q iq (calc_intW.F:647-767): [w][G] over NGpsi(q)."""
    L, s, b = c["L"], c["s"], c["b"]
    ntk, ntg, nwf = c["ntk"], c["ntg"], c["nwf"]
    sk0, ng0, kg0, nw = s["sk0"], s["ng0"], s["kg0"], s["nw"]
    q = c["sq"][iq]
    q0 = q[0] == 0.0 and q[1] == 0.0 and q[2] == 0.0
    n = c["ngp"][iq]
    lg = c["lg0"][iq]
    qg = np.linalg.norm((q[None, :] + lg) @ b, axis=1)
    nog0 = int(np.where((lg == 0).all(1))[0][0]) if q0 else -1
    def per_k(k):
        # M_ww,k for every w (the C calls release the GIL: k points run in threads)
        out = np.zeros((nwf, n), np.complex128)
        m = np.zeros(ntg, np.complex128)
        sg = np.zeros(3, np.int32)
        kq = k if q0 else L.ro_search_kq(ntk, ptr(sk0), q[0], q[1], q[2], k, ptr(sg))
        kgk = np.ascontiguousarray(kg0[k, :ng0[k]])
        kgq = np.ascontiguousarray(kg0[kq, :ng0[kq]])
        for iw in range(nwf):
            ck = np.ascontiguousarray(wan[k, iw, :ng0[k]])
            cq = np.ascontiguousarray(wan[kq, iw, :ng0[kq]])
            L.ro_pairdens(int(ng0[k]), ptr(kgk), ptr(ck), int(ng0[kq]), ptr(kgq), ptr(cq),
                          nw[0], nw[1], nw[2], n, ptr(lg), ptr(sg), ptr(m))
            if q0:
                m[nog0] = 0.0
            out[iw] = m[:n]
        return out

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        parts = list(ex.map(per_k, range(ntk)))
    rho = np.zeros((nwf, ntg), np.complex128)
    for p in parts:  # the k order of the reference's sum
        rho[:, :n] += p
    for g in range(n):
        if g != nog0:
            rho[:, g] = rho[:, g] / qg[g]
    return rho / float(ntk)


def _func_q(c, w, rho, iq, epsmk):
    """func and funcw of full-mesh q iq (calc_intW.F:833-912)."""
    L, nwf, nen, ntg = c["L"], c["nwf"], c["nen"], c["ntg"]
    r = np.ascontiguousarray(rho)
    f = np.zeros((nwf, nwf), np.complex128)
    L.ro_w3d_func(nwf, c["ngp"][iq], ntg, ptr(r), ptr(f))
    _epsmk(c, w, iq, epsmk)
    fw = np.zeros((nen, nwf, nwf), np.complex128)
    L.ro_w3d_funcw(nwf, c["nge"][iq], c["ngp"][iq], ntg, ptr(r), c["ntgq"], nen, ptr(epsmk), ptr(fw))
    return f, fw


def funcs_at(w, wan, sqi, epsinv, em, ecut_eps, iqs):
    """func and funcw (calc_intW.F:647-912) at the full-mesh q indices iqs only:
    dict(func [len(iqs)][j][i], funcw [len(iqs)][ie][j][i], sq, trsq, numirrq).
    epsinv may be a function of the irreducible index; iqs may be a function
    of the setup dict (keys sq, trsq, numirrq, nge) returning the indices."""
    c = setup_w3d(w, wan, sqi, epsinv, em, ecut_eps)
    if callable(iqs):
        iqs = iqs(c)
    epsmk = np.zeros((c["nen"], c["ntgq"], c["ntgq"]), np.complex64)
    func, funcw = [], []
    for iq in iqs:
        f, fw = _func_q(c, w, _rho_q(c, wan, iq), iq, epsmk)
        func.append(f)
        funcw.append(fw)
    return dict(func=np.array(func), funcw=np.array(funcw), sq=np.array(c["sq"]), iqs=list(iqs),
                trsq=list(c["trsq"]), numirrq=list(c["numirrq"]), nge=c["nge"])


def run_w3d(w, wan, sqi, epsinv, em, ecut_eps):
    """calc_w3d: rho, func, funcw, V, W (calc_intW.F:647-985)."""
    c = setup_w3d(w, wan, sqi, epsinv, em, ecut_eps)
    L = c["L"]
    ntg, ntq, nwf, nen = c["ntg"], c["ntq"], c["nwf"], c["nen"]
    sq, lg0, ngp, nge, na, vol = c["sq"], c["lg0"], c["ngp"], c["nge"], c["na"], c["vol"]
    ntgq = c["ntgq"]
    rho = np.zeros((ntq, nwf, ntg), np.complex128)
    func = np.zeros((ntq, nwf, nwf), np.complex128)
    funcw = np.zeros((ntq, nen, nwf, nwf), np.complex128)
    epsmk = np.zeros((nen, ntgq, ntgq), np.complex64)
    for iq in range(ntq):
        rho[iq] = _rho_q(c, wan, iq)
        func[iq], funcw[iq] = _func_q(c, w, rho[iq], iq, epsmk)
    chead, cheadw = _heads(c, w)
    n1, n2, n3 = (2 * x + 1 for x in na)
    vmat = np.zeros((n3, n2, n1, nwf, nwf), np.complex128)
    wmat = np.zeros((nen, n3, n2, n1, nwf, nwf), np.complex128)
    sqa = np.ascontiguousarray(np.array(sq), np.float64)
    L.ro_w3d_vw(nwf, ntq, ptr(sqa), ptr(func), ptr(funcw), nen, na[0], na[1], na[2], vol, chead,
                ptr(np.ascontiguousarray(cheadw)), ptr(vmat), ptr(wmat))
    return dict(func=func, funcw=funcw, vmat=vmat, wmat=wmat, chead=chead, cheadw=cheadw,
                rho=rho, sq=sqa, lg0=lg0, nge=nge, ngp=ngp, na=na, trsq=c["trsq"],
                numirrq=c["numirrq"])


def run_j3d(w, wan, sqi, epsinv, em, ecut_eps, rlist):
    """calc_j3d: per R of rlist, rho_ij(q+G) = (1/NTK) sum_k e^{-i2pi k.R}
    M_(jw at k, iw at k+q)(G)/|q+G|, func/funcw of each pair, X and J
    (calc_intJ.F:606-904).  Returns X [r][j][i], J [r][ie][j][i]."""
    c = setup_w3d(w, wan, sqi, epsinv, em, ecut_eps)
    L, s = c["L"], c["s"]
    ntk, ntg, ntq, nwf, nen = c["ntk"], c["ntg"], c["ntq"], c["nwf"], c["nen"]
    sq, lg0, ngp, nge, vol, b = c["sq"], c["lg0"], c["ngp"], c["nge"], c["vol"], c["b"]
    sk0, ng0, kg0, nw = s["sk0"], s["ng0"], s["kg0"], s["nw"]
    chead, cheadw = _heads(c, w)
    epsmk = np.zeros((nen, c["ntgq"], c["ntgq"]), np.complex64)
    X = np.zeros((len(rlist), nwf, nwf), np.complex128)
    J = np.zeros((len(rlist), nen, nwf, nwf), np.complex128)
    m = np.zeros(ntg, np.complex128)
    for ir, R in enumerate(rlist):
        func = np.zeros((ntq, nwf, nwf), np.complex128)
        funcw = np.zeros((ntq, nen, nwf, nwf), np.complex128)
        for iq in range(ntq):
            q = sq[iq]
            q0 = q[0] == 0.0 and q[1] == 0.0 and q[2] == 0.0
            n = ngp[iq]
            lg = lg0[iq]
            qg = np.linalg.norm((q[None, :] + lg) @ b, axis=1)
            nog0 = int(np.where((lg == 0).all(1))[0][0]) if q0 else -1
            rho = np.zeros((nwf, nwf, ntg), np.complex128)  # [iw][jw][G]
            for k in range(ntk):
                pf = np.exp(-1j * 2 * np.pi * float(sk0[k] @ np.asarray(R, np.float64)))
                sg = np.zeros(3, np.int32)
                kq = k if q0 else L.ro_search_kq(ntk, ptr(sk0), q[0], q[1], q[2], k, ptr(sg))
                kgk = np.ascontiguousarray(kg0[k, :ng0[k]])
                kgq = np.ascontiguousarray(kg0[kq, :ng0[kq]])
                for iw in range(nwf):
                    for jw in range(nwf):
                        ck = np.ascontiguousarray(wan[k, jw, :ng0[k]])
                        cq = np.ascontiguousarray(wan[kq, iw, :ng0[kq]])
                        L.ro_pairdens(int(ng0[k]), ptr(kgk), ptr(ck), int(ng0[kq]), ptr(kgq), ptr(cq),
                                      nw[0], nw[1], nw[2], n, ptr(lg), ptr(sg), ptr(m))
                        if q0:
                            m[nog0] = 0.0
                            rho[iw, jw, :n] += m[:n] * pf
                        else:
                            rho[iw, jw, :n] += m[:n] * pf / qg
            if q0:
                for g in range(n):
                    rho[:, :, g] = rho[:, :, g] / (qg[g] if g != nog0 else 1.0) / float(ntk)
            else:
                rho /= float(ntk)
            _epsmk(c, w, iq, epsmk)
            for iw in range(nwf):
                for jw in range(nwf):
                    rp = np.ascontiguousarray(rho[iw, jw])
                    f = np.zeros(1, np.complex128)
                    L.ro_w3d_func(1, n, ntg, ptr(rp), ptr(f))
                    func[iq, jw, iw] = f[0]
                    fw = np.zeros(nen, np.complex128)
                    L.ro_w3d_funcw(1, nge[iq], n, ntg, ptr(rp), c["ntgq"], nen, ptr(epsmk), ptr(fw))
                    funcw[iq, :, jw, iw] = fw
        X[ir] = 2.0 * 2 * np.pi * func.sum(0) / vol / ntq
        J[ir] = 2.0 * 2 * np.pi * funcw.sum(0) / vol / ntq
        if tuple(R) == (0, 0, 0):
            for iw in range(nwf):
                X[ir, iw, iw] += chead
                J[ir, :, iw, iw] += cheadw
    return dict(X=X, J=J, chead=chead, cheadw=cheadw, trsq=c["trsq"])
