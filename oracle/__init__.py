"""ctypes access to the oracle (C restatement) and to the compiled reference.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports this.

* ``lib()``  -> oracle/_build/librespack_oracle.so (built from oracle.c)
* ``ref()``  -> oracle/_ref/librespack_ref.so (RESPACK's own subroutines, built
  from the reference sources by ``make ref``; absent where the reference is not
  mounted).  Loaded with RTLD_LAZY because chiqw_sub.F's LAPACK-calling
  routines are present but never called.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "_build", "librespack_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "librespack_ref.so")
REF_INT_SO = os.path.join(_HERE, "_ref", "librespack_ref_int.so")

_lib = None
_ref = None
_ref_int = None

P = C.c_void_p
I32 = C.c_int
D = C.c_double


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} not built (make -C oracle oracle)")
        L = C.CDLL(ORACLE_SO)
        L.ro_algn235.restype = I32
        L.ro_algn235.argtypes = [I32]
        L.ro_gen_grid.argtypes = [I32, D, P]
        L.ro_make_kg0.restype = I32
        L.ro_make_kg0.argtypes = [I32, P, P, P, D, D, D, D, P]
        L.ro_search_kq.restype = I32
        L.ro_search_kq.argtypes = [I32, P, D, D, D, I32, P]
        L.ro_unfold_k.restype = I32
        L.ro_unfold_k.argtypes = [I32, I32, P, P, P, P, P, P, P]
        L.ro_est_nkbi.argtypes = [I32, P, P]
        L.ro_make_index_kpt.argtypes = [I32, I32, I32, I32, P, P]
        L.ro_judge_fermi.argtypes = [I32, I32, P, D, P, P, P, P, P]
        L.ro_make_c0.argtypes = [I32, I32, P, P, P, P, I32, I32, I32, I32, P, P, P]
        L.ro_ttrhdrn_mkidx.argtypes = [I32, I32, I32, P, P, P, P]
        L.ro_ttrhdrn_chi0.argtypes = [D, D, I32, I32, I32, P, P, P, D, D, D, D, I32, P, P]
        L.ro_ttrhdrn_simple.argtypes = [D, D, I32, I32, I32, P, P, P, P]
        L.ro_ttrhdrn_calc.argtypes = [D, D, P, P, I32, P, P, I32]
        L.ro_ttrhdrn_calc_quad.argtypes = [D, D, P, P, I32, P, P]
        L.ro_ttrhdrn_chi0_closed.argtypes = L.ro_ttrhdrn_chi0.argtypes
        L.ro_ttrhdrn_chi0_quad.argtypes = L.ro_ttrhdrn_chi0.argtypes
        L.ro_tet_series.argtypes = [P, I32, P, P, P]
        L.ro_tet_sub.argtypes = [I32, P, I32, P, P]
        L.ro_dft3.argtypes = [I32, I32, I32, I32, P, P]
        L.ro_pairdens.argtypes = [I32, P, P, I32, P, P, I32, I32, I32, I32, P, P, P]
        L.ro_vmab.argtypes = [I32, P, P, P, P, P, P, P]
        L.ro_vmaa.argtypes = [I32, P, P, P, P, P, P, P, P]
        L.ro_zinv.restype = I32
        L.ro_zinv.argtypes = [I32, P]
        L.ro_eps_rpa.restype = I32
        L.ro_eps_rpa.argtypes = [I32, P, P, P, P, P, P, P]
        L.ro_eps_rpa_q0.restype = I32
        L.ro_eps_rpa_q0.argtypes = [I32, P, P, P, P, I32, D, D, D, P, P]
        L.ro_kcheck.argtypes = [P, P]
        L.ro_kcheck_trs.argtypes = [P, P]
        L.ro_search_list.argtypes = [P, P]
        L.ro_chiqw_q.restype = I32
        L.ro_chiqw_q.argtypes = [I32, I32, P, P, I32, P, I32, P, P, P, P, P, I32, P, P, P, P,
                                 I32, I32, I32, I32, P, I32, I32, I32, P, P, D, D, D, D, I32,
                                 P, D, P, P, P, I32]
        L.ro_make_lg0.restype = I32
        L.ro_make_lg0.argtypes = [I32, P, P, P, D, D, D, D, D, P, P]
        L.ro_make_eps.argtypes = [I32, I32, I32, I32, P, P, P, P, I32, I32, I32, I32, P, P, P]
        L.ro_w3d_func.argtypes = [I32, I32, I32, P, P]
        L.ro_w3d_funcw.argtypes = [I32, I32, I32, I32, P, I32, I32, P, P]
        L.ro_w3d_vw.argtypes = [I32, I32, P, P, P, I32, I32, I32, I32, D, D, P, P, P]
        L.ro_chiqw_q0.restype = I32
        L.ro_chiqw_q0.argtypes = [I32, I32, P, P, I32, P, I32, P, P, P, P, P, I32, P, P,
                                  I32, I32, I32, I32, P, I32, P, P, P, I32, I32, I32, P, P,
                                  D, D, D, D, I32, P, D, P, I32]
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        # RTLD_GLOBAL: MKL dlopens its CPU-specific kernels, which resolve symbols
        # of libmkl_core loaded with this library
        R = C.CDLL(REF_SO, mode=os.RTLD_LAZY | os.RTLD_GLOBAL)
        R.rpref_gen_grid.argtypes = [I32, D, P]
        R.rpref_algn235.restype = I32
        R.rpref_algn235.argtypes = [I32]
        R.rpref_make_kg0.argtypes = [I32, P, P, P, D, D, D, D, P, P]
        R.rpref_search_kq.argtypes = [I32, P, D, D, D, I32, P, P]
        R.rpref_search_list.argtypes = [P, P]
        R.rpref_kcheck.argtypes = [P, P, I32]
        R.rpref_est_ntk.argtypes = [I32, I32, P, P, P]
        R.rpref_est_nkbi.argtypes = [I32, P, P]
        R.rpref_make_index_kpt.argtypes = [I32, I32, I32, I32, P, P]
        R.rpref_judge_fermi.argtypes = [I32, I32, P, D, P, P, P, P, P]
        R.rpref_make_c0.argtypes = [I32, I32, I32, P, P, P, P, I32, I32, I32, I32, P, P, P]
        R.rpref_mkidx.argtypes = [I32, I32, I32, P, P, P, P]
        R.rpref_chi0.argtypes = [D, D, I32, I32, I32, P, P, P, D, D, D, D, I32, P, P]
        R.rpref_simple.argtypes = [D, D, I32, I32, I32, P, P, P, P]
        R.rpref_fft3.argtypes = [I32, I32, I32, I32, P]
        R.rpref_pairdens.argtypes = [I32, I32, P, P, P, P, I32, I32, I32, I32, I32, P, I32, P, P]
        R.rpref_vmab.argtypes = [I32, I32, P, P, P, P, I32, P, P, P, P]
        R.rpref_vmaa.argtypes = [I32, I32, P, P, P, P, I32, P, P, P, P, P]
        R.rpref_eps_rpa.argtypes = [I32, I32, P, D, D, D, P, P, P, P, P]
        R.rpref_tet_sub.argtypes = [I32, P, I32, P, P]
        R.rpref_eps_rpa_q0.argtypes = [I32, I32, P, P, P, P, P, I32, D, D, D, D, P]
        _ref = R
    return _ref


def have_ref_int() -> bool:
    return os.path.exists(REF_INT_SO)


def ref_int():
    """calc_int_sub.F subroutines (make_LG0, make_eps, calc_InterStateMatrix)."""
    global _ref_int
    if _ref_int is None:
        R = C.CDLL(REF_INT_SO, mode=os.RTLD_LAZY)
        R.rpref_make_lg0.argtypes = [I32, P, P, P, D, D, D, D, D, P, P, P]
        R.rpref_make_eps.argtypes = [I32, I32, I32, I32, I32, P, P, P, P, I32, I32, I32, I32, P, P, P]
        R.rpref_int_pairdens.argtypes = [I32, I32, P, P, P, P, I32, I32, I32, I32, I32, P, I32, P, P]
        _ref_int = R
    return _ref_int


_ref_mvmc = {}


def have_ref_mvmc() -> bool:
    return all(os.path.exists(os.path.join(_HERE, "_ref", f"librespack_ref_mvmc{x}.so")) for x in "wj")


def ref_mvmc(kind: str):
    """The reference's mVMC export writer wrt_mvmc of wrt_mvmc_w3d.f90 (kind
    "w") or wrt_mvmc_j3d.f90 (kind "j"): rpref_wrt_mvmc(nkb1, nkb2, nkb3, NTK,
    Na1, Na2, Na3, NWF, cut (Ha), HR (NWF, NWF, -Na1:Na1, -Na2:Na2, -Na3:Na3)
    complex, Ha) writes its .def files into the current directory."""
    if kind not in _ref_mvmc:
        R = C.CDLL(os.path.join(_HERE, "_ref", f"librespack_ref_mvmc{kind}.so"))
        R.rpref_wrt_mvmc.argtypes = [I32, I32, I32, I32, I32, I32, I32, I32, D, P]
        _ref_mvmc[kind] = R
    return _ref_mvmc[kind]


# ----------------------------------------------------------------- helpers

def gen_grid(nen: int, emax: float) -> np.ndarray:
    em = np.zeros(nen, np.complex128)
    lib().ro_gen_grid(nen, emax, ptr(em))
    return em


def make_kg0(b: np.ndarray, gcut: float, q, ntg: int = 200000) -> np.ndarray:
    b = np.ascontiguousarray(b, np.float64)
    kg = np.zeros((ntg, 3), np.int32)
    n = lib().ro_make_kg0(ntg, ptr(b[0]), ptr(b[1]), ptr(b[2]), gcut, q[0], q[1], q[2], ptr(kg))
    if n > ntg:
        raise ValueError("ntg too small")
    return kg[:n].copy()


def mkidx(nkb, b: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(b, np.float64)
    idx = np.zeros((6 * nkb[0] * nkb[1] * nkb[2], 4), np.int32)
    lib().ro_ttrhdrn_mkidx(nkb[0], nkb[1], nkb[2], ptr(b[0]), ptr(b[1]), ptr(b[2]), ptr(idx))
    return idx


def ttrhdrn_chi0(nkb, idx, e1, e2, ef, delt, dsgn, em, dmna=1e-3, dmnr=1e-3, xo=None,
                 mode="stable"):
    """ttrhdrn_chi0 accumulated into xo [nk3][ne].  mode "stable": the series
    where it converges, else the closed forms (the oracle's evaluation);
    "closed": the reference's closed forms everywhere in double; "quad": the
    closed forms in __float128 (the arbiter)."""
    fn = {"stable": "ro_ttrhdrn_chi0", "closed": "ro_ttrhdrn_chi0_closed",
          "quad": "ro_ttrhdrn_chi0_quad"}[mode]
    nk3 = nkb[0] * nkb[1] * nkb[2]
    em = np.ascontiguousarray(em, np.complex128)
    if xo is None:
        xo = np.zeros((nk3, len(em)), np.complex128)
    e1 = np.ascontiguousarray(e1, np.complex128)
    e2 = np.ascontiguousarray(e2, np.complex128)
    getattr(lib(), fn)(dmna, dmnr, nkb[0], nkb[1], nkb[2], ptr(idx), ptr(e1), ptr(e2),
                          ef, ef, delt, dsgn, len(em), ptr(em), ptr(xo))
    return xo


def zinv(a: np.ndarray) -> np.ndarray:
    f = np.asfortranarray(a, np.complex128).copy(order="F")
    info = lib().ro_zinv(f.shape[0], f.ctypes.data_as(C.c_void_p))
    if info:
        raise np.linalg.LinAlgError(f"singular at {info}")
    return f
