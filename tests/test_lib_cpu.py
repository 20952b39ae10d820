"""CPU-side checks of librespack_gpu.so: it loads, exports every symbol that
include/respack_gpu.h declares, its host-side setup helpers agree with the
oracle, and it fails loudly (no fallback) where no GPU is visible."""
import os
import re

import numpy as np
import pytest

import oracle
from respackdev_amd import _lib
from respackdev_amd._lib import ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "respack_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|const char\*)\s+(rp_\w+)\(", txt, re.M)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTS) == syms


def test_host_helpers_match_oracle():
    L = _lib.lib()
    for n in range(1, 200):
        assert L.rp_algn235(n) == oracle.lib().ro_algn235(n)
    for nen in (14, 16, 32):
        a = np.zeros(nen, np.complex128)
        assert L.rp_gen_grid(nen, 50 / 27.21151, ptr(a)) == 0
        np.testing.assert_array_equal(a, oracle.gen_grid(nen, 50 / 27.21151))
    b = np.array([[0.9, 0.0, 0.0], [0.3, 0.85, 0.0], [0.1, 0.2, 0.95]])
    for q in [(0, 0, 0.25), (0.5, -1 / 3, 0.1)]:
        qq = np.array(q, np.float64)
        kg = np.zeros((20000, 3), np.int32)
        ng = np.zeros(1, np.int32)
        assert L.rp_make_kg0(20000, ptr(np.ascontiguousarray(b)), 9.0, ptr(qq), ptr(kg), ptr(ng)) == 0
        np.testing.assert_array_equal(kg[: ng[0]], oracle.make_kg0(b, 9.0, q))
    idx = np.zeros((6 * 48, 4), np.int32)
    assert L.rp_ttrhdrn_mkidx(4, 4, 3, ptr(np.ascontiguousarray(b)), ptr(idx)) == 0
    np.testing.assert_array_equal(idx, oracle.mkidx((4, 4, 3), b))


def test_fails_loudly_without_device():
    import ctypes as C
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    h = _lib.P()
    rc = _lib.lib().rp_init(0, C.byref(h))
    assert rc != 0
    assert _lib.lib().rp_last_error()


def test_dir_wfn_roundtrip(tmp_path):
    from respackdev_amd import gen
    from respackdev_amd.chiqw import from_system, read_dir_wfn
    s = gen.make_system(a=7.0, nk=2, nocc=2, nvir=3, ecut_psi=4.5, seed=7)
    gen.write_dir_wfn(s, str(tmp_path))
    w1 = read_dir_wfn(str(tmp_path))
    w2 = from_system(s)
    np.testing.assert_array_equal(w1.kgi, w2.kgi)
    np.testing.assert_array_equal(w1.wfn, w2.wfn)
    np.testing.assert_allclose(w1.eig, w2.eig, rtol=1e-15)
    np.testing.assert_allclose(w1.ski, w2.ski, rtol=0, atol=1e-15)
    assert abs(w1.ef - w2.ef) < 1e-15
