"""rp_dyson_noalloc (the Dyson step with its scratch in a fixed device array,
for callers that cannot allocate) against rp_dyson on the same prepared q:
chi and eps^-1 in device memory the test allocates itself (HIP runtime
through ctypes), two calls in a row, the second with a smaller chi sphere,
both at q != 0 and at q = 0 of a metal (three directions, the Drude terms in
the scratch).  The two entry points run the same kernels on the same inputs,
so the results are equal bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from respackdev_amd import _lib, gen
from respackdev_amd.chiqw import Chiqw, from_system

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime (test plumbing)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.L.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _noalloc(g, hip, chi):
    chi = np.ascontiguousarray(chi, np.complex128)
    d_chi, d_eps = hip.alloc(chi.nbytes), hip.alloc(chi.nbytes)
    try:
        hip.h2d(d_chi, chi)
        _lib.check(_lib.lib().rp_dyson_noalloc(g._h, d_chi, d_eps))
        out = np.empty_like(chi)
        hip.d2h(out, d_eps)
        return out
    finally:
        hip.free(d_chi)
        hip.free(d_eps)


@pytest.mark.parametrize("name,cases", [
    ("c1_plain_fp64", [((0.0, 0.0, 0.25), 6.9), ((0.25, 0.0, 0.25), 4.0)]),
    ("metal_tiny", [((0.0, 0.0, 0.0), 2.0), ((0.0, 0.0, 1.0 / 3.0), 1.2)]),
])
def test_dyson_noalloc_matches_dyson(name, cases):
    wl = gen.WORKLOADS[name]
    keys = ["a", "nk", "nocc", "nvir", "ecut_psi", "seed"] + (["n_cross"] if "n_cross" in wl else [])
    w = from_system(gen.make_system(**{k: wl[k] for k in keys}))
    hip = Hip()
    g = Chiqw(0)
    try:
        g.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        sizes = []
        for q, ecut in cases:
            g.set_window(ecut, 0)
            r = g.run_q(q, "fp64", want_chi=True)
            sizes.append(r["chi"].shape[-1])
            e = _noalloc(g, hip, r["chi"])
            print(f"{name} q={q} NG {sizes[-1]} dirs {r['chi'].shape[0]}: max |diff| "
                  f"{np.abs(e - r['epsinv']).max():.3e}", flush=True)
            np.testing.assert_array_equal(e, r["epsinv"])
        assert sizes[1] < sizes[0]
    finally:
        g.close()


def test_dyson_noalloc_rejects_oversized():
    wl = gen.WORKLOADS["tiny"]
    w = from_system(gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")}))
    g = Chiqw(0)
    try:
        g.open(w, ecut_for_eps=wl["ecut_eps"], num_freq_grid=wl["nen"])
        g.prepare_q(wl["q"])
        # null device pointers are refused before anything runs
        assert _lib.lib().rp_dyson_noalloc(g._h, None, None) != 0
        assert b"null device pointer" in _lib.lib().rp_last_error()
    finally:
        g.close()
