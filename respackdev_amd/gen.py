"""Seeded synthetic inputs in RESPACK's ``dir-wfn`` / ``dir-wan`` file formats.

The files are what ``util/qe2respack/qe2respack.f90:118-194`` and
``util/xtapp2respack/wfn2respack.F90:103-199`` write, so ``calc_chiqw`` (the
reference's or this package's) reads them unchanged:

* ``dat.bandcalc``   Ecut_psi (Ry), E_F (Ha), E_tot
* ``dat.lattice``    a1, a2, a3 (bohr)
* ``dat.symmetry``   nsym, nnp, then per op 9 ints rg and 3 ints pg
* ``dat.sample-k``   Nk_irr, then k (fractional)
* ``dat.nkm``        plane waves per k
* ``dat.kg``         per k: NG, then G triplets
* ``dat.eigenvalue`` NTB, then Nk_irr*NTB eigenvalues (Ha), k outer
* ``dat.wfn``        Fortran unformatted: ncomp=1, then per k per band one
                     record of NG complex(8)

System: simple cubic cell (edge ``a`` bohr), identity symmetry only, a
Gamma-centred ``nk^3`` mesh folded to (-1/2, 1/2] of which only the
time-reversal-irreducible half is listed (the reference regenerates -k as the
conjugate, ``chiqw.F:458-473``).  Coefficients are complex Gaussians scaled by
exp(-|k+G|^2/8) and orthonormalised per k; eigenvalues are uniform in
[-10, 0] eV (occupied) and [1, 40] eV (empty), sorted per k; E_F = 0.5 eV.
Random numbers come from numpy's counter-based Philox generator.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

AU_EV = 27.21151  # chiqw/config.h:19


@dataclass
class System:
    a: float
    nk: int
    nocc: int
    nvir: int
    ecut_psi: float
    seed: int
    ef: float                      # Ha
    lattice: np.ndarray            # (3,3) rows a1,a2,a3
    bvec: np.ndarray               # (3,3) rows b1,b2,b3
    ski: np.ndarray                # (nk_irr,3)
    kg: list = field(default_factory=list)     # per k: (NG,3) int32
    coef: list = field(default_factory=list)   # per k: (NTB,NG) complex128
    eig: np.ndarray | None = None  # (nk_irr, NTB) Ha

    @property
    def ntb(self) -> int:
        return self.nocc + self.nvir

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.lattice)))


def _mesh_order(n: int) -> list[int]:
    """0, 1, .., floor(n/2), -1, .., so that +k precedes -k."""
    pos = list(range(0, n // 2 + 1))
    neg = [-i for i in range(1, n - len(pos) + 1)]
    return pos + neg


def _fold(x: float) -> float:
    # k = n/N folded into (-1/2, 1/2]
    while x > 0.5 + 1e-9:
        x -= 1.0
    while x <= -0.5 + 1e-9:
        x += 1.0
    return x


def irreducible_kmesh(nk: int) -> np.ndarray:
    order = _mesh_order(nk)
    ks: list[tuple[float, float, float]] = []
    seen: set[tuple[int, int, int]] = set()
    for n1 in order:
        for n2 in order:
            for n3 in order:
                k = (_fold(n1 / nk), _fold(n2 / nk), _fold(n3 / nk))
                key = tuple(int(round(v * nk)) for v in k)
                mk = tuple(int(round(_fold(-v) * nk)) for v in k)
                if key in seen or mk in seen:
                    continue
                seen.add(key)
                ks.append(k)
    return np.array(ks, dtype=np.float64)


def _sphere(kf: np.ndarray, bvec: np.ndarray, ecut: float, gmax: int) -> tuple[np.ndarray, np.ndarray]:
    r = np.arange(-gmax, gmax + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)  # g1 outer, g3 inner
    kc = (kf[None, :] + g) @ bvec
    e = np.einsum("ij,ij->i", kc, kc)
    sel = e <= ecut
    return g[sel].astype(np.int32), e[sel]


def make_system(a: float = 7.0, nk: int = 4, nocc: int = 8, nvir: int = 24,
                ecut_psi: float = 13.9, seed: int = 1000, ef_ev: float = 0.5,
                gap_ev: float = 1.0, emax_ev: float = 40.0, n_cross: int = 0,
                rough: bool = False) -> System:
    """n_cross > 0 makes a metal: the top n_cross occupied and the lowest
    n_cross empty bands cross E_F.  Their energies are, per k and sorted,
    E_F + d_n + w_n (cos 2pi k1 + cos 2pi k2 + cos 2pi k3)/3 with d_n in
    +-0.6 eV and w_n in [1.5, 2.5] eV (smooth bands), or uniform in
    E_F +- 2 eV (rough=True: a Fermi surface that cuts many tetrahedra).  The
    other bands then keep 2.5 eV from E_F, so the band classes stay in index
    order: fully occupied, partially occupied, empty."""
    lat = np.eye(3) * a
    bvec = 2.0 * np.pi * np.linalg.inv(lat).T  # rows b1, b2, b3
    ski = irreducible_kmesh(nk)
    gmax = int(np.ceil(np.sqrt(ecut_psi) / (2 * np.pi / a))) + 2
    # keep Ecut away from every |k+G|^2 so the reference's re-derived cutoff
    # (max|k+G|^2 + 1e-8, chiqw.F:349) selects the same spheres
    allk = np.array([[_fold(n1 / nk), _fold(n2 / nk), _fold(n3 / nk)]
                     for n1 in range(nk) for n2 in range(nk) for n3 in range(nk)])
    r = np.arange(-gmax, gmax + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    ek = np.concatenate([np.einsum("ij,ij->i", (k + g) @ bvec, (k + g) @ bvec) for k in allk])
    for _ in range(100):
        if np.all(np.abs(ek - ecut_psi) > 1e-5):
            break
        ecut_psi += 3e-5
    rng = np.random.Generator(np.random.Philox(seed))
    sysm = System(a=a, nk=nk, nocc=nocc, nvir=nvir, ecut_psi=ecut_psi, seed=seed,
                  ef=ef_ev / AU_EV, lattice=lat, bvec=bvec, ski=ski)
    ntb = nocc + nvir
    eig = np.empty((len(ski), ntb))
    if n_cross > 0:
        if not (0 < n_cross <= min(nocc, nvir)):
            raise ValueError("n_cross must be at most min(nocc, nvir)")
        d_n = rng.uniform(-0.6, 0.6, 2 * n_cross)
        w_n = rng.uniform(1.5, 2.5, 2 * n_cross)
    for ik, kf in enumerate(ski):
        kg, e2 = _sphere(kf, bvec, ecut_psi, gmax)
        ng = len(kg)
        if ng < ntb:
            raise ValueError(f"{ng} plane waves cannot hold {ntb} bands")
        z = rng.standard_normal((ng, ntb)) + 1j * rng.standard_normal((ng, ntb))
        z *= np.exp(-e2 / 8.0)[:, None]
        q, _ = np.linalg.qr(z)
        sysm.kg.append(kg)
        sysm.coef.append(np.ascontiguousarray(q.T))
        if n_cross == 0:
            eo = np.sort(rng.uniform(-10.0, 0.0, nocc))
            ev = np.sort(rng.uniform(gap_ev, emax_ev, nvir))
        else:
            eo = np.sort(rng.uniform(-10.0, ef_ev - 2.5, nocc))
            ev = np.sort(rng.uniform(max(gap_ev, ef_ev + 2.5), emax_ev, nvir))
            if rough:
                ec = np.sort(rng.uniform(ef_ev - 2.0, ef_ev + 2.0, 2 * n_cross))
            else:
                c = np.cos(2 * np.pi * np.asarray(kf)).sum() / 3.0
                ec = np.sort(ef_ev + d_n + w_n * c)
            eo[nocc - n_cross:] = ec[:n_cross]
            ev[:n_cross] = ec[n_cross:]
        eig[ik] = np.concatenate([eo, ev]) / AU_EV
    sysm.eig = eig
    return sysm


def free_electron_system(a: float = 7.0, nk: int = 6, nbands: int = 12, ecut_psi: float = 6.0,
                         ef_ev: float = 9.0) -> System:
    """The empty lattice: per k the nbands lowest plane waves, each band one
    unit coefficient on its G, with E = |k+G|^2 / 2 (Ha).  A metal whose
    Drude plasma frequency is known in closed form, w_p^2 = 4 pi n with
    n = k_F^3 / (3 pi^2) and k_F = sqrt(2 E_F)."""
    lat = np.eye(3) * a
    bvec = 2.0 * np.pi * np.linalg.inv(lat).T
    ski = irreducible_kmesh(nk)
    gmax = int(np.ceil(np.sqrt(ecut_psi) / (2 * np.pi / a))) + 2
    sysm = System(a=a, nk=nk, nocc=nbands // 2, nvir=nbands - nbands // 2, ecut_psi=ecut_psi,
                  seed=0, ef=ef_ev / AU_EV, lattice=lat, bvec=bvec, ski=ski)
    eig = np.empty((len(ski), nbands))
    for ik, kf in enumerate(ski):
        kg, e2 = _sphere(kf, bvec, ecut_psi, gmax)
        if len(kg) < nbands:
            raise ValueError(f"{len(kg)} plane waves cannot hold {nbands} bands")
        order = np.argsort(e2, kind="stable")[:nbands]
        c = np.zeros((nbands, len(kg)), np.complex128)
        c[np.arange(nbands), order] = 1.0
        sysm.kg.append(kg)
        sysm.coef.append(c)
        eig[ik] = e2[order] / 2.0
    sysm.eig = eig
    return sysm


def full_kmesh(ski: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The full mesh in the reference's unfolding order for the identity
    symmetry (calc_intW.F:208-262): each listed k, then its time-reversal
    partner -k when new.  Returns (sk0 (NTK,3), trs (NTK,) +1/-1)."""
    sk0, trs = [], []

    def known(k):
        return any(np.all(np.abs(np.asarray(x) - k) < 1e-4) for x in sk0)

    for k in ski:
        kk = np.array([_fold(v) for v in k])
        if not known(kk):
            sk0.append(kk)
            trs.append(1)
        mk = np.array([_fold(-v) for v in k])
        if not known(mk):
            sk0.append(mk)
            trs.append(-1)
    return np.array(sk0), np.array(trs, np.int32)


def full_kg(sysm: System) -> list:
    """G lists of the full mesh as the reference builds them (calc_intW.F:277-330):
    the sphere at k, or for a time-reversal partner the negated sphere at -k."""
    sk0, trs = full_kmesh(sysm.ski)
    gmax = int(np.ceil(np.sqrt(sysm.ecut_psi) / (2 * np.pi / sysm.a))) + 2
    out = []
    for k, t in zip(sk0, trs):
        kg, _ = _sphere(k * t, sysm.bvec, sysm.ecut_psi, gmax)
        out.append(kg * t)
    return out


# angular factors of the model Wannier orbitals, functions of p = k+G (Cartesian)
_SHAPES = {
    "s": lambda p: np.ones(len(p)),
    "px": lambda p: p[:, 0], "py": lambda p: p[:, 1], "pz": lambda p: p[:, 2],
    "dxy": lambda p: p[:, 0] * p[:, 1], "dyz": lambda p: p[:, 1] * p[:, 2],
    "dzx": lambda p: p[:, 2] * p[:, 0],
    "dx2y2": lambda p: p[:, 0] ** 2 - p[:, 1] ** 2,
    "dz2": lambda p: 3 * p[:, 2] ** 2 - np.einsum("ij,ij->i", p, p),
}


def make_wannier(sysm: System, nwf: int, seed: int, alpha: float = 1.0, spread: float = 0.5):
    """Model Wannier functions for dir-wan/dat.wan: Gaussian orbitals
    exp(-alpha r^2) x angular factor (d-like for nwf = 5, else s, p, d in turn)
    centred within `spread` bohr of the origin, as plane-wave coefficients
    c_w(k+G) ~ exp(-|k+G|^2/(4 alpha)) Y_w(k+G) exp(-i (k+G).tau_w) on every
    full-mesh k, normalised per k.  Returns (wan list per k of (nwf, NG0) complex,
    centres (nwf, 3) bohr)."""
    sk0, _ = full_kmesh(sysm.ski)
    return wannier_on(sk0, full_kg(sysm), sysm.bvec, nwf, seed, alpha, spread)


def wannier_on(sk0, kgs, bvec: np.ndarray, nwf: int, seed: int, alpha: float = 1.0,
               spread: float = 0.5):
    """The model Wannier functions of make_wannier on any full k mesh: sk0
    (NTK, 3) fractional and kgs the G list of each k (the plane waves k + G)."""
    names = (["dxy", "dyz", "dzx", "dx2y2", "dz2"] if nwf == 5 else
             ["s", "px", "py", "pz", "dxy", "dyz", "dzx", "dx2y2", "dz2"][:nwf])
    if len(names) != nwf:
        raise ValueError("at most 9 model Wannier functions")
    rng = np.random.Generator(np.random.Philox(seed))
    tau = rng.uniform(-spread, spread, (nwf, 3)) / np.sqrt(3.0)
    wan = []
    for k, kg in zip(sk0, kgs):
        p = (np.asarray(k)[None, :] + kg) @ bvec
        e2 = np.einsum("ij,ij->i", p, p)
        c = np.empty((nwf, len(kg)), np.complex128)
        for w, nm in enumerate(names):
            v = np.exp(-e2 / (4.0 * alpha)) * _SHAPES[nm](p) * np.exp(-1j * (p @ tau[w]))
            n = np.linalg.norm(v)
            c[w] = v / n if n > 0 else v
        wan.append(c)
    return wan, tau


def write_dir_wan(wan: list, centres: np.ndarray, root: str) -> str:
    """dir-wan/dat.wan (unformatted: NWF, then per full-mesh k one record
    ((C0(G,w), G), w)) and dat.wan-center (wannier.F:2127-2177)."""
    d = os.path.join(root, "dir-wan")
    os.makedirs(d, exist_ok=True)
    nwf = len(centres)
    with open(os.path.join(d, "dat.wan"), "wb") as f:
        _fortran_record(f, struct.pack("<i", nwf))
        for c in wan:
            _fortran_record(f, np.ascontiguousarray(c, dtype=np.complex128).tobytes())
    with open(os.path.join(d, "dat.wan-center"), "w") as f:
        f.write("# Wannier centres (bohr)\n# x y z\n")
        for t in centres:
            f.write(" " + " ".join(f"{v:.16E}" for v in t) + "\n")
    return d


def write_dir_wan_crpa(ns, nb, umat, root: str) -> str:
    """dir-wan/dat.ns-nb (per full-mesh k: bands below the window, bands in
    it) and dat.umat (NWF, then per k and window band the NWF complex
    <psi_b|w> as Fortran list-directed (re,im) pairs), as the reference's
    wannier program writes them (wannier.F:567-572, 2193-2217) and
    calc_chiqw reads them (chiqw.F:870-895).  umat: (NTK, NWF, Mb)."""
    d = os.path.join(root, "dir-wan")
    os.makedirs(d, exist_ok=True)
    umat = np.asarray(umat, np.complex128)
    with open(os.path.join(d, "dat.ns-nb"), "w") as f:
        for a, b in zip(ns, nb):
            f.write(f" {int(a):8d} {int(b):8d}\n")
    with open(os.path.join(d, "dat.umat"), "w") as f:
        f.write(f" {umat.shape[1]:8d}\n")
        for k in range(umat.shape[0]):
            for ib in range(int(nb[k])):
                f.write(" " + " ".join(f"({v.real:.17E},{v.imag:.17E})" for v in umat[k, :, ib]) + "\n")
    return d


def _fortran_record(fh, payload: bytes) -> None:
    n = struct.pack("<i", len(payload))
    fh.write(n)
    fh.write(payload)
    fh.write(n)


def write_dir_wfn(sysm: System, root: str) -> str:
    d = os.path.join(root, "dir-wfn")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "dat.bandcalc"), "w") as f:
        f.write(f" {sysm.ecut_psi:.16E}\n {sysm.ef:.16E}\n {0.0:.16E}\n")
    with open(os.path.join(d, "dat.lattice"), "w") as f:
        for row in sysm.lattice:
            f.write(" " + " ".join(f"{v:.16E}" for v in row) + "\n")
    with open(os.path.join(d, "dat.symmetry"), "w") as f:
        f.write(" 1\n 1\n")
        f.write("    1    0    0    0    1    0    0    0    1\n")
        f.write("    0    0    0\n")
    with open(os.path.join(d, "dat.sample-k"), "w") as f:
        f.write(f" {len(sysm.ski)}\n")
        for k in sysm.ski:
            f.write(" " + " ".join(f"{v:.16E}" for v in k) + "\n")
    with open(os.path.join(d, "dat.nkm"), "w") as f:
        for kg in sysm.kg:
            f.write(f" {len(kg)}\n")
    with open(os.path.join(d, "dat.kg"), "w") as f:
        for kg in sysm.kg:
            f.write(f" {len(kg)}\n")
            for g in kg:
                f.write(f" {g[0]:6d} {g[1]:6d} {g[2]:6d}\n")
    with open(os.path.join(d, "dat.eigenvalue"), "w") as f:
        f.write(f" {sysm.ntb}\n")
        for row in sysm.eig:
            for v in row:
                f.write(f" {v:.16E}\n")
    with open(os.path.join(d, "dat.wfn"), "wb") as f:
        _fortran_record(f, struct.pack("<i", 1))
        for c in sysm.coef:
            for band in c:
                _fortran_record(f, np.ascontiguousarray(band, dtype=np.complex128).tobytes())
    return d


# Named workloads (sizes follow the reference's own derivation rules: the FFT
# grid comes from algn235(int(2 sqrt(Ecut)/d)+1), the frequency grid from
# gen_grid with E_max = band spread).
WORKLOADS = {
    # 4^3 k, 8 occ + 24 empty, ~300 PW (9^3 grid), NG ~100, q = (0,0,1/4)
    "c1_plain_fp64": dict(a=7.0, nk=4, nocc=8, nvir=24, ecut_psi=13.9, seed=1000,
                          ecut_eps=6.9, nen=16, q=(0.0, 0.0, 0.25)),
    # 6^3 k, 20 occ + 180 empty, ~1000 PW (15^3 grid), NG ~600, 32 frequencies
    "c2_precision": dict(a=7.0, nk=6, nocc=20, nvir=180, ecut_psi=30.97, seed=1001,
                         ecut_eps=22.2, nen=32, q=(0.0, 0.0, 1.0 / 6.0)),
    # small parity case used by tests
    "tiny": dict(a=7.0, nk=2, nocc=2, nvir=3, ecut_psi=4.5, seed=7,
                 ecut_eps=2.0, nen=14, q=(0.0, 0.0, 0.5)),
    # calc_w3d parity case: 3^3 mesh (time-reversal partners present), 3 Wannier functions
    "w3d_tiny": dict(a=7.0, nk=3, nocc=2, nvir=3, ecut_psi=4.5, seed=11,
                     ecut_eps=2.0, nen=14, q=(0.0, 0.0, 1.0 / 3.0), nwf=3),
    # metals: bands crossing E_F (partially occupied; Fermi-surface tetrahedron cuts,
    # intraband Drude term at q = 0)
    "metal_tiny": dict(a=7.0, nk=3, nocc=3, nvir=4, ecut_psi=4.5, seed=21, n_cross=1,
                       ecut_eps=2.0, nen=14, q=(0.0, 0.0, 1.0 / 3.0)),
    "metal_c1": dict(a=7.0, nk=4, nocc=8, nvir=24, ecut_psi=13.9, seed=1004, n_cross=2,
                     ecut_eps=6.9, nen=16, q=(0.0, 0.0, 0.25)),
    # batching (C3): every q of a 3^3 mesh, per-q chi cutoff 17.5-24.5 Ry (NG ~420-700)
    # and per-q windows of 150-200 empty bands (C3_WINDOWS)
    "c3_batch": dict(a=7.0, nk=3, nocc=8, nvir=200, ecut_psi=30.97, seed=1002,
                     ecut_eps=24.5, nen=32, q=(0.0, 0.0, 1.0 / 3.0)),
    # calc_int scale (C4): 6^3 mesh, 5 d-like Wannier functions, NG ~600
    "c4_calc_int": dict(a=7.0, nk=6, nocc=20, nvir=180, ecut_psi=30.97, seed=1003,
                        ecut_eps=22.2, nen=32, q=(0.0, 0.0, 1.0 / 6.0), nwf=5),
}


def c3_windows(nk: int = 3):
    """The batched configuration's q list (every point of the nk^3 mesh, folded)
    with its per-q chi cutoff (Ry) and empty-band window."""
    qs = [(_fold(n1 / nk), _fold(n2 / nk), _fold(n3 / nk))
          for n1 in range(nk) for n2 in range(nk) for n3 in range(nk)]
    nq = len(qs)
    return [(q, 17.5 + 7.0 * i / (nq - 1), 150 + int(round(50.0 * i / (nq - 1))))
            for i, q in enumerate(qs)]
