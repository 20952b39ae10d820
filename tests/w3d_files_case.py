"""Input files of the calc_w3d / calc_j3d known-answer case: the w3d_tiny
system (simple cubic a = 7 bohr, 3^3 k mesh, time-reversal partners
unfolded), three model Wannier functions, and a dir-eps tree written in the
reference's formats with seeded random eps^-1 (not the identity) at every
irreducible q and three different directions at Gamma (dat.epsqw.600/601/602,
which calc_int averages).  The same files are written here and, for the
fixtures under tests/golden/w3d_ref, for the compiled reference programs.
"""
import os
import struct

import numpy as np

import oracle
from respackdev_amd import gen

NWF = 3
NEN = 14
ECUT_EPS = 2.0
EPS_SEED = 2024
# namelist of both programs: static W (calc_ifreq = 1 is w = 0) and the
# exchange terms of the lattice vectors with components in [-1, 1] x [0, 1] x [0, 0]
NAMELIST = ("&param_calc_int\n calc_ifreq = 1\n ix_intJ_min = -1\n ix_intJ_max = 1\n"
            " iy_intJ_min = 0\n iy_intJ_max = 1\n iz_intJ_min = 0\n iz_intJ_max = 0\n"
            " wcut_mvmc = 0.05\n jcut_mvmc = 0.0001\n/\n")


def _record(f, data: bytes):
    f.write(struct.pack("<i", len(data)))
    f.write(data)
    f.write(struct.pack("<i", len(data)))


def system():
    wl = gen.WORKLOADS["w3d_tiny"]
    return gen.make_system(**{k: wl[k] for k in ("a", "nk", "nocc", "nvir", "ecut_psi", "seed")})


def eps_inputs(s):
    """(sqi, em, eps) with eps per irreducible q a list of directions, each
    [ie][jg][ig] complex128 (3 at Gamma, else 1)."""
    em = oracle.gen_grid(NEN, 50.0 / gen.AU_EV)
    sqi = np.array([[gen._fold(x) for x in k] for k in s.ski])
    rng = np.random.default_rng(EPS_SEED)
    eps = []
    for q in sqi:
        n = len(oracle.make_kg0(s.bvec, ECUT_EPS, q))
        ndir = 3 if np.all(np.abs(q) < 1e-8) else 1
        dirs = []
        for _ in range(ndir):
            e = 0.08 * (rng.standard_normal((NEN, n, n)) + 1j * rng.standard_normal((NEN, n, n)))
            e[:, np.arange(n), np.arange(n)] += 1.0
            dirs.append(e)
        eps.append(dirs)
    return sqi, em, eps


def write_inputs(root):
    """Writes dir-wfn, dir-wan, dir-eps (and an empty dir-mvmc) under root."""
    s = system()
    gen.write_dir_wfn(s, root)
    wan, tau = gen.make_wannier(s, NWF, 5)
    gen.write_dir_wan(wan, tau, root)
    sqi, em, eps = eps_inputs(s)
    d = os.path.join(root, "dir-eps")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "dat.chi_cutoff"), "w") as f:
        f.write(f" {ECUT_EPS:.16f}\n")
    with open(os.path.join(d, "dat.wgrid"), "w") as f:
        f.write(f" {NEN}\n")
        for z in em:
            f.write(f"{z.real:15.10f}{z.imag:15.10f}\n")
    for iq, (q, dirs) in enumerate(zip(sqi, eps)):
        qd = os.path.join(d, f"q{iq + 1:03d}")
        os.makedirs(qd, exist_ok=True)
        with open(os.path.join(qd, "dat.sq"), "w") as f:
            f.write(" " + " ".join(f"{x:.16f}" for x in q) + "\n")
        for ix, e in enumerate(dirs):
            with open(os.path.join(qd, f"dat.epsqw.{600 + ix}"), "wb") as f:
                _record(f, np.ascontiguousarray(e, np.complex128).tobytes())
        with open(os.path.join(qd, "dat.log.400"), "w") as f:
            f.write(" #\n")
    os.makedirs(os.path.join(root, "dir-mvmc"), exist_ok=True)
    return s, wan, sqi, em, eps


def output_files(root):
    """Relative paths of every file the programs wrote."""
    out = []
    for sub in ("dir-intW", "dir-intJ", "dir-mvmc"):
        p = os.path.join(root, sub)
        if os.path.isdir(p):
            out += [os.path.join(sub, f) for f in sorted(os.listdir(p))]
    return out
