"""Every output file of the drop-in calc_w3d and calc_j3d (respackdev_amd/bin)
on the inputs of tests/w3d_files_case.py: seeded random eps^-1 (not the
identity) at every irreducible q, three different directions at Gamma
(dat.epsqw.600/601/602, averaged by the reader, calc_intW.F:595-612), static W
(calc_ifreq = 1) and the exchange terms of nine lattice vectors.

* dat.Vmat, dat.Wmat, dat.Xmat, dat.Jmat: every element within 1e-6 eV of the
  oracle on the same files (the GPU sums the pair densities over k before the
  transform, the reference after single-precision rounding: ~1e-8 relative).
* dat.VvsR.III, dat.WvsR.III: per (R, j) the distance |R + tau_j - tau_i| in
  Angstrom (bohr = 0.529177, config_w3d.h:4) from dat.lattice and
  dat.wan-center, and V/W of the oracle (3F12.5); dat.UvsE.III-JJJ and
  dat.JvsE.III-JJJ: the frequencies of dat.wgrid and the onsite W/J (4F12.5).
* dir-mvmc/*.def: the same text as the reference's own writer wrt_mvmc
  (wrt_mvmc_w3d.f90, wrt_mvmc_j3d.f90, compiled from the reference sources
  into oracle/_ref) given the oracle's W and J, numbers within 1e-6 eV.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import w3d_files_case as case
from respackdev_amd.chiqw import from_system
from respackdev_amd.w3d import AU, read_dir_eps_grid

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "respackdev_amd", "bin")
BOHR = 0.529177  # calc_int/config_w3d.h:4
RLIST = [(a1, a2, 0) for a1 in (-1, 0, 1) for a2 in (0, 1)]  # ia1 outer, as calc_intJ.F


def _numbers(tok):
    try:
        return float(tok.replace("D", "E"))
    except ValueError:
        return None


def _lines(text_or_list):
    lines = text_or_list.splitlines() if isinstance(text_or_list, str) else list(text_or_list)
    while lines and not lines[-1].strip():
        lines.pop()
    return lines


def compare_text(got_path, want_lines, tol, what):
    """Line by line (trailing blank lines aside): the same tokens, numbers within tol."""
    got = _lines(open(got_path).read())
    want_lines = _lines(want_lines)
    assert len(got) == len(want_lines), (what, len(got), len(want_lines))
    worst = 0.0
    for n, (g, w) in enumerate(zip(got, want_lines)):
        gt, wt = g.split(), w.split()
        assert len(gt) == len(wt), (what, n, g, w)
        for a, b in zip(gt, wt):
            x, y = _numbers(a), _numbers(b)
            if x is None or y is None:
                assert a == b, (what, n, g, w)
            else:
                worst = max(worst, abs(x - y))
    assert worst <= tol, (what, worst)
    return worst


def _mat_lines(title, cols, m, rs):
    """dat.Vmat-style text: m[(a1,a2,a3)] is (NWF, NWF) [i][j] complex (eV)."""
    out = [title, "#1:R1, 2:R2, 3:R3 (lattice vector)", cols]
    for r in rs:
        out.append(" ".join(str(x) for x in r))
        for i in range(m[r].shape[0]):
            for j in range(m[r].shape[1]):
                out.append(f"{i + 1} {j + 1} {m[r][i, j].real:.10f} {m[r][i, j].imag:.10f}")
        out.append("")
    return out


def _ref_def_files(kind, hr, cut_ha, tmp):
    """The .def files the reference writer writes for hr (Ha) and cut (Ha)."""
    d = tmp / f"ref_{kind}"
    d.mkdir()
    here = os.getcwd()
    os.chdir(d)
    try:
        oracle.ref_mvmc(kind).rpref_wrt_mvmc(3, 3, 3, 27, 1, 1, 1, case.NWF, cut_ha,
                                             np.ascontiguousarray(hr).ctypes.data)
    finally:
        os.chdir(here)
    return {f: _lines(open(d / f).read()) for f in sorted(os.listdir(d))}


@pytest.mark.skipif(not oracle.have_ref_mvmc(), reason="reference writer build absent")
def test_host_output_files(tmp_path):
    from oracle import pipeline_w3d
    s, wan, sqi, em, eps = case.write_inputs(str(tmp_path))
    for prog in ("calc_w3d", "calc_j3d"):
        r = subprocess.run([os.path.join(BIN, prog)], input=case.NAMELIST, cwd=str(tmp_path),
                           capture_output=True, text=True, timeout=300)
        print(prog, r.stdout[-800:], r.stderr[-800:], flush=True)
        assert r.returncode == 0
    # the oracle on the same files (frequencies as calc_int reads them back)
    ecut, em_file = read_dir_eps_grid(str(tmp_path))
    w = from_system(s)
    ntg = max(c.shape[1] for c in wan)
    W = np.zeros((len(wan), case.NWF, ntg), np.complex128)
    for k, c in enumerate(wan):
        W[k, :, :c.shape[1]] = c
    eavg = [sum(d) / len(d) for d in eps]
    o = pipeline_w3d.run_w3d(w, W, sqi, eavg, em_file, ecut)
    oj = pipeline_w3d.run_j3d(w, W, sqi, eavg, em_file, ecut, RLIST)
    na = 1
    rs = [(a1, a2, a3) for a1 in range(-na, na + 1) for a2 in range(-na, na + 1)
          for a3 in range(-na, na + 1)]
    ix = lambda r: (r[2] + na, r[1] + na, r[0] + na)  # noqa: E731
    vm = {r: o["vmat"][ix(r)].T * AU for r in rs}           # [i][j] eV
    wm = {r: o["wmat"][0][ix(r)].T * AU for r in rs}
    xm = {r: oj["X"][n].T * AU for n, r in enumerate(RLIST)}
    jm = {r: oj["J"][n, 0].T * AU for n, r in enumerate(RLIST)}
    iw_dir, ij_dir = tmp_path / "dir-intW", tmp_path / "dir-intJ"
    errs = {}
    errs["Vmat"] = compare_text(iw_dir / "dat.Vmat", _mat_lines(
        "#Bare Coulomb: V", "#1:i, 2:j, 3:Re(V_ij) [eV], 4:Im(V_ij) [eV]", vm, rs), 1e-6, "Vmat")
    errs["Wmat"] = compare_text(iw_dir / "dat.Wmat", _mat_lines(
        "#Screened Coulomb: W at omega=calc_ifreq", "#1:i, 2:j, 3:Re(W_ij) [eV], 4:Im(W_ij) [eV]",
        wm, rs), 1e-6, "Wmat")
    errs["Xmat"] = compare_text(ij_dir / "dat.Xmat", _mat_lines(
        "#Bare exchamge: X", "#1:i, 2:j, 3:Re(X_ij) [eV], 4:Im(X_ij) [eV]", xm, RLIST),
        1e-6, "Xmat")
    errs["Jmat"] = compare_text(ij_dir / "dat.Jmat", _mat_lines(
        "#Screened exchamge: J at omega=calc_ifreq", "#1:i, 2:j, 3:Re(J_ij) [eV], 4:Im(J_ij) [eV]",
        jm, RLIST), 1e-6, "Jmat")
    # distances from dat.lattice and dat.wan-center
    lat = np.asarray(s.lattice)
    tau = np.loadtxt(tmp_path / "dir-wan" / "dat.wan-center", comments="#")
    for i in range(case.NWF):
        for name, m, head in (("VvsR", vm, "#R dependence of bare Coulomb"),
                              ("WvsR", wm, "#R dependence of screened Coulomb at omega=calc_ifreq")):
            col = "V" if name == "VvsR" else "W"
            want = [head, f"#1:R [AA], 2:Re[{col}(R)] [eV], 3:Im[{col}(R)] [eV]"]
            for r in rs:
                for j in range(case.NWF):
                    d = np.linalg.norm(np.asarray(r, float) @ lat + tau[j] - tau[i]) * BOHR
                    want.append(f"{d:.5f} {m[r][i, j].real:.5f} {m[r][i, j].imag:.5f}")
            errs[name] = max(errs.get(name, 0.0),
                             compare_text(iw_dir / f"dat.{name}.{i + 1:03d}", want, 1.5e-5, name))
        for j in range(i, case.NWF):
            for name, arr, head, cols in (
                    ("UvsE", o["wmat"][:, na, na, na, j, i], "#Omega dependence of onsite Coulomb",
                     "#1:Re(w) [eV], 2:Im(w) [eV], 3:Re[U(w)] [eV], 4:Im[U(w)] [eV]"),
                    ("JvsE", oj["J"][RLIST.index((0, 0, 0)), :, j, i],
                     "#Omega dependence of screened exchange J",
                     "#Re(w) [eV], Im(w) [eV], Re[J(w)] [eV], Im[J(w)] [eV]")):
                want = [head, cols] + [f"{z.real * AU:.5f} {z.imag * AU:.5f} {v.real * AU:.5f} "
                                       f"{v.imag * AU:.5f}" for z, v in zip(em_file, arr)]
                d = iw_dir if name == "UvsE" else ij_dir
                errs[name] = max(errs.get(name, 0.0),
                                 compare_text(d / f"dat.{name}.{i + 1:03d}-{j + 1:03d}", want, 1.5e-5,
                                              name))
    # dir-mvmc against the reference's own writer
    jfull = np.zeros_like(o["wmat"][0])
    for n, r in enumerate(RLIST):
        jfull[ix(r)] = oj["J"][n, 0]
    ref = {**_ref_def_files("w", o["wmat"][0], 0.05 / AU, tmp_path),
           **_ref_def_files("j", jfull, 0.0001 / AU, tmp_path)}
    assert sorted(ref) == sorted(os.listdir(tmp_path / "dir-mvmc"))
    for f, want in ref.items():
        errs[f] = compare_text(tmp_path / "dir-mvmc" / f, want, 1e-6, f)
    print("max |diff| per file: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
