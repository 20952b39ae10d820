"""Build recipe of librespack_gpu.so (gfx950) and the drop-in Fortran hosts.

``python -m respackdev_amd.build`` compiles every HIP/C++ source under
respackdev_amd/csrc with hipcc into respackdev_amd/librespack_gpu.so (in tree,
so it travels with the repository snapshot), then the Fortran drop-in programs
under respackdev_amd/host with amdflang into respackdev_amd/bin/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librespack_gpu.so")
BIN = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("RP_FC", "/opt/rocm/llvm/bin/amdflang")

# -packed-fp32-ops: no v_pk_*_f32 (they have given wrong results on this pool)
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
            "-Wno-unused-function", "-Wno-unused-variable",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["rp_setup.cpp", "rp_capi.cpp", "rp_w3d.cpp", "rp_batch.cpp", "rp_pairdens.hip", "rp_tetra.hip",
           "rp_chi_gemm.hip", "rp_chi_bf16.hip", "rp_dyson.hip", "rp_w3d.hip",
           "rp_mlp.cpp", "rp_mlp.hip", "rp_noalloc.hip"]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_lib(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC, *HIPFLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_hosts(verbose: bool = True) -> list[str]:
    src_dir = os.path.join(HERE, "host")
    if not os.path.isdir(src_dir) or not os.path.exists(FC):
        return []
    os.makedirs(BIN, exist_ok=True)
    mod_dir = os.path.join(OBJ, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    iface = os.path.join(os.path.dirname(HERE), "include", "respack_gpu.f90")
    common = os.path.join(src_dir, "rp_host_io.f90")
    out = []
    progs = [f for f in sorted(os.listdir(src_dir)) if f.startswith("calc_") and f.endswith(".f90")]
    for p in progs:
        exe = os.path.join(BIN, p[:-4])
        cmd = [FC, "-O2", "-J", mod_dir, iface, common, os.path.join(src_dir, p), "-o", exe,
               f"-L{HERE}", "-lrespack_gpu", f"-Wl,-rpath,{HERE}"]
        _run(cmd, cwd=mod_dir)
        out.append(exe)
        if verbose:
            print(f"built {exe}")
    return out


def build_all(verbose: bool = True) -> None:
    build_lib(verbose)
    build_hosts(verbose)


if __name__ == "__main__":
    build_all()
    sys.exit(0)
