"""Build the in-tree native artefacts.

* ``lib/libevp1d.so``      HIP kernels + C ABI, exact build (-ffp-contract=off)
* ``lib/libevp1d_fma.so``  same sources, contraction on (fused multiply-add)
* ``lib/libevp1d_probe.so`` HBM streaming probe (bench.py --stream-probe)
* ``lib/evp1d_fortran/``   Fortran host layer (ISO_C_BINDING module, the
                           evp1d driver API module and its driver programs),
                           built with flang when it is available

Everything is compiled for gfx950 only.  Run ``python -m evp1d.build`` from
``cice-evp1d_amd/`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent.parent          # cice-evp1d_amd/
REPO = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIB = PKG / "lib"
INCLUDE = REPO / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
ARCH = "gfx950"

# hipcc flags shared by both flavours.  -packed-fp32-ops keeps the compiler
# from forming packed FP32 instructions; the kernels are FP64, the flag is a
# guard for any FP32 code that may appear.
HIP_COMMON = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-fast-math", "-fno-gpu-approx-transcendentals",
    f"-I{INCLUDE}",
]
FLAVOURS = {
    "libevp1d.so": ["-ffp-contract=off"],
    "libevp1d_fma.so": ["-ffp-contract=fast", "-DEVP1D_FMA=1"],
}
HIP_SOURCES = ["evp1d_device.hip", "evp2d_device.hip", "evp1d_index.cpp", "evp1d_tiles.cpp"]
PROBE = ("libevp1d_probe.so", ["evp1d_stream_probe.hip"])

FORTRAN_MODULES = ["evp1d_capi.F90", "cice_standins.F90", "ice_dyn_evp1d.F90"]
FORTRAN_PROGRAMS = {
    "evp1d_v2_driver": "evp1d_v2_driver.F90",
    "evp1d_api_driver": "evp1d_api_driver.F90",
    "evp1d_v0_driver": "evp1d_v0_driver.F90",
}


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_hip(verbose: bool = True) -> list[Path]:
    LIB.mkdir(parents=True, exist_ok=True)
    out = []
    for name, extra in FLAVOURS.items():
        target = LIB / name
        tmp = LIB / (name + ".tmp")
        cmd = [HIPCC, *HIP_COMMON, *extra, "-o", tmp, *[CSRC / s for s in HIP_SOURCES]]
        _run(cmd)
        os.replace(tmp, target)
        out.append(target)
    name, srcs = PROBE
    tmp = LIB / (name + ".tmp")
    _run([HIPCC, *HIP_COMMON, "-o", tmp, *[CSRC / s for s in srcs]])
    os.replace(tmp, LIB / name)
    out.append(LIB / name)
    return out


def have_flang() -> bool:
    return Path(FLANG).exists() or shutil.which("flang") is not None


def build_fortran() -> Path | None:
    """Compile the Fortran host layer against lib/libevp1d.so."""
    if not have_flang():
        print("flang not found: Fortran host layer not built", file=sys.stderr)
        return None
    missing = [s for s in FORTRAN_MODULES + list(FORTRAN_PROGRAMS.values()) if not (FSRC / s).exists()]
    if missing:
        print(f"Fortran sources missing ({missing}): host layer not built", file=sys.stderr)
        return None
    outdir = LIB / "evp1d_fortran"
    outdir.mkdir(parents=True, exist_ok=True)
    fflags = ["-O2", "-cpp", "-ffp-contract=off", f"-J{outdir}"]
    objs = []
    for src in FORTRAN_MODULES:
        obj = outdir / (Path(src).stem + ".o")
        _run([FLANG, *fflags, "-c", FSRC / src, "-o", obj])
        objs.append(obj)
    rpath = os.path.relpath(LIB, outdir)
    for prog, src in FORTRAN_PROGRAMS.items():
        obj = outdir / (Path(src).stem + ".o")
        _run([FLANG, *fflags, "-c", FSRC / src, "-o", obj])
        _run([FLANG, "-o", outdir / prog, obj, *objs, f"-L{LIB}", "-levp1d",
              f"-Wl,-rpath,$ORIGIN/{rpath}"])
    return outdir


def build_all() -> None:
    build_hip()
    build_fortran()


if __name__ == "__main__":
    build_all()
