"""Build the spexhip shared library (HIP, gfx950) and the Fortran interface.

Run as ``python -m project1_amd.build``; ``__graft_entry__.build()`` calls
:func:`build_all`.  Objects go to ``project1_amd/_build/``; the library is
``project1_amd/libspexhip.so`` (in-tree so it travels with the sources).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libspexhip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPEXHIP_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
    # packed FP32 VALU ops stay disabled in every device pass
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
    )
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


# Files whose MFMA results feed further MFMAs or VALU work (chained DFT
# passes, per-plane accumulators): MFMA destinations and C operands in
# VGPRs, which removes the AGPR<->VGPR copies the default form inserts.  The
# int8 rank-k kernel keeps the default (its 36 accumulators live in AGPRs).
VGPR_FORM = {"wfprod_pw"}


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    extra = ["-mllvm", "-amdgpu-mfma-vgpr-form"] if src.stem in VGPR_FORM else []
    _run([HIPCC, *HIP_FLAGS, *extra, "-c", str(src), "-o", str(obj)])
    return obj


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


FDRIVER = PKG / "fortran" / "spexhip_fdriver"


def build_fortran() -> Path | None:
    """Compile the ISO_C_BINDING interface module and the Fortran driver
    (amdflang), linked against libspexhip.so, if the compiler is present."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    src = PKG / "fortran" / "m_spexhip.f90"
    drv = PKG / "fortran" / "fdriver.f90"
    if not (Path(fc).exists() and src.exists()):
        return None
    BUILD.mkdir(exist_ok=True)
    obj = BUILD / "m_spexhip.o"
    _run([fc, "-c", "-O2", "-fPIC", f"-J{BUILD}", str(src), "-o", str(obj)])
    _run([fc, "-O2", f"-I{BUILD}", str(drv), str(obj), f"-L{PKG}", "-lspexhip",
          f"-Wl,-rpath,$ORIGIN/..", "-o", str(FDRIVER)])
    return FDRIVER


COR_LIB = PKG / "libspexcor.so"
HDF5_ROOT = Path(os.environ.get("SPEXCOR_HDF5", "/opt/conda"))


def build_cor() -> Path | None:
    """The spex.cor checkpoint library (host C++ over the HDF5 C API), a
    library of its own so that libspexhip.so does not depend on HDF5; built
    when the HDF5 headers are present."""
    if not (HDF5_ROOT / "include" / "hdf5.h").exists():
        return None
    src = PKG / "cor" / "spexcor.cpp"
    tmp = COR_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{HDF5_ROOT / 'include'}",
          str(src), "-o", str(tmp), f"-L{HDF5_ROOT / 'lib'}", "-lhdf5",
          f"-Wl,-rpath,{HDF5_ROOT / 'lib'}"])
    os.replace(tmp, COR_LIB)
    return COR_LIB


def build_all() -> Path:
    lib = build_lib()
    build_fortran()
    build_cor()
    return lib


if __name__ == "__main__":
    print(build_all())
