"""Build steps of the package: the gfx950 HIP library and the Fortran shim.

`build_library()` compiles the translation units of monc_amd/csrc with hipcc
and links them into monc_amd/libmonc_tvd.so (in-tree, so it travels with the source tree);
`build_fortran()` compiles the iso_c_binding interface and the drop-in driver
with amdflang when that compiler is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmonc_tvd.so"
FORTRAN = PKG / "fortran"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 arithmetic is kept off for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
# translation units of the library, compiled in parallel and linked together
SOURCES = ["tvd_lib.hip", "tvd_ensemble.hip"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")
    out = _filter(proc.stdout)
    if out:
        sys.stderr.write(out + "\n")


def _filter(out: str) -> str:
    # keep compiler warnings visible, but drop the known host-pass notice
    return "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())


def build_library(extra_flags: list[str] | None = None) -> Path:
    """Compile the HIP library for gfx950 (always rebuilds): every translation
    unit in parallel, then one shared link."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *(extra_flags or []), "-I", str(ROOT / "include"), "-c", "-o", str(obj),
               str(CSRC / src)]
        procs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = None
    for cmd, _, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = failed or f"build step failed ({p.returncode}): {' '.join(cmd)}"
        elif _filter(out):
            sys.stderr.write(_filter(out) + "\n")
    if failed:
        raise RuntimeError(failed)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *[str(o) for _, o, _ in procs]])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> Path | None:
    """Compile the Fortran interface module, the shim core and the drop-in driver; None if no amdflang."""
    if not shutil.which(FLANG) and not Path(FLANG).exists():
        return None
    out = FORTRAN / "_build"
    out.mkdir(exist_ok=True)
    srcs = [FORTRAN / "monc_tvd_interface.F90", FORTRAN / "tvd_shim_core.F90", FORTRAN / "tvd_dropin_driver.F90"]
    objs = []
    for s in srcs:
        o = out / (s.stem + ".o")
        _run([FLANG, "-O2", "-cpp", "-c", "-module-dir", str(out), "-o", str(o), str(s)])
        objs.append(o)
    exe = out / "tvd_dropin_driver"
    _run([FLANG, "-o", str(exe), *map(str, objs), "-L", str(PKG), "-lmonc_tvd",
          f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../.."])
    return exe


if __name__ == "__main__":
    print(build_library())
    print(build_fortran())
