"""Build the native library libheomq.so (HIP for gfx950) in-tree.

hipcc cross-compiles for gfx950 without a GPU, so this runs anywhere the ROCm
toolchain is installed.  Packed-FP32 instructions are disabled for all device
code (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR.parent / "csrc"
FORTRAN = PKG_DIR.parent / "fortran"
INCLUDE = PKG_DIR.parent.parent / "include"
LIB_PATH = PKG_DIR / "libheomq.so"
FLIB_PATH = PKG_DIR / "libheomq_fortran.so"

SOURCES = ["heomq_model.cpp", "heomq_apply.hip", "heomq_packed.hip", "heomq_solver.hip", "heomq_capi.cpp",
           "heomq_solve.cpp", "heomq_batch.cpp", "heomq_jacobi.cpp", "heomq_spa.hip", "heomq_field.hip"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libheomq")


def _flang() -> str | None:
    for cand in ("/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if cand and Path(cand).exists():
            return cand
    return None


def build(verbose: bool = False, jobs: int = 4) -> Path:
    """Compile every source and link libheomq.so next to this file."""
    hipcc = _hipcc()
    objdir = PKG_DIR.parent / "_build"
    objdir.mkdir(exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}",
              "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
    procs, objs = [], []
    for src in SOURCES:
        obj = objdir / (src.rsplit(".", 1)[0] + ".o")
        cmd = [hipcc, *common, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        objs.append(obj)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{out}")
        if verbose and out.strip():
            print(out)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    link = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(link)}\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB_PATH)
    build_fortran()
    return LIB_PATH


def build_fortran() -> Path | None:
    """Compile the ISO_C_BINDING drop-in shim (the reference's subroutine names)."""
    flang = _flang()
    src = FORTRAN / "heomq_iface.f90"
    if flang is None or not src.exists():
        return None
    objdir = PKG_DIR.parent / "_build"
    objdir.mkdir(exist_ok=True)
    tmp = FLIB_PATH.with_suffix(".so.tmp")
    cmd = [flang, "-O2", "-fPIC", "-shared", f"-J{objdir}", str(src), "-o", str(tmp),
           f"-L{PKG_DIR}", "-lheomq", f"-Wl,-rpath,{PKG_DIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran shim build failed: {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    os.replace(tmp, FLIB_PATH)
    return FLIB_PATH


SAN_DIR = PKG_DIR.parent / "_build" / "san"
SAN_LIB = SAN_DIR / "libheomq_san.so"


def asan_runtime() -> Path:
    """The compiler's AddressSanitizer runtime, preloaded into the process that
    loads the sanitized library."""
    r = subprocess.run([_hipcc(), "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    p = Path(r.stdout.strip())
    if r.returncode != 0 or not p.is_absolute() or not p.exists():
        raise RuntimeError("the compiler's ASan runtime (libclang_rt.asan-x86_64.so) was not found")
    return p


def build_sanitized(jobs: int = 8) -> Path:
    """libheomq with its host code under AddressSanitizer and UBSan (device
    code unchanged: the flags reach the host pass only), for the CPU checks of
    the host logic: table builders, the reference loader, layout and
    eigensystem builders.  Load it with HEOMQ_LIBRARY=<path> and the runtime of
    asan_runtime() in LD_PRELOAD."""
    hipcc = _hipcc()
    SAN_DIR.mkdir(parents=True, exist_ok=True)
    common = ["-O1", "-g", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}",
              "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
              "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]
    procs, objs = [], []
    for src in SOURCES:
        obj = SAN_DIR / (src.rsplit(".", 1)[0] + ".o")
        cmd = [hipcc, *common, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        objs.append(obj)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            next(p for _, p in procs if p.poll() is None).wait()
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{out}")
    link = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objs), "-o", str(SAN_LIB)]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(link)}\n{r.stdout}{r.stderr}")
    return SAN_LIB


if __name__ == "__main__":
    import sys
    print(build_sanitized() if "--sanitize" in sys.argv else build(verbose=True))
