"""The host logic of libheomq under AddressSanitizer and UBSan.

The library is rebuilt with its host code instrumented (device code is
unchanged; no GPU is needed) and the native CPU tests -- model and table
builders, the reference-array loader and its rejections, the packed layout
and the block-Jacobi eigensystem -- run against it in a child process with the
sanitizer runtime preloaded.  Any heap or stack overflow, use after free or
undefined behaviour aborts that process.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_native_host_code_is_clean_under_asan_and_ubsan():
    from heomq import _build
    try:
        lib = _build.build_sanitized()
        rt = _build.asan_runtime()
    except RuntimeError as e:   # no ROCm toolchain on this host
        pytest.skip(str(e))
    env = dict(os.environ, HEOMQ_LIBRARY=str(lib), LD_PRELOAD=str(rt),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        str(ROOT / "tests" / "test_native.py")], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=1200)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert r.returncode == 0 and " passed" in out
