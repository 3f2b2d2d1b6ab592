"""The text records of curr.data / popu.data (format 518 of main/main.f90:
f12.4, 2x, 32(e15.6e3, 1x)) written by heomq.propagate, checked character
for character against a Fortran compiler's own formatted output."""
import shutil
import subprocess
from pathlib import Path

import pytest

from heomq.propagate import format_518, fortran_e

VALUES = [0.0, 1.0, -1.0, 0.1, 0.09999996, 0.99999996, 123456.789, -0.000123456789, 1e-300, 2.5e200,
          5464.725531624135, -3.14159265358979, 1.9999999999998401, 7.0e-7]
TIMES = [0.0, 1.25, 12.3456, 1234.5678, -0.5]


def _flang():
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang"), shutil.which("gfortran")):
        if c and Path(c).exists():
            return c
    return None


def test_e_format_basics():
    assert fortran_e(0.0) == "  0.000000E+000"
    assert fortran_e(1.0) == "  0.100000E+001"
    assert fortran_e(-0.5) == " -0.500000E+000"
    assert len(format_518(1.0, [1.0, 2.0])) == 12 + 2 + 2 * 16


@pytest.mark.skipif(_flang() is None, reason="no Fortran compiler in this toolchain")
def test_format_518_matches_a_fortran_compiler(tmp_path):
    lit = lambda v: f"{v:.17e}".replace("e", "d")
    assign = "\n".join(f"v({k + 1}) = {lit(v)}" for k, v in enumerate(VALUES))
    lines = "\n".join(f"write(6,518) {lit(t)}, (v(i), i=1,{len(VALUES)})" for t in TIMES)
    src = f"""
program fmt_check
implicit none
integer :: i
real*8 :: v({len(VALUES)})
{assign}
{lines}
518 format(f12.4, 2x, 32(e15.6e3, 1x))
end program
"""
    (tmp_path / "fmt_check.f90").write_text(src)
    subprocess.run([_flang(), "-O0", "fmt_check.f90", "-o", "fmt_check"], cwd=tmp_path, check=True,
                   capture_output=True)
    out = subprocess.run(["./fmt_check"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    got = out.rstrip("\n").split("\n")
    want = [format_518(t, VALUES) for t in TIMES]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g.rstrip() == w.rstrip()
