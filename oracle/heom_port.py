"""ctypes binding of the C port of the reference apply (heom_apply_port.c) --
TEST AND BASELINE INFRASTRUCTURE ONLY.

The port runs the dense TFQMR-convention apply of
tfqmr_double/calc_ax_tfqmr_omp.f90 on the reference's module tables with
OpenMP threads; it serves as the CPU baseline beside the GPU and is checked
against the numpy oracle.
"""
from __future__ import annotations

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np

from . import heom_oracle as O
from .ref_tables import reference_arrays

HERE = Path(__file__).resolve().parent
SRC = HERE / "heom_apply_port.c"
LIB = HERE / "libheom_port.so"


def build() -> Path:
    """gcc -O3 -fopenmp of the port (generic x86-64 code: it runs on any host)."""
    cc = shutil.which("gcc") or "gcc"
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "-O3", "-fopenmp", "-fPIC", "-shared", str(SRC), "-o", str(tmp), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"port build failed: {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    tmp.replace(LIB)
    return LIB


class Tables(C.Structure):
    _fields_ = [("nrho", C.c_int32), ("ntier", C.c_int32), ("nvar", C.c_int32), ("norbs", C.c_int32),
                ("nspin", C.c_int32), ("nunk", C.c_int64)] + \
               [(nm, C.c_void_p) for nm in ("indextable", "nfirst", "ifirst", "indexcoef", "itypecoef",
                                            "ifactfcoef", "ifactrcoef", "index_coef_ref", "indexdraw", "cb",
                                            "cd", "gdraw", "hs", "ams", "sgn", "spin", "orb", "dbsqrt")]


class Port:
    """The C port bound to one oracle model's tables."""

    def __init__(self, model: O.Model):
        if not LIB.exists():
            build()
        self.lib = C.CDLL(str(LIB))
        self.lib.heom_port_apply.argtypes = [C.POINTER(Tables), C.c_void_p, C.c_void_p, C.c_int]
        self.lib.heom_port_apply.restype = None
        self.lib.heom_port_threads.restype = C.c_int
        self.lib.heom_port_tfqmr.argtypes = [C.POINTER(Tables), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int, C.c_int, C.c_int, C.c_int]
        self.lib.heom_port_tfqmr.restype = C.c_double
        self.model = model
        arr = reference_arrays(model)
        keep = {}

        def ptr(name, a, dt):
            a = np.ascontiguousarray(np.asarray(a).astype(dt, copy=False))
            keep[name] = a
            return a.ctypes.data

        t = Tables()
        t.nrho, t.ntier, t.nvar, t.norbs, t.nspin = arr["nrho"], arr["ntier"], model.M, arr["norbs"], arr["nspin"]
        t.nunk = arr["nunk"]
        for nm, dt in (("indextable", np.int32), ("nfirst", np.int64), ("ifirst", np.int64),
                       ("indexcoef", np.int64), ("itypecoef", np.int8), ("ifactfcoef", np.int8),
                       ("ifactrcoef", np.int8), ("index_coef_ref", np.int64), ("indexdraw", np.int32)):
            setattr(t, nm, ptr(nm, arr[nm], dt))
        t.cb = ptr("cb", model.cb, np.complex128)
        t.cd = ptr("cd", model.cd, np.complex128)
        t.gdraw = ptr("gdraw", model.gdraw, np.complex128)
        t.hs = ptr("hs", np.asfortranarray(model.H).ravel(order="F"), np.complex128)
        ams = np.zeros((model.norbs, 2, model.d, model.d))
        for m in range(model.norbs):
            for s in range(2):
                ams[m, s] = model.ams[m][s].T          # C order [m][s][j][i] == column-major (i, j)
        t.ams = ptr("ams", ams.ravel(), np.float64)
        t.sgn = ptr("sgn", model.sgn, np.int32)
        t.spin = ptr("spin", model.spin, np.int32)
        t.orb = ptr("orb", model.orb, np.int32)
        t.dbsqrt = ptr("dbsqrt", model.dbsqrt, np.float64) if model.p.lscale else None
        self._keep = keep
        self.tables = t

    def threads(self) -> int:
        return int(self.lib.heom_port_threads())

    def apply(self, x, trace_row="last"):
        """y = L x on math-indexed (N, d, d) arrays (converted to the reference layout)."""
        mode = {None: 0, "last": 1, "rdo": 2}[trace_row]
        xf = np.ascontiguousarray(O.to_fortran(x))
        yf = np.zeros_like(xf)
        self.lib.heom_port_apply(C.byref(self.tables), xf.ctypes.data, yf.ctypes.data, mode)
        return O.from_fortran(yf)

    def apply_fortran(self, xf, yf, trace_row="last"):
        """The same on arrays already in the reference layout (no conversion; for timing)."""
        mode = {None: 0, "last": 1, "rdo": 2}[trace_row]
        self.lib.heom_port_apply(C.byref(self.tables), xf.ctypes.data, yf.ctypes.data, mode)

    def steady_rhs(self, trace_row="last"):
        """(x0, b) of solve_tfqmr.f90:97-179 in the reference layout as real
        vectors: x0 = identity / d on tier 0, b = e_trace - A x0."""
        m = self.model
        d = m.d
        x0 = np.zeros((m.N, d, d), complex)
        x0[0] = np.eye(d) / d
        e = np.zeros_like(x0)
        e[-1 if trace_row == "last" else 0, d - 1, d - 1] = 1.0
        b = O.to_fortran(e - self.apply(x0, trace_row=trace_row))
        return O.to_fortran(x0).view(np.float64).ravel(), np.ascontiguousarray(b).view(np.float64).ravel()

    def tfqmr(self, b, nit, parallel=False, trace_row="last", x=None, work=None, state=None, resume=False):
        """nit TFQMR iterations on A dx = b (real vectors in the reference
        layout), the vector passes serial (the reference's DUTFX) or
        OpenMP-parallel; from dx = 0, or (resume) continuing the recurrences
        kept in x, work and state.  Returns (dx, quasi-residual bound)."""
        mode = {None: 0, "last": 1, "rdo": 2}[trace_row]
        n = b.size
        x = np.zeros(n) if x is None else x
        work = np.zeros(8 * n) if work is None else work
        state = np.zeros(5) if state is None else state
        bound = self.lib.heom_port_tfqmr(C.byref(self.tables), b.ctypes.data, x.ctypes.data, work.ctypes.data,
                                         state.ctypes.data, int(nit), int(bool(parallel)), mode, int(not resume))
        return x, bound
