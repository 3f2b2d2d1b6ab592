"""heomq -- MI355X-native HEOM Liouvillian engine (host binding).

Python face of libheomq.so.  The model tables are built natively
(``Model``), uploaded once to the GPU (``Context``), and every apply / solve
runs as HIP kernels on torch-owned device buffers.  Vectors are complex128
(or complex64) tensors of shape ``(N, d, d)`` in the reference's Fortran
element order: ``v[a, j, i]`` is element (i, j) of ADO a, i.e. the memory of
``rcgtmp(dim0, dim0, nunk)`` in tfqmr_double/calc_ax_tfqmr_omp.f90.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (HAM_DQD, HAM_SIAM, TRACE_LAST, TRACE_NONE, TRACE_RDO, HeomqError, check, lib)

__all__ = ["Params", "Model", "Context", "siam", "two_impurity", "HeomqError",
           "TRACE_NONE", "TRACE_LAST", "TRACE_RDO", "HBAR", "JCONST"]

HBAR = 0.658211928
JCONST = 1.602177e5


@dataclass
class Params:
    """Physical model (the reference's positional input + $para1/$para4)."""
    norbs: int = 1
    nalf: int = 2
    nmats: int = 8
    ntop: int = 3
    bandwidth: tuple = (10.0, 10.0)
    linewidth: tuple = ((0.1,), (0.1,))
    temperature: tuple = (0.05, 0.05)
    engyshift: tuple = ((0.0, 0.0), (0.0, 0.0))
    ham: int = HAM_SIAM
    hpar: tuple = ()
    lsdraw: bool = False
    steady: bool = True
    lscale: bool = False        # ADOs stored scaled by prod_k |cb_k|^(-1/2) (the reference's lscale)

    def to_c(self) -> _lib.CParams:
        p = _lib.CParams()
        p.norbs, p.nspin, p.nalf, p.nmats, p.ntop = self.norbs, 2, self.nalf, self.nmats, self.ntop
        p.lsdraw, p.ham, p.steady = int(self.lsdraw), int(self.ham), int(self.steady)
        p.lscale = int(self.lscale)
        for a in range(min(self.nalf, 4)):
            p.bandwidth[a] = self.bandwidth[a]
            p.temperature[a] = self.temperature[a]
            for m in range(min(self.norbs, len(self.linewidth[a]), 4)):
                p.linewidth[a][m] = self.linewidth[a][m]
            for s in range(2):
                p.engyshift[a][s] = self.engyshift[a][s]
        for i, v in enumerate(self.hpar[:16]):
            p.hpar[i] = v
        return p


def siam(eps=-1.0, U=2.0, gamma=0.1, W=10.0, T=0.05, V=0.0, nk=8, ntop=3, gl=None, gr=None,
         lsdraw=False, lscale=False) -> Params:
    """Single-impurity Anderson model; bias split +V/2 on the left lead, -V/2 on the right."""
    gl = gamma if gl is None else gl
    gr = gamma if gr is None else gr
    return Params(norbs=1, nalf=2, nmats=nk, ntop=ntop, bandwidth=(W, W), linewidth=((gl,), (gr,)),
                  temperature=(T, T), engyshift=((V / 2, V / 2), (-V / 2, -V / 2)), ham=HAM_SIAM,
                  hpar=(eps, eps, U, 0.0), lsdraw=lsdraw, lscale=lscale)


def two_impurity(eps=-1.5, U=3.0, t=0.5, u12=0.0, gamma=0.2, W=10.0, T=0.1, V=0.4, nk=6, ntop=3,
                 lsdraw=False, lscale=False) -> Params:
    """Two-impurity model ($para4): both dots coupled to both leads with the same Gamma."""
    return Params(norbs=2, nalf=2, nmats=nk, ntop=ntop, bandwidth=(W, W),
                  linewidth=((gamma, gamma), (gamma, gamma)), temperature=(T, T),
                  engyshift=((V / 2, V / 2), (-V / 2, -V / 2)), ham=HAM_DQD,
                  hpar=(eps, eps, eps, eps, U, U, u12, t), lsdraw=lsdraw, lscale=lscale)


class Batch:
    """Many problems of one hierarchy shape solved together on one GPU.

    Problems must share norbs, nalf, nmats (Pade terms), ntop, lsdraw and the
    Hamiltonian kind; energies, Gamma, temperature and bias may differ.
    """

    def __init__(self, params_list, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("heomq.Batch needs a GPU (HIP device); none is visible")
        self.torch = torch
        self.params = list(params_list)
        arr = (_lib.CParams * len(self.params))(*[p.to_c() for p in self.params])
        self._h = C.c_void_p()
        torch.cuda.set_device(device)
        check(lib().heomq_batch_create(arr, len(self.params), device, C.byref(self._h)))
        info = np.zeros(8, np.int64)
        check(lib().heomq_batch_info(self._h, info.ctypes.data_as(_lib.I64)))
        self.nprob, self.N, self.packed, self.d = (int(v) for v in info[:4])
        self.device = device
        _LIVE_BATCH.add(self)

    def close(self):
        if self._h:
            self.torch.cuda.synchronize(self.device)
            lib().heomq_batch_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply_packed(self, xp, yp, trace=TRACE_RDO):
        """y_p = Pi A_p x_p for every problem; xp, yp: (nprob, packed) complex128 device tensors."""
        t = self.torch
        for v in (xp, yp):
            if v.dtype != t.complex128 or not v.is_cuda or not v.is_contiguous() or v.numel() != self.nprob * self.packed:
                raise HeomqError(_lib.BAD_DIM, "expected (nprob, packed) complex128 device tensors")
        check(lib().heomq_batch_apply_packed(self._h, C.c_void_p(xp.data_ptr()), C.c_void_p(yp.data_ptr()), trace,
                                             C.c_void_p(t.cuda.current_stream(self.device).cuda_stream)))
        return yp

    def solve(self, tol=1e-10, maxiter=20000, raise_on_fail=True):
        """Steady states of all problems: (rdo (P, d, d) math-indexed, current (P, nalf, 2), occ (P, norbs, 2), stats)."""
        P, d = self.nprob, self.d
        nalf = self.params[0].nalf
        norbs = self.params[0].norbs
        rdo = np.zeros((P, d, d), np.complex128)
        cur = np.zeros((P, nalf, 2))
        occ = np.zeros((P, norbs, 2))
        stats = np.zeros((P, 4))
        rc = lib().heomq_batch_solve(self._h, tol, maxiter, _dp(rdo.view(np.float64)), _dp(cur), _dp(occ), _dp(stats),
                                     C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream))
        if rc != _lib.OK and raise_on_fail:
            check(rc)
        rdo = np.ascontiguousarray(rdo.transpose(0, 2, 1))   # column-major blocks -> [p, i, j]
        info = dict(iters=stats[:, 0].astype(int), relres=stats[:, 1], status=stats[:, 2].astype(int),
                    restarts=stats[:, 3].astype(int), rc=rc)
        return rdo, cur, occ, info


class Sweep:
    """Problems of several hierarchy shapes (Pade terms, top tier) solved
    together on one GPU, one launch per solver step over all of them.

    Problems must share norbs, nalf and d; up to 16 shapes.  Outputs come in
    the order of ``params_list``.
    """

    def __init__(self, params_list, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("heomq.Sweep needs a GPU (HIP device); none is visible")
        self.torch = torch
        self.params = list(params_list)
        arr = (_lib.CParams * len(self.params))(*[p.to_c() for p in self.params])
        self._h = C.c_void_p()
        torch.cuda.set_device(device)
        check(lib().heomq_sweep_create(arr, len(self.params), device, C.byref(self._h)))
        info = np.zeros(8, np.int64)
        check(lib().heomq_sweep_info(self._h, info.ctypes.data_as(_lib.I64)))
        self.nprob, self.nshapes, self.packed, self.d, self.apply_blocks, self.vector_blocks = (int(v) for v in info[:6])
        P = self.nprob
        self.shape = np.zeros(P, np.int32)
        self.offset = np.zeros(P, np.int64)
        self.length = np.zeros(P, np.int64)
        check(lib().heomq_sweep_problems(self._h, self.shape.ctypes.data_as(_lib.I32),
                                         self.offset.ctypes.data_as(_lib.I64), self.length.ctypes.data_as(_lib.I64)))
        self.device = device
        _LIVE_BATCH.add(self)

    def close(self):
        if self._h:
            self.torch.cuda.synchronize(self.device)
            lib().heomq_sweep_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply_packed(self, xp, yp):
        """y = Pi A x for every problem on flat-layout complex128 device vectors
        (problem p at ``offset[p]``, ``length[p]`` elements), trace row on tier 0."""
        t = self.torch
        for v in (xp, yp):
            if v.dtype != t.complex128 or not v.is_cuda or not v.is_contiguous() or v.numel() != self.packed:
                raise HeomqError(_lib.BAD_DIM, "expected flat complex128 device vectors of the sweep's size")
        check(lib().heomq_sweep_apply_packed(self._h, C.c_void_p(xp.data_ptr()), C.c_void_p(yp.data_ptr()), TRACE_RDO,
                                             C.c_void_p(t.cuda.current_stream(self.device).cuda_stream)))
        return yp

    def solve(self, tol=1e-10, maxiter=20000, raise_on_fail=True):
        """Steady states of all problems, as Batch.solve."""
        P, d = self.nprob, self.d
        nalf = self.params[0].nalf
        norbs = self.params[0].norbs
        rdo = np.zeros((P, d, d), np.complex128)
        cur = np.zeros((P, nalf, 2))
        occ = np.zeros((P, norbs, 2))
        stats = np.zeros((P, 4))
        rc = lib().heomq_sweep_solve(self._h, tol, maxiter, _dp(rdo.view(np.float64)), _dp(cur), _dp(occ), _dp(stats),
                                     C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream))
        if rc != _lib.OK and raise_on_fail:
            check(rc)
        rdo = np.ascontiguousarray(rdo.transpose(0, 2, 1))
        info = dict(iters=stats[:, 0].astype(int), relres=stats[:, 1], status=stats[:, 2].astype(int),
                    restarts=stats[:, 3].astype(int), rc=rc)
        return rdo, cur, occ, info


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_lib.D)


_LIVE_CTX: "weakref.WeakSet" = weakref.WeakSet()
_LIVE_MODELS: "weakref.WeakSet" = weakref.WeakSet()
_LIVE_BATCH: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _release_all():
    """Free batches and device contexts (then models) while the HIP runtime is still up."""
    for b in list(_LIVE_BATCH):
        b.close()
    for c in list(_LIVE_CTX):
        c.close()
    for m in list(_LIVE_MODELS):
        m.close()


class Model:
    """Host tables: bath coefficients, operators, hierarchy, coupling table."""

    def __init__(self, params: Params | None = None, _handle=None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            p = params.to_c()
            check(lib().heomq_model_build(C.byref(p), C.byref(self._h)))
        self.params = params
        info = np.zeros(8, np.int64)
        first = np.zeros(16, np.int64)
        check(lib().heomq_model_info(self._h, info.ctypes.data_as(_lib.I64), first.ctypes.data_as(_lib.I64)))
        self.d, self.M, self.L, self.N, self.ncor, self.nopr = (int(v) for v in info[:6])
        self.n_lower, self.n_upper = int(info[6]), int(info[7])
        self.first = first[: self.L + 2].copy()
        # orbitals and leads from the model itself (a model loaded from the
        # reference's arrays has no parameters): d = 4**norbs, nopr = 2 norbs nspin nalf
        self.norbs = int(round(np.log(self.d) / np.log(4)))
        self.nalf = self.nopr // (2 * self.norbs * 2)
        _LIVE_MODELS.add(self)

    def close(self):
        if self._h:
            lib().heomq_model_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_reference(cls, **arrays) -> "Model":
        """Build from the reference's own module arrays (see include/heomq.h).

        ``ntable`` and ``lall`` (the lengths the reference keeps in its common
        block info3) default to the lengths of ``indextable`` and ``indexcoef``;
        ``dbsqrt`` (the reference's lscale factors) is optional.
        """
        a = arrays
        h = C.c_void_p()
        keep = []

        def ptr(name, dtype, ctype):
            v = np.ascontiguousarray(a[name], dtype=dtype)
            keep.append(v)
            return v.ctypes.data_as(ctype)

        check(lib().heomq_model_from_reference(
            int(a["nrho"]), int(a["nunk"]), int(a["ntier"]), int(a["norbs"]), int(a["nspin"]),
            int(a["nalf"]), int(a["ncor"]), int(a["igroundsteady"]),
            int(a.get("ntable", np.size(a["indextable"]))), int(a.get("lall", np.size(a["indexcoef"]))),
            ptr("indextable", np.int32, _lib.I32), ptr("nfirst", np.int64, _lib.I64),
            ptr("nlast", np.int64, _lib.I64), ptr("ifirst", np.int64, _lib.I64),
            ptr("indexcoef", np.int64, _lib.I64), ptr("itypecoef", np.int8, _lib.I8),
            ptr("ifactfcoef", np.int8, _lib.I8), ptr("ifactrcoef", np.int8, _lib.I8),
            ptr("index_coef_ref", np.int64, _lib.I64), ptr("indexdraw", np.int32, _lib.I32),
            ptr("cb", np.complex128, _lib.D), ptr("cd", np.complex128, _lib.D),
            ptr("cgamma", np.complex128, _lib.D), ptr("dipm", np.float64, _lib.D),
            ptr("ilead", np.int32, _lib.I32), ptr("mspin", np.int32, _lib.I32),
            ptr("morbs", np.int32, _lib.I32), ptr("mpm", np.int32, _lib.I32),
            ptr("eleadinfty", np.float64, _lib.D), ptr("hs", np.complex128, _lib.D),
            ptr("amsall", np.float64, _lib.D),
            ptr("dbsqrt", np.float64, _lib.D) if a.get("dbsqrt") is not None else None, C.byref(h)))
        return cls(None, _handle=h)

    def set_params(self, params: Params):
        p = params.to_c()
        check(lib().heomq_model_set_params(self._h, C.byref(p)))
        self.params = params

    def coefs(self):
        cb = np.zeros(self.M, np.complex128)
        cd = np.zeros(self.M, np.complex128)
        g = np.zeros(self.M, np.complex128)
        check(lib().heomq_model_coefs(self._h, _dp(cb.view(np.float64)), _dp(cd.view(np.float64)),
                                      _dp(g.view(np.float64))))
        return cb, cd, g

    def operators(self):
        Hf = np.zeros(self.d * self.d, np.complex128)
        nops = self.nopr // (2 * self.nalf)
        ams = np.zeros(self.d * self.d * nops, np.float64)
        check(lib().heomq_model_operators(self._h, _dp(Hf.view(np.float64)), _dp(ams)))
        H = Hf.reshape(self.d, self.d).T.copy()                     # column-major -> H[i, j]
        A = ams.reshape(nops, self.d, self.d).transpose(0, 2, 1).copy()   # A[m*nspin+s][i, j]
        return H, A

    def number_operator(self, orb: int, spin: int) -> np.ndarray:
        """n = c^+ c of one orbital and spin (math-indexed d x d), the
        operators calcdhs builds its dot fields from."""
        _, A = self.operators()
        c = A[orb * 2 + spin]
        return c.conj().T @ c

    def tier(self, n: int) -> np.ndarray:
        cnt = int(self.first[n + 1] - self.first[n])
        out = np.zeros((cnt, n), np.int32)
        check(lib().heomq_model_tier(self._h, n, out.ctypes.data_as(_lib.I32)))
        return out

    def coupling(self):
        lower = np.zeros((self.n_lower_total(), 4), np.int32)
        upper = np.zeros((int(self.first[self.L]) * self.M, 4), np.int32)
        check(lib().heomq_model_coupling(self._h, lower.ctypes.data_as(_lib.I32), upper.ctypes.data_as(_lib.I32)))
        return lower, upper

    def packed_layout(self) -> dict:
        """The symmetry-packed layout the device solvers use, built on the host."""
        info = np.zeros(12, np.int64)
        check(lib().heomq_model_packed_layout(self._h, info.ctypes.data_as(_lib.I64)))
        keys = ("packed", "sectors", "n_conserved", "sz_conserved", "dense", "n_heavy", "light_begin",
                "light_blocks", "max_ados_per_light_block", "empty_ados", "max_light_couplings", "map_entries")
        return dict(zip(keys, (int(v) for v in info)))

    def spa_pattern(self, dsmall=1e-32, seed=0):
        """The sparse-mode pattern fermion/prelude_spa.f90:532-749 builds, in the
        reference's module sparsemod arrays (1-based): (nnz_spa, ind_spa,
        irow_spa, icol_spa).  dsmall is the reference's dsmall_spa (1e-32; 1e-12
        with lproj); seed fixes the mixing coefficients."""
        n = np.int64(0)
        lunk = np.zeros(1, np.int64)
        nnz = np.zeros(self.N, np.int32)
        ind = np.zeros(self.N, np.int64)
        check(lib().heomq_spa_build_pattern(self._h, dsmall, seed, 0, lunk.ctypes.data_as(_lib.I64),
                                            nnz.ctypes.data_as(_lib.I32), ind.ctypes.data_as(_lib.I64), None, None))
        n = int(lunk[0])
        ir = np.zeros(max(n, 1), np.int32)
        ic = np.zeros(max(n, 1), np.int32)
        check(lib().heomq_spa_build_pattern(self._h, dsmall, seed, n, lunk.ctypes.data_as(_lib.I64), None, None,
                                            ir.ctypes.data_as(_lib.I32), ic.ctypes.data_as(_lib.I32)))
        return nnz, ind, ir[:n], ic[:n]

    def ado_scale(self) -> np.ndarray:
        """Per ADO, the factor between the physical ADO and the stored one:
        prod_k |cb_k|^(1/2) over its drawers with lscale (the stored ADOs are
        rho_K / that product), 1 without."""
        s = np.ones(self.N)
        if self.params is None or not self.params.lscale:
            return s
        cbs = np.abs(self.coefs()[0])     # scaled cb: |cb| / |cb|^(1/2) = |cb|^(1/2)
        for n in range(1, self.L + 1):
            s[self.first[n]:self.first[n + 1]] = np.prod(cbs[self.tier(n)], axis=1)
        return s

    def filter_pattern(self, x, thresh, pattern=None):
        """The ADO norm filter: the sparse pattern (default: spa_pattern()) with
        every ADO whose physical Frobenius norm in the state x is below thresh
        left empty; tier 0 is always kept.  x is a state as the context holds
        it ((N, d, d), Fortran element order); with lscale its ADOs are scaled
        back by ado_scale() first.
        Returns (nnz_spa, ind_spa, irow_spa, icol_spa, kept) with kept the
        boolean mask of ADOs that hold elements."""
        nnz, ind, ir, ic = self.spa_pattern() if pattern is None else pattern
        xt = x if isinstance(x, np.ndarray) else x.reshape(self.N, -1).cpu().numpy()
        norms = np.linalg.norm(xt.reshape(self.N, -1), axis=1) * self.ado_scale()
        keep = norms >= thresh
        keep[0] = True
        sel = np.repeat(keep, nnz)
        nnz2 = np.where(keep, nnz, 0).astype(np.int32)
        ind2 = (1 + np.concatenate([[0], np.cumsum(nnz2)[:-1]])).astype(np.int64)
        return nnz2, ind2, ir[sel], ic[sel], keep & (nnz > 0)

    def jacobi_info(self):
        """(info dict, eigenvalues of H, eigenvectors of H) of the block-Jacobi
        iteration, built on the host (prelude_jacobi.f90:43-113)."""
        info = np.zeros(5)
        ev = np.zeros(self.d)
        u = np.zeros(self.d * self.d, np.complex128)
        check(lib().heomq_model_jacobi_info(self._h, _dp(info), _dp(ev), _dp(u.view(np.float64))))
        keys = ("ejac", "gap", "cb_term", "nnzjac", "group_max")
        return dict(zip(keys, (float(v) for v in info))), ev, u.reshape(self.d, self.d, order="F")

    def n_lower_total(self) -> int:
        return self.n_lower


class Context:
    """Device-resident tables on one GPU plus the solver workspace."""

    def __init__(self, model: Model, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("heomq.Context needs a GPU (HIP device); none is visible")
        self.torch = torch
        self.model = model
        self.device = device
        self._h = C.c_void_p()
        torch.cuda.set_device(device)
        check(lib().heomq_ctx_create(model._h, device, C.byref(self._h)))
        self.N, self.d = model.N, model.d
        self._ptotal = self.packed_info()["packed"]
        _LIVE_CTX.add(self)

    def close(self):
        if self._h:
            self.torch.cuda.synchronize(self.device)
            lib().heomq_ctx_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- helpers ----
    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def zeros(self, dtype=None):
        dtype = dtype or self.torch.complex128
        return self.torch.zeros((self.N, self.d, self.d), dtype=dtype, device=f"cuda:{self.device}")

    def _check_vec(self, v, dtype):
        t = self.torch
        if v.dtype != dtype or not v.is_cuda or not v.is_contiguous() or v.numel() != self.N * self.d * self.d:
            raise HeomqError(_lib.BAD_DIM, f"expected contiguous {dtype} device tensor with {self.N * self.d * self.d} elements")
        return C.c_void_p(v.data_ptr())

    def set_precond(self, mode: int, shift: float | None = None):
        """Left preconditioning of forward steady solves: -1 automatic (Jacobi
        scaling for diagonal H, block Jacobi otherwise), 0 none, 1 Jacobi
        scaling, 2 block Jacobi with the given shift (None / 0: default)."""
        check(lib().heomq_ctx_set_precond(self._h, int(mode)))
        if shift is not None:
            check(lib().heomq_ctx_set_precond_shift(self._h, float(shift)))

    def set_shadow_seed(self, seed: int):
        """Seed of TFQMR's +-1 shadow vector at each solve's first start."""
        check(lib().heomq_ctx_set_shadow_seed(self._h, int(seed)))

    def set_presmooth(self, enable: bool):
        """Two block-Jacobi sweeps on the start of block-Jacobi-preconditioned
        steady solves (default on)."""
        check(lib().heomq_ctx_set_presmooth(self._h, int(enable)))

    def set_precond_fusion(self, enable: bool):
        """Block-Jacobi-preconditioned applies as one fused launch where the
        apply form allows it (default on; the same bits as the two kernels)."""
        check(lib().heomq_ctx_set_precond_fusion(self._h, int(enable)))

    def set_trun_der(self, enable: bool = True):
        """The derivative terminator of the top tier ($jobinfo ltrun_der,
        calc_ax_tfqmr_omp.f90:362-478): the forward packed applies and the
        steady TFQMR solve then include it; the other solvers and propagators
        refuse to run while it is set."""
        check(lib().heomq_ctx_set_trun_der(self._h, int(enable)))

    def autotune(self, reps: int = 20) -> dict:
        """Time the packed apply's forms on this GPU and keep the fastest per
        operator; returns the choices and times (us; None for forms the layout
        does not take)."""
        forms = np.zeros(2, np.int32)
        times = np.full(16, np.nan)
        check(lib().heomq_ctx_autotune(self._h, int(reps), forms.ctypes.data_as(_lib.I32), _dp(times)))
        clean = [None if np.isnan(v) else float(v) for v in times]
        return dict(forward=int(forms[0]), transposed=int(forms[1]),
                    times_us={"forward": clean[:8], "transposed": clean[8:]})

    def apply_form(self):
        """The packed apply's forms in use (forward, transposed)."""
        forms = np.zeros(2, np.int32)
        check(lib().heomq_ctx_get_apply_form(self._h, forms.ctypes.data_as(_lib.I32)))
        return int(forms[0]), int(forms[1])

    def set_apply_form(self, forward: int, transposed: int | None = None):
        """Fix the packed apply's form (bits as autotune); returns the forms in
        use (a form the layout does not take falls back per part)."""
        check(lib().heomq_ctx_set_apply_form(self._h, int(forward), int(forward if transposed is None else transposed)))
        forms = np.zeros(2, np.int32)
        check(lib().heomq_ctx_get_apply_form(self._h, forms.ctypes.data_as(_lib.I32)))
        return int(forms[0]), int(forms[1])

    def set_graphs(self, enable: bool):
        """Replay solver iterations from captured graphs (default) or launch eagerly."""
        check(lib().heomq_ctx_set_graphs(self._h, int(enable)))

    def refresh(self, model: Model | None = None):
        check(lib().heomq_ctx_refresh(self._h, (model or self.model)._h))

    # ---- applies ----
    def apply(self, x, y=None, trace=TRACE_LAST):
        t = self.torch
        y = self.zeros() if y is None else y
        check(lib().heomq_apply(self._h, self._check_vec(x, t.complex128), self._check_vec(y, t.complex128),
                                trace, self._stream()))
        return y

    def apply_transpose(self, x, y=None, trace=TRACE_LAST):
        t = self.torch
        y = self.zeros() if y is None else y
        check(lib().heomq_apply_transpose(self._h, self._check_vec(x, t.complex128),
                                          self._check_vec(y, t.complex128), trace, self._stream()))
        return y

    def apply_f32(self, x, y=None, trace=TRACE_LAST):
        t = self.torch
        y = self.zeros(t.complex64) if y is None else y
        check(lib().heomq_apply_f32(self._h, self._check_vec(x, t.complex64), self._check_vec(y, t.complex64),
                                    trace, self._stream()))
        return y

    def apply_packed(self, x, y=None, trace=TRACE_RDO, transpose=False):
        """Projected apply through the symmetry-packed layout (dense in/out)."""
        t = self.torch
        y = self.zeros() if y is None else y
        check(lib().heomq_apply_packed(self._h, self._check_vec(x, t.complex128), self._check_vec(y, t.complex128),
                                       trace, int(transpose), self._stream()))
        return y

    def packed_zeros(self, dtype=None):
        n = self.packed_info()["packed"]
        return self.torch.zeros(n, dtype=dtype or self.torch.complex128, device=f"cuda:{self.device}")

    def _check_packed(self, v, dtype):
        if v.dtype != dtype or not v.is_cuda or not v.is_contiguous() or v.numel() != self._ptotal:
            raise HeomqError(_lib.BAD_DIM, f"expected contiguous {dtype} device tensor with {self._ptotal} elements")
        return C.c_void_p(v.data_ptr())

    def apply_packed_precond(self, xp, yp=None, trace=TRACE_RDO, form=1):
        """y = M^-1 (Pi A x) on packed vectors with the block-Jacobi preconditioner
        (form 1: apply, then the one-kernel inverse; 2: two-kernel transform; 3: LDS-window kernel;
        4: apply and inverse fused in one launch)."""
        t = self.torch
        yp = self.packed_zeros() if yp is None else yp
        check(lib().heomq_apply_packed_precond(self._h, self._check_packed(xp, t.complex128),
                                               self._check_packed(yp, t.complex128), trace, int(form),
                                               self._stream()))
        return yp

    def apply_packed_raw(self, xp, yp, trace=TRACE_RDO, transpose=False):
        t = self.torch
        if xp.dtype == t.complex64:
            check(lib().heomq_apply_packed_raw_f32(self._h, self._check_packed(xp, t.complex64),
                                                   self._check_packed(yp, t.complex64), trace, int(transpose),
                                                   self._stream()))
        else:
            check(lib().heomq_apply_packed_raw(self._h, self._check_packed(xp, t.complex128),
                                               self._check_packed(yp, t.complex128), trace, int(transpose),
                                               self._stream()))
        return yp

    def pack(self, x, xp=None):
        t = self.torch
        xp = self.packed_zeros() if xp is None else xp
        check(lib().heomq_pack(self._h, self._check_vec(x, t.complex128), self._check_packed(xp, t.complex128),
                               self._stream()))
        return xp

    def unpack(self, xp, x=None):
        t = self.torch
        x = self.zeros() if x is None else x
        check(lib().heomq_unpack(self._h, self._check_packed(xp, t.complex128), self._check_vec(x, t.complex128),
                                 self._stream()))
        return x

    def packed_info(self):
        info = np.zeros(8, np.int64)
        check(lib().heomq_packed_info(self._h, info.ctypes.data_as(_lib.I64)))
        return dict(packed=int(info[0]), sectors=int(info[1]), conserves_n=bool(info[2]),
                    conserves_sz=bool(info[3]), dense=int(info[4]))

    def apply_dparam(self, x, dH, dcb, dcd, dgam, y=None):
        t = self.torch
        y = self.zeros() if y is None else y
        dH = np.asfortranarray(dH, dtype=np.complex128)
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (dcb, dcd, dgam)]
        check(lib().heomq_apply_dparam(self._h, self._check_vec(x, t.complex128), self._check_vec(y, t.complex128),
                                       _dp(dH.reshape(-1, order="F").view(np.float64)),
                                       *[_dp(a.view(np.float64)) for a in arrs], self._stream()))
        return y

    def apply_bicg(self, x, y=None):
        """BiCG-convention apply (calc_ax_bicg_omp); x's tier 0 is rebuilt in place from its lower triangle."""
        t = self.torch
        y = self.zeros() if y is None else y
        check(lib().heomq_apply_bicg(self._h, self._check_vec(x, t.complex128), self._check_vec(y, t.complex128),
                                     self._stream()))
        return y

    def apply_transpose_bicg(self, w, z=None):
        """Transpose of the BiCG-convention apply (calc_atrx_bicg_omp); w's tier 0 is reduced in place."""
        t = self.torch
        z = self.zeros() if z is None else z
        check(lib().heomq_apply_transpose_bicg(self._h, self._check_vec(w, t.complex128),
                                               self._check_vec(z, t.complex128), self._stream()))
        return z

    def solve_bicg(self, x=None, crit=1e-10, maxiter=20000, raise_on_fail=True):
        """Steady state by BiCG (solvebicg.f90 convention: absolute ||r|| <= crit)."""
        t = self.torch
        x = self.initial_guess() if x is None else x
        stats = np.zeros(_lib.STATS_LEN)
        rc = lib().heomq_solve_bicg(self._h, self._check_vec(x, t.complex128), crit, maxiter, _dp(stats),
                                    self._stream())
        info = dict(iters=int(stats[0]), applies=int(stats[1]), resid=stats[2], seconds=stats[3], status=rc,
                    packed=bool(stats[4]))
        if rc not in (_lib.OK, _lib.NOT_CONVERGED) or (rc == _lib.NOT_CONVERGED and raise_on_fail):
            check(rc)
        return x, info

    def set_spa_pattern(self, nnz_spa, ind_spa, irow_spa, icol_spa):
        """The reference's sparse (COO per ADO) pattern, 1-based as module sparsemod holds it."""
        nnz = np.ascontiguousarray(nnz_spa, np.int32)
        ind = np.ascontiguousarray(ind_spa, np.int64)
        ir = np.ascontiguousarray(irow_spa, np.int32)
        ic = np.ascontiguousarray(icol_spa, np.int32)
        if nnz.size != self.N or ind.size != self.N or ir.size != ic.size:
            raise HeomqError(_lib.BAD_DIM, "nnz_spa / ind_spa need N entries, irow_spa / icol_spa lunk_spa each")
        self._spa_lunk = int(ir.size)
        self._spa_pat = (nnz.copy(), ir.copy(), ic.copy())
        check(lib().heomq_spa_pattern(self._h, ir.size, nnz.ctypes.data_as(_lib.I32), ind.ctypes.data_as(_lib.I64),
                                      ir.ctypes.data_as(_lib.I32), ic.ctypes.data_as(_lib.I32)))

    def apply_spa(self, x_spa, y_spa=None):
        """calc_ax_tfqmr_spa_omp on device vectors: y (lunk_spa + 1) from x (lunk_spa)."""
        t = self.torch
        n = getattr(self, "_spa_lunk", -1)
        if y_spa is None:
            y_spa = t.zeros(n + 1, dtype=t.complex128, device=f"cuda:{self.device}")
        for v, m in ((x_spa, n), (y_spa, n + 1)):
            if v.dtype != t.complex128 or not v.is_cuda or not v.is_contiguous() or v.numel() != m:
                raise HeomqError(_lib.BAD_DIM, "apply_spa: contiguous complex128 device vectors of lunk_spa / lunk_spa + 1")
        check(lib().heomq_apply_spa(self._h, C.c_void_p(x_spa.data_ptr()), C.c_void_p(y_spa.data_ptr()),
                                    self._stream()))
        return y_spa

    def spa_initial_guess(self):
        """rho_spa of the trace-one start: identity / d on tier 0's diagonal
        elements of the registered pattern, every other element zero."""
        t = self.torch
        n = getattr(self, "_spa_lunk", -1)
        if n < 0:
            raise HeomqError(_lib.BAD_ARG, "no sparse pattern: call set_spa_pattern first")
        nnz0 = int(self._spa_pat[0][0])
        ir, ic = self._spa_pat[1][:nnz0], self._spa_pat[2][:nnz0]
        x = np.zeros(n, np.complex128)
        x[:nnz0][ir == ic] = 1.0 / self.model.d
        return t.from_numpy(x).to(f"cuda:{self.device}")

    def solve_steady_spa(self, x_spa=None, tol=1e-12, maxiter=20000, raise_on_fail=True):
        """The steady state on the registered sparse pattern (solve_tfqmr with
        lsparse = T): TFQMR on P A P^T over rho_spa, preconditioned as the
        packed solve.  Returns (x_spa, stats dict)."""
        t = self.torch
        x_spa = self.spa_initial_guess() if x_spa is None else x_spa
        n = getattr(self, "_spa_lunk", -1)
        if x_spa.dtype != t.complex128 or not x_spa.is_cuda or not x_spa.is_contiguous() or x_spa.numel() != n:
            raise HeomqError(_lib.BAD_DIM, "solve_steady_spa: a contiguous complex128 device vector of lunk_spa")
        stats = np.zeros(_lib.STATS_LEN)
        rc = lib().heomq_spa_solve_steady(self._h, C.c_void_p(x_spa.data_ptr()), tol, maxiter, _dp(stats),
                                          self._stream())
        info = dict(iters=int(stats[0]), applies=int(stats[1]), relres=stats[2], seconds=stats[3], status=rc,
                    precond=int(stats[4]))
        if rc not in (_lib.OK, _lib.NOT_CONVERGED) or (rc == _lib.NOT_CONVERGED and raise_on_fail):
            check(rc)
        return x_spa, info

    def calcderiv_spa(self, x_spa, tt=0.0, y_spa=None):
        """calcderiv_spa_omp(tt) on device vectors: y = P L(tt) P^T x over the
        sparse pattern (lunk_spa each), the lead field (set_field) at tt."""
        t = self.torch
        n = getattr(self, "_spa_lunk", -1)
        if y_spa is None:
            y_spa = t.zeros(max(n, 0), dtype=t.complex128, device=f"cuda:{self.device}")
        for v in (x_spa, y_spa):
            if v.dtype != t.complex128 or not v.is_cuda or not v.is_contiguous() or v.numel() != n:
                raise HeomqError(_lib.BAD_DIM, "calcderiv_spa: contiguous complex128 device vectors of lunk_spa")
        check(lib().heomq_calcderiv_spa_at(self._h, float(tt), C.c_void_p(x_spa.data_ptr()),
                                           C.c_void_p(y_spa.data_ptr()), self._stream()))
        return y_spa

    def jacobi_info(self) -> dict:
        info = np.zeros(5)
        check(lib().heomq_jacobi_info(self._h, _dp(info)))
        return dict(zip(("ejac", "gap", "cb_term", "nnzjac", "group_max"), (float(v) for v in info)))

    def jacobi_step(self, x, y=None, ejac=0.0):
        """One block-Jacobi sweep (calc_ax_jacobi_omp): y = new iterate from x."""
        t = self.torch
        y = self.zeros() if y is None else y
        check(lib().heomq_jacobi_step(self._h, self._check_vec(x, t.complex128), self._check_vec(y, t.complex128),
                                      float(ejac), self._stream()))
        return y

    def solve_jacobi(self, x=None, crit=1e-8, maxiter=20000, ejac=0.0, raise_on_fail=True):
        """Steady state by the block-Jacobi iteration (solve_jacobi.f90: absolute
        trace-row residual <= crit, checked every 50 sweeps)."""
        t = self.torch
        x = self.initial_guess() if x is None else x
        stats = np.zeros(_lib.STATS_LEN)
        rc = lib().heomq_solve_jacobi(self._h, self._check_vec(x, t.complex128), crit, maxiter, float(ejac),
                                      _dp(stats), self._stream())
        info = dict(iters=int(stats[0]), applies=int(stats[1]), resid=stats[2], seconds=stats[3], ejac=stats[4],
                    best=stats[5], status=rc)
        if rc not in (_lib.OK, _lib.NOT_CONVERGED) or (rc == _lib.NOT_CONVERGED and raise_on_fail):
            check(rc)
        return x, info

    # ---- solvers ----
    def random_state(self, seed):
        """Dense state whose tier-0 block is a seeded random density matrix,
        G G^H with G complex normal, reduced to the elements the symmetry
        sectors allow (its (N, S_z) blocks, so it stays positive) and
        normalised to trace 1; every higher tier is zero."""
        t = self.torch
        rng = np.random.default_rng(seed)
        g = rng.standard_normal((self.d, self.d)) + 1j * rng.standard_normal((self.d, self.d))
        r = g @ g.conj().T
        x = self.zeros()
        x[0] = t.from_numpy(np.ascontiguousarray(r.T)).to(x.device)   # block (i, j) stored at [j, i]
        x = self.unpack(self.pack(x))
        x[0] /= t.diagonal(x[0]).sum()
        return x

    def initial_guess(self):
        x = self.zeros()
        x[0] = self.torch.eye(self.d, dtype=self.torch.complex128, device=x.device) / self.d
        return x

    def solve_steady(self, x=None, b=None, tol=1e-12, maxiter=20000, trace=TRACE_RDO, transpose=False,
                     raise_on_fail=True):
        """TFQMR steady state on the device; returns (x, stats dict): iters,
        applies, relres, seconds, status, precond (the left preconditioner
        the iteration ended with) and fused_iteration (block-Jacobi
        iterations ran with their vector work in the operator launches)."""
        t = self.torch
        x = self.initial_guess() if x is None else x
        stats = np.zeros(_lib.STATS_LEN)
        bp = self._check_vec(b, t.complex128) if b is not None else None
        rc = lib().heomq_solve_steady(self._h, self._check_vec(x, t.complex128), bp, tol, maxiter, trace,
                                      int(transpose), _dp(stats), self._stream())
        info = dict(iters=int(stats[0]), applies=int(stats[1]), relres=stats[2], seconds=stats[3], status=rc,
                    precond=int(stats[4]), fused_iteration=bool(stats[6]))
        if rc not in (_lib.OK, _lib.NOT_CONVERGED) or (rc == _lib.NOT_CONVERGED and raise_on_fail):
            check(rc)
        return x, info

    def solve_steady_mixed(self, x=None, b=None, tol=1e-12, inner_tol=1e-5, maxiter=20000, trace=TRACE_RDO,
                           raise_on_fail=True):
        """FP32 TFQMR inner solves inside FP64 iterative refinement."""
        t = self.torch
        x = self.initial_guess() if x is None else x
        stats = np.zeros(_lib.STATS_MIXED_LEN)
        bp = self._check_vec(b, t.complex128) if b is not None else None
        rc = lib().heomq_solve_steady_mixed(self._h, self._check_vec(x, t.complex128), bp, tol, inner_tol,
                                            maxiter, trace, _dp(stats), self._stream())
        outer = int(stats[4])
        info = dict(iters=int(stats[0]), applies=int(stats[1]), relres=stats[2], seconds=stats[3], outer=outer,
                    history=stats[8:8 + min(outer + 1, 40)].tolist(), inner=stats[48:48 + min(outer, 16)].tolist(),
                    fp64_iters=int(stats[6]), status=rc)
        if rc not in (_lib.OK, _lib.NOT_CONVERGED) or (rc == _lib.NOT_CONVERGED and raise_on_fail):
            raise HeomqError(rc, lib().heomq_last_error().decode() + f" {info}")
        return x, info

    def rk4(self, x, dt, nstep, t0=0.0):
        """RK4 (td/intstp.f90) from time t0; the lead field, when set, enters at each stage time."""
        check(lib().heomq_rk4_at(self._h, self._check_vec(x, self.torch.complex128), float(t0), dt, nstep,
                                 self._stream()))
        return x

    def set_system_field(self, terms=()):
        """Time-dependent system Hamiltonian hs + dhs(t) of the propagators
        (main/calcdhs.f90): terms = [(D, kind, params), ...] with D a Hermitian
        d x d operator (math-indexed D[i, j]), kind and up to four params of
        the time function (heomq.h lists them; amplitudes in eV, times in fs).
        No terms removes it."""
        terms = list(terms)
        n = len(terms)
        d = self.d
        ops = np.zeros((max(n, 1), d, d), np.complex128)
        kinds = np.zeros(max(n, 1), np.int32)
        pars = np.zeros((max(n, 1), 4))
        for k, (D, kind, par) in enumerate(terms):
            ops[k] = np.asarray(D, np.complex128).T   # memory column-major
            kinds[k] = int(kind)
            par = tuple(par)
            pars[k, : len(par)] = par
        check(lib().heomq_ctx_set_system_field(self._h, n, _dp(ops.view(np.float64)),
                                               kinds.ctypes.data_as(_lib.I32), _dp(pars)))

    def set_field(self, fieldtype, tchar=None, amps_dc=None, t_off=1e99):
        """Time-dependent lead shift eshift(alpha, s; t) (main/getengyshift.f90): fieldtype 0
        amps (1 - exp(-t/tchar)), 1 amps sin(2 pi t/tchar) + amps_dc, -1 none; tchar in fs and
        amps_dc in eV, shaped (nalf, nspin); the amplitudes are the model's engyshift."""
        tc = None if tchar is None else np.ascontiguousarray(tchar, np.float64).ravel()
        dc = None if amps_dc is None else np.ascontiguousarray(amps_dc, np.float64).ravel()
        check(lib().heomq_ctx_set_field(self._h, int(fieldtype), None if tc is None else _dp(tc),
                                        None if dc is None else _dp(dc), float(t_off)))
        self._field_type = int(fieldtype)

    def field_type(self) -> int:
        """The lead field's fieldtype (-1: none)."""
        return getattr(self, "_field_type", -1)

    def field_shifts(self, tt: float) -> np.ndarray:
        """eshift(alpha, s; tt) of the lead field in eV, (nalf, nspin); zeros without one."""
        nalf = self.model.nalf
        es = np.zeros(nalf * 2)
        check(lib().heomq_ctx_field_shifts(self._h, float(tt), _dp(es)))
        return es.reshape(nalf, 2)

    def calcderiv(self, x, tt=0.0, y=None):
        """dx/dt at time tt (calcderiv_omp(tt)): the apply with no trace row and gamma_K(tt)."""
        t = self.torch
        y = self.zeros() if y is None else y
        check(lib().heomq_calcderiv_at(self._h, float(tt), self._check_vec(x, t.complex128),
                                       self._check_vec(y, t.complex128), self._stream()))
        return y

    def chebyshev(self, x, dt, nstep):
        """Chebyshev steps (td/chebyshev.f90) on a dense state, in place."""
        stats = np.zeros(_lib.STATS_CHEBY_LEN)
        check(lib().heomq_chebyshev(self._h, self._check_vec(x, self.torch.complex128), dt, nstep, _dp(stats),
                                    self._stream()))
        return x, dict(ncheby=int(stats[0]), delta=stats[1], packed=bool(stats[2]))

    def rk4_packed(self, xp, dt, nstep, t0=0.0):
        """RK4 (td/intstp.f90) on a symmetry-packed state, in place."""
        check(lib().heomq_rk4_packed(self._h, self._check_packed(xp, self.torch.complex128), float(t0), dt, nstep,
                                     self._stream()))
        return xp

    def chebyshev_packed(self, xp, dt, nstep):
        """Chebyshev steps (td/chebyshev.f90) on a symmetry-packed state, in place."""
        stats = np.zeros(_lib.STATS_CHEBY_LEN)
        check(lib().heomq_chebyshev_packed(self._h, self._check_packed(xp, self.torch.complex128), dt, nstep,
                                           _dp(stats), self._stream()))
        return xp, dict(ncheby=int(stats[0]), delta=stats[1], packed=True)

    def observables(self, x):
        cur = np.zeros((self.model.nalf, 2))   # current(nalf, nspin) and occupation(norbs, nspin)
        occ = np.zeros((self.model.norbs, 2))
        check(lib().heomq_observables(self._h, self._check_vec(x, self.torch.complex128), _dp(cur), _dp(occ),
                                      self._stream()))
        return cur, occ

    # ---- reference-named host drop-ins (numpy, Fortran order) ----
    def calc_ax_tfqmr_omp(self, dim0: int, rcgtmp: np.ndarray, rcgrhs: np.ndarray):
        """rcgrhs = A rcgtmp with the trace row (calc_ax_tfqmr_omp.f90:1-510)."""
        check(lib().heomq_calc_ax_tfqmr(self._h, dim0, _dp(rcgtmp.view(np.float64)), _dp(rcgrhs.view(np.float64))))

    def calcderiv_omp(self, rhotmp: np.ndarray, rhorhs: np.ndarray):
        """rhorhs = L rhotmp, no trace row (td/calcderiv_omp.f90)."""
        check(lib().heomq_calcderiv(self._h, _dp(rhotmp.view(np.float64)), _dp(rhorhs.view(np.float64))))
