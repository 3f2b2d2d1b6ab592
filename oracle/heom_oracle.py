"""CPU oracle for the HEOM Liouvillian apply path -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  It is an independent numpy restatement of the reference algorithm
(HEOM-QUICK2, Fortran) for one hot path: the action of the fermionic HEOM
Liouvillian on the auxiliary-density-operator (ADO) hierarchy, its real
transpose, the steady-state solve built on it and the observables read from it.

Every function cites the reference file:line it restates.  How each function
is pinned (closed forms, brute force, invariants) is listed in
``tests/test_oracle.py``; the reference itself cannot be built in this image
(it needs Intel MKL), so no function is pinned against compiled reference
output.

Conventions
-----------
* Drawers (dissipation modes) are 0-based ids ``k = ((((sgn*nspin + s)*norbs
  + m)*nalf + a)*ncor + j)`` -- sign slowest, exponent fastest, the order of
  ``main/evaluatepara.f90:1701-1760``.  ``bar(k)`` flips the sign
  (``evaluatepara.f90:1766-1772``).
* An ADO of tier n is a sorted drawer tuple; tiers are stored tier-major, each
  tier in lexicographic order (``fermion/sortindextable3.f90``).
* Matrices in this module are indexed mathematically, ``X[a, i, j]`` for ADO a.
  The Fortran layout of the reference (``x(i, j, a)``, column-major) is the
  same memory as a C-ordered ``(N, d, d)`` array indexed ``[a, j, i]``;
  :func:`to_fortran` / :func:`from_fortran` convert.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

HBAR = 0.658211928          # include/sizes:23 (meV*ps with funits=1)
JCONST = 1.602177e5         # include/sizes:27


# --------------------------------------------------------------------------
# Pade spectrum decomposition of the Fermi function
# --------------------------------------------------------------------------
def pade_psd(n: int):
    """Pade [N-1/N] poles/residues of 1/(1+e^x).  Restates main/zpout_psd.f90:34-114.

    Returns (eta, xi): f(x) ~= 1/2 - sum_j 2*eta_j*x / (x^2 + xi_j^2).
    """
    dim0, dim1 = 2 * n, 2 * n - 1
    a0 = np.zeros((dim0, dim0))
    for i in range(dim0 - 1):           # 1-based (2i-1)(2j-1) off-diagonals
        v = 1.0 / math.sqrt((2.0 * (i + 1) - 1.0) * (2.0 * (i + 2) - 1.0))
        a0[i, i + 1] = a0[i + 1, i] = v
    a1 = np.zeros((dim1, dim1))
    for i in range(dim1 - 1):           # 1-based (2i+1)(2j+1)
        v = 1.0 / math.sqrt((2.0 * (i + 1) + 1.0) * (2.0 * (i + 2) + 1.0))
        a1[i, i + 1] = a1[i + 1, i] = v
    e0 = np.linalg.eigvalsh(a0)[:n]     # ascending; the n negative ones
    e1 = np.linalg.eigvalsh(a1)[:n - 1]
    x0 = 4.0 / e0 ** 2                  # eigXi  (zpout_psd.f90:79-81)
    x1 = 4.0 / e1 ** 2                  # eigZeta
    eta = np.ones(n)
    for j in range(n):                  # zpout_psd.f90:88-96
        for k in range(j):
            eta[j] *= (x1[k] - x0[j]) / (x0[k] - x0[j])
        for k in range(j + 1, n):
            eta[j] *= (x1[k - 1] - x0[j]) / (x0[k] - x0[j])
    eta = 0.5 * n * (2 * n + 1) * eta   # zpout_psd.f90:99
    xi = np.sqrt(x0)
    return eta, xi


def fermi_pade(z, eta, xi):
    """Pade Fermi function at complex argument.  Restates main/fermifunc.f90:18-23."""
    z = np.asarray(z, dtype=complex)
    out = 0.5 + 0j * z
    for e, x in zip(eta, xi):
        out = out - e / (z + 1j * x) - e / (z - 1j * x)
    return out


# --------------------------------------------------------------------------
# Model parameters and derived tables
# --------------------------------------------------------------------------
@dataclass
class Params:
    norbs: int = 1
    nspin: int = 2
    nalf: int = 2
    nmats: int = 8                      # Pade terms Nk (one Drude pole is added)
    bandwidth: tuple = (10.0, 10.0)     # W per lead
    linewidth: tuple = ((0.1,), (0.1,))  # Gamma[lead][orbital]
    temperature: tuple = (0.05, 0.05)   # k_B T per lead
    engyshift: tuple = ((0.0, 0.0), (0.0, 0.0))  # bias per [lead][spin]
    ham: str = "siam"                   # "siam" ($para1) or "dqd" ($para4)
    hpar: dict = field(default_factory=dict)
    ntop: int = 3                       # top tier L (reference ntier = L+1)
    lsdraw: bool = False                # allow repeated drawers (reference default True)
    steady: bool = True                 # igroundsteady=1: bias enters gamma
    lscale: bool = False                # ADOs stored scaled by prod_k |cb_k|^(-1/2) ($jobinfo lscale)


def siam_params(eps=-1.0, U=2.0, gamma=0.1, W=10.0, T=0.05, V=0.0, nk=8, ntop=3,
                gl=None, gr=None, lsdraw=False):
    """Single-impurity Anderson model, bias split +V/2 (left) / -V/2 (right)."""
    gl = gamma if gl is None else gl
    gr = gamma if gr is None else gr
    return Params(norbs=1, nspin=2, nalf=2, nmats=nk, bandwidth=(W, W),
                  linewidth=((gl,), (gr,)), temperature=(T, T),
                  engyshift=((V / 2, V / 2), (-V / 2, -V / 2)), ham="siam",
                  hpar=dict(eup=eps, edown=eps, uu=U), ntop=ntop, lsdraw=lsdraw)


def dqd_params(eps=-1.5, U=3.0, t=0.5, u12=0.0, gamma=0.2, W=10.0, T=0.1, V=0.4,
               nk=6, ntop=3, lsdraw=False):
    """Two-impurity (serial double dot, both dots coupled to both leads)."""
    return Params(norbs=2, nspin=2, nalf=2, nmats=nk, bandwidth=(W, W),
                  linewidth=((gamma, gamma), (gamma, gamma)), temperature=(T, T),
                  engyshift=((V / 2, V / 2), (-V / 2, -V / 2)), ham="dqd",
                  hpar=dict(e1up=eps, e1down=eps, e2up=eps, e2down=eps, uu1=U, uu2=U,
                            u12=u12, t12=t), ntop=ntop, lsdraw=lsdraw)


# ---- fermion operators ----------------------------------------------------
def _decode(idx1, nrho, norbs):
    """Occupations (up, down) per orbital of 1-based Fock index (swapspin.f90:24-55)."""
    occ = []
    table = {0: [0, 0], 1: [1, 0], 2: [0, 1], 3: [1, 1]}
    for ni in range(1, norbs + 1):
        len0 = nrho // 4 ** (norbs - ni + 1)
        occ.append(list(table[(idx1 - 1) // len0 % 4]))
    return occ


def _encode(occ, nrho, norbs):
    nj = 1
    for ni in range(norbs, 0, -1):
        len0 = nrho // 4 ** (norbs - ni + 1)
        u, dn = occ[ni - 1]
        nj += len0 * {(0, 0): 0, (1, 0): 1, (0, 1): 2, (1, 1): 3}[(u, dn)]
    return nj


def _swapspin(nrho, norbs, ir, ic, iorbs):
    """main/swapspin.f90:1-143 (1-based indices in/out)."""
    out, sg = [], 1
    for idx in (ir, ic):
        occ = _decode(idx, nrho, norbs)
        if occ[iorbs - 1] == [1, 1]:
            sg = -sg
        occ[iorbs - 1] = occ[iorbs - 1][::-1]
        out.append(_encode(occ, nrho, norbs))
    return out[0], out[1], sg


def _swapms(nrho, norbs, ir, ic, iorbs, ispin):
    """main/swapms.f90:144-229, nspin=2 branch (swap orbitals iorbs-1, iorbs of one spin)."""
    out, sg = [], 1
    s = ispin - 1
    for idx in (ir, ic):
        occ = _decode(idx, nrho, norbs)
        if occ[iorbs - 2][s] == 1 and occ[iorbs - 1][s] == 1:
            sg = -sg
        occ[iorbs - 2][s], occ[iorbs - 1][s] = occ[iorbs - 1][s], occ[iorbs - 2][s]
        out.append(_encode(occ, nrho, norbs))
    return out[0], out[1], sg


def annihilators(norbs, nspin=2):
    """c_{m,s} in the occupation basis, ams[m][s] (d x d real).

    Restates main/calcams.f90:79-103 (c_{1,up}) and main/buildoperator.f90:168-226
    (the others by spin/orbital swaps with fermionic signs); amsall = amsori
    (main/checkanticomm.f90:69).
    """
    assert nspin == 2
    nrho = 4 ** norbs
    c1 = np.zeros((nrho, nrho))
    len0 = nrho // 4 ** norbs
    for ni in range(1, 4 ** (norbs - 1) + 1):
        irow0 = (ni - 1) * len0 * 4 + 1
        irow1 = irow0 + len0 * 2
        icol0 = irow0 + len0
        icol1 = irow0 + len0 * 3
        for nj in range(len0):
            c1[irow0 + nj - 1, icol0 + nj - 1] = 1.0
            c1[irow1 + nj - 1, icol1 + nj - 1] = 1.0
    ams = [[None, None] for _ in range(norbs)]
    ams[0][0] = c1
    ent = {1: [(r + 1, c + 1, 1) for r, c in zip(*np.nonzero(c1))]}
    # c_{1,down} (buildoperator.f90:190-199)
    ent2 = []
    for (r, c, s) in ent[1]:
        r2, c2, sg = _swapspin(nrho, norbs, r, c, 1)
        ent2.append((r2, c2, sg * s))
    cur = {1: ent[1], 2: ent2}
    ams[0][1] = _mat(ent2, nrho)
    ispin = 2
    for ni in range(2, norbs + 1):           # buildoperator.f90:202-226
        new = []
        for (r, c, s) in cur[ispin]:
            r2, c2, sg = _swapms(nrho, norbs, r, c, ni, ispin)
            new.append((r2, c2, sg * s))
        cur[ispin] = new
        ams[ni - 1][ispin - 1] = _mat(new, nrho)
        ispin = 3 - ispin
        new = []
        for (r, c, s) in cur[3 - ispin]:
            r2, c2, sg = _swapspin(nrho, norbs, r, c, ni)
            new.append((r2, c2, sg * s))
        cur[ispin] = new
        ams[ni - 1][ispin - 1] = _mat(new, nrho)
    return ams


def _mat(ent, nrho):
    m = np.zeros((nrho, nrho))
    for r, c, s in ent:
        m[r - 1, c - 1] = s
    return m


def hamiltonian(p: Params, ams):
    """System Hamiltonian divided by hbar.

    SIAM: main/calchs.f90:368-387 (para1); two-impurity: calchs.f90:514-567 (para4).
    """
    d = ams[0][0].shape[0]
    h = np.zeros((d, d), dtype=complex)
    n = lambda m, s: ams[m][s].T @ ams[m][s]
    hp = p.hpar
    if p.ham == "siam":
        h += hp.get("eup", 0.0) * n(0, 0) + hp.get("edown", 0.0) * n(0, 1)
        h += hp.get("uu", 0.0) * n(0, 0) @ n(0, 1)
        sf = ams[0][0].T @ ams[0][1]
        h += hp.get("tupdn", 0.0) * (sf + sf.T)
    elif p.ham == "dqd":
        h += hp["e1up"] * n(0, 0) + hp["e1down"] * n(0, 1)
        h += hp["uu1"] * n(0, 0) @ n(0, 1)
        h += hp["e2up"] * n(1, 0) + hp["e2down"] * n(1, 1)
        h += hp["uu2"] * n(1, 0) @ n(1, 1)
        n1 = n(0, 0) + n(0, 1)
        n2 = n(1, 0) + n(1, 1)
        h += hp.get("u12", 0.0) * n1 @ n2
        hop = (ams[0][0].T @ ams[1][0] + ams[1][0].T @ ams[0][0]
               + ams[0][1].T @ ams[1][1] + ams[1][1].T @ ams[0][1])
        h += hp["t12"] * hop
    else:
        raise ValueError(p.ham)
    return h / HBAR


class Model:
    """All tables the apply reads (the reference's module matmod, restated)."""

    def __init__(self, p: Params):
        self.p = p
        self.norbs, self.nspin, self.nalf = p.norbs, p.nspin, p.nalf
        self.ncor = 1 + p.nmats                      # one Drude pole (evaluatepara.f90:1166-1190)
        self.nopr = 2 * p.norbs * p.nspin * p.nalf   # evaluatepara.f90:1205
        self.M = self.nopr * self.ncor               # evaluatepara.f90:1231
        self.d = (2 * p.nspin) ** p.norbs            # evaluatepara.f90:1232
        self.L = p.ntop
        self._drawers()
        self._bath()
        self.ams = annihilators(p.norbs, p.nspin)
        self.H = hamiltonian(p, self.ams)
        self._hierarchy()
        self._coupling()

    # --- drawer attributes (evaluatepara.f90:1701-1772) ---
    def _drawers(self):
        p, M = self.p, None
        rows = []
        for sg in range(2):
            for s in range(p.nspin):
                for m in range(p.norbs):
                    for a in range(p.nalf):
                        for j in range(self.ncor):
                            rows.append((sg, s, m, a, j))
        self.attr = np.array(rows, dtype=np.int64)
        self.sgn, self.spin, self.orb, self.alf, self.cor = self.attr.T
        half = len(rows) // 2
        self.bar = np.array([k + half if k < half else k - half for k in range(len(rows))])
        self.dipm = np.where(self.sgn == 0, 1.0, -1.0)
        self.opr = np.arange(len(rows)) // self.ncor      # operator id (look4drawer.f90:24)

    # --- bath coefficients (evaluatepara.f90:1897-1982, 2210-2227, 2297, 2431) ---
    def _bath(self):
        p = self.p
        eta, xi = pade_psd(p.nmats)
        self.pade = (eta, xi)
        cb = np.zeros(self.M, complex)
        cd = np.zeros(self.M, complex)
        cg = np.zeros(self.M, complex)
        for k in range(self.M):
            sg, s, m, a, j = self.attr[k]
            W = p.bandwidth[a]
            T = p.temperature[a]
            zxip = p.linewidth[a][m] * W ** 2 * 0.5                 # :1511
            dpm = 1.0 if sg == 0 else -1.0
            if j == 0:                                              # Drude pole :1920-1953
                cg[k] = complex(-W, dpm * 0.0) / HBAR
                f1 = fermi_pade((1.0 / T) * complex(dpm * 0.0, W), eta, xi)
                cb[k] = zxip / W * f1 / HBAR ** 2
                f2 = fermi_pade((1.0 / T) * complex(-dpm * 0.0, W), eta, xi)
                cd[k] = np.conj(zxip / W * f2) / HBAR ** 2
            else:                                                   # Pade poles :1964-1982
                jm = j - 1
                pole = 1j * xi[jm] if sg == 0 else -1j * xi[jm]     # zpout_psd.f90:110-114
                pole_b = -pole
                coef = -eta[jm]
                cg[k] = dpm * 1j * T * pole / HBAR
                cb[k] = 1j * 2.0 * coef * zxip * T / ((pole * T - 0.0) ** 2 + W ** 2) / HBAR ** 2
                cd[k] = np.conj(1j * 2.0 * coef * zxip * T) / np.conj((pole_b * T - 0.0) ** 2 + W ** 2) / HBAR ** 2
        # lscale (evaluatepara.f90:2927-2990): every ADO is stored as
        # rho_K / prod_{k in K} |cb_k|^(1/2); dbsqrt_k = |cb_k|^(1/2) of the
        # unscaled cb, cb and cd divided by it (1/dbsqrt = 0 where |cb| <=
        # dpico = 1e-12), and every upper coupling multiplied by it
        # (calc_ax_tfqmr_omp.f90:120-122,325-328,388-390)
        self.dbsqrt = np.ones(self.M)
        if p.lscale:
            self.dbsqrt = np.sqrt(np.abs(cb))
            inv = np.where(np.abs(cb) > 1e-12, 1.0 / np.where(self.dbsqrt > 0, self.dbsqrt, 1.0), 0.0)
            cb = cb * inv
            cd = cd * inv
        self.cb, self.cd, self.cgamma = cb, cd, cg
        self.eleadinfty = np.array(p.engyshift, dtype=float) / HBAR   # fieldtype 0, t -> inf
        gk = cg.copy()
        if p.steady:                                              # calc_ax_tfqmr_omp.f90:190-193
            gk = gk + 1j * self.dipm * self.eleadinfty[self.alf, self.spin]
        self.gdraw = gk

    # --- hierarchy enumeration (ball2drawer.f90, findoprindex.f90, sortindextable3.f90) ---
    def _recorded_opset(self, ops):
        half = self.nopr // 2
        comp = tuple(sorted(o + half if o < half else o - half for o in ops))
        return tuple(ops) <= comp

    def _hierarchy(self):
        M, L = self.M, self.L
        tiers = [np.zeros((1, 0), dtype=np.int64)]
        gen = itertools.combinations_with_replacement if self.p.lsdraw else itertools.combinations
        for n in range(1, L + 1):
            keep = []
            for K in gen(range(M), n):
                ops = tuple(k // self.ncor for k in K)
                if self._recorded_opset(ops):
                    keep.append(K)
            tiers.append(np.array(keep, dtype=np.int64).reshape(-1, n))
        self.tiers = tiers
        self.first = np.cumsum([0] + [len(t) for t in tiers])
        self.N = int(self.first[-1])
        self.index = [{tuple(K): i for i, K in enumerate(t)} for t in tiers]

    def tuple_of(self, a):
        n = int(np.searchsorted(self.first, a, side="right") - 1)
        return n, tuple(self.tiers[n][a - self.first[n]])

    @staticmethod
    def _parity(seq):
        inv = 0
        for i in range(len(seq)):
            for j in range(i + 1, len(seq)):
                if seq[i] > seq[j]:
                    inv += 1
        return -1 if inv % 2 else 1

    def lookup(self, K):
        """Stored ADO for drawer tuple K: (global idx, itype, fact) or None.

        Restates main/findindex.f90:51-76 (direct record, then the reversed,
        sign-flipped complement sorted with swap parity `fact`).
        """
        n = len(K)
        Ks = tuple(sorted(K))
        i = self.index[n].get(Ks)
        if i is not None:
            return int(self.first[n] + i), 0, 1
        comp = [int(self.bar[k]) for k in reversed(Ks)]
        fact = self._parity(comp)
        i = self.index[n].get(tuple(sorted(comp)))
        if i is None:
            return None
        return int(self.first[n] + i), 1, fact

    # --- coupling table (buildcoefindex.f90:203-362, findlowertier/findhighertier) ---
    def _coupling(self):
        M, L = self.M, self.L
        lower, upper = [], []
        for n in range(0, L + 1):
            T = self.tiers[n]
            lo = np.zeros((len(T), n, 3), dtype=np.int64)      # (nbr, itype, ff, fr) packed below
            lof = np.zeros((len(T), n, 2), dtype=np.int64)
            for a, K in enumerate(T):
                K = tuple(K)
                for r in range(n):
                    rest = K[:r] + K[r + 1:]
                    nb, it, fact = self.lookup(rest) if n > 1 else (0, 0, 1)
                    ff = 1 if r % 2 == 0 else -1                    # buildcoefindex.f90:267-271
                    fr = 1 if (n - 1 - r) % 2 == 0 else -1          # :272-276
                    if it == 1:
                        ff *= fact
                        fr *= fact
                    lo[a, r] = (nb, it, 0)
                    lof[a, r] = (ff, fr)
            lower.append((lo[..., :2].copy(), lof))
            if n < L:
                up = np.full((len(T), M, 2), -1, dtype=np.int64)
                upf = np.zeros((len(T), M, 2), dtype=np.int64)
                for a, K in enumerate(T):
                    K = tuple(K)
                    for e in range(M):
                        if not self.p.lsdraw and e in K:
                            continue                                 # Pauli: no such ADO
                        fr_list = list(K) + [e]
                        ff_list = [e] + list(K)
                        ffact = self._parity(ff_list)                # findhighertier.f90:52-65
                        rfact = self._parity(fr_list)                # :69-82
                        res = self.lookup(tuple(K) + (e,))
                        if res is None:
                            raise RuntimeError("higher-tier ADO not found")
                        nb, it, fact = res
                        up[a, e] = (nb, it)
                        upf[a, e] = (ffact * fact, rfact * fact)
                upper.append((up, upf))
        self.lower, self.upper = lower, upper

    def sizes(self):
        return [len(t) for t in self.tiers]


# --------------------------------------------------------------------------
# The apply and its transpose
# --------------------------------------------------------------------------
def _cop(model, k, left):
    """Operator multiplying a neighbour for drawer k (lower-tier term)."""
    c = model.ams[model.orb[k]][model.spin[k]]
    return c.T if model.sgn[k] == 0 else c


def apply(model: Model, x, trace_row="last", H=None, gdraw=None, cb=None, cd=None, trun_der=False):
    """y = L x, the dense TFQMR-convention apply.

    Restates tfqmr_double/calc_ax_tfqmr_omp.f90: same tier :198-203, lower
    tier :221-271, upper tier :275-357, tier 0 :99-149 (written through the
    generic upper-tier rule with complement lookups, which is the same sum),
    trace row :492-495.  ``trace_row`` is "last" (reference: element (d,d)
    of the last ADO), "rdo" (element (d,d) of the tier-0 block, a row that is
    linearly dependent in L for every truncation tier), or None
    (calcderiv_omp: no normalisation row).  Optional H/gdraw/cb/cd override
    the model's coefficients (used for parameter derivatives).  ``trun_der``
    adds the derivative terminator of the top tier (:362-478,
    :func:`trun_der_term`).
    """
    H = model.H if H is None else H
    gdraw = model.gdraw if gdraw is None else gdraw
    cb = model.cb if cb is None else cb
    cd = model.cd if cd is None else cd
    d = model.d
    y = np.zeros_like(x)
    for n in range(model.L + 1):
        s0, s1 = model.first[n], model.first[n + 1]
        X = x[s0:s1]
        K = model.tiers[n]
        gam = gdraw[K].sum(axis=1) if n > 0 else np.zeros(len(K), complex)
        yt = -1j * (H @ X - X @ H) + gam[:, None, None] * X
        lo, lof = model.lower[n]
        for r in range(n):
            nb, it = lo[:, r, 0], lo[:, r, 1]
            ff, fr = lof[:, r, 0], lof[:, r, 1]
            Xn = x[nb].copy()
            Xn[it == 1] = np.conj(np.transpose(Xn[it == 1], (0, 2, 1)))
            kr = K[:, r]
            C = np.stack([model.ams[model.orb[k]][model.spin[k]] for k in kr]) if len(kr) else np.zeros((0, d, d))
            op = np.where((model.sgn[kr] == 0)[:, None, None], np.transpose(C, (0, 2, 1)), C)
            a = (-1j * ff * cb[kr])[:, None, None]
            b = (1j * fr * cd[kr])[:, None, None]
            yt += a * (op @ Xn) + b * (Xn @ op)
        if n < model.L:
            up, upf = model.upper[n]
            for e in range(model.M):
                nb, it = up[:, e, 0], up[:, e, 1]
                ok = nb >= 0
                if not ok.any():
                    continue
                Xn = np.zeros((len(K), d, d), complex)
                Xn[ok] = x[nb[ok]]
                m = ok & (it == 1)
                Xn[m] = np.conj(np.transpose(Xn[m], (0, 2, 1)))
                c = model.ams[model.orb[e]][model.spin[e]]
                op = c if model.sgn[e] == 0 else c.T                  # :278-279
                ff = upf[:, e, 0][:, None, None] * model.dbsqrt[e]   # lscale, :325-328
                fr = upf[:, e, 1][:, None, None] * model.dbsqrt[e]
                yt += -1j * ff * (op @ Xn) + 1j * fr * (Xn @ op)
        y[s0:s1] = yt
    if trun_der:
        y += trun_der_term(model, x, H=H, gdraw=gdraw, cb=cb, cd=cd)
    if trace_row == "last":
        y[-1, d - 1, d - 1] = np.trace(x[0])
    elif trace_row == "rdo":
        y[0, d - 1, d - 1] = np.trace(x[0])
    elif trace_row is not None:
        raise ValueError(trace_row)
    return y


def liouvillian_eigensystem(H):
    """fermion/prelude_trun.f90:41-78: the eigenvalues and eigenvectors of the
    commutator superoperator X -> H X - X H (zhil2liou 'l' minus 'r', column-
    major vec(X)), as zheev gives them for the Hermitian d^2 x d^2 matrix."""
    d = H.shape[0]
    lsys = np.kron(np.eye(d), H) - np.kron(H.T, np.eye(d))
    return np.linalg.eigh(lsys)


def trun_der_term(model: Model, x, H=None, gdraw=None, cb=None, cd=None, eig=None):
    """The derivative terminator (ltrun_der) of calc_ax_tfqmr_omp.f90:362-478:
    every top-tier ADO K receives, from each drawer j, the upper coupling of
    the tier-(L+1) ADO V = K + j that is not stored but set to the steady
    solution of its own equation without its upper couplings,
    rho_V = (i L_sys - gamma_V)^-1 B_V, B_V being V's lower-coupling terms
    from the stored tier-L ADOs V - m (:397-423; findlowertier, add2list) and
    the resolvent taken in the eigenbasis of L_sys = [H, .] (:426-447:
    -B/gamma_V plus the nonzero-eigenvalue correction; -i/lambda only when
    |gamma_V| <= dpico).  The coupling back to K is the upper-tier rule with
    the signs of j's place in V (:376-385, :462-467; dbsqrt with lscale).
    With lsdraw the reference loops over every drawer (a repeated one gives a
    V whose lower terms cancel); without it, a drawer already in K is skipped
    (no such V exists, and the reference's lookup of it would stop)."""
    H = model.H if H is None else H
    gdraw = model.gdraw if gdraw is None else gdraw
    cb = model.cb if cb is None else cb
    cd = model.cd if cd is None else cd
    d, L, M = model.d, model.L, model.M
    lam, U = liouvillian_eigensystem(H) if eig is None else eig
    nz = np.abs(lam) > 1e-12
    y = np.zeros_like(x)
    if L < 1:
        raise ValueError("the derivative terminator needs a top tier L >= 1")

    def vec(A):
        return A.reshape(-1, order="F")

    for a, K in enumerate(model.tiers[L]):
        g = int(model.first[L] + a)
        K = [int(k) for k in K]
        gam_k = gdraw[K].sum()
        for e in range(M):
            if not model.p.lsdraw and e in K:
                continue
            mj = next((i for i, k in enumerate(K) if k > e), len(K)) + 1    # add2list: after equal drawers
            V = K[:mj - 1] + [e] + K[mj - 1:]
            B = np.zeros((d, d), complex)
            for mi in range(1, L + 2):
                nn = V[mi - 1]
                res = model.lookup(tuple(V[:mi - 1] + V[mi:]))
                if res is None:
                    raise RuntimeError("findlowertier: lower neighbour of the virtual ADO not stored")
                lnj, it, fact = res
                ff = (-1) ** (mi - 1)
                fr = (-1) ** (L + 1 - mi)
                Xn = x[lnj]
                if it == 1:
                    ff, fr = ff * fact, fr * fact
                    Xn = Xn.conj().T
                c = model.ams[model.orb[nn]][model.spin[nn]]
                op = c.T if model.sgn[nn] == 0 else c
                B += cb[nn] * (-1j * ff) * (op @ Xn) + cd[nn] * (1j * fr) * (Xn @ op)
            gv = gam_k + gdraw[e]
            w = U.conj().T @ vec(B)
            if abs(gv) > 1e-12:
                lam0 = np.where(nz, lam, 0.0)
                w = w / (1j * lam0 - gv)
            else:
                w = np.where(nz, -1j / np.where(nz, lam, 1.0), 0.0) * w
            rv = (U @ w).reshape(d, d, order="F")
            c = model.ams[model.orb[e]][model.spin[e]]
            op = c if model.sgn[e] == 0 else c.T
            sc = model.dbsqrt[e]
            f1 = (-1) ** (mj - 1) * sc
            f2 = (-1) ** (L + 1 - mj) * sc
            y[g] += -1j * f1 * (op @ rv) + 1j * f2 * (rv @ op)
    return y


def apply_transpose(model: Model, w, trace_row="last"):
    """z = L^T w, the real transpose of :func:`apply` (Re<w, Lx> = Re<L^T w, x>).

    Scatter form of the adjoint of each term of calc_ax_tfqmr_omp.f90: a term
    y_K += a*A*X*B contributes conj(a)*A^H*w_K*B^H to x_J (X = x_J) or
    a*B*w_K^H*A (X = x_J^H).  The reference's own transpose
    (bicg/calc_atrx_bicg_omp.f90) is the same operator in BiCG's unknown
    convention.
    """
    H, d = model.H, model.d
    w = w.copy()
    tr = 0j
    if trace_row in ("last", "rdo"):
        pos = -1 if trace_row == "last" else 0
        tr = w[pos, d - 1, d - 1]
        w[pos, d - 1, d - 1] = 0.0
    z = np.zeros_like(w)
    for n in range(model.L + 1):
        s0, s1 = model.first[n], model.first[n + 1]
        W = w[s0:s1]
        K = model.tiers[n]
        gam = model.gdraw[K].sum(axis=1) if n > 0 else np.zeros(len(K), complex)
        z[s0:s1] += 1j * (H @ W - W @ H) + np.conj(gam)[:, None, None] * W
        lo, lof = model.lower[n]
        for r in range(n):
            nb, it = lo[:, r, 0], lo[:, r, 1]
            ff, fr = lof[:, r, 0], lof[:, r, 1]
            kr = K[:, r]
            C = np.stack([model.ams[model.orb[k]][model.spin[k]] for k in kr]) if len(kr) else np.zeros((0, d, d))
            op = np.where((model.sgn[kr] == 0)[:, None, None], np.transpose(C, (0, 2, 1)), C)
            opH = np.transpose(op, (0, 2, 1))
            a = (-1j * ff * model.cb[kr])[:, None, None]
            b = (1j * fr * model.cd[kr])[:, None, None]
            direct = np.conj(a) * (opH @ W) + np.conj(b) * (W @ opH)
            Wh = np.conj(np.transpose(W, (0, 2, 1)))
            dag = a * (Wh @ op) + b * (op @ Wh)
            contrib = np.where((it == 0)[:, None, None], direct, dag)
            np.add.at(z, nb, contrib)
        if n < model.L:
            up, upf = model.upper[n]
            Wh = np.conj(np.transpose(W, (0, 2, 1)))
            for e in range(model.M):
                nb, it = up[:, e, 0], up[:, e, 1]
                ok = nb >= 0
                if not ok.any():
                    continue
                c = model.ams[model.orb[e]][model.spin[e]]
                op = c if model.sgn[e] == 0 else c.T
                ff = upf[:, e, 0][:, None, None] * model.dbsqrt[e]
                fr = upf[:, e, 1][:, None, None] * model.dbsqrt[e]
                a, b = -1j * ff, 1j * fr
                direct = np.conj(a) * (op.T @ W) + np.conj(b) * (W @ op.T)
                dag = a * (Wh @ op) + b * (op @ Wh)
                contrib = np.where((it == 0)[:, None, None], direct, dag)
                np.add.at(z, nb[ok], contrib[ok])
    if trace_row in ("last", "rdo"):
        z[0] += tr * np.eye(d)
    return z


# --------------------------------------------------------------------------
# Layout helpers, observables
# --------------------------------------------------------------------------
def to_fortran(x):
    """(N, d, d) math-indexed -> complex128 array in the reference's x(i,j,a) memory order."""
    return np.ascontiguousarray(np.transpose(x, (0, 2, 1)))


def from_fortran(xf):
    return np.ascontiguousarray(np.transpose(xf, (0, 2, 1)))


def seeded_ado_vector(model: Model, seed: int):
    """Complex-normal ADO vector; tier-0 block Hermitianised and trace-normalised."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((model.N, model.d, model.d)) + 1j * rng.standard_normal((model.N, model.d, model.d))
    r0 = x[0]
    while True:
        r0 = 0.5 * (r0 + r0.conj().T)
        if abs(np.trace(r0)) >= 0.1:
            break
        r0 = rng.standard_normal((model.d, model.d)) + 1j * rng.standard_normal((model.d, model.d))
    x[0] = r0 / np.trace(r0)
    return x


def current(model: Model, x):
    """Lead currents I[alpha, spin] = -2 Im sum_{m,j} tr(c_ms rho_(+,m,s,alpha,j)) * jconst.

    Restates td/calcj.f90:15-37,68 and utility/zams_trace_ab.f90.
    """
    I = np.zeros((model.nalf, model.nspin))
    for k in range(model.M):
        if model.sgn[k] != 0:
            continue
        nb, it, fact = model.lookup((k,))
        X = x[nb] if it == 0 else fact * x[nb].conj().T
        c = model.ams[model.orb[k]][model.spin[k]]
        # lscale: the stored tier-1 ADO is rho / |cb|^(1/2) of its own drawer
        ks = k if it == 0 else int(model.bar[k])
        I[model.alf[k], model.spin[k]] += -2.0 * np.imag(np.trace(c @ X)) * model.dbsqrt[ks]
    return I * JCONST


def occupation(model: Model, x):
    """<n_{m,s}> = tr(c^+ c rho0) (main/calcocc.f90:20-36)."""
    occ = np.zeros((model.norbs, model.nspin))
    for m in range(model.norbs):
        for s in range(model.nspin):
            c = model.ams[m][s]
            occ[m, s] = np.real(np.trace(c.T @ c @ x[0]))
    return occ


# --------------------------------------------------------------------------
# Solvers
# --------------------------------------------------------------------------
def as_real(x):
    return np.ascontiguousarray(to_fortran(x)).view(np.float64).ravel()


def from_real(v, model):
    return from_fortran(v.view(np.complex128).reshape(model.N, model.d, model.d))


def tfqmr(matvec, b, tol=1e-12, maxiter=20000, x0=None, max_restarts=3, project=None,
          restart_every=None):
    """Transpose-free QMR (Freund, SIAM J. Sci. Comput. 14 (1993) 470), real vectors.

    Textbook form; the reference drives the same method through its reverse-
    communication kernel (tfqmr_double/dutfx.f90) with a random shadow vector
    (dutfx.f90:178).  Stops when the true residual ||b - A x|| <= tol*||b||.
    When the quasi-residual bound is met but the true residual is not (the
    attainable-accuracy floor), the iteration restarts from the true residual
    at most ``max_restarts`` times.  ``project`` (optional) maps the iterate
    onto an invariant subspace that holds the solution; it is applied at every
    restart, and ``restart_every`` forces a restart after that many
    iterations.  Returns (x, iters, relres).
    """
    x = np.zeros_like(b) if x0 is None else x0.copy()
    nb = np.linalg.norm(b)
    rt = np.random.default_rng(1).standard_normal(b.shape)
    total = 0
    best = np.inf
    nrest = 0
    while True:
        if project is not None:
            x = project(x)
        r = b - matvec(x)
        rr = np.linalg.norm(r)
        if rr <= tol * nb or total >= maxiter:
            return x, total, rr / nb
        if rr >= best:
            nrest += 1
            if nrest > max_restarts:
                return x, total, rr / nb
        best = min(best, rr)
        start = total
        w = r.copy()
        y1 = r.copy()
        v = matvec(y1)
        u1 = v.copy()
        d = np.zeros_like(b)
        tau = rr
        theta = eta = 0.0
        rho = rt @ r
        while total < maxiter:
            total += 1
            sigma = rt @ v
            alpha = rho / sigma
            y2 = y1 - alpha * v
            u2 = matvec(y2)
            for (yj, uj) in ((y1, u1), (y2, u2)):
                w = w - alpha * uj
                d = yj + (theta ** 2 * eta / alpha) * d
                theta = np.linalg.norm(w) / tau
                c = 1.0 / math.sqrt(1.0 + theta ** 2)
                tau = tau * theta * c
                eta = c ** 2 * alpha
                x = x + eta * d
            if tau * math.sqrt(2 * (total - start) + 1) <= tol * nb:
                break
            if restart_every and total - start >= restart_every:
                break
            rho_n = rt @ w
            beta = rho_n / rho
            rho = rho_n
            y1 = w + beta * y2
            u1 = matvec(y1)
            v = u1 + beta * (u2 + beta * v)



def steady_state(model: Model, tol=1e-12, maxiter=20000, trace_row="rdo", trun_der=False):
    """Steady state: solve A x = e_trace with A the trace-row apply.

    solve_tfqmr.f90:97-179 solves A*dx = e_trace - A*x0 from a physical
    trace-one guess x0 and sets rho = x0 + dx; the same is done here with
    x0 = identity/d on tier 0.  Starting from a trace-one guess keeps the
    Krylov space inside the Hermitian, parity-allowed, traceless sector.
    ``trun_der`` solves with the derivative terminator of the top tier.
    """
    d = model.d
    nvec = model.N * d * d * 2
    if trun_der:
        eig = liouvillian_eigensystem(model.H)
        base = apply
        apply_ = lambda m, v, trace_row: base(m, v, trace_row=trace_row) + trun_der_term(m, v, eig=eig)
    else:
        apply_ = lambda m, v, trace_row: apply(m, v, trace_row=trace_row)
    x0 = np.zeros((model.N, d, d), complex)
    x0[0] = np.eye(d) / d
    e = np.zeros((model.N, d, d), complex)
    e[-1 if trace_row == "last" else 0, d - 1, d - 1] = 1.0
    b = as_real(e - apply_(model, x0, trace_row=trace_row))
    proj = lambda v: as_real(physical_projection(model, from_real(v, model)))
    if trace_row == "rdo":
        # iterate with Pi*A: A maps the physical sector into itself and the
        # steady state is its unique solution there
        mv = lambda v: proj(as_real(apply_(model, from_real(v, model), trace_row=trace_row)))
        xr, it, rel = tfqmr(mv, b, tol=tol, maxiter=maxiter, project=proj)
    else:
        mv = lambda v: as_real(apply_(model, from_real(v, model), trace_row=trace_row))
        xr, it, rel = tfqmr(mv, b, tol=tol, maxiter=maxiter, project=proj, restart_every=100)
    return x0 + from_real(xr, model), it, rel


def fock_quantum_numbers(model: Model):
    """Electron number N and 2*S_z of each occupation-basis state (calcams.f90 ordering)."""
    N, S2 = [], []
    for i in range(model.d):
        n = s2 = 0
        v = i
        for _ in range(model.norbs):
            dig = v % 4
            up, dn = dig in (1, 3), dig in (2, 3)
            n += up + dn
            s2 += up - dn
            v //= 4
        N.append(n)
        S2.append(s2)
    return np.array(N), np.array(S2)


def fock_parity(model: Model):
    return fock_quantum_numbers(model)[0] % 2


def ado_sectors(model: Model):
    """(dN, d2Sz) carried by every ADO: sum over its drawers of q = +1 (sigma=+) / -1,
    times +-1 for the spin in the second component."""
    q = np.where(model.sgn == 0, 1, -1)
    qs = q * np.where(model.spin == 0, 1, -1)
    dN = np.zeros(model.N, np.int64)
    dS = np.zeros(model.N, np.int64)
    for n in range(1, model.L + 1):
        K = model.tiers[n]
        dN[model.first[n]:model.first[n + 1]] = q[K].sum(axis=1)
        dS[model.first[n]:model.first[n + 1]] = qs[K].sum(axis=1)
    return dN, dS


def conserved(model: Model):
    """Which of N, S_z the Hamiltonian conserves (H never mixes states that differ)."""
    N, S2 = fock_quantum_numbers(model)
    nz = np.abs(model.H) > 0
    return bool(np.all(N[:, None] == N[None, :]) or np.all((N[:, None] == N[None, :]) | ~nz)), \
        bool(np.all((S2[:, None] == S2[None, :]) | ~nz))


def physical_projection(model: Model, x, hermitian=True):
    """Project onto the sector that holds every physical ADO vector.

    Each c (c^+) lowers (raises) N by one and 2S_z by +-1, H conserves N (and
    S_z when it has no spin-flip term), so element (i, j) of ADO K can be
    nonzero only if N_i - N_j = dN_K and 2(S_z,i - S_z,j) = d2Sz_K (the sums of
    q = +-1 over K's drawers); the tier-0 block is Hermitian.  The Liouvillian
    maps this sector into itself.
    """
    N, S2 = fock_quantum_numbers(model)
    cN, cS = conserved(model)
    dN, dS = ado_sectors(model)
    y = x.copy()
    if hermitian:
        y[0] = 0.5 * (y[0] + y[0].conj().T)
    okN = (N[:, None] - N[None, :])[None] == dN[:, None, None]
    if not cN:
        okN = ((N[:, None] - N[None, :])[None] - dN[:, None, None]) % 2 == 0
    ok = okN
    if cS:
        ok = ok & ((S2[:, None] - S2[None, :])[None] == dS[:, None, None])
    return y * ok


def dense_matrix(model: Model, trace_row="last", transpose=False):
    """Full real matrix of the apply (tiny models only), column by column."""
    n = model.N * model.d * model.d * 2
    A = np.zeros((n, n))
    f = apply_transpose if transpose else apply
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = as_real(f(model, from_real(e, model), trace_row=trace_row))
    return A


# --------------------------------------------------------------------------
# BiCG convention (bicg/calc_ax_bicg_omp.f90, calc_atrx_bicg_omp.f90, solvebicg.f90)
# --------------------------------------------------------------------------
def _herm_from_lower(x0):
    """rho0 rebuilt from its lower triangle: real diagonal, upper = conj(lower)^T
    (calc_ax_bicg_omp.f90:70-77; math indices, row i > column j is lower)."""
    L = np.tril(x0, -1)
    return L + L.conj().T + np.diag(np.real(np.diag(x0)))


def apply_bicg(model: Model, x):
    """BiCG-convention apply (calc_ax_bicg_omp.f90): tier 0 of x is
    Hermitianised from its lower triangle (the caller's x is left untouched
    here; the reference overwrites it), y = L x, y(1,1) of tier 0 = tr rho0
    (:321-324), then the upper triangle of y's tier 0 is zeroed and its
    diagonal made real (:330-337)."""
    xh = x.copy()
    xh[0] = _herm_from_lower(x[0])
    y = apply(model, xh, trace_row=None)
    y[0, 0, 0] = np.trace(xh[0])
    y[0] = np.tril(y[0], -1) + np.diag(np.real(np.diag(y[0])))
    return y


def apply_transpose_bicg(model: Model, w):
    """The real transpose of :func:`apply_bicg` (the operator of
    calc_atrx_bicg_omp.f90): w's tier 0 reduced to its lower triangle with a
    real diagonal (:103-110), element (1,1) taken out of L^T's input and added
    to the diagonal (the trace row), then the transpose of the lower-triangle
    Hermitianisation folds the result back onto the lower triangle."""
    d = model.d
    wp = w.copy()
    wp[0] = np.tril(w[0], -1) + np.diag(np.real(np.diag(w[0])))
    t = np.real(wp[0, 0, 0])
    wp[0, 0, 0] = 0
    z = apply_transpose(model, wp, trace_row=None)
    z[0] += t * np.eye(d)
    z0 = z[0]
    z[0] = np.tril(z0, -1) + np.tril(z0.T.conj(), -1) + np.diag(np.real(np.diag(z0)))
    return z


def bicg(model: Model, tol=1e-10, maxiter=20000, x0=None):
    """Steady state by BiCG in the reference's convention (solvebicg.f90):
    b = e_(1,1) of tier 0, shadow s0 = r0, absolute stop ||r|| <= tol
    (:124-150); the returned rho0 is Hermitianised from its lower triangle
    (:108-113).  Returns (x, iters, ||r||)."""
    d = model.d
    x = np.zeros((model.N, d, d), complex) if x0 is None else x0.copy()
    if x0 is None:
        x[0] = np.eye(d) / d
    x[0] = np.tril(x[0], -1) + np.diag(np.real(np.diag(x[0])))
    dot = lambda a, b: float(np.real(np.vdot(a, b)))
    r = -apply_bicg(model, x)
    r[0, 0, 0] += 1.0
    if np.sqrt(dot(r, r)) <= tol:
        x[0] = _herm_from_lower(x[0])
        return x, 0, np.sqrt(dot(r, r))
    s, dk, f = r.copy(), r.copy(), r.copy()
    it = 0
    while True:
        q = apply_bicg(model, dk)
        rho = dot(s, r)
        alpha = rho / dot(f, q)
        x = x + alpha * dk
        r = r - alpha * q
        s = s - alpha * apply_transpose_bicg(model, f)
        err = np.sqrt(dot(r, r))
        if err <= tol or it >= maxiter:
            x[0] = _herm_from_lower(x[0])
            return x, it + 1, err
        beta = dot(s, r) / rho
        dk = r + beta * dk
        f = s + beta * f
        it += 1


# --------------------------------------------------------------------------
# Sensitivities of the steady current (adjoint; finite differences)
# --------------------------------------------------------------------------
def current_weights(model: Model, lead=0):
    """g with I_lead (summed over spin) = Re <g, x> for every x, found by
    probing :func:`current` (real-linear in x) on each element of the tier-1
    ADOs, the only ones it reads (td/calcj.f90:15-37)."""
    g = np.zeros((model.N, model.d, model.d), complex)
    e = np.zeros_like(g)
    for a in range(model.first[1], model.first[2]):
        for i in range(model.d):
            for j in range(model.d):
                for ph in (1.0, 1j):
                    e[a, i, j] = ph
                    g[a, i, j] += ph * current(model, e)[lead].sum()
                    e[a, i, j] = 0
    return g


def _projected(model: Model, transpose=False):
    proj = lambda v: as_real(physical_projection(model, from_real(v, model)))
    f = apply_transpose if transpose else apply
    mv = lambda v: proj(as_real(f(model, from_real(proj(v), model), trace_row="rdo")))
    return mv, proj


def current_gradient_adjoint(params_fn, theta, lead=0, tol=1e-13):
    """dI_lead/dtheta from one transposed solve.

    A(theta) x = e_trace with the theta-independent trace row; I = Re<g, x>;
    A^T lam = g (the transposed projected operator Pi A^T Pi, built on
    apply_transpose -- the role calc_atrx_bicg_omp plays in the reference's
    only transposed path); dI/dtheta_i = -Re<lam, Pi (dA/dtheta_i) x> with the
    trace-row element of the derivative zero.  Every theta_i used here enters
    A affinely (H linear in the level energies and U, cb/cd linear in Gamma,
    gamma affine in the bias), so dA/dtheta_i x = A(theta + e_i) x - A(theta) x
    exactly.  Returns (I, grad)."""
    theta = np.asarray(theta, float)
    model = Model(params_fn(theta))
    x, _, _ = steady_state(model, tol=tol)
    g = current_weights(model, lead)
    mvT, proj = _projected(model, transpose=True)
    lam_r, _, _ = tfqmr(mvT, proj(as_real(g)), tol=tol, project=proj)
    lam = from_real(lam_r, model)
    d = model.d
    ax = apply(model, x, trace_row=None)
    grad = np.zeros(len(theta))
    for i in range(len(theta)):
        tp = theta.copy()
        tp[i] += 1.0
        z = physical_projection(model, apply(Model(params_fn(tp)), x, trace_row=None) - ax)
        z[0, d - 1, d - 1] = 0.0
        grad[i] = -np.real(np.vdot(lam, z))
    return float(current(model, x)[lead].sum()), grad


def current_gradient_fd(params_fn, theta, lead=0, h=1e-5, tol=1e-13):
    """Central finite differences of I_lead with a full steady solve per point."""
    theta = np.asarray(theta, float)
    grad = np.zeros(len(theta))
    for i in range(len(theta)):
        vals = []
        for sgn in (1.0, -1.0):
            tp = theta.copy()
            tp[i] += sgn * h
            m = Model(params_fn(tp))
            x, _, _ = steady_state(m, tol=tol)
            vals.append(current(m, x)[lead].sum())
        grad[i] = (vals[0] - vals[1]) / (2 * h)
    return grad


# --------------------------------------------------------------------------
# Propagation (fixed H)
# --------------------------------------------------------------------------
def rk4(model: Model, x, dt, nstep):
    """Classic RK4 on dx/dt = L x (td/intstp.f90:20-100): k1..k4 from
    calcderiv at x, x + dt/2 k1, x + dt/2 k2, x + dt k3."""
    x = x.copy()
    for _ in range(nstep):
        k1 = apply(model, x, trace_row=None)
        k2 = apply(model, x + 0.5 * dt * k1, trace_row=None)
        k3 = apply(model, x + 0.5 * dt * k2, trace_row=None)
        k4 = apply(model, x + dt * k3, trace_row=None)
        x = x + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return x


def engy_shift(model: Model, t, fieldtype, tchar, amps_dc=None, t_off=1e99):
    """main/getengyshift.f90 for every (lead, spin), in 1/fs: fieldtype 0
    amps (1 - exp(-t / tchar)), 1 amps sin(2 pi t / tchar) + amps_dc (eV in),
    -1 zero; zero for t >= t_off.  amps = engyshift / hbar, the amplitude the
    steady bias uses (evaluatepara.f90:2296,2431)."""
    amps = model.eleadinfty
    if t >= t_off or fieldtype == -1:
        return np.zeros_like(amps)
    tchar = np.asarray(tchar, float)
    if fieldtype == 0:
        return amps * (1.0 - np.exp(-t / tchar))
    if fieldtype == 1:
        dc = 0.0 if amps_dc is None else np.asarray(amps_dc, float) / HBAR
        return amps * np.sin(2.0 * np.pi * t / tchar) + dc
    raise ValueError("fieldtype")


def gdraw_at(model: Model, eshift):
    """Per-drawer rate with a lead shift (td/calcderiv_omp.f90:144-148):
    cgamma + i dipm eshift(alpha, s)."""
    return model.cgamma + 1j * model.dipm * eshift[model.alf, model.spin]


def dhs_coef(kind, p, t):
    """The scalar time functions of main/calcdhs.f90's options (p[0] in eV):
    0 dedot (1 - exp(-t/aD)) (:93-94); 1 dedot sin(wD t) (:95-96); 2 dedot
    cos(w t) (:160-161); 3 egate on ton_gate < t < toff_gate (:112-115); 4 on
    tson <= t < tsoff (:267-270); 5 gpara1 + gpara2 cos(gpara3 t + gpara4)
    (:120-128); 6 dedot until t = wdot (:140-142); 7 the Hubbard dot field of
    :46-58; 8 a term present for t > t_on (tdot12, B fields)."""
    if kind == 0:
        return p[0] * (1.0 - np.exp(-t / p[1]))
    if kind == 1:
        return p[0] * np.sin(p[1] * t)
    if kind == 2:
        return p[0] * np.cos(p[1] * t)
    if kind == 3:
        return p[0] if p[1] < t < p[2] else 0.0
    if kind == 4:
        return p[0] if p[1] <= t < p[2] else 0.0
    if kind == 5:
        return p[0] + p[1] * np.cos(p[2] * t + p[3])
    if kind == 6:
        return p[0] if t < p[1] else 0.0
    if kind == 7:
        if t < p[2]:
            return p[0] * np.exp(-p[1] * t)
        if t < p[3]:
            return -p[0] * np.exp(-p[1] * (t - p[2]))
        return p[0] * np.exp(-p[1] * (t - p[3]))
    if kind == 8:
        return p[0] if t > p[1] else 0.0
    raise ValueError("kind")


def h_at(model: Model, t, terms):
    """hs + dhs(t) (td/calcderiv_omp.f90:60-68), dhs = sum f_k(t) D_k / hbar
    for terms [(D math-indexed, kind, params)]."""
    H = model.H.astype(complex).copy()
    for D, kind, p in terms:
        p = list(p) + [0.0] * (4 - len(p))
        H = H + dhs_coef(kind, p, t) * np.asarray(D, complex) / HBAR
    return H


def rk4_field(model: Model, x, t0, dt, nstep, field, terms=()):
    """td/intstp.f90 with the derivative calcderiv_omp(tt) at t, t + dt/2 (twice)
    and t + dt; field = (fieldtype, tchar, amps_dc, t_off), terms the system
    field of h_at."""
    f = lambda t, v: apply(model, v, trace_row=None, gdraw=gdraw_at(model, engy_shift(model, t, *field)),
                           H=h_at(model, t, terms) if terms else None)
    x = x.copy()
    for step in range(nstep):
        t = t0 + step * dt
        k1 = f(t, x)
        k2 = f(t + dt / 2, x + dt / 2 * k1)
        k3 = f(t + dt / 2, x + dt / 2 * k2)
        k4 = f(t + dt, x + dt * k3)
        x = x + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return x


def ado_gamma(model: Model):
    """gamma_K = sum of the drawers' gamma (incl. steady bias) per ADO."""
    g = np.zeros(model.N, complex)
    for n in range(1, model.L + 1):
        g[model.first[n]:model.first[n + 1]] = model.gdraw[model.tiers[n]].sum(axis=1)
    return g


def cheby_delta(model: Model):
    """Spectral width of the real part of diag(L) over the hierarchy, the
    tier-0 block (gamma = 0) included (td/prechebyshev.f90:71-109).  For a
    real H the commutator adds nothing to Re diag(L)."""
    re = ado_gamma(model).real
    return max(0.0, re.max()) - min(0.0, re.min())


def cheby_terms(arg, dpre=1e-21):
    """Bessel weights J_0 .. J_{n-1}(arg), n the first index whose |J_n| /
    sum_{m<=n} |J_m| < dpre (td/prechebyshev.f90:122-134)."""
    from scipy.special import jv
    J, tot, k = [], 0.0, 0
    while True:
        j = float(jv(k, arg))
        J.append(j)
        tot += abs(j)
        k += 1
        if abs(j) / tot < dpre:
            return np.array(J)


def chebyshev(model: Model, x, dt, nstep):
    """Chebyshev step x <- sum_k c_k J_k(delta dt) P_k (td/chebyshev.f90:126-200):
    P_0 = x, P_1 = L x / delta, P_k = P_{k-2} + (2/delta) L P_{k-1},
    c_0 = 1, c_k = 2.  P_k = (-i)^k T_k(iL/delta) x, so the sum is the
    Jacobi-Anger series of exp(L dt)."""
    delta = cheby_delta(model)
    J = cheby_terms(delta * dt)
    x = x.copy()
    for _ in range(nstep):
        pm2 = x
        pm1 = apply(model, pm2, trace_row=None) / delta
        out = J[0] * pm2 + 2 * J[1] * pm1
        for k in range(2, len(J)):
            pk = pm2 + (2.0 / delta) * apply(model, pm1, trace_row=None)
            out = out + 2 * J[k] * pk
            pm2, pm1 = pm1, pk
        x = out
    return x, len(J), delta


# --------------------------------------------------------------------------
# Sparse (COO per ADO) storage: tfqmr_double/calc_ax_tfqmr_spa_omp.f90
# --------------------------------------------------------------------------
def spa_pattern_from_mask(mask):
    """The reference's sparse pattern arrays (module sparsemod, 1-based) for a
    boolean (N, d, d) mask: nnz_spa and ind_spa per ADO, irow_spa / icol_spa
    per element, each ADO's elements in column-major order as
    fermion/prelude_spa.f90:550-561 lists them."""
    N, d, _ = mask.shape
    nnz = np.zeros(N, np.int32)
    ind = np.zeros(N, np.int64)
    rows, cols = [], []
    pos = 0
    for a in range(N):
        jj, ii = np.nonzero(mask[a].T)          # column-major: column slowest
        nnz[a] = len(ii)
        ind[a] = pos + 1 if len(ii) else 0
        rows.extend(ii + 1)
        cols.extend(jj + 1)
        pos += len(ii)
    return nnz, ind, np.array(rows, np.int32), np.array(cols, np.int32)


def calc_exp_hmat(H, alpha, kmax):
    """utility/calc_exp_hmat.f90: 1 + sum_{k=1}^{kmax} (alpha H)^k / k!."""
    out = np.eye(H.shape[0], dtype=complex)
    pw = np.eye(H.shape[0], dtype=complex)
    f = 1.0
    for k in range(1, kmax + 1):
        f /= k
        pw = alpha * (pw @ H)
        out = out + f * pw
    return out


def spa_pattern_build(model: Model, dsmall=1e-32, seed=0):
    """The sparse-mode pattern of fermion/prelude_spa.f90:532-749, in the
    module sparsemod arrays (nnz_spa, ind_spa, irow_spa, icol_spa; 1-based,
    ADO after ADO as :741-742).

    Tier 0: the elements of E = calc_exp_hmat(H, 0.31415926, 10) above dsmall
    (:231-238; 10 and 11 terms must agree, :233-237), E its values.  Tier n:
    for every lower neighbour J of K (drawer k removed), Y_L = op_k X and
    Y_R = X op_k with X = rho_J (rho_J^H for itype 1) and op_k = c^+ for
    sigma = +, c for sigma = - (:640-686, zamsmm_coo); Z += m1 E Y_L + m2 E Y_R
    with random complex normals m1, m2 (:676-683, mkl_zcoomm with the tier-0
    values).  K holds the elements of Z above dsmall (zmat_dns2coo, column
    major) and Z is its value for the next tier; a drawer whose lead-orbital
    linewidth is below dsmall empties K (:650-653, :733-737).
    """
    d = model.d
    E = calc_exp_hmat(model.H, 0.31415926, 10)
    if not np.array_equal(np.abs(E) > dsmall, np.abs(calc_exp_hmat(model.H, 0.31415926, 11)) > dsmall):
        raise RuntimeError("prelude_spa: sparsity analysis failed")
    rng = np.random.default_rng(seed)
    vals = [None] * model.N
    masks = np.zeros((model.N, d, d), bool)
    masks[0] = np.abs(E) > dsmall
    vals[0] = np.where(masks[0], E, 0)
    for n in range(1, model.L + 1):
        lo, _ = model.lower[n]
        for a, K in enumerate(model.tiers[n]):
            g = model.first[n] + a
            Z = np.zeros((d, d), complex)
            empty = False
            for r in range(n):
                k = int(K[r])
                if abs(model.p.linewidth[model.alf[k]][model.orb[k]]) < dsmall:
                    empty = True
                    continue
                X = vals[lo[a, r, 0]]
                if lo[a, r, 1] == 1:
                    X = X.conj().T
                c = model.ams[model.orb[k]][model.spin[k]]
                op = c.T if model.sgn[k] == 0 else c
                while True:
                    m1, m2 = rng.standard_normal(2) + 1j * rng.standard_normal(2)
                    if abs(m1) > 1e-6 and abs(m2) > 1e-6:
                        break
                Z += m1 * (E @ (op @ X)) + m2 * (E @ (X @ op))
            if not empty:
                masks[g] = np.abs(Z) > dsmall
            vals[g] = np.where(masks[g], Z, 0)
    nnz, ind, irow, icol = spa_pattern_from_mask(masks)
    ind = np.concatenate([[1], 1 + np.cumsum(nnz)[:-1]]).astype(np.int64)
    return nnz, ind, irow, icol


def apply_spa(model: Model, pattern, x_spa):
    """calc_ax_tfqmr_spa_omp: the operator on COO storage keeps only the
    pattern's elements of the result (zmat_extract_coo), y = P A P^T x
    (trace row none), and appends tr rho0 over the tier-0 diagonal elements
    of the pattern as element lunk_spa + 1 (:1061-1069)."""
    nnz, ind, irow, icol = pattern
    d = model.d
    x = np.zeros((model.N, d, d), complex)
    a_of = np.repeat(np.arange(model.N), nnz)
    x[a_of, irow - 1, icol - 1] = x_spa
    y = apply(model, x, trace_row=None)
    out = np.empty(len(x_spa) + 1, complex)
    out[:-1] = y[a_of, irow - 1, icol - 1]
    diag0 = (a_of == 0) & (irow == icol)
    out[-1] = x_spa[diag0].sum()
    return out


def calcderiv_spa(model: Model, pattern, x_spa, gdraw=None):
    """td/calcderiv_spa_omp.f90: rhotmp_spa -> rhorhs_spa.  Each ADO expanded
    from COO (zmat_coo2dns), the derivative of the dense hierarchy applied
    (no trace row; H fixed, the drawers' rates with the lead shift at t in
    ``gdraw``, :84-98), and the pattern's elements kept (zmat_extract_coo)."""
    nnz, ind, irow, icol = pattern
    d = model.d
    x = np.zeros((model.N, d, d), complex)
    a_of = np.repeat(np.arange(model.N), nnz)
    x[a_of, irow - 1, icol - 1] = x_spa
    y = apply(model, x, trace_row=None, gdraw=gdraw)
    return y[a_of, irow - 1, icol - 1]


# --------------------------------------------------------------------------
# Block-Jacobi iteration (jacobi/prelude_jacobi.f90, calc_ax_jacobi_omp.f90,
# solve_jacobi.f90; J. Chem. Phys. 147, 044105 (2017), Eq. 5)
# --------------------------------------------------------------------------
def liouville_commutator(model: Model):
    """d^2 x d^2 matrix of X -> HX - XH on column-major vec(X): zhil2liou('l')
    minus zhil2liou('r') (prelude_jacobi.f90:43-45)."""
    d = model.d
    return np.kron(np.eye(d), model.H) - np.kron(model.H.T, np.eye(d))


def jacobi_ejac(model: Model):
    """Default shift (prelude_jacobi.f90:48-113): the eigenvalues of the
    Hermitian commutator matrix with |lambda| > dpico = 1e-12 (:62-70), their
    spread, sqrt((ntier - 1) * max |cb|) (:88-105), and the larger of the two.
    Returns (ejac, spread, cb_term, number of nonzero eigenvalues)."""
    lam = np.linalg.eigvalsh(liouville_commutator(model))
    nz = lam[np.abs(lam) > 1e-12]
    gap = float(nz.max() - nz.min()) if nz.size else 0.0
    cbt = math.sqrt(model.L * float(np.abs(model.cb).max()))
    return max(gap, cbt), gap, cbt, int(nz.size)


def same_tier_matrix(model: Model, gamma):
    """S_K on column-major vec(X): -i[H, X] + gamma_K X (calc_ax_tfqmr_omp.f90:180-203)."""
    return -1j * liouville_commutator(model) + gamma * np.eye(model.d ** 2)


def _jacobi_blocks(model: Model, ejac):
    """Per-ADO (S_K, (ejac - S_K)^-1), shared between ADOs of equal gamma_K."""
    key = ("_jacobi_blocks", ejac)
    cached = getattr(model, "_jac_cache", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    d = model.d
    gam = ado_gamma(model)
    mats, blocks = {}, []
    for K in range(model.N):
        g = complex(gam[K])
        if abs(ejac - g) <= 1e-12:
            raise ValueError("ejac equals an ADO's gamma_K")
        if g not in mats:
            S = same_tier_matrix(model, g)
            mats[g] = (S, np.linalg.inv(ejac * np.eye(d * d) - S))
        blocks.append(mats[g])
    model._jac_cache = (key, blocks)
    return blocks


def jacobi_sweep(model: Model, x, ejac):
    """One sweep, calc_ax_jacobi_omp.f90: per ADO K the right-hand side
    R_K = ejac x_K + C_K(x) (:63-88 tier 0, :130-272 higher tiers; C_K the
    lower- and upper-tier couplings, here (A_0 x)_K - S_K x_K with A_0 the
    apply without a trace row), then x_K <- (ejac - S_K)^-1 R_K (:95-108,
    :276-298, there through the eigenvectors of [H, .]; here a dense solve).
    Rates within 1e-12 of ejac are not supported (the reference's
    1/(i lambda) branch, :285-287)."""
    d = model.d
    ax = apply(model, x, trace_row=None)
    blocks = _jacobi_blocks(model, ejac)
    out = np.empty_like(x)
    for K in range(model.N):
        S, Minv = blocks[K]
        xv = x[K].reshape(-1, order="F")
        cK = ax[K].reshape(-1, order="F") - S @ xv
        out[K] = (Minv @ (ejac * xv + cK)).reshape(d, d, order="F")
    return out


def block_jacobi_precondition(model: Model, y, shift):
    """M^-1 y with M = shift - S_K on every ADO K (the sweep operator of
    calc_ax_jacobi_omp.f90 used as a preconditioner), tier 0 made Hermitian."""
    d = model.d
    blocks = _jacobi_blocks(model, shift)
    z = np.empty_like(y)
    for K in range(model.N):
        z[K] = (blocks[K][1] @ y[K].reshape(-1, order="F")).reshape(d, d, order="F")
    z[0] = 0.5 * (z[0] + z[0].conj().T)
    return z


def jacobi_residual(model: Model, x):
    """||A x - e||_2 on the (Re, Im) vector with calc_ax_tfqmr_omp's trace row
    (solve_jacobi.f90:186-190: the real part of the last element minus 1)."""
    r = apply(model, x, trace_row="last")
    r[-1, model.d - 1, model.d - 1] -= 1.0
    return float(np.linalg.norm(as_real(r)))


def jacobi_solve(model: Model, x0, crit=1e-10, maxiter=20000, ejac=None):
    """solve_jacobi.f90: rho0 made Hermitian from its lower triangle with a real
    diagonal (:70-76) and required to have unit trace (:77-85); sweeps with the
    residual checked after sweeps 1, 51, 101, ... (:183-229) until it is
    <= crit or maxiter.  Returns (x, sweeps, residual)."""
    x = x0.copy()
    x[0] = _herm_from_lower(x[0])
    if abs(np.trace(x[0]) - 1.0) > 1e-12:
        raise ValueError("rho is not normalized")
    ejac = jacobi_ejac(model)[0] if ejac is None else ejac
    res = jacobi_residual(model, x)
    if res <= crit:
        return x, 0, res
    it = 0
    while True:
        x = jacobi_sweep(model, x, ejac)
        if it % 50 == 0:
            res = jacobi_residual(model, x)
            if res <= crit:
                break
        if it >= maxiter:
            break
        it += 1
    return x, it, res


# --------------------------------------------------------------------------
# Closed form at U = 0: Landauer with the same Lorentzian and Pade Fermi function
# --------------------------------------------------------------------------
def landauer_siam(p: Params, npts=400001):
    """Noninteracting single level: current per spin into the dot from each lead
    (same units as :func:`current`) and occupation per spin.

    Lead alpha has linewidth Gamma_a W^2/((w-V_a)^2+W^2) and Fermi function
    f_pade((w-V_a)/T), i.e. the rigidly shifted band that the bias gauge term
    i*dipm*V/hbar in gamma_K represents.  Integrated on omega = s*tan(theta).
    """
    eta, xi = pade_psd(p.nmats)
    eps = p.hpar["eup"]
    th = np.linspace(-np.pi / 2, np.pi / 2, npts)[1:-1]
    sc = max(p.bandwidth) * 0.5
    w = sc * np.tan(th)
    jac = sc / np.cos(th) ** 2
    G = []
    S = np.zeros_like(w, dtype=complex)
    F = []
    for a in range(p.nalf):
        W = p.bandwidth[a]
        V = p.engyshift[a][0]
        ga = p.linewidth[a][0]
        Ga = ga * W ** 2 / ((w - V) ** 2 + W ** 2)
        S += 0.5 * ga * W / (w - V + 1j * W)
        G.append(Ga)
        F.append(np.real(fermi_pade((w - V) / p.temperature[a], eta, xi)))
    den = np.abs(w - eps - S) ** 2
    h = th[1] - th[0]
    # composite Simpson on the uniform theta grid
    wts = np.ones_like(th)
    wts[1:-1:2] = 4
    wts[2:-1:2] = 2
    wts *= h / 3
    trans = G[0] * G[1] / den
    i_l = np.sum(wts * jac * trans * (F[0] - F[1])) / (2 * np.pi) / HBAR
    n_occ = np.sum(wts * jac * (G[0] * F[0] + G[1] * F[1]) / den) / (2 * np.pi)
    return i_l * JCONST, n_occ


def landauer_dqd(p: Params, npts=400001):
    """Noninteracting two-orbital system ($para4 with uu1 = uu2 = u12 = 0):
    current per spin into the dots from each lead and occupation per orbital
    and spin, from the 2 x 2 retarded Green's function.

    Each (lead alpha, orbital m) pair is its own bath (the drawers carry one
    orbital index, evaluatepara.f90:1701-1760, and the correlation functions
    are diagonal in it), so the self-energy of lead alpha is
    diag_m(Gamma_am W^2 / 2 / (w - V_a + i W)) with the Lorentzian width of
    :func:`landauer_siam`.  One-body h per spin: [[e1, t12], [t12, e2]]
    (main/calchs.f90:514-567).  I_L = int tr(G_L G G_R G^H)(f_L - f_R) / 2 pi,
    n_m = int [G (G_L f_L + G_R f_R) G^H]_mm / 2 pi (Caroli / Keldysh).
    Returns (I[alpha, s] in the units of :func:`current`, occ[m, s]).
    """
    assert p.norbs == 2 and p.nalf == 2
    eta, xi = pade_psd(p.nmats)
    hp = p.hpar
    th = np.linspace(-np.pi / 2, np.pi / 2, npts)[1:-1]
    sc = max(p.bandwidth) * 0.5
    w = sc * np.tan(th)
    jac = sc / np.cos(th) ** 2
    h = th[1] - th[0]
    wts = np.ones_like(th)
    wts[1:-1:2] = 4
    wts[2:-1:2] = 2
    wts *= h / 3 * jac / (2 * np.pi)
    cur = np.zeros((2, 2))
    occ = np.zeros((2, 2))
    eps = {0: (hp["e1up"], hp["e2up"]), 1: (hp["e1down"], hp["e2down"])}
    for s in range(2):
        hs = np.array([[eps[s][0], hp["t12"]], [hp["t12"], eps[s][1]]], dtype=complex)
        Ginv = w[:, None, None] * np.eye(2)[None] - hs[None]
        Gam, F = [], []
        for a in range(2):
            W = p.bandwidth[a]
            V = p.engyshift[a][s]
            g = np.zeros((len(w), 2, 2))
            for m in range(2):
                ga = p.linewidth[a][m]
                Ginv[:, m, m] -= 0.5 * ga * W / (w - V + 1j * W)
                g[:, m, m] = ga * W ** 2 / ((w - V) ** 2 + W ** 2)
            Gam.append(g)
            F.append(np.real(fermi_pade((w - V) / p.temperature[a], eta, xi)))
        G = np.linalg.inv(Ginv)
        GH = np.conj(np.transpose(G, (0, 2, 1)))
        trans = np.real(np.einsum("wij,wjk,wkl,wli->w", Gam[0], G, Gam[1], GH))
        cur[0, s] = np.sum(wts * trans * (F[0] - F[1])) / HBAR
        cur[1, s] = -cur[0, s]
        lesser = G @ (Gam[0] * F[0][:, None, None] + Gam[1] * F[1][:, None, None]) @ GH
        for m in range(2):
            occ[m, s] = np.sum(wts * np.real(lesser[:, m, m]))
    return cur * JCONST, occ


def ado_scale(model: Model):
    """prod_{k in K} |cb_k|^(1/2) per ADO (1 without lscale): rho_K = s_K * stored_K."""
    s = np.ones(model.N)
    for n in range(1, model.L + 1):
        s[model.first[n]:model.first[n + 1]] = np.prod(model.dbsqrt[model.tiers[n]], axis=1)
    return s
