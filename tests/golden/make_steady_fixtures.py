"""Write the steady-state fixtures of the benchmark workloads (tests/golden/*.json).

TEST INFRASTRUCTURE: run in a CPU container, never by the product.  For each
workload the steady state is solved with the oracle's operator -- the
projected trace-row apply Pi A Pi of tfqmr_double/calc_ax_tfqmr_omp.f90 as the
OpenMP C port computes it (oracle/heom_apply_port.c, checked against the numpy
oracle by tests/test_oracle.py) -- by restarted GMRES (Saad & Schultz 1986;
Givens rotations, modified Gram-Schmidt with one reorthogonalisation), right-
preconditioned by the per-ADO inverse (shift - S_K)^-1 applied in the
eigenbasis of H.  GMRES shares no code and no failure mode with the device
TFQMR it checks.  It stops at the tolerance or when a restart cycle no longer
halves the true residual: the FP64 floor of the system, where the residual
of the computed x is as large as the rounding error of forming A x.  The
fixture holds the tier-0 block (RDO), the lead currents (td/calcj.f90 units)
and the occupations (main/calcocc.f90), plus the true residual
||b - A x|| / ||b|| the solve reached, the change of the RDO over the last
restart cycle (its error estimate), and the rounding noise of the residual:
||r_port - r_numpy|| / ||b|| for the same x, the residual formed by the C port
and by the numpy oracle's apply (two summation orders of one operator).

    python tests/golden/make_steady_fixtures.py [name ...]
"""
from __future__ import annotations

import json
import math
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import heom_oracle as O  # noqa: E402
from oracle.heom_port import Port  # noqa: E402

# The workloads: single-impurity Anderson model, Pade Nk = 8, L = 3, V = 0;
# two-impurity model, d = 16, Nk = 6, L = 3, V = 0.4; three bias points of the
# single-impurity sweep in [-1, 1].
WORKLOADS = {
    "siam_nk8_L3": dict(kind="siam", eps=-1.0, U=2.0, gamma=0.1, W=10.0, T=0.05, V=0.0, nk=8, ntop=3),
    "dqd_nk6_L3": dict(kind="dqd", eps=-1.5, U=3.0, t=0.5, gamma=0.2, W=10.0, T=0.1, V=0.4, nk=6, ntop=3),
    # the same system with the ADOs stored scaled (the reference's lscale)
    "dqd_nk6_L3_lscale": dict(kind="dqd", eps=-1.5, U=3.0, t=0.5, gamma=0.2, W=10.0, T=0.1, V=0.4, nk=6, ntop=3,
                              lscale=True),
    "siam_nk8_L3_V-1": dict(kind="siam", eps=-1.0, U=2.0, gamma=0.1, W=10.0, T=0.05, V=-1.0, nk=8, ntop=3),
    "siam_nk8_L3_V0.5": dict(kind="siam", eps=-1.0, U=2.0, gamma=0.1, W=10.0, T=0.05, V=0.5, nk=8, ntop=3),
}


def params_of(w):
    import dataclasses
    w = dict(w)
    kind = w.pop("kind")
    lscale = w.pop("lscale", False)
    p = O.siam_params(**w) if kind == "siam" else O.dqd_params(**w)
    return dataclasses.replace(p, lscale=lscale)


class Operator:
    """Pi A Pi on the packed physical sector, as real vectors (Re, Im)."""

    def __init__(self, model):
        self.m = model
        self.port = Port(model)
        d = model.d
        ones = np.ones((model.N, d, d), complex)
        self.idx = np.flatnonzero(np.abs(O.physical_projection(model, ones, hermitian=False)) > 0)
        self.n = self.idx.size
        self.dense = np.zeros((model.N, d, d), complex)
        # eigenbasis of H for the block-Jacobi preconditioner
        ev, U = np.linalg.eigh(model.H)
        self.ev, self.U = ev, U
        self.gam = O.ado_gamma(model)
        self.shift = 0.05 * O.jacobi_ejac(model)[0]

    def unpack(self, v):
        x = np.zeros(self.m.N * self.m.d * self.m.d, complex)
        x[self.idx] = v[: self.n] + 1j * v[self.n:]
        x = x.reshape(self.m.N, self.m.d, self.m.d)
        x[0] = 0.5 * (x[0] + x[0].conj().T)
        return x

    def pack(self, x):
        x = x.copy()
        x[0] = 0.5 * (x[0] + x[0].conj().T)
        f = x.reshape(-1)[self.idx]
        return np.concatenate([f.real, f.imag])

    def apply_dense(self, x):
        return self.port.apply(x, trace_row="rdo")

    def matvec(self, v):
        return self.pack(self.apply_dense(self.unpack(v)))

    def precond(self, v):
        x = self.unpack(v)
        U, Uh = self.U, self.U.conj().T
        xt = Uh[None] @ x @ U[None]
        den = self.shift - self.gam[:, None, None] + 1j * (self.ev[:, None] - self.ev[None, :])[None]
        return self.pack(U[None] @ (xt / den) @ Uh[None])


def gmres(op, b, x0, tol, restart=60, maxit=3000, log=print):
    """Right-preconditioned restarted GMRES on real vectors; stops on the true
    residual (tolerance, or a cycle that does not halve it).  Returns (x, matvecs,
    relres, x of the previous cycle)."""
    x = x0.copy()
    xprev = x0.copy()
    nb = np.linalg.norm(b)
    total = 0
    last = np.inf
    while True:
        r = b - op.matvec(x)
        beta = np.linalg.norm(r)
        log(f"  restart at {total} matvecs: true relres {beta / nb:.3e}")
        if beta <= tol * nb or total >= maxit or beta > 0.5 * last:
            return x, total, beta / nb, xprev
        last = beta
        xprev = x.copy()
        V = [r / beta]
        Z = []
        Hm = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k = 0
        for k in range(restart):
            z = op.precond(V[k])
            w = op.matvec(z)
            total += 1
            Z.append(z)
            for _ in range(2):                      # MGS with one reorthogonalisation
                for j in range(k + 1):
                    h = V[j] @ w
                    Hm[j, k] += h
                    w = w - h * V[j]
            Hm[k + 1, k] = np.linalg.norm(w)
            V.append(w / Hm[k + 1, k])
            for j in range(k):                      # apply the previous rotations
                t = cs[j] * Hm[j, k] + sn[j] * Hm[j + 1, k]
                Hm[j + 1, k] = -sn[j] * Hm[j, k] + cs[j] * Hm[j + 1, k]
                Hm[j, k] = t
            den = math.hypot(Hm[k, k], Hm[k + 1, k])
            cs[k], sn[k] = Hm[k, k] / den, Hm[k + 1, k] / den
            Hm[k, k] = den
            Hm[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            if abs(g[k + 1]) <= 0.1 * tol * nb or total >= maxit:
                break
        kk = k + 1
        yk = np.linalg.solve(np.triu(Hm[:kk, :kk]), g[:kk])
        for j in range(kk):
            x = x + yk[j] * Z[j]


def solve(name, tol=1e-13, log=print):
    w = WORKLOADS[name]
    p = params_of(w)
    t0 = time.time()
    m = O.Model(p)
    op = Operator(m)
    log(f"{name}: N = {m.N}, packed = {op.n}, threads = {op.port.threads()}")
    d = m.d
    x0 = np.zeros((m.N, d, d), complex)
    x0[0] = np.eye(d) / d
    e = np.zeros_like(x0)
    e[0, d - 1, d - 1] = 1.0
    b = op.pack(e)
    x, it, rel, xprev = gmres(op, b, op.pack(x0), tol, log=log)
    xs = op.unpack(x)
    r = np.linalg.norm(b - op.matvec(x)) / np.linalg.norm(b)
    rdo_change = float(np.abs(xs[0] - op.unpack(xprev)[0]).max())
    # the same residual through the numpy oracle's apply (another summation order)
    r_np = op.pack(O.apply(m, xs, trace_row="rdo"))
    noise = float(np.linalg.norm(op.matvec(x) - r_np) / np.linalg.norm(b))
    log(f"  RDO change over the last cycle {rdo_change:.2e}, residual rounding noise {noise:.2e}")
    cur = O.current(m, xs)
    occ = O.occupation(m, xs)
    out = dict(workload=name, params=w, N=int(m.N), d=int(d), packed=int(op.n),
               rdo_re=xs[0].real.tolist(), rdo_im=xs[0].imag.tolist(), current=cur.tolist(),
               occupation=occ.tolist(), relres=float(r), matvecs=int(it),
               rdo_change_last_cycle=rdo_change, residual_rounding_noise=noise,
               trace=[float(np.trace(xs[0]).real), float(np.trace(xs[0]).imag)],
               seconds=round(time.time() - t0, 1),
               solver="right-preconditioned GMRES(60), C-port operator Pi A Pi, trace row on tier 0")
    (HERE / f"steady_{name}.json").write_text(json.dumps(out, indent=1) + "\n")
    log(f"{name}: relres {r:.3e} after {it} matvecs, {out['seconds']} s")
    return out


if __name__ == "__main__":
    names = sys.argv[1:] or list(WORKLOADS)
    for nm in names:
        solve(nm, log=lambda s: print(s, flush=True))
