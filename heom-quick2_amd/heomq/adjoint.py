"""Adjoint sensitivities of the steady-state lead current.

The reference has no gradient path: it obtains dI/dV by finite differences
over separate runs (tools/make_input_didv/makeinput.f90, deltav = 5e-4).
Here the steady state A(theta) x = e_trace and the current I = Re<g, x> are
differentiated exactly:

    A^T lam = g                       (one transposed solve on the device)
    dI/dtheta_i = -Re<lam, Pi dA/dtheta_i x>

with Pi the projection the solvers iterate with (N / S_z selection rules and
a Hermitian tier-0 block) and the trace row (theta-independent) removed from
dA/dtheta_i x.  The transposed solve uses the real-transpose apply -- the
position of bicg/calc_atrx_bicg_omp.f90 -- inside TFQMR; dA/dtheta_i x is one
launch of the derivative apply (``heomq_apply_dparam``) whose coefficients are
the exact derivatives of H, cb, cd and gamma.  All five parameters
(eps_d, U, Gamma_L, Gamma_R, V) enter those coefficients affinely (H is
linear in the level energies and U; cb, cd are linear in Gamma through
zxip = Gamma W^2 / 2, main/evaluatepara.f90:1511; the steady bias adds
i dipm V_alpha / hbar to gamma, calc_ax_tfqmr_omp.f90:180-194), so each
derivative is the difference of the coefficient sets at theta + e_i and
theta, which is exact.

``SteadyCurrent`` wraps this as a ``torch.autograd.Function``: forward solves
for I_L(theta), backward returns grad * dI_L/dtheta from the one adjoint solve.
"""
from __future__ import annotations

import numpy as np

from . import Context, JCONST, Model, Params, siam

THETA_NAMES = ("eps_d", "U", "Gamma_L", "Gamma_R", "V")


def siam_theta(theta, **kw) -> Params:
    """SIAM parameters at theta = (eps_d, U, Gamma_L, Gamma_R, V); other
    keywords (nk, ntop, T, W, lsdraw) go to :func:`heomq.siam`."""
    e, U, gl, gr, V = (float(t) for t in theta)
    return siam(eps=e, U=U, gl=gl, gr=gr, V=V, **kw)


def current_weights(model: Model, ctx: Context, lead: int = 0):
    """Device vector g (Fortran element order) with I_lead = Re<g, x> summed over
    spin, from the tier-1 lookups and c operators (td/calcj.f90:15-37,68)."""
    torch = ctx.torch
    M, d = model.M, model.d
    nspin, norbs = 2, model.norbs
    nalf = model.nalf
    ncor = M // (2 * nspin * norbs * nalf)
    _, upper = model.coupling()
    _, A = model.operators()                      # A[m*nspin + s][i, j] = c_{ms}(i, j)
    n1 = int(model.first[2]) if model.L >= 1 else 1
    g = np.zeros((n1, d, d), np.complex128)       # math-indexed g[a][i, j]
    for k in range(M // 2):                       # sigma = + drawers (sign slowest)
        r = k
        j = r % ncor; r //= ncor
        a = r % nalf; r //= nalf
        m = r % norbs; r //= norbs
        s = r
        if a != lead:
            continue
        nb, itype, ff = int(upper[k, 0]), int(upper[k, 1]), int(upper[k, 2])
        c = A[m * nspin + s]
        if itype == 0:      # X = x_nb:        I += -2 J sum c[i,k] Im x[k,i]
            g[nb] += 1j * (-2.0 * JCONST) * c.T
        else:               # X = ff x_nb^H:   I += 2 J ff sum c[i,k] Im x[i,k]
            g[nb] += 1j * (2.0 * JCONST * ff) * c
    out = ctx.zeros()
    out[:n1] = torch.from_numpy(np.ascontiguousarray(g.transpose(0, 2, 1))).to(out.device)
    return out


class SteadyStateProblem:
    """One model shape (hierarchy, tiers) with parameters theta that change
    only coefficients; owns the device context and an auxiliary host model
    used to form exact coefficient derivatives."""

    def __init__(self, theta, make_params=siam_theta, lead=0, tol=1e-12, maxiter=20000, device=0, **kw):
        self.make_params = lambda th: make_params(th, **kw)
        self.lead = lead
        self.tol = tol
        self.maxiter = maxiter
        self.theta = np.array(theta, float)
        self.model = Model(self.make_params(self.theta))
        self.aux = Model(self.make_params(self.theta))
        self.ctx = Context(self.model, device)
        self.g = current_weights(self.model, self.ctx, lead)
        self.last = {}

    def set_theta(self, theta):
        theta = np.array(theta, float)
        if not np.array_equal(theta, self.theta):
            self.model.set_params(self.make_params(theta))
            self.ctx.refresh(self.model)
            self.theta = theta

    def current(self, x):
        t = self.ctx.torch
        return float(t.real(t.vdot(self.g.flatten(), x.flatten())).item())

    def solve(self, theta=None, x0=None):
        """Steady state and I_lead at theta."""
        if theta is not None:
            self.set_theta(theta)
        x, st = self.ctx.solve_steady(x=x0, tol=self.tol, maxiter=self.maxiter)
        self.last["forward"] = st
        return x, self.current(x)

    def _coef_set(self, theta):
        self.aux.set_params(self.make_params(theta))
        cb, cd, gam = self.aux.coefs()
        H, _ = self.aux.operators()
        return H, cb, cd, gam

    def coef_derivatives(self, i):
        """(dH, dcb, dcd, dgam) of parameter i: exact, as the coefficients are affine in it."""
        base = self._coef_set(self.theta)
        tp = self.theta.copy()
        tp[i] += 1.0
        plus = self._coef_set(tp)
        return tuple(p - b for p, b in zip(plus, base))

    def gradient(self, x):
        """dI_lead/dtheta at the current theta for the steady state x (device)."""
        ctx = self.ctx
        t = ctx.torch
        lam, st = ctx.solve_steady(x=ctx.zeros(), b=self.g, tol=self.tol, maxiter=self.maxiter, transpose=True)
        self.last["adjoint"] = st
        d = self.model.d
        grad = np.zeros(len(self.theta))
        for i in range(len(self.theta)):
            dH, dcb, dcd, dgam = self.coef_derivatives(i)
            z = ctx.apply_dparam(x, dH, dcb, dcd, dgam)
            z[0] = 0.5 * (z[0] + z[0].conj().transpose(0, 1))   # Pi on tier 0 (x's layout is [j, i]; the sum is symmetric)
            z[0, d - 1, d - 1] = 0.0                             # the trace row does not depend on theta
            grad[i] = -float(t.real(t.vdot(lam.flatten(), z.flatten())).item())
        return grad

    def fd_gradient(self, h=1e-5):
        """Central finite differences (two full solves per parameter)."""
        theta0 = self.theta.copy()
        grad = np.zeros(len(theta0))
        for i in range(len(theta0)):
            vals = []
            for sgn in (1.0, -1.0):
                tp = theta0.copy()
                tp[i] += sgn * h
                vals.append(self.solve(tp)[1])
            grad[i] = (vals[0] - vals[1]) / (2 * h)
        self.set_theta(theta0)
        return grad


def _autograd_function():
    import torch

    class SteadyCurrent(torch.autograd.Function):
        """I_lead(theta) at the steady state; backward by one adjoint solve."""

        @staticmethod
        def forward(fctx, theta, problem: SteadyStateProblem):
            th = theta.detach().cpu().numpy().astype(float)
            x, cur = problem.solve(th)
            fctx.problem = problem
            fctx.theta = th
            fctx.x = x
            return torch.tensor(cur, dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad_out):
            fctx.problem.set_theta(fctx.theta)
            g = fctx.problem.gradient(fctx.x)
            return grad_out * torch.from_numpy(g), None

    return SteadyCurrent


_FN = None


def steady_current(theta, problem: SteadyStateProblem):
    """Differentiable I_lead(theta): ``steady_current(theta, problem).backward()``."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(theta, problem)


def bias_sweep(problem: SteadyStateProblem, voltages, theta=None):
    """I_lead and dI_lead/dtheta at each bias point (warm-started solves); the
    last component of each gradient is the differential conductance dI/dV."""
    theta = problem.theta.copy() if theta is None else np.array(theta, float)
    cur, grads, x = [], [], None
    for V in voltages:
        th = theta.copy()
        th[4] = V
        x, I = problem.solve(th, x0=x.clone() if x is not None else None)
        cur.append(I)
        grads.append(problem.gradient(x))
    return np.array(cur), np.array(grads)
