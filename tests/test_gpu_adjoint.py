"""Adjoint sensitivities of the steady current on the GPU.

The device adjoint (transposed packed TFQMR + derivative apply) is checked
against the oracle's adjoint (itself pinned to central finite differences in
tests/test_oracle.py), against central finite differences of device solves,
and through torch.autograd.  Also covers the derivative apply and coefficient
refresh that the adjoint relies on.
"""
import numpy as np
import pytest

import heomq
from heomq import adjoint as ADJ
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

THETA = np.array([-0.7, 1.5, 0.1, 0.15, 0.3])
KW = dict(nk=2, ntop=3, T=0.2)
_cache = {}


def problem():
    if "p" not in _cache:
        _cache["p"] = ADJ.SteadyStateProblem(THETA, tol=1e-13, **KW)
    return _cache["p"]


def oracle_params(th):
    return O.siam_params(eps=th[0], U=th[1], gl=th[2], gr=th[3], V=th[4], **KW)


def to_dev(torch, x):
    return torch.from_numpy(O.to_fortran(x)).to("cuda:0")


def from_dev(t):
    return O.from_fortran(t.cpu().numpy())


def test_derivative_apply_matches_oracle():
    pr = problem()
    pr.set_theta(THETA)
    o0 = O.Model(oracle_params(THETA))
    x = O.seeded_ado_vector(o0, 51)
    xd = to_dev(pr.ctx.torch, x)
    a0 = O.apply(o0, x, trace_row=None)
    for i in range(5):
        print("parameter", ADJ.THETA_NAMES[i], flush=True)
        tp = THETA.copy()
        tp[i] += 1.0
        ref = O.apply(O.Model(oracle_params(tp)), x, trace_row=None) - a0
        got = from_dev(pr.ctx.apply_dparam(xd, *pr.coef_derivatives(i)))
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(a0).max()


def test_current_weights_match_observables():
    pr = problem()
    pr.set_theta(THETA)
    o = O.Model(oracle_params(THETA))
    xd = to_dev(pr.ctx.torch, O.seeded_ado_vector(o, 52))
    cur, _ = pr.ctx.observables(xd)
    assert abs(pr.current(xd) - cur[0].sum()) <= 1e-12 * abs(cur).max()


def test_refresh_updates_packed_operator():
    # set_params + refresh must reach every table the packed solver reads (H included)
    pr = problem()
    th = THETA + np.array([0.3, -0.4, 0.02, -0.03, 0.1])
    pr.set_theta(th)
    o = O.Model(oracle_params(th))
    x = O.physical_projection(o, O.seeded_ado_vector(o, 53))
    y = from_dev(pr.ctx.apply_packed(to_dev(pr.ctx.torch, x), trace=heomq.TRACE_RDO))
    yo = O.physical_projection(o, O.apply(o, x, trace_row="rdo"))
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    pr.set_theta(THETA)


def test_adjoint_gradient_matches_oracle_and_fd():
    pr = problem()
    pr.set_theta(THETA)
    x, I = pr.solve(THETA)
    g = pr.gradient(x)
    print("forward", pr.last["forward"], "adjoint", pr.last["adjoint"], flush=True)
    Io, go = O.current_gradient_adjoint(oracle_params, THETA)
    print("I", I, Io, "grad", g, go, flush=True)
    assert abs(I - Io) <= 1e-10 * abs(Io)
    assert np.abs(g - go).max() <= 1e-8 * np.abs(go).max()
    gf = pr.fd_gradient(h=1e-5)
    print("fd", gf, flush=True)
    assert np.abs(g - gf).max() <= 1e-6 * np.abs(gf).max()


def test_autograd_function():
    import torch
    pr = problem()
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    I = ADJ.steady_current(th, pr)
    I.backward()
    x, _ = pr.solve(THETA)
    assert np.abs(th.grad.numpy() - pr.gradient(x)).max() <= 1e-9 * np.abs(th.grad.numpy()).max()
