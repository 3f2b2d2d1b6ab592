"""The derivative terminator of the top tier (ltrun_der) on the GPU.

The oracle restates the reference's loop (tfqmr_double/calc_ax_tfqmr_omp.f90
:362-478): per top-tier ADO and drawer, the virtual tier-(L+1) ADO from its
lower couplings and the resolvent in the Liouvillian eigenbasis, coupled back
up.  The device instead builds the hierarchy one tier higher and eliminates
its top tier in every apply.  Checked: the apply against the oracle; the
steady state against the steady state of the hierarchy one tier higher, which
the terminator reproduces exactly on tiers 0..L (a Schur complement), both
against the oracle and, at the single-impurity workload's bath (Pade Nk = 8),
against the device's own solve of the larger hierarchy.
"""
import dataclasses

import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

SMALL = {
    "siam_nk2_L2": (dict(nk=2, ntop=2, V=0.3, U=1.5, T=0.2), "siam"),
    "dqd_nk1_L1": (dict(nk=1, ntop=1, T=0.3), "dqd"),
}


def models(name):
    kw, kind = SMALL[name]
    if kind == "siam":
        return heomq.siam(**kw), O.siam_params(**kw)
    return heomq.two_impurity(**kw), O.dqd_params(**kw)


def physical_vector(o, seed):
    """A vector of the physical subspace (parity-allowed elements, the pair
    rule on self-paired ADOs): a random mix of A^k on a tier-0 state."""
    d = o.d
    v = np.zeros((o.N, d, d), complex)
    v[0] = np.diag(np.arange(1.0, d + 1))
    x = np.zeros_like(v)
    rng = np.random.default_rng(seed)
    for _ in range(o.L + 3):
        v = O.apply(o, v, trace_row=None)
        v /= np.linalg.norm(v)
        x += rng.standard_normal() * v
    return x


def to_dev(torch, x):
    return torch.from_numpy(O.to_fortran(x)).to("cuda:0")


def from_dev(t):
    return O.from_fortran(t.cpu().numpy())


@pytest.mark.parametrize("name", list(SMALL))
def test_trun_apply_matches_oracle(name):
    import torch
    pn, po = models(name)
    ctx = heomq.Context(heomq.Model(pn))
    o = O.Model(po)
    ctx.set_trun_der(True)
    x = physical_vector(o, 0x7E12)
    y = from_dev(ctx.apply_packed(to_dev(torch, x), trace=heomq.TRACE_RDO))
    yo = O.physical_projection(o, O.apply(o, x, trace_row="rdo", trun_der=True))
    yp = O.physical_projection(o, O.apply(o, x, trace_row="rdo"))
    err = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print(f"{name}: terminator apply rel err {err:.2e}, terminator share "
          f"{np.linalg.norm(yo - yp) / np.linalg.norm(yo):.2e}", flush=True)
    assert err <= 1e-12
    assert np.linalg.norm(yo - yp) >= 1e-6 * np.linalg.norm(yo)   # the term is really there
    ctx.set_trun_der(False)
    y0 = from_dev(ctx.apply_packed(to_dev(torch, x), trace=heomq.TRACE_RDO))
    assert np.linalg.norm(y0 - yp) <= 1e-12 * np.linalg.norm(yp)
    ctx.close()


def test_trun_steady_state_is_the_next_tier_against_oracle():
    pn, po = models("siam_nk2_L2")
    ctx = heomq.Context(heomq.Model(pn))
    ctx.set_trun_der(True)
    x, st = ctx.solve_steady(tol=1e-12)
    print("siam nk2 L2 with terminator:", st, flush=True)
    assert st["status"] == 0
    o1 = O.Model(dataclasses.replace(po, ntop=po.ntop + 1))
    x1, _, _ = O.steady_state(o1, tol=1e-13)
    xd = from_dev(x)
    assert np.abs(xd[0] - x1[0]).max() <= 1e-10
    assert np.abs(xd - x1[:len(xd)]).max() <= 1e-9
    # and it is not the plain tier-L state
    o = O.Model(po)
    xp, _, _ = O.steady_state(o, tol=1e-13)
    assert np.abs(xp[0] - x1[0]).max() >= 1e-6
    ctx.close()


def test_trun_steady_state_single_impurity_bath():
    # the single-impurity workload's bath (Pade Nk = 8, M = 72) at L = 2 with
    # the terminator against the device's own L = 3 solve
    kw = dict(eps=-1.0, U=2.0, gamma=0.1, W=10.0, T=0.05, V=0.4, nk=8)
    ctx = heomq.Context(heomq.Model(heomq.siam(**kw, ntop=2)))
    ctx.set_trun_der(True)
    x, st = ctx.solve_steady(tol=1e-12)
    print("siam nk8 L2 with terminator:", st, flush=True)
    ctx3 = heomq.Context(heomq.Model(heomq.siam(**kw, ntop=3)))
    x3, st3 = ctx3.solve_steady(tol=1e-12)
    r, r3 = x[0].cpu().numpy(), x3[0].cpu().numpy()
    c, c3 = ctx.observables(x)[0], ctx3.observables(x3)[0]
    print(f"RDO diff {np.abs(r - r3).max():.2e}, current {c.sum():.10e} vs {c3.sum():.10e}", flush=True)
    assert st["status"] == 0 and st3["status"] == 0
    assert np.abs(r - r3).max() <= 1e-10
    assert np.abs(c - c3).max() <= 1e-9 * np.abs(c3).max()
    ctx.close()
    ctx3.close()


def test_trun_refuses_the_other_paths():
    pn, _ = models("siam_nk2_L2")
    ctx = heomq.Context(heomq.Model(pn))
    ctx.set_trun_der(True)
    with pytest.raises(heomq.HeomqError, match="derivative terminator"):
        ctx.solve_bicg(crit=1e-10, maxiter=10)
    with pytest.raises(heomq.HeomqError, match="derivative terminator"):
        ctx.rk4(ctx.initial_guess(), 0.01, 1)
    ctx.set_trun_der(False)
    x, st = ctx.solve_bicg(crit=1e-10, maxiter=5000, raise_on_fail=False)
    assert st["iters"] > 0
    ctx.close()


def test_trun_needs_a_model_built_from_parameters():
    from ref_arrays import reference_arrays
    _, po = models("siam_nk2_L2")
    o = O.Model(po)
    ctx = heomq.Context(heomq.Model.from_reference(**reference_arrays(o)))
    with pytest.raises(heomq.HeomqError, match="parameters"):
        ctx.set_trun_der(True)
    ctx.close()
