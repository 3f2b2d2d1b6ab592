"""Three orbitals, d = 64, through the reference-array loader.

The drop-in loader takes the reference's module arrays for any number of
orbitals (nrho = 4**norbs); the model builder from parameters has the one-
and two-orbital Hamiltonians only.  Here the oracle's tables of a
three-orbital model (every orbital coupled to both leads, the on-site
energy and interaction on orbital 0; Pade Nk = 1, so M = 48) go through
heomq_model_from_reference, and the device's dense apply and transpose
(every trace-row convention), the packed apply and the steady state are
checked against the oracle.
"""
import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O
from ref_arrays import reference_arrays

pytestmark = pytest.mark.gpu

_cache = {}


def params(L):
    return O.Params(norbs=3, nspin=2, nalf=2, nmats=1, bandwidth=(10.0, 10.0),
                    linewidth=((0.1, 0.08, 0.12), (0.1, 0.08, 0.12)), temperature=(0.2, 0.2),
                    engyshift=((0.1, 0.1), (-0.1, -0.1)), ham="siam",
                    hpar=dict(eup=-0.5, edown=-0.5, uu=1.0), ntop=L)


def setup(L):
    if L not in _cache:
        import torch
        o = O.Model(params(L))
        m = heomq.Model.from_reference(**reference_arrays(o))
        ctx = heomq.Context(m)
        _cache[L] = (o, m, ctx, torch)
    return _cache[L]


def to_dev(torch, x):
    return torch.from_numpy(O.to_fortran(x)).to("cuda:0")


def from_dev(t):
    return O.from_fortran(t.cpu().numpy())


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("L", [1, 2])
@pytest.mark.parametrize("trace", ["last", "rdo", None])
def test_d64_dense_apply_and_transpose(L, trace):
    o, m, ctx, torch = setup(L)
    print(f"d {m.d} N {m.N} L {L} trace {trace}", flush=True)
    assert m.d == 64 and o.d == 64
    tr = {"last": heomq.TRACE_LAST, "rdo": heomq.TRACE_RDO, None: heomq.TRACE_NONE}[trace]
    x = O.seeded_ado_vector(o, 31)
    assert rel(from_dev(ctx.apply(to_dev(torch, x), trace=tr)), O.apply(o, x, trace_row=trace)) <= 1e-12
    w = O.seeded_ado_vector(o, 32)
    assert rel(from_dev(ctx.apply_transpose(to_dev(torch, w), trace=tr)),
               O.apply_transpose(o, w, trace_row=trace)) <= 1e-12


@pytest.mark.parametrize("L", [1, 2])
def test_d64_packed_apply(L):
    o, m, ctx, torch = setup(L)
    x = O.physical_projection(o, O.seeded_ado_vector(o, 33))
    for tr in (False, True):
        y = from_dev(ctx.apply_packed(to_dev(torch, x), trace=heomq.TRACE_RDO, transpose=tr))
        f = O.apply_transpose if tr else O.apply
        assert rel(y, O.physical_projection(o, f(o, x, trace_row="rdo"))) <= 1e-12


def test_d64_steady_state():
    o, m, ctx, torch = setup(1)
    x, st = ctx.solve_steady(tol=1e-12)
    print(st, flush=True)
    assert st["status"] == 0
    xo, _, _ = O.steady_state(o, tol=1e-13)
    xs = from_dev(x)
    assert np.abs(xs[0] - xo[0]).max() <= 1e-10
    cur, occ = ctx.observables(x)
    assert np.abs(cur - O.current(o, xo)).max() <= 1e-9 * np.abs(O.current(o, xo)).max()
    assert np.abs(occ - O.occupation(o, xo)).max() <= 1e-10
