"""Batched steady states on the GPU.

A batch of SIAM problems of one hierarchy shape (bias, level, temperature and
Gamma differing) is solved in one set of launches; every problem must agree
with its own single-problem device solve and with the oracle's steady state.
"""
import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

POINTS = [dict(V=-0.8, eps=-1.2, T=0.1), dict(V=0.0, eps=-1.0, T=0.2), dict(V=0.5, eps=-0.3, T=0.05),
          dict(V=1.2, eps=-1.6, T=0.15, gl=0.15), dict(V=-0.3, eps=0.0, T=0.08)]
SHAPE = dict(nk=3, ntop=3, U=1.5)


def test_batch_apply_matches_single():
    import torch
    pl = [heomq.siam(**SHAPE, **pt) for pt in POINTS]
    b = heomq.Batch(pl)
    xs = torch.zeros((b.nprob, b.packed), dtype=torch.complex128, device="cuda:0")
    ref = []
    for p, pp in enumerate(pl):
        o = O.Model(O.siam_params(**SHAPE, **POINTS[p]))
        x = O.physical_projection(o, O.seeded_ado_vector(o, 80 + p))
        ctx = heomq.Context(heomq.Model(pp))
        xp = ctx.pack(torch.from_numpy(O.to_fortran(x)).to("cuda:0"))
        xs[p] = xp
        ref.append(ctx.apply_packed_raw(xp, ctx.packed_zeros(), trace=heomq.TRACE_RDO).clone())
        ctx.close()
    ys = torch.zeros_like(xs)
    b.apply_packed(xs, ys)
    for p in range(b.nprob):
        err = (ys[p] - ref[p]).abs().max().item() / ref[p].abs().max().item()
        print("problem", p, "rel err", err, flush=True)
        assert err <= 1e-13


def test_batch_solve_matches_oracle():
    pl = [heomq.siam(**SHAPE, **pt) for pt in POINTS]
    b = heomq.Batch(pl)
    rdo, cur, occ, info = b.solve(tol=1e-12)
    print(info, flush=True)
    assert (info["status"] == 0).all() and (info["relres"] <= 1e-12).all()
    # currents carry jconst = 1.6e5: compared relative to the batch's current scale
    # (the V = 0 point has I ~ 0)
    iscale = np.abs(cur).max()
    for p, pt in enumerate(POINTS):
        o = O.Model(O.siam_params(**SHAPE, **pt))
        xo, _, _ = O.steady_state(o, tol=1e-13)
        assert np.abs(rdo[p] - xo[0]).max() <= 1e-10
        Io = O.current(o, xo)
        assert np.abs(cur[p] - Io).max() <= 1e-10 * iscale
        assert np.abs(occ[p] - O.occupation(o, xo)).max() <= 1e-10
    # a second solve on the same batch (cached graph) gives the same states
    rdo2, _, _, _ = b.solve(tol=1e-12)
    assert np.abs(rdo2 - rdo).max() <= 1e-14


def test_batch_rejects_mixed_shapes():
    with pytest.raises(heomq.HeomqError):
        heomq.Batch([heomq.siam(nk=3, ntop=3), heomq.siam(nk=2, ntop=3)])


def test_batch_scaled_problems_match_oracle():
    # the reference's lscale (ADOs stored scaled by prod |cb_k|^(-1/2), each
    # problem with its own factors): states and currents are the unscaled ones
    pl = [heomq.siam(**SHAPE, **pt, lscale=True) for pt in POINTS[:3]]
    b = heomq.Batch(pl)
    rdo, cur, occ, info = b.solve(tol=1e-12)
    print(info, flush=True)
    assert (info["status"] == 0).all()
    iscale = np.abs(cur).max()
    for p, pt in enumerate(POINTS[:3]):
        o = O.Model(O.siam_params(**SHAPE, **pt))
        xo, _, _ = O.steady_state(o, tol=1e-13)
        assert np.abs(rdo[p] - xo[0]).max() <= 1e-10
        assert np.abs(cur[p] - O.current(o, xo)).max() <= 1e-10 * iscale
