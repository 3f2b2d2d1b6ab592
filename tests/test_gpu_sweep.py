"""Sweeps: problems of several hierarchy shapes in one launch per solver step.

Single-impurity problems of uneven shape (Pade terms Nk, top tier L), given
in an interleaved order, are grouped by the library and advanced together:
one flat-grid apply and one flat-grid launch of each TFQMR vector kernel per
step.  Checked against the single-problem device apply, the oracle's steady
states (small shapes), the per-shape batches, and, for the sweep workload's
extreme shapes (Nk = 4..8, L = 2..4, lscale), by properties: unit trace, a
Hermitian tier-0 block, I_L = -I_R, and the single-problem solve.
"""
import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

SMALL = [dict(nk=2, ntop=2, V=-0.8, eps=-1.2, T=0.1), dict(nk=3, ntop=3, V=0.5, eps=-0.3, T=0.05),
         dict(nk=2, ntop=3, V=0.0, eps=-1.0, T=0.2), dict(nk=3, ntop=3, V=1.2, eps=-1.6, T=0.15, gl=0.15),
         dict(nk=2, ntop=2, V=-0.3, eps=0.0, T=0.08), dict(nk=3, ntop=1, V=0.7, eps=-0.6, T=0.12),
         dict(nk=2, ntop=3, V=0.9, eps=-0.9, T=0.1)]


def _params(pt, **kw):
    return heomq.siam(U=1.5, **pt, **kw)


def test_sweep_layout_groups_shapes_largest_first():
    w = heomq.Sweep([_params(pt) for pt in SMALL])
    print("shapes", w.shape, "offsets", w.offset, "lengths", w.length, flush=True)
    assert w.nprob == len(SMALL) and w.nshapes == 4
    # problems of one shape share a length; shapes are numbered largest first
    for g in range(w.nshapes):
        assert len(set(w.length[w.shape == g])) == 1
    by_shape = [w.length[w.shape == g][0] for g in range(w.nshapes)]
    assert by_shape == sorted(by_shape, reverse=True)
    # the flat layout tiles [0, packed) without overlap
    order = np.argsort(w.offset)
    assert w.offset[order[0]] == 0
    assert np.all(w.offset[order[1:]] == (w.offset + w.length)[order[:-1]])
    assert (w.offset + w.length).max() == w.packed
    w.close()


def test_sweep_apply_matches_single():
    import torch
    pl = [_params(pt) for pt in SMALL]
    w = heomq.Sweep(pl)
    xs = torch.zeros(w.packed, dtype=torch.complex128, device="cuda:0")
    ref = []
    for p, pt in enumerate(SMALL):
        o = O.Model(O.siam_params(U=1.5, **pt))
        x = O.physical_projection(o, O.seeded_ado_vector(o, 90 + p))
        ctx = heomq.Context(heomq.Model(pl[p]))
        xp = ctx.pack(torch.from_numpy(O.to_fortran(x)).to("cuda:0"))
        assert xp.numel() == w.length[p]
        xs[w.offset[p]:w.offset[p] + w.length[p]] = xp
        ref.append(ctx.apply_packed_raw(xp, ctx.packed_zeros(), trace=heomq.TRACE_RDO).clone())
        ctx.close()
    ys = torch.zeros_like(xs)
    w.apply_packed(xs, ys)
    for p in range(w.nprob):
        y = ys[w.offset[p]:w.offset[p] + w.length[p]]
        err = (y - ref[p]).abs().max().item() / ref[p].abs().max().item()
        print("problem", p, "shape", w.shape[p], "rel err", err, flush=True)
        assert err <= 1e-13
    w.close()


def test_sweep_solve_matches_oracle_and_batches():
    pl = [_params(pt) for pt in SMALL]
    w = heomq.Sweep(pl)
    rdo, cur, occ, info = w.solve(tol=1e-12)
    print(info, flush=True)
    assert (info["status"] == 0).all() and (info["relres"] <= 1e-12).all()
    iscale = np.abs(cur).max()
    for p, pt in enumerate(SMALL):
        o = O.Model(O.siam_params(U=1.5, **pt))
        xo, _, _ = O.steady_state(o, tol=1e-13)
        assert np.abs(rdo[p] - xo[0]).max() <= 1e-10
        assert np.abs(cur[p] - O.current(o, xo)).max() <= 1e-10 * iscale
        assert np.abs(occ[p] - O.occupation(o, xo)).max() <= 1e-10
    # the same problems as one batch per shape
    for g in range(w.nshapes):
        idx = np.nonzero(w.shape == g)[0]
        b = heomq.Batch([pl[i] for i in idx])
        rb, cb, _, ib = b.solve(tol=1e-12)
        print("shape", g, "batch iters", ib["iters"], "sweep iters", info["iters"][idx], flush=True)
        assert np.abs(rb - rdo[idx]).max() <= 1e-12
        b.close()
    # a second solve (cached graph) reaches the same states
    rdo2, _, _, info2 = w.solve(tol=1e-12)
    assert np.abs(rdo2 - rdo).max() <= 1e-14 and np.array_equal(info2["iters"], info["iters"])
    w.close()


def test_sweep_rejects_unsupported_mixes():
    with pytest.raises(heomq.HeomqError, match="same d"):
        heomq.Sweep([heomq.siam(nk=2, ntop=2), heomq.two_impurity(nk=2, ntop=2)])
    many = [heomq.siam(nk=nk, ntop=L) for nk in range(1, 6) for L in range(1, 5)]   # 20 shapes
    with pytest.raises(heomq.HeomqError, match="shapes"):
        heomq.Sweep(many)


def test_sweep_extreme_shapes_properties():
    # the sweep workload's corner shapes (Nk = 4 and 8, L = 2 and 4), lscale,
    # two points each: trace one, Hermitian tier 0, I_L = -I_R, and the
    # single-problem device solve
    pts = []
    rng = np.random.default_rng(7)
    for nk, L in ((4, 2), (8, 4), (8, 2), (4, 4)):
        for _ in range(2):
            pts.append(dict(nk=nk, ntop=L, V=float(rng.uniform(-2, 2)), eps=float(rng.uniform(-2, 0)),
                            T=float(rng.uniform(0.02, 0.2)), U=2.0, gamma=0.1, W=10.0, lscale=True))
    pl = [heomq.siam(**pt) for pt in pts]
    w = heomq.Sweep(pl)
    rdo, cur, occ, info = w.solve(tol=1e-10)
    print(info, flush=True)
    assert (info["status"] == 0).all()
    iscale = np.abs(cur).max()
    for p in range(len(pts)):
        r = rdo[p]
        assert abs(np.trace(r) - 1) <= 1e-10
        assert np.abs(r - r.conj().T).max() <= 1e-12
        assert abs(cur[p][0].sum() + cur[p][1].sum()) <= 1e-8 * iscale
        ctx = heomq.Context(heomq.Model(pl[p]))
        x, st = ctx.solve_steady(tol=1e-10)
        r1 = x[0].cpu().numpy().T
        print("point", p, "sweep vs single", np.abs(r1 - r).max(), st["iters"], info["iters"][p], flush=True)
        assert np.abs(r1 - r).max() <= 1e-8
        ctx.close()
    w.close()
