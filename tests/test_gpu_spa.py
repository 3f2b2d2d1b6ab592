"""GPU parity of the sparse-storage drop-in (calc_ax_tfqmr_spa_omp).

Patterns in the reference's own module layout (nnz_spa, ind_spa, irow_spa,
icol_spa; 1-based), built from masks: the full d x d blocks, the physical
sector (what an exact run's pattern can hold), a random sub-pattern of it,
and the pattern the native prelude_spa restatement builds (also for a model
with a decoupled lead, whose left-lead ADOs are empty).  The device result is
checked against the oracle's P A P^T x plus the appended trace; at the
two-impurity workload's full size against the OpenMP C port of the dense
reference apply restricted to the pattern.
"""
import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "siam_nk2_L3": (lambda: heomq.siam(nk=2, ntop=3, V=0.3, U=1.5), lambda: O.siam_params(nk=2, ntop=3, V=0.3, U=1.5)),
    "dqd_nk2_L2": (lambda: heomq.two_impurity(nk=2, ntop=2), lambda: O.dqd_params(nk=2, ntop=2)),
    "siam_gl0": (lambda: heomq.siam(nk=2, ntop=3, V=0.3, U=1.5, gl=0.0),
                 lambda: O.siam_params(nk=2, ntop=3, V=0.3, U=1.5, gl=0.0)),
}

_cache = {}


def setup(name):
    if name not in _cache:
        import torch
        pf, of = CASES[name]
        m = heomq.Model(pf())
        ctx = heomq.Context(m)
        o = O.Model(of())
        _cache[name] = (m, ctx, o, torch)
    return _cache[name]


def make_mask(o, kind, seed=7):
    if kind == "full":
        return np.ones((o.N, o.d, o.d), bool)
    mask = O.physical_projection(o, np.ones((o.N, o.d, o.d), complex), hermitian=False) != 0
    if kind == "sub":
        mask &= np.random.default_rng(seed).random(mask.shape) < 0.7
        mask[0] |= np.eye(o.d, dtype=bool)
    return mask


@pytest.mark.parametrize("name", ["siam_nk2_L3", "dqd_nk2_L2"])
@pytest.mark.parametrize("kind", ["full", "physical", "sub", "built"])
def test_apply_spa_matches_oracle(name, kind):
    m, ctx, o, torch = setup(name)
    print(f"case {name} pattern {kind}", flush=True)
    pat = m.spa_pattern() if kind == "built" else O.spa_pattern_from_mask(make_mask(o, kind))
    ctx.set_spa_pattern(*pat)
    rng = np.random.default_rng(11)
    n = len(pat[2])
    xs = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = ctx.apply_spa(torch.from_numpy(xs).to("cuda:0")).cpu().numpy()
    yo = O.apply_spa(o, pat, xs)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()


def test_spa_pattern_rejects_bad_tables():
    m, ctx, o, torch = setup("siam_nk2_L3")
    nnz, ind, irow, icol = O.spa_pattern_from_mask(make_mask(o, "physical"))
    bad_row = irow.copy()
    bad_row[3] = o.d + 1
    with pytest.raises(heomq.HeomqError, match="irow"):
        ctx.set_spa_pattern(nnz, ind, bad_row, icol)
    dup = icol.copy()
    k = int(ind[0]) - 1
    dup[k + 1], bad = icol[k], irow.copy()
    bad[k + 1] = irow[k]
    with pytest.raises(heomq.HeomqError, match="twice"):
        ctx.set_spa_pattern(nnz, ind, bad, dup)
    last = int(np.nonzero(nnz)[0][-1])
    short = ind.copy()
    short[last] = len(irow) + 1  # the last stored ADO's elements start past lunk_spa
    with pytest.raises(heomq.HeomqError, match="outside"):
        ctx.set_spa_pattern(nnz, short, irow, icol)


def test_apply_spa_built_pattern_with_a_decoupled_lead():
    m, ctx, o, torch = setup("siam_gl0")
    pat = m.spa_pattern()
    assert (pat[0] == 0).any()   # the left-lead ADOs hold nothing
    ctx.set_spa_pattern(*pat)
    rng = np.random.default_rng(12)
    n = len(pat[2])
    xs = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = ctx.apply_spa(torch.from_numpy(xs).to("cuda:0")).cpu().numpy()
    yo = O.apply_spa(o, pat, xs)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()


def test_apply_spa_two_impurity_workload_against_c_port():
    # d = 16, M = 112, L = 3: the built pattern (1.8 M elements), the apply on
    # rho_spa against the C port's dense apply (no trace row) at the pattern
    from oracle.heom_port import Port
    from ref_arrays import reference_arrays
    o = O.Model(O.dqd_params())
    m = heomq.Model.from_reference(**reference_arrays(o))
    ctx = heomq.Context(m)
    pat = heomq.Model(heomq.two_impurity()).spa_pattern()
    nnz, ind, irow, icol = pat
    ctx.set_spa_pattern(*pat)
    x = O.seeded_ado_vector(o, 0x5EED0013)
    a_of = np.repeat(np.arange(o.N), nnz)
    xs = x[a_of, irow - 1, icol - 1]
    import torch
    y = ctx.apply_spa(torch.from_numpy(xs).to("cuda:0")).cpu().numpy()
    xm = np.zeros_like(x)
    xm[a_of, irow - 1, icol - 1] = xs
    yd = Port(o).apply(xm, trace_row=None)
    yo = yd[a_of, irow - 1, icol - 1]
    print(f"two-impurity sparse apply: lunk_spa = {len(xs)}", flush=True)
    assert np.linalg.norm(y[:-1] - yo) <= 1e-12 * np.linalg.norm(yo)
    assert abs(y[-1] - np.trace(xm[0])) <= 1e-12 * abs(np.trace(xm[0]))
    ctx.close()

