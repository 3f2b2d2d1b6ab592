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



def _spa_to_dense(o, pat, xs):
    nnz, _, irow, icol = pat
    x = np.zeros((o.N, o.d, o.d), complex)
    x[np.repeat(np.arange(o.N), nnz), irow - 1, icol - 1] = xs
    return x


@pytest.mark.parametrize("name", ["siam_nk2_L3", "dqd_nk2_L2"])
def test_solve_steady_spa_matches_oracle(name):
    # solve_tfqmr with lsparse = T on the built pattern: Jacobi scaling for
    # the diagonal H, block Jacobi for the double dot
    m, ctx, o, torch = setup(name)
    pat = m.spa_pattern()
    ctx.set_spa_pattern(*pat)
    xs, st = ctx.solve_steady_spa(tol=1e-12)
    print(name, st, flush=True)
    assert st["status"] == 0 and st["precond"] == (1 if name.startswith("siam") else 2)
    x = _spa_to_dense(o, pat, xs.cpu().numpy())
    xo, _, _ = O.steady_state(o, tol=1e-13)
    assert np.abs(x[0] - xo[0]).max() <= 1e-10
    assert np.abs(O.current(o, x) - O.current(o, xo)).max() <= 1e-9 * np.abs(O.current(o, xo)).max()


def test_solve_steady_spa_two_impurity_workload_and_norm_filter():
    # the two-impurity workload (lscale) on its built pattern: the sparse
    # solve reaches the fixture's state; then the ADO norm filter (physical
    # Frobenius norm below 1e-8 dropped, from the full state) and a solve on
    # the filtered pattern, whose RDO error against the unfiltered fixture is
    # reported and bounded
    import json
    from pathlib import Path
    import torch
    fix = json.loads((Path(__file__).resolve().parent / "golden" / "steady_dqd_nk6_L3.json").read_text())
    rho = np.array(fix["rdo_re"]) + 1j * np.array(fix["rdo_im"])
    m = heomq.Model(heomq.two_impurity(lscale=True))
    ctx = heomq.Context(m)
    pat = m.spa_pattern()
    ctx.set_spa_pattern(*pat)
    xs, st = ctx.solve_steady_spa(tol=1e-10)
    print("sparse solve, full pattern:", st, flush=True)
    r = xs[:int(pat[0][0])].cpu().numpy()
    nnz0, ir, ic = int(pat[0][0]), pat[2][:int(pat[0][0])] - 1, pat[3][:int(pat[0][0])] - 1
    r0 = np.zeros((m.d, m.d), complex)
    r0[ir, ic] = r
    assert st["status"] == 0 and np.abs(r0 - rho).max() <= 1e-10
    x, _ = ctx.solve_steady(tol=1e-10)
    # at 1e-8 no ADO that holds elements is dropped (the smallest physical ADO
    # norm of this state is above it); at 1e-5 a quarter of the elements go
    # and the RDO moves by a few 1e-2 (measured 3.2e-2): the filter's error,
    # reported against the unfiltered fixture
    for thresh, bound in ((1e-8, 1e-10), (1e-5, 1e-1)):
        fp = m.filter_pattern(x, thresh)
        kept = fp[4]
        ctx.set_spa_pattern(*fp[:4])
        xf, stf = ctx.solve_steady_spa(tol=1e-10)
        nf = int(fp[0][0])
        rf = np.zeros((m.d, m.d), complex)
        rf[fp[2][:nf] - 1, fp[3][:nf] - 1] = xf[:nf].cpu().numpy()
        err = np.abs(rf - rho).max()
        print(f"filter {thresh:.0e}: {kept.sum()} of {m.N} ADOs, {len(fp[2])} of {len(pat[2])} elements; "
              f"RDO error vs the unfiltered fixture {err:.2e}; {stf}", flush=True)
        assert stf["status"] == 0 and err <= bound
    ctx.close()
