"""GPU parity of the time-dependent lead shift in the propagators:
calcderiv_omp(tt) and the RK4 of td/intstp.f90 with eshift(alpha, s; t) of
main/getengyshift.f90 (fieldtype 0 and 1), against the oracle."""
import dataclasses

import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "siam_nk2_L2": (lambda: heomq.siam(nk=2, ntop=2, V=0.6, U=1.5, T=0.2),
                    lambda: O.siam_params(nk=2, ntop=2, V=0.6, U=1.5, T=0.2)),
    "dqd_nk1_L2": (lambda: heomq.two_impurity(nk=1, ntop=2), lambda: O.dqd_params(nk=1, ntop=2)),
}

_cache = {}


def setup(name):
    if name not in _cache:
        import torch
        pf, of = CASES[name]
        m = heomq.Model(dataclasses.replace(pf(), steady=False))
        ctx = heomq.Context(m)
        o = O.Model(dataclasses.replace(of(), steady=False))
        _cache[name] = (m, ctx, o, torch)
    return _cache[name]


def to_dev(torch, x):
    return torch.from_numpy(O.to_fortran(x)).to("cuda:0")


def from_dev(t):
    return O.from_fortran(t.cpu().numpy())


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


FIELDS = {
    "switch_on": (0, np.array([[3.0, 3.0], [5.0, 5.0]]), None, 1e99),
    "ac_dc": (1, np.array([[2.0, 2.5], [3.0, 3.5]]), np.array([[0.05, 0.05], [-0.02, -0.02]]), 1e99),
    "off": (0, np.array([[3.0, 3.0], [5.0, 5.0]]), None, 0.5),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("fname", list(FIELDS))
def test_calcderiv_at_time(name, fname):
    m, ctx, o, torch = setup(name)
    ft, tc, dc, toff = FIELDS[fname]
    print(f"case {name} field {fname}", flush=True)
    ctx.set_field(ft, tc, dc, toff)
    try:
        x = O.seeded_ado_vector(o, 17)
        for tt in (0.0, 0.7, 2.3):
            y = from_dev(ctx.calcderiv(to_dev(torch, x), tt=tt))
            yo = O.apply(o, x, trace_row=None, gdraw=O.gdraw_at(o, O.engy_shift(o, tt, ft, tc, dc, toff)))
            assert rel(y, yo) <= 1e-12
    finally:
        ctx.set_field(-1)


@pytest.mark.parametrize("name", list(CASES))
def test_rk4_with_field(name):
    m, ctx, o, torch = setup(name)
    ft, tc, dc, toff = FIELDS["ac_dc"]
    ctx.set_field(ft, tc, dc, toff)
    try:
        x0 = O.physical_projection(o, O.seeded_ado_vector(o, 18))
        x = from_dev(ctx.rk4(to_dev(torch, x0), 0.02, 10, t0=0.3))
        xo = O.rk4_field(o, x0, 0.3, 0.02, 10, (ft, tc, dc, toff))
        assert rel(x, xo) <= 1e-12
    finally:
        ctx.set_field(-1)


def test_field_rejected_on_a_steady_model():
    ctx = heomq.Context(heomq.Model(heomq.siam(nk=1, ntop=1)))
    with pytest.raises(heomq.HeomqError, match="steady"):
        ctx.set_field(0, np.ones((2, 2)))
    ctx.close()


@pytest.mark.parametrize("name", list(CASES))
def test_calcderiv_spa_at_time(name):
    """calcderiv_spa_omp(tt): the sparse derivative over a pattern smaller
    than the physical sector, with the lead shift at tt."""
    m, ctx, o, torch = setup(name)
    ft, tc, dc, toff = FIELDS["ac_dc"]
    mask = O.physical_projection(o, np.ones((o.N, o.d, o.d), complex), hermitian=False) != 0
    mask &= np.random.default_rng(11).random(mask.shape) < 0.75
    mask[0] |= np.eye(o.d, dtype=bool)
    pat = O.spa_pattern_from_mask(mask)
    ctx.set_spa_pattern(*pat)
    rng = np.random.default_rng(12)
    xs = rng.standard_normal(len(pat[2])) + 1j * rng.standard_normal(len(pat[2]))
    ctx.set_field(ft, tc, dc, toff)
    try:
        for tt in (0.0, 1.1):
            print(f"case {name} calcderiv_spa tt {tt}", flush=True)
            y = ctx.calcderiv_spa(torch.from_numpy(xs).to("cuda:0"), tt=tt).cpu().numpy()
            yo = O.calcderiv_spa(o, pat, xs, gdraw=O.gdraw_at(o, O.engy_shift(o, tt, ft, tc, dc, toff)))
            assert rel(y, yo) <= 1e-12
    finally:
        ctx.set_field(-1)


def _sys_terms(m, o):
    """calcdhs-style terms: a gated level shift of dot 1 (egate window), an
    ac drive on the last orbital, and for two dots a switched-on hopping
    t12 (c1^+ c2 + h.c. per spin); operators from the device model,
    checked against the oracle's."""
    n = lambda orb: m.number_operator(orb, 0) + m.number_operator(orb, 1)
    terms = [(n(0), 3, (0.25, 0.2, 1.5)), (n(m.params.norbs - 1), 5, (0.05, 0.1, 2.0, 0.3))]
    if m.params.norbs == 2:
        _, A = m.operators()
        hop = sum(A[0 * 2 + s].conj().T @ A[1 * 2 + s] for s in range(2))
        terms.append((hop + hop.conj().T, 8, (0.15, 0.05)))
    no = sum(o.ams[0][s].T @ o.ams[0][s] for s in range(2))
    assert np.abs(no - n(0)).max() == 0.0
    return terms


@pytest.mark.parametrize("name", list(CASES))
def test_calcderiv_with_system_field(name):
    m, ctx, o, torch = setup(name)
    terms = _sys_terms(m, o)
    ft, tc, dc, toff = FIELDS["switch_on"]
    ctx.set_system_field(terms)
    ctx.set_field(ft, tc, dc, toff)
    try:
        x = O.seeded_ado_vector(o, 19)
        for tt in (0.0, 0.1, 0.9, 2.0):
            print(f"case {name} system field tt {tt}", flush=True)
            y = from_dev(ctx.calcderiv(to_dev(torch, x), tt=tt))
            yo = O.apply(o, x, trace_row=None, gdraw=O.gdraw_at(o, O.engy_shift(o, tt, ft, tc, dc, toff)),
                         H=O.h_at(o, tt, terms))
            assert rel(y, yo) <= 1e-12
        x0 = O.physical_projection(o, O.seeded_ado_vector(o, 20))
        xr = from_dev(ctx.rk4(to_dev(torch, x0), 0.02, 8, t0=0.1))
        xo = O.rk4_field(o, x0, 0.1, 0.02, 8, (ft, tc, dc, toff), terms)
        assert rel(xr, xo) <= 1e-12
    finally:
        ctx.set_system_field([])
        ctx.set_field(-1)


def test_system_field_rejects_non_hermitian():
    m, ctx, o, torch = setup("siam_nk2_L2")
    bad = np.zeros((m.d, m.d), complex)
    bad[0, 1] = 1.0
    with pytest.raises(heomq.HeomqError):
        ctx.set_system_field([(bad, 2, (0.1, 1.0))])
    with pytest.raises(heomq.HeomqError):
        ctx.set_system_field([(np.eye(m.d), 9, (0.1,))])


def test_propagate_writes_reference_records(tmp_path):
    """heomq.propagate: RK4 steps with the lead field, one curr.data and
    popu.data line per step in the reference's column layout (nalf = 2)."""
    from heomq.propagate import propagate
    m, ctx, o, torch = setup("dqd_nk1_L2")
    ft, tc, dc, toff = FIELDS["switch_on"]
    ctx.set_field(ft, tc, dc, toff)
    try:
        x0 = O.physical_projection(o, O.seeded_ado_vector(o, 21))
        x = to_dev(torch, x0)
        xr = x.clone()
        t_end = propagate(ctx, x, 0.1, 0.02, 3, curr=tmp_path / "curr.data", popu=tmp_path / "popu.data")
        assert abs(t_end - 0.16) < 1e-12
        curr = [list(map(float, l.split())) for l in (tmp_path / "curr.data").read_text().splitlines()]
        popu = [list(map(float, l.split())) for l in (tmp_path / "popu.data").read_text().splitlines()]
        assert len(curr) == 3 and len(popu) == 3 and len(curr[0]) == 9 and len(popu[0]) == 6
        for k in range(3):
            t = 0.1 + 0.02 * (k + 1)
            ctx.rk4(xr, 0.02, 1, t0=t - 0.02)
            cur, occ = ctx.observables(xr)
            es = O.engy_shift(o, t, ft, tc, dc, toff) * heomq.HBAR
            want_c = [t, es[0, 0], es[1, 0], cur[0].sum(), cur[1].sum(), cur.sum(), occ[:, 0].sum(),
                      occ[:, 1].sum(), occ.sum()]
            want_p = [t, es[0, 0], es[1, 0], occ[:, 0].sum(), occ[:, 1].sum(), occ.sum()]
            for got, want in ((curr[k], want_c), (popu[k], want_p)):
                assert abs(got[0] - want[0]) <= 1e-4
                for g, w in zip(got[1:], want[1:]):
                    # e15.6e3 keeps six significant digits: half a unit of the last one
                    ulp = 10.0 ** (np.floor(np.log10(abs(w))) + 1 - 6) if w != 0 else 1e-300
                    assert abs(g - w) <= 0.5 * ulp * (1 + 1e-9) + 1e-12 * abs(w)
    finally:
        ctx.set_field(-1)
