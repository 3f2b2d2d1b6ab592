"""GPU parity: the HIP apply / transpose / solvers against the CPU oracle.

Every comparison runs the native library through its C ABI on cuda:0 and
checks it against oracle/heom_oracle.py on the same seeded inputs.  FP64
applies are held to 1e-12 relative, steady-state RDO and
current to 1e-10 against the oracle's own TFQMR solution.
"""
import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "siam_nk3_L3": (lambda: heomq.siam(nk=3, ntop=3, V=0.3, U=1.5), lambda: O.siam_params(nk=3, ntop=3, V=0.3, U=1.5)),
    "siam_nk2_L2_rep": (lambda: heomq.siam(nk=2, ntop=2, V=-0.4, lsdraw=True), lambda: O.siam_params(nk=2, ntop=2, V=-0.4, lsdraw=True)),
    "dqd_nk2_L2": (lambda: heomq.two_impurity(nk=2, ntop=2), lambda: O.dqd_params(nk=2, ntop=2)),
    # the reference's lscale: ADOs stored scaled by prod |cb_k|^(-1/2)
    "dqd_nk2_L2_lscale": (lambda: heomq.two_impurity(nk=2, ntop=2, lscale=True),
                          lambda: _lscale(O.dqd_params(nk=2, ntop=2))),
}


def _lscale(p):
    import dataclasses
    return dataclasses.replace(p, lscale=True)

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


def to_dev(torch, x):
    return torch.from_numpy(O.to_fortran(x)).to("cuda:0")


def from_dev(t):
    return O.from_fortran(t.cpu().numpy())


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("trace", ["last", "rdo", None])
def test_apply_fp64(name, trace):
    m, ctx, o, torch = setup(name)
    print(f"case {name} trace {trace}", flush=True)
    x = O.seeded_ado_vector(o, 11)
    tr = {"last": heomq.TRACE_LAST, "rdo": heomq.TRACE_RDO, None: heomq.TRACE_NONE}[trace]
    y = from_dev(ctx.apply(to_dev(torch, x), trace=tr))
    yo = O.apply(o, x, trace_row=trace)
    assert rel(y, yo) <= 1e-12


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("trace", ["last", "rdo", None])
def test_apply_transpose_fp64(name, trace):
    m, ctx, o, torch = setup(name)
    w = O.seeded_ado_vector(o, 12)
    tr = {"last": heomq.TRACE_LAST, "rdo": heomq.TRACE_RDO, None: heomq.TRACE_NONE}[trace]
    z = from_dev(ctx.apply_transpose(to_dev(torch, w), trace=tr))
    zo = O.apply_transpose(o, w, trace_row=trace)
    assert rel(z, zo) <= 1e-12


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("trace", ["rdo", None])
@pytest.mark.parametrize("transpose", [False, True])
def test_apply_packed(name, trace, transpose):
    """Projected operator on the symmetry-packed layout vs Pi(A(Pi x)) from the oracle."""
    m, ctx, o, torch = setup(name)
    x = O.physical_projection(o, O.seeded_ado_vector(o, 14))
    tr = heomq.TRACE_RDO if trace else heomq.TRACE_NONE
    y = from_dev(ctx.apply_packed(to_dev(torch, x), trace=tr, transpose=transpose))
    f = O.apply_transpose if transpose else O.apply
    yo = O.physical_projection(o, f(o, x, trace_row=trace))
    assert rel(y, yo) <= 1e-12
    info = ctx.packed_info()
    assert info["conserves_n"] and info["conserves_sz"]
    assert info["packed"] == int((np.abs(O.physical_projection(o, np.ones((o.N, o.d, o.d)) + 0j)) > 0).sum())


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2"])
@pytest.mark.parametrize("form", range(8))
def test_apply_packed_every_form(name, form):
    """Every form of the packed apply (grouped / sliced heavy rows, sector-
    batched / packed-order top tier, one wave per small heavy ADO) computes
    the same projected operator, forward and transposed; the wave form only
    takes layouts whose heavy ADOs hold at most 16 elements (the single
    impurity's), others fall back."""
    m, ctx, o, torch = setup(name)
    used = ctx.set_apply_form(form)
    print(f"{name}: form {form} -> {used}", flush=True)
    if name.startswith("siam"):
        assert used[0] == form
    else:
        assert used[0] == form & 3
    x = O.physical_projection(o, O.seeded_ado_vector(o, 15))
    for transpose in (False, True):
        y = from_dev(ctx.apply_packed(to_dev(torch, x), trace=heomq.TRACE_RDO, transpose=transpose))
        f = O.apply_transpose if transpose else O.apply
        yo = O.physical_projection(o, f(o, x, trace_row="rdo"))
        assert rel(y, yo) <= 1e-12
    ctx.set_apply_form(2)


@pytest.mark.parametrize("name", list(CASES))
def test_apply_fp32(name):
    m, ctx, o, torch = setup(name)
    x = O.seeded_ado_vector(o, 13)
    xd = to_dev(torch, x).to(torch.complex64)
    y = from_dev(ctx.apply_f32(xd, trace=heomq.TRACE_RDO)).astype(np.complex128)
    yo = O.apply(o, O.from_fortran(xd.cpu().numpy()).astype(np.complex128), trace_row="rdo")
    assert rel(y, yo) <= 2e-6


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2", "dqd_nk2_L2_lscale"])
def test_steady_state_fp64(name):
    m, ctx, o, torch = setup(name)
    xs, st = ctx.solve_steady(tol=1e-12, raise_on_fail=False)
    print(name, "device", st, flush=True)
    assert st["status"] == 0
    xs = from_dev(xs)
    xo, it, relres = O.steady_state(o, tol=1e-12)
    print(name, st, it, relres, flush=True)
    assert np.abs(xs[0] - xo[0]).max() <= 1e-10
    cur, occ = ctx.observables(to_dev(torch, xs))
    assert np.abs(cur - O.current(o, xo)).max() <= 1e-10 * max(1.0, np.abs(O.current(o, xo)).max())
    assert np.abs(occ - O.occupation(o, xo)).max() <= 1e-10


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2"])
def test_autotuned_apply_matches_oracle(name):
    # every form the autotuner may keep computes the same operator
    m, ctx, o, torch = setup(name)
    res = ctx.autotune(reps=5)
    print(name, res, flush=True)
    assert 0 <= res["forward"] <= 7 and 0 <= res["transposed"] <= 7
    if not name.startswith("siam"):   # the wave form takes small sectors only
        assert res["forward"] < 4 and res["transposed"] < 4
    timed = [t for t in res["times_us"]["forward"] if t is not None]
    assert len(timed) == (8 if name.startswith("siam") else 4) and min(timed) > 0
    x = O.physical_projection(o, O.seeded_ado_vector(o, 17))
    for tr in (False, True):
        y = from_dev(ctx.apply_packed(to_dev(torch, x), trace=heomq.TRACE_RDO, transpose=tr))
        f = O.apply_transpose if tr else O.apply
        assert rel(y, O.physical_projection(o, f(o, x, trace_row="rdo"))) <= 1e-12


def test_steady_state_mixed():
    m, ctx, o, torch = setup("siam_nk3_L3")
    x64, _ = ctx.solve_steady(tol=1e-12)
    xm, st = ctx.solve_steady_mixed(tol=1e-12, inner_tol=1e-5)
    print(st, flush=True)
    assert st["outer"] >= 2 and st["relres"] <= 1e-12
    assert (xm[0] - x64[0]).abs().max().item() <= 1e-10


def test_dot_product_transpose_device():
    m, ctx, o, torch = setup("dqd_nk2_L2")
    x = to_dev(torch, O.seeded_ado_vector(o, 21))
    w = to_dev(torch, O.seeded_ado_vector(o, 22))
    lhs = torch.vdot(w.flatten(), ctx.apply(x, trace=heomq.TRACE_LAST).flatten()).real.item()
    rhs = torch.vdot(ctx.apply_transpose(w, trace=heomq.TRACE_LAST).flatten(), x.flatten()).real.item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2"])
def test_rk4_device(name):
    """td/intstp.f90 RK4 on the device vs the oracle's RK4 (same steps)."""
    m, ctx, o, torch = setup(name)
    x = O.seeded_ado_vector(o, 41)
    xd = ctx.rk4(to_dev(torch, x), 0.02, 3)
    assert rel(from_dev(xd), O.rk4(o, x, 0.02, 3)) <= 1e-12


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2"])
def test_chebyshev_device(name):
    """td/chebyshev.f90 step on the device vs the oracle's recurrence."""
    m, ctx, o, torch = setup(name)
    x = O.seeded_ado_vector(o, 42)
    xd, info = ctx.chebyshev(to_dev(torch, x), 0.02, 2)
    xo, ncheby, delta = O.chebyshev(o, x, 0.02, 2)
    assert info["ncheby"] == ncheby and abs(info["delta"] - delta) <= 1e-12 * delta
    assert rel(from_dev(xd), xo) <= 1e-12


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2", "dqd_nk2_L2_lscale"])
def test_rk4_in_sector_takes_the_packed_layout(name):
    """A state with nothing outside the symmetry sectors is propagated on the
    packed layout (heomq_rk4_at routes it; heomq_rk4_packed takes it packed):
    both equal the oracle's dense RK4 of td/intstp.f90."""
    m, ctx, o, torch = setup(name)
    x = O.physical_projection(o, O.seeded_ado_vector(o, 43), hermitian=False)
    xo = O.rk4(o, x, 0.02, 3)
    assert rel(from_dev(ctx.rk4(to_dev(torch, x), 0.02, 3)), xo) <= 1e-12
    xp = ctx.pack(to_dev(torch, x))
    ctx.rk4_packed(xp, 0.02, 3)
    assert rel(from_dev(ctx.unpack(xp)), xo) <= 1e-12


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2", "dqd_nk2_L2_lscale"])
def test_chebyshev_in_sector_takes_the_packed_layout(name):
    """td/chebyshev.f90 steps on an in-sector state: the dense entry point
    reports the packed layout, and it and the packed entry point equal the
    oracle's recurrence; a state with an off-sector element stays dense."""
    m, ctx, o, torch = setup(name)
    x = O.physical_projection(o, O.seeded_ado_vector(o, 44))
    xo, ncheby, delta = O.chebyshev(o, x, 0.02, 2)
    xd, info = ctx.chebyshev(to_dev(torch, x), 0.02, 2)
    assert info["packed"] and info["ncheby"] == ncheby
    assert rel(from_dev(xd), xo) <= 1e-12
    xp, info = ctx.chebyshev_packed(ctx.pack(to_dev(torch, x)), 0.02, 2)
    assert rel(from_dev(ctx.unpack(xp)), xo) <= 1e-12
    y = O.seeded_ado_vector(o, 45)
    _, info = ctx.chebyshev(to_dev(torch, y), 0.02, 1)
    assert not info["packed"]


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2"])
def test_chebyshev_fused_term_is_the_two_pass_result(name):
    """Each Chebyshev term k >= 2 runs as one launch (the apply's output folded
    into P_k = P_{k-2} + (2/delta) L P_{k-1} and x += 2 J_k P_k) by default:
    the same bits as the apply followed by the two vector passes."""
    m, ctx, o, torch = setup(name)
    x = O.physical_projection(o, O.seeded_ado_vector(o, 46))
    xp = ctx.pack(to_dev(torch, x))
    a = xp.clone()
    ctx.chebyshev_packed(a, 0.02, 3)
    ctx.set_precond_fusion(False)
    try:
        b = xp.clone()
        ctx.chebyshev_packed(b, 0.02, 3)
    finally:
        ctx.set_precond_fusion(True)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    xo, _, _ = O.chebyshev(o, x, 0.02, 3)
    assert rel(from_dev(ctx.unpack(a)), xo) <= 1e-12


def test_packed_propagation_conserves_trace_and_hermiticity():
    """Size-independent properties of the propagation at the two-impurity
    workload's full size (d=16, Nk=6, L=3, 117,321 ADOs) from a seeded random
    tier-0 density matrix with (N, S_z) block structure: tr rho0 stays 1 and
    rho0 Hermitian over 20 Chebyshev steps and 20 RK4 steps."""
    ctx = heomq.Context(heomq.Model(heomq.two_impurity(lscale=True)))
    x = ctx.random_state(7)
    r0 = x[0].cpu().numpy().T
    assert abs(np.trace(r0) - 1.0) <= 1e-14 and np.linalg.eigvalsh(r0).min() > 0
    dt = 0.02
    for step in ("chebyshev", "rk4"):
        y = x.clone()
        if step == "chebyshev":
            _, info = ctx.chebyshev(y, dt, 20)
            assert info["packed"]
        else:
            ctx.rk4(y, dt, 20)
        rho = y[0].cpu().numpy().T
        assert abs(np.trace(rho) - 1.0) <= 1e-12
        assert np.abs(rho - rho.conj().T).max() <= 1e-12 * np.abs(rho).max()
        assert np.abs(rho - r0).max() > 1e-6   # it moved
    ctx.close()


def test_graph_replays_match_eager():
    """The solver's captured chunk graphs, replayed many times with host
    synchronisations between replays, give the same iterates as launching the
    same kernels eagerly; a second solve (cached graph) and a transposed solve
    on the same context agree too."""
    m, ctx, o, torch = setup("siam_nk3_L3")
    ctx.set_graphs(False)
    xe, se = ctx.solve_steady(tol=1e-12)
    be = O.seeded_ado_vector(o, 61)
    be[0] = 0
    le, _ = ctx.solve_steady(x=ctx.zeros(), b=to_dev(torch, be), tol=1e-12, transpose=True)
    ctx.set_graphs(True)
    for rep in range(2):
        xg, sg = ctx.solve_steady(tol=1e-12)
        print("eager", se, "graph", sg, flush=True)
        assert sg["iters"] == se["iters"]
        assert (xg - xe).abs().max().item() <= 1e-14
        lg, _ = ctx.solve_steady(x=ctx.zeros(), b=to_dev(torch, be), tol=1e-12, transpose=True)
        assert (lg - le).abs().max().item() <= 1e-14 * max(le.abs().max().item(), 1.0)


@pytest.mark.parametrize("name", list(CASES))
def test_bicg_convention_applies(name):
    """calc_ax_bicg_omp / calc_atrx_bicg_omp conventions on the device vs the oracle."""
    m, ctx, o, torch = setup(name)
    x = O.seeded_ado_vector(o, 71)
    x[0] += np.triu(np.full((o.d, o.d), 0.2 - 0.5j), 1)
    xd = to_dev(torch, x)
    y = from_dev(ctx.apply_bicg(xd))
    assert rel(y, O.apply_bicg(o, x)) <= 1e-12
    xh = from_dev(xd)                                  # input rebuilt in place from its lower triangle
    assert np.abs(xh[0] - O._herm_from_lower(x[0])).max() <= 1e-15
    w = O.seeded_ado_vector(o, 72)
    z = from_dev(ctx.apply_transpose_bicg(to_dev(torch, w)))
    assert rel(z, O.apply_transpose_bicg(o, w)) <= 1e-12


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2"])
def test_bicg_steady_state(name):
    """solvebicg on the device reaches the unique steady state (the oracle's
    TFQMR solution; the oracle's own BiCG is pinned to it on the CPU)."""
    m, ctx, o, torch = setup(name)
    x, st = ctx.solve_bicg(crit=1e-10, maxiter=5000)
    print(name, st, flush=True)
    assert st["status"] == 0 and st["resid"] <= 1e-10
    assert st["packed"]   # the identity guess is in the sectors: the packed iteration
    xb = from_dev(x)
    assert np.abs(xb[0] - xb[0].conj().T).max() == 0.0
    xt, _, _ = O.steady_state(o, tol=1e-13)
    assert np.abs(xb[0] - xt[0]).max() <= 1e-9
    # a guess with an element outside the sectors iterates on the dense layout
    # and reaches the same state
    g = ctx.initial_guess()
    g[1, 0, 0] = 1e-3   # element (1,1) of the first tier-1 ADO, whose sector excludes it
    assert not ctx.unpack(ctx.pack(g)).equal(g)
    x2, st2 = ctx.solve_bicg(g, crit=1e-10, maxiter=5000)
    assert st2["status"] == 0 and not st2["packed"]
    assert np.abs(from_dev(x2)[0] - xt[0]).max() <= 1e-9


@pytest.mark.parametrize("name", ["siam_nk3_L3", "dqd_nk2_L2"])
def test_jacobi_scaling_on_and_off(name):
    """The steady state does not depend on the iteration's diagonal scaling
    (automatic for diagonal H; forced on for the two-impurity H with hopping)."""
    m, ctx, o, torch = setup(name)
    xo, _, _ = O.steady_state(o, tol=1e-13)
    its = {}
    for mode in (0, 1):
        ctx.set_precond(mode)
        x, st = ctx.solve_steady(tol=1e-12)
        its[mode] = st["iters"]
        assert st["status"] == 0
        assert np.abs(from_dev(x)[0] - xo[0]).max() <= 1e-10
    ctx.set_precond(-1)
    print(name, "iterations plain / scaled:", its, flush=True)
