"""The fused block-Jacobi-preconditioned apply: one launch in which every
apply workgroup (a heavy ADO, or a sector-batched run of top-tier ADOs)
applies the inverse (shift - S_K)^-1 to its own outputs before writing them.
It must give the same bits as the apply followed by the one-kernel inverse:
on small models (and against the oracle's dense per-ADO solve), at the
two-impurity workload's full size, and through a whole preconditioned
steady solve.  With the fusion on, the solver also carries TFQMR's vector
work in those launches; its states are checked against the five-kernel
iteration and the oracle's steady state.
"""
import numpy as np
import pytest

import heomq
from oracle import heom_oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "siam_nk2_L3": (lambda: heomq.siam(nk=2, ntop=3, V=0.3, U=1.5), lambda: O.siam_params(nk=2, ntop=3, V=0.3, U=1.5)),
    "dqd_nk2_L2": (lambda: heomq.two_impurity(nk=2, ntop=2), lambda: O.dqd_params(nk=2, ntop=2)),
    "dqd_nk2_L3": (lambda: heomq.two_impurity(nk=2, ntop=3), lambda: O.dqd_params(nk=2, ntop=3)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fused_precond_apply_is_the_two_kernel_result(name):
    import torch
    pf, of = CASES[name]
    ctx = heomq.Context(heomq.Model(pf()))
    o = O.Model(of())
    print(f"case {name}", flush=True)
    shift = 0.05 * O.jacobi_ejac(o)[0]
    ctx.set_precond(-1, shift)
    prev = ctx.apply_form()
    ctx.set_apply_form(2, prev[1])   # sector-batched top tier, sliced heavy rows: the fusable form
    x = O.physical_projection(o, O.seeded_ado_vector(o, 47))
    xp = ctx.pack(torch.from_numpy(O.to_fortran(x)).to("cuda:0"))
    y1 = ctx.apply_packed_precond(xp, form=1)
    y4 = ctx.apply_packed_precond(xp, form=4)
    torch.cuda.synchronize()
    assert torch.equal(y1, y4)
    y = O.from_fortran(ctx.unpack(y4).cpu().numpy())
    yo = O.block_jacobi_precondition(o, O.physical_projection(o, O.apply(o, x, trace_row="rdo")), shift)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    # a form the fused kernel does not take is refused, not approximated
    ctx.set_apply_form(0, prev[1])
    with pytest.raises(heomq.HeomqError):
        ctx.apply_packed_precond(xp, form=4)
    ctx.set_apply_form(*prev)
    ctx.close()


def test_fused_precond_two_impurity_workload():
    # d = 16, Nk = 6, L = 3, lscale: the fused apply against the two kernels
    # on a random packed vector, then the block-Jacobi-preconditioned steady
    # solve with the fusion on and off
    import torch
    ctx = heomq.Context(heomq.Model(heomq.two_impurity(lscale=True)))
    prev = ctx.apply_form()
    ctx.set_apply_form(2, prev[1])
    g = torch.Generator(device="cuda:0").manual_seed(5)
    n = ctx.packed_zeros().numel()
    xp = torch.complex(torch.randn(n, generator=g, device="cuda:0", dtype=torch.float64),
                       torch.randn(n, generator=g, device="cuda:0", dtype=torch.float64))
    y1 = ctx.apply_packed_precond(xp, form=1)
    y4 = ctx.apply_packed_precond(xp, form=4)
    torch.cuda.synchronize()
    assert torch.equal(y1, y4)
    ctx.set_precond_fusion(True)
    x_on, st_on = ctx.solve_steady(tol=1e-10)
    ctx.set_precond_fusion(False)
    x_off, st_off = ctx.solve_steady(tol=1e-10)
    ctx.set_precond_fusion(True)
    print("fused", st_on, "two kernels", st_off, flush=True)
    assert st_on["status"] == 0 and st_off["status"] == 0
    assert st_on["fused_iteration"] and not st_off["fused_iteration"]
    # the fused iteration has the five-kernel iteration's iterates in exact
    # arithmetic; its reductions are summed in another order (per operator
    # workgroup) and sigma is one recurrence step from a directly summed rt.v,
    # so the iteration counts of the two finite-precision runs differ a little
    # (152 and 162 on this system)
    assert abs(st_on["iters"] - st_off["iters"]) <= 0.15 * st_off["iters"]
    rel = ((x_on - x_off).abs().max() / x_off.abs().max()).item()
    print("relative difference of the two states", rel, flush=True)
    assert rel <= 1e-7
    ctx.set_apply_form(*prev)
    ctx.close()


@pytest.mark.parametrize("name", ["siam_nk2_L3", "dqd_nk2_L2", "dqd_nk2_L3"])
def test_fused_iteration_steady_state_against_the_oracle(name):
    # block Jacobi on the fusable form: the iteration's vector work rides in
    # the two preconditioned launches.  The state's residual through the
    # oracle's operator (one apply: the oracle's own steady solve takes
    # minutes at L = 3), and the state against the five-kernel iteration's.
    import torch
    pf, of = CASES[name]
    ctx = heomq.Context(heomq.Model(pf()))
    o = O.Model(of())
    print(f"case {name}", flush=True)
    ctx.set_precond(2)
    ctx.set_apply_form(2, ctx.apply_form()[1])
    ctx.set_precond_fusion(True)
    x, st = ctx.solve_steady(tol=1e-12)
    print(st, flush=True)
    assert st["status"] == 0 and st["precond"] == 2 and st["fused_iteration"]
    xs = O.from_fortran(x.cpu().numpy())
    e = np.zeros_like(xs)
    e[0, o.d - 1, o.d - 1] = 1.0
    r = e - O.physical_projection(o, O.apply(o, xs, trace_row="rdo"))
    assert np.abs(r).max() <= 1e-10
    ctx.set_precond_fusion(False)
    x5, st5 = ctx.solve_steady(tol=1e-12)
    assert st5["status"] == 0 and not st5["fused_iteration"]
    assert ((x - x5).abs().max() / x5.abs().max()).item() <= 1e-9
    # a second fused solve from the state (after the start's smoothing sweeps)
    ctx.set_precond_fusion(True)
    x2, st2 = ctx.solve_steady(x=x.clone(), tol=1e-12)
    assert st2["status"] == 0
    assert ((x2 - x).abs().max() / x.abs().max()).item() <= 1e-9
    ctx.close()
