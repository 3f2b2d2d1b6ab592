"""Pins of the CPU oracle (oracle/heom_oracle.py) by closed forms, brute force and invariants.

The reference (Fortran + Intel MKL) cannot be built in this image and holds no
golden vectors, so each oracle function is checked against mathematics a
plausible mistake would break:

* Pade decomposition       -> the Fermi function itself; pole count/positivity.
* fermion operators        -> canonical anticommutation relations.
* hierarchy enumeration    -> brute-force counts; every tuple stored or its
                              Hermitian partner stored, never both for odd tiers.
* coupling signs           -> permutation parity recomputed by brute force.
* apply                    -> trace preservation, Hermiticity of (Lx)_0,
                              dot-product test against the transpose, dense
                              brute-force transpose on a tiny model.
* steady state (U = 0)     -> Landauer closed form with the same Lorentzian
                              band and the same Pade Fermi function (the
                              noninteracting HEOM is exact from tier 2 on).
* steady state (U > 0)     -> tr rho = 1, Hermitian, I_L = -I_R, I(V=0) = 0,
                              <n> = 1 at particle-hole symmetry.
* two-impurity ($para4)    -> U = 0: Landauer with the 2 x 2 retarded Green's
                              function (symmetric and asymmetric dots, leads
                              and spins); H's spectrum in closed form; <N> = 2
                              at the particle-hole point eps = -U/2 - u12.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import heom_oracle as O


# ---------------------------------------------------------------- Pade / bath
@pytest.mark.parametrize("n", [1, 2, 4, 8, 12])
def test_pade_matches_fermi(n):
    eta, xi = O.pade_psd(n)
    assert np.all(eta > 0) and np.all(np.diff(xi) > 0)
    x = np.linspace(-2.0 * n, 2.0 * n, 401)
    f = 1.0 / (1.0 + np.exp(x))
    err = np.abs(O.fermi_pade(x, eta, xi).real - f).max()
    # [N-1/N] Pade is accurate well inside |x| < 2N; bound tightens with N
    assert err < {1: 0.2, 2: 0.05, 4: 2e-3, 8: 1e-5, 12: 1e-7}[n]


def test_pade_first_pole_near_matsubara():
    # the lowest Pade pole converges to the lowest Matsubara frequency pi
    eta, xi = O.pade_psd(10)
    assert abs(xi[0] - math.pi) < 1e-10 and abs(eta[0] - 1.0) < 1e-10


def test_bath_coefficients_reproduce_correlation():
    # C(t) = sum_k cb_k exp(gamma_k t) against direct quadrature of
    # (1/pi) int J(w) f_pade(w/T) e^{i w t} dw / hbar^2 (Lorentzian J)
    p = O.siam_params(nk=6, T=0.3, W=4.0, gamma=0.7)
    m = O.Model(p)
    eta, xi = m.pade
    W, T, G = 4.0, 0.3, 0.7
    w = np.linspace(-2000, 2000, 4_000_001)
    J = 0.5 * G * W ** 2 / (w ** 2 + W ** 2)
    f = O.fermi_pade(w / T, eta, xi).real
    for t in (0.05, 0.3, 1.0):
        # the Pade Fermi function tends to 1/2 at large |w|; that part is done in closed form
        tail = 0.25 * G * W * np.exp(-W * t / O.HBAR)
        ref = (np.trapezoid(J * (f - 0.5) * np.exp(1j * w * t / O.HBAR), w) / np.pi + tail) / O.HBAR ** 2
        sel = (m.sgn == 0) & (m.alf == 0) & (m.spin == 0)
        ours = np.sum(m.cb[sel] * np.exp(m.cgamma[sel] * t))
        assert abs(ours - ref) < 2e-5 * abs(ref)


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize("norbs", [1, 2, 3])
def test_anticommutators(norbs):
    ams = O.annihilators(norbs)
    ops = [ams[m][s] for m in range(norbs) for s in range(2)]
    d = ops[0].shape[0]
    for i, a in enumerate(ops):
        for j, b in enumerate(ops):
            assert np.abs(a @ b + b @ a).max() == 0
            assert np.abs(a @ b.T + b.T @ a - (i == j) * np.eye(d)).max() == 0


def test_siam_operator_convention():
    # c_up = |0><u| + |d><2|, c_down = |0><d| - |u><2|  (calcams.f90 + swapspin.f90)
    ams = O.annihilators(1)
    cu, cd = ams[0]
    assert cu[0, 1] == 1 and cu[2, 3] == 1 and np.count_nonzero(cu) == 2
    assert cd[0, 2] == 1 and cd[1, 3] == -1 and np.count_nonzero(cd) == 2


# ---------------------------------------------------------------- hierarchy
@pytest.mark.parametrize("nk,ntop,lsdraw", [(1, 3, False), (2, 3, False), (1, 3, True), (2, 2, True)])
def test_hierarchy_counts_bruteforce(nk, ntop, lsdraw):
    m = O.Model(O.siam_params(nk=nk, ntop=ntop, lsdraw=lsdraw))
    gen = itertools.combinations_with_replacement if lsdraw else itertools.combinations
    for n in range(1, ntop + 1):
        stored = set(map(tuple, m.tiers[n]))
        allk = set(gen(range(m.M), n))
        comp = {tuple(sorted(m.bar[k] for k in K)) for K in stored}
        assert stored | comp == allk                     # every ADO reachable
        if n % 2 == 1:
            assert not (stored & comp)                    # odd tiers: exactly one of each pair
            assert 2 * len(stored) == len(allk)
        if not lsdraw:
            assert len(allk) == math.comb(m.M, n)


@pytest.mark.parametrize("p", [O.siam_params(nk=2, ntop=3), O.dqd_params(nk=1, ntop=3)], ids=["siam", "dqd"])
def test_coupling_signs_bruteforce(p):
    m = O.Model(p)

    def perm_sign(seq):
        s = 1
        seq = list(seq)
        for i in range(len(seq)):
            for j in range(i + 1, len(seq)):
                if seq[i] > seq[j]:
                    s = -s
        return s

    for n in range(0, m.L):
        up, upf = m.upper[n]
        for a, K in enumerate(m.tiers[n][:40]):
            K = tuple(K)
            for e in range(m.M):
                nb = up[a, e, 0]
                if e in K:
                    assert nb == -1
                    continue
                target = tuple(sorted(K + (e,)))
                _, T = m.tuple_of(nb)
                if up[a, e, 1] == 0:
                    assert T == target and upf[a, e, 0] == perm_sign((e,) + K)
                else:
                    # stored partner: reversed, sign-flipped, sorted with parity
                    rev = [m.bar[k] for k in reversed(target)]
                    assert tuple(sorted(rev)) == T
                    assert upf[a, e, 0] == perm_sign((e,) + K) * perm_sign(rev)
                    assert upf[a, e, 1] == perm_sign(K + (e,)) * perm_sign(rev)


# ---------------------------------------------------------------- apply
SMALL = [O.siam_params(nk=2, ntop=3, V=0.3, U=1.5), O.dqd_params(nk=1, ntop=2)]


@pytest.mark.parametrize("p", SMALL)
def test_trace_preservation_and_hermiticity(p):
    m = O.Model(p)
    x = O.seeded_ado_vector(m, 3)
    y = O.apply(m, x, trace_row=None)
    assert abs(np.trace(y[0])) < 1e-12 * np.abs(y).max()
    assert np.abs(y[0] - y[0].conj().T).max() < 1e-12 * np.abs(y).max()


@pytest.mark.parametrize("p", SMALL)
@pytest.mark.parametrize("trace", [None, "last", "rdo"])
def test_dot_product_transpose(p, trace):
    m = O.Model(p)
    x = O.seeded_ado_vector(m, 4)
    w = O.seeded_ado_vector(m, 5)
    lhs = np.real(np.vdot(w, O.apply(m, x, trace_row=trace)))
    rhs = np.real(np.vdot(O.apply_transpose(m, w, trace_row=trace), x))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_dense_transpose_bruteforce():
    m = O.Model(O.siam_params(nk=1, ntop=2, V=0.2))
    A = O.dense_matrix(m, trace_row="last")
    AT = O.dense_matrix(m, trace_row="last", transpose=True)
    assert np.abs(A.T - AT).max() < 1e-13 * np.abs(A).max()


def test_apply_linear_and_real_linear():
    m = O.Model(SMALL[0])
    x = O.seeded_ado_vector(m, 6)
    z = O.seeded_ado_vector(m, 7)
    y = O.apply(m, 2.5 * x - 0.5 * z)
    assert np.abs(y - (2.5 * O.apply(m, x) - 0.5 * O.apply(m, z))).max() < 1e-11 * np.abs(y).max()
    # not complex-linear: the Hermitian-partner terms conjugate their input
    yi = O.apply(m, 1j * x, trace_row=None)
    assert np.abs(yi - 1j * O.apply(m, x, trace_row=None)).max() > 1e-3


# ---------------------------------------------------------------- steady state
@pytest.mark.parametrize("ntop", [2, 3])
def test_noninteracting_landauer(ntop):
    p = O.siam_params(nk=2, ntop=ntop, U=0.0, eps=-0.3, V=0.5, gamma=0.3, T=0.2)
    m = O.Model(p)
    x, it, rel = O.steady_state(m, tol=1e-12)
    assert rel <= 1e-12
    i_l, n_occ = O.landauer_siam(p)
    I = O.current(m, x)
    assert np.abs(I[0] - i_l).max() <= 1e-9 * abs(i_l)
    assert np.abs(I[1] + i_l).max() <= 1e-9 * abs(i_l)
    assert np.abs(O.occupation(m, x) - n_occ).max() <= 1e-9


def test_landauer_zero_bias():
    p = O.siam_params(nk=2, ntop=2, U=0.0, eps=-0.3, V=0.0, gamma=0.3, T=0.2)
    i_l, _ = O.landauer_siam(p)
    assert abs(i_l) < 1e-9


def test_interacting_invariants():
    p = O.siam_params(nk=2, ntop=3, U=2.0, eps=-1.0, V=0.0, gamma=0.2, T=0.1)
    m = O.Model(p)
    x, it, rel = O.steady_state(m, tol=1e-12)
    assert abs(np.trace(x[0]) - 1) < 1e-11
    assert np.abs(x[0] - x[0].conj().T).max() < 1e-12
    assert np.abs(O.current(m, x)).max() < 1e-7                  # V = 0
    assert abs(O.occupation(m, x).sum() - 1.0) < 1e-10           # particle-hole symmetric
    p2 = O.siam_params(nk=2, ntop=3, U=2.0, eps=-0.7, V=0.4, gamma=0.2, T=0.1)
    m2 = O.Model(p2)
    x2, _, _ = O.steady_state(m2, tol=1e-12)
    I = O.current(m2, x2)
    assert np.abs(I[0] + I[1]).max() < 1e-8 * np.abs(I).max()  # current conservation


def _dqd_asym(nk=1, ntop=2):
    # every orbital, spin and lead distinct: e1 != e2, up != down, Gamma per
    # (lead, orbital) all different, an asymmetric bias split
    return O.Params(norbs=2, nspin=2, nalf=2, nmats=nk, bandwidth=(8.0, 12.0),
                    linewidth=((0.3, 0.12), (0.18, 0.25)), temperature=(0.2, 0.15),
                    engyshift=((0.35, 0.3), (-0.2, -0.25)), ham="dqd",
                    hpar=dict(e1up=-0.3, e1down=-0.1, e2up=0.2, e2down=0.05, uu1=0.0, uu2=0.0,
                              u12=0.0, t12=0.4), ntop=ntop)


@pytest.mark.parametrize("case", ["sym_t", "sym_t0", "asym", "asym_nk2"])
def test_two_impurity_noninteracting_landauer(case):
    # the noninteracting HEOM is exact from tier 2 on, for any number of
    # orbitals: currents and occupations equal the 2 x 2 Green's-function result
    if case == "sym_t":
        p = O.dqd_params(nk=1, ntop=2, U=0.0, eps=-0.3, t=0.4, V=0.5, gamma=0.3, T=0.2)
    elif case == "sym_t0":
        p = O.dqd_params(nk=1, ntop=2, U=0.0, eps=-0.3, t=0.0, V=0.5, gamma=0.3, T=0.2)
    else:
        p = _dqd_asym(nk=2 if case == "asym_nk2" else 1)
    m = O.Model(p)
    x, it, rel = O.steady_state(m, tol=1e-12)
    assert rel <= 1e-12
    I, n = O.landauer_dqd(p)
    Io = O.current(m, x)
    assert np.abs(Io - I).max() <= 1e-9 * np.abs(I).max()
    assert np.abs(O.occupation(m, x) - n).max() <= 1e-9


def test_two_impurity_hamiltonian_spectrum_closed_form():
    # equal levels eps, on-dot U, inter-dot u12, hopping t: the Fock spectrum
    # sector by sector (N = 0..4), from second quantisation by hand
    eps, U, u12, t = -0.7, 2.3, 0.6, 0.45
    p = O.dqd_params(eps=eps, U=U, t=t, u12=u12)
    m = O.Model(p)
    ev = np.sort(np.linalg.eigvalsh(m.H * O.HBAR))
    r = math.sqrt(((U - u12) / 2) ** 2 + 4 * t * t)
    ref = [0.0]
    ref += [eps - t, eps + t] * 2                                        # N = 1
    ref += [2 * eps + u12] * 3 + [2 * eps + U]                           # N = 2: triplet, odd singlet
    ref += [2 * eps + (U + u12) / 2 - r, 2 * eps + (U + u12) / 2 + r]    # N = 2: even singlets
    ref += [3 * eps + U + 2 * u12 - t, 3 * eps + U + 2 * u12 + t] * 2    # N = 3
    ref += [4 * eps + 2 * U + 4 * u12]                                   # N = 4
    assert np.abs(ev - np.sort(ref)).max() < 1e-12


def test_two_impurity_particle_hole_point():
    # eps = -U/2 - u12 with symmetric leads and bias: <N> = 2 exactly, both
    # dots and spins equal, I_L = -I_R
    U, u12 = 2.0, 0.5
    p = O.dqd_params(nk=1, ntop=2, U=U, u12=u12, eps=-U / 2 - u12, t=0.3, V=0.4, gamma=0.2, T=0.2)
    m = O.Model(p)
    x, _, rel = O.steady_state(m, tol=1e-12)
    n = O.occupation(m, x)
    assert abs(n.sum() - 2.0) < 1e-10
    assert np.abs(n - n[0, 0]).max() < 1e-10
    I = O.current(m, x)
    assert np.abs(I[0] + I[1]).max() < 1e-8 * np.abs(I).max()
    p2 = O.dqd_params(nk=1, ntop=2, U=U, u12=u12, eps=-U / 2, t=0.3, V=0.4, gamma=0.2, T=0.2)
    m2 = O.Model(p2)
    x2, _, _ = O.steady_state(m2, tol=1e-12)
    assert O.occupation(m2, x2).sum() < 2.0 - 1e-3          # off the point it moves


def _scaled(p):
    import dataclasses
    return dataclasses.replace(p, lscale=True)


@pytest.mark.parametrize("p", [O.siam_params(nk=2, ntop=3, V=0.3, U=1.5), O.dqd_params(nk=1, ntop=2)],
                         ids=["siam", "dqd"])
def test_lscale_is_the_similarity_of_the_ado_scaling(p):
    # with lscale the operator acting on stored ADOs rho_K / s_K is exactly
    # S^-1 L S (S = diag(s_K), s_K = prod |cb_k|^(1/2)): coefficients, not the
    # vector, carry the scaling, so this pins the rescaled cb, cd and the
    # sqrt|cb| factor on every upper coupling
    m, ms = O.Model(p), O.Model(_scaled(p))
    s = O.ado_scale(ms)[:, None, None]
    assert np.abs(s - 1).max() > 0.1
    x = O.seeded_ado_vector(m, 4)
    for tr in (None, "rdo"):
        y = O.apply(m, x, trace_row=tr) / s
        ys = O.apply(ms, x / s, trace_row=tr)
        assert np.abs(ys - y).max() <= 1e-13 * np.abs(y).max()
        z = O.apply_transpose(m, x, trace_row=tr) * s
        zs = O.apply_transpose(ms, x * s, trace_row=tr)
        assert np.abs(zs - z).max() <= 1e-13 * np.abs(z).max()


def test_lscale_steady_state_is_the_unscaled_one():
    p = O.dqd_params(nk=1, ntop=2, U=2.0, t=0.3, V=0.3, T=0.2)
    m, ms = O.Model(p), O.Model(_scaled(p))
    x, _, _ = O.steady_state(m, tol=1e-12)
    xs, _, _ = O.steady_state(ms, tol=1e-12)
    assert np.abs(xs[0] - x[0]).max() <= 1e-10
    assert np.abs(xs * O.ado_scale(ms)[:, None, None] - x).max() <= 1e-9
    assert np.abs(O.current(ms, xs) - O.current(m, x)).max() <= 1e-10 * np.abs(O.current(m, x)).max()


def test_trace_row_placements_agree_for_odd_top_tier():
    m = O.Model(O.siam_params(nk=2, ntop=3, V=0.3, U=1.5))
    xr, _, _ = O.steady_state(m, tol=1e-12, trace_row="rdo")
    xl, _, _ = O.steady_state(m, tol=1e-12, trace_row="last")
    assert np.abs(xr[0] - xl[0]).max() < 1e-10


def _tiny():
    return O.Model(O.siam_params(nk=1, ntop=2, V=0.3, U=1.5, T=0.2))


def test_rk4_is_fourth_order_taylor_polynomial():
    # for a linear ODE one RK4 step is exactly sum_{k<=4} (h A)^k / k! applied to x
    o = _tiny()
    A = O.dense_matrix(o, trace_row=None)
    x = O.seeded_ado_vector(o, 3)
    h = 0.05
    v = O.as_real(x)
    ref, term = v.copy(), v.copy()
    for k in range(1, 5):
        term = h * (A @ term) / k
        ref = ref + term
    got = O.as_real(O.rk4(o, x, h, 1))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_chebyshev_matches_matrix_exponential():
    # the Jacobi-Anger series of exp(L dt) is exact for any spectrum.  The
    # damping puts L's spectrum on the real axis, where the Chebyshev
    # polynomials of iL/delta grow like (1 + sqrt 2)^k, so the 1e-21 Bessel-tail
    # cut leaves round-off only for delta*dt of order one (the reference's
    # recommended step); at delta*dt = 15 the same cut is visibly short.
    from scipy.linalg import expm
    o = _tiny()
    A = O.dense_matrix(o, trace_row=None)
    x = O.seeded_ado_vector(o, 4)
    for dt in (0.01, 0.05):
        ref = expm(A * dt) @ O.as_real(x)
        got, ncheby, delta = O.chebyshev(o, x, dt, 2)
        ref = expm(A * dt) @ ref
        assert ncheby > delta * dt
        assert np.abs(O.as_real(got) - ref).max() <= 1e-13 * np.abs(ref).max()
        # trace of the RDO is conserved by the exact propagator
        assert abs(np.trace(got[0]) - np.trace(x[0])) <= 1e-13
    got, _, _ = O.chebyshev(o, x, 0.5, 1)
    assert np.abs(O.as_real(got) - expm(A * 0.5) @ O.as_real(x)).max() > 1e-6


def test_cheby_delta_is_width_of_real_diagonal():
    # brute force over the dense matrix's diagonal (real part of L's diagonal)
    o = _tiny()
    A = O.dense_matrix(o, trace_row=None)
    dg = np.diag(A)
    assert abs(O.cheby_delta(o) - (dg.max() - min(dg.min(), 0.0))) <= 1e-12 * abs(dg).max()


def test_bessel_weights_sum_rule():
    # J_0(z) + 2 sum_k J_2k(z) = 1 (generating function at theta = 0)
    J = O.cheby_terms(7.3)
    assert abs(J[0] + 2 * J[2::2].sum() - 1.0) <= 1e-14


def _siam_theta(th, **kw):
    return O.siam_params(eps=th[0], U=th[1], gl=th[2], gr=th[3], V=th[4], **kw)


def test_parameters_enter_affinely():
    # the adjoint's exact derivative rule dA x = A(theta + e_i) x - A(theta) x
    # needs A affine in each of (eps, U, Gamma_L, Gamma_R, V)
    th = np.array([-0.7, 1.5, 0.1, 0.15, 0.3])
    pf = lambda t: _siam_theta(t, nk=1, ntop=2, T=0.2)
    m0 = O.Model(pf(th))
    x = O.seeded_ado_vector(m0, 5)
    a0 = O.apply(m0, x, trace_row=None)
    for i in range(5):
        t1, t2 = th.copy(), th.copy()
        t1[i] += 1.0
        t2[i] += 2.5
        d1 = O.apply(O.Model(pf(t1)), x, trace_row=None) - a0
        d2 = O.apply(O.Model(pf(t2)), x, trace_row=None) - a0
        assert np.abs(d2 - 2.5 * d1).max() <= 1e-12 * max(np.abs(a0).max(), 1.0)


def test_current_weights_reproduce_current():
    m = O.Model(O.siam_params(nk=1, ntop=2, V=0.3))
    g = O.current_weights(m, 0)
    x = O.seeded_ado_vector(m, 6)
    assert abs(np.real(np.vdot(g, x)) - O.current(m, x)[0].sum()) <= 1e-9 * np.abs(g).max()


def test_adjoint_gradient_matches_finite_differences():
    th = np.array([-0.7, 1.5, 0.1, 0.15, 0.3])
    pf = lambda t: _siam_theta(t, nk=1, ntop=2, T=0.2)
    I, ga = O.current_gradient_adjoint(pf, th)
    gf = O.current_gradient_fd(pf, th, h=1e-5)
    assert np.abs(ga - gf).max() <= 1e-6 * np.abs(gf).max()


def test_bicg_transpose_is_exact_transpose():
    # brute force: the dense matrices of the BiCG-convention apply and of its
    # transpose are transposes of each other
    o = O.Model(O.siam_params(nk=1, ntop=1, V=0.3, U=1.5))
    n = o.N * o.d * o.d * 2
    A = np.zeros((n, n))
    AT = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = O.as_real(O.apply_bicg(o, O.from_real(e, o)))
        AT[:, j] = O.as_real(O.apply_transpose_bicg(o, O.from_real(e, o)))
    assert np.abs(A.T - AT).max() <= 1e-12 * np.abs(A).max()


def test_bicg_apply_agrees_with_tfqmr_apply_on_physical_states():
    # for a Hermitian rho0 the two conventions differ only in where the trace
    # row sits and in which tier-0 triangle is kept
    o = O.Model(O.siam_params(nk=2, ntop=2, V=0.3, U=1.5))
    x = O.seeded_ado_vector(o, 8)
    yb = O.apply_bicg(o, x)
    yn = O.apply(o, x, trace_row=None)
    assert np.abs(yb[1:] - yn[1:]).max() <= 1e-13 * np.abs(yn).max()
    low = np.tril(np.ones((o.d, o.d), bool), -1)
    assert np.abs(yb[0][low] - yn[0][low]).max() <= 1e-13 * np.abs(yn).max()
    assert abs(yb[0, 0, 0] - np.trace(x[0])) <= 1e-14


def test_bicg_steady_state_matches_landauer():
    # U = 0 closed form through the BiCG path (solvebicg.f90 convention)
    p = O.siam_params(nk=3, ntop=2, V=0.4, U=0.0, eps=-0.3, T=0.2)
    o = O.Model(p)
    x, it, err = O.bicg(o, tol=1e-12, maxiter=3000)
    assert err <= 1e-12
    xt, _, _ = O.steady_state(o, tol=1e-13)
    assert np.abs(x[0] - xt[0]).max() <= 1e-9
    IL, n = O.landauer_siam(p)                  # per spin
    assert np.abs(O.current(o, x)[0] - IL).max() <= 1e-8 * abs(IL)


@pytest.mark.parametrize("p", [O.siam_params(nk=2, ntop=3, V=0.3, U=1.5), O.dqd_params(nk=1, ntop=2),
                               O.siam_params(nk=2, ntop=2, lsdraw=True)])
def test_c_port_matches_numpy_oracle(p):
    # the OpenMP C port of calc_ax_tfqmr_omp (the CPU baseline) on the
    # reference's own table layout agrees with the pinned numpy oracle
    from oracle import heom_port as P
    o = O.Model(p)
    port = P.Port(o)
    x = O.seeded_ado_vector(o, 3)
    for tr in ("last", "rdo", None):
        y, yo = port.apply(x, tr), O.apply(o, x, trace_row=tr)
        assert np.abs(y - yo).max() <= 1e-14 * np.abs(yo).max()


def test_c_port_tfqmr_solves_the_steady_state():
    # the CPU baseline's TFQMR loop (serial and parallel vector passes) drives
    # the C-port apply to the steady state: residual of A (x0 + dx) = e_trace
    # and the RDO of the oracle's own solve (odd top tier: the reference's
    # last-ADO trace row is a consistent normalisation there)
    from oracle.heom_port import Port
    o = O.Model(O.siam_params(nk=1, ntop=3, V=0.3, U=1.5, T=0.2))
    port = Port(o)
    x0, b = port.steady_rhs("last")
    for par in (False, True):
        dx, bound = port.tfqmr(b, 200, parallel=par)
        x = O.from_fortran((x0 + dx).view(np.complex128).reshape(o.N, o.d, o.d))
        r = O.apply(o, x, trace_row="last")
        r[-1, o.d - 1, o.d - 1] -= 1.0
        assert np.linalg.norm(r) <= 1e-9 and bound <= 1e-8
        xo, _, _ = O.steady_state(o, tol=1e-12)
        assert np.abs(x[0] - xo[0]).max() <= 1e-9


# ---- block-Jacobi iteration (jacobi/) ----
@pytest.mark.parametrize("p", [O.siam_params(nk=2, ntop=2, V=0.3, U=1.5), O.dqd_params(nk=1, ntop=2)])
def test_jacobi_ejac_closed_form(p):
    # the eigenvalues of X -> HX - XH are E_i - E_j (E: eigenvalues of H), so
    # the spread of the nonzero ones is 2 (E_max - E_min)
    o = O.Model(p)
    ej, gap, cbt, nnz = O.jacobi_ejac(o)
    E = np.linalg.eigvalsh(o.H)
    lam = (E[:, None] - E[None, :]).ravel()
    assert abs(gap - 2 * (E.max() - E.min())) <= 1e-12 * gap
    assert nnz == int((np.abs(lam) > 1e-12).sum())
    assert abs(cbt - np.sqrt(o.L * np.abs(o.cb).max())) <= 1e-15 * cbt
    assert ej == max(gap, cbt)


@pytest.mark.parametrize("p", [O.siam_params(nk=2, ntop=2, V=0.3, U=1.5), O.dqd_params(nk=1, ntop=2)])
def test_jacobi_same_tier_block_is_the_applys_diagonal_block(p):
    # S_K is the block of the apply that maps ADO K to itself: apply a vector
    # supported on ADO K alone and read ADO K back
    o = O.Model(p)
    gam = O.ado_gamma(o)
    rng = np.random.default_rng(5)
    for K in [0, 1, int(o.first[2]), o.N - 1]:
        x = np.zeros((o.N, o.d, o.d), complex)
        x[K] = rng.standard_normal((o.d, o.d)) + 1j * rng.standard_normal((o.d, o.d))
        y = O.apply(o, x, trace_row=None)[K]
        S = O.same_tier_matrix(o, gam[K])
        assert np.abs(y.reshape(-1, order="F") - S @ x[K].reshape(-1, order="F")).max() <= 1e-12 * np.abs(y).max()


def test_jacobi_sweep_fixed_point_and_convergence():
    # the steady state is a fixed point of a sweep, and the iteration reaches
    # it from the reference's start (identity / d on tier 0)
    o = O.Model(O.siam_params(nk=2, ntop=2, V=0.3, U=1.5))
    xs, _, _ = O.steady_state(o, tol=1e-13)
    ej = O.jacobi_ejac(o)[0]
    assert np.abs(O.jacobi_sweep(o, xs, ej) - xs).max() <= 1e-10 * np.abs(xs).max()
    x0 = np.zeros((o.N, o.d, o.d), complex)
    x0[0] = np.eye(o.d) / o.d
    x, it, res = O.jacobi_solve(o, x0, crit=1e-10, maxiter=2000)
    assert res <= 1e-10 and it < 2000
    assert np.abs(x[0] - xs[0]).max() <= 1e-9


def test_jacobi_sweep_shift_independent_fixed_point():
    # a different shift changes the iterates, never the fixed point: a sweep
    # with a wrong sign of the commutator or of gamma_K in S_K would move it
    o = O.Model(O.dqd_params(nk=1, ntop=2))
    xs, _, _ = O.steady_state(o, tol=1e-13)
    for ej in (0.5, 3.0):
        assert np.abs(O.jacobi_sweep(o, xs, ej) - xs).max() <= 1e-10 * np.abs(xs).max()


def test_block_jacobi_preconditioner_inverts_the_same_tier_blocks():
    # M^-1 applied to M y returns y (tier 0 Hermitian), with M built from the
    # apply's own diagonal blocks
    o = O.Model(O.dqd_params(nk=1, ntop=2))
    shift = 0.05 * O.jacobi_ejac(o)[0]
    y = O.physical_projection(o, O.seeded_ado_vector(o, 9))
    gam = O.ado_gamma(o)
    My = np.empty_like(y)
    for K in range(o.N):
        S = O.same_tier_matrix(o, gam[K])
        My[K] = ((shift * np.eye(o.d ** 2) - S) @ y[K].reshape(-1, order="F")).reshape(o.d, o.d, order="F")
    z = O.block_jacobi_precondition(o, My, shift)
    assert np.abs(z - y).max() <= 1e-12 * np.abs(y).max()


# ---- sparse (COO per ADO) storage ----
def test_apply_spa_full_pattern_is_the_dense_apply():
    o = O.Model(O.siam_params(nk=2, ntop=2, V=0.3, U=1.5))
    pat = O.spa_pattern_from_mask(np.ones((o.N, o.d, o.d), bool))
    x = O.seeded_ado_vector(o, 21)
    xs = np.array([x[a, i - 1, j - 1] for a, i, j in zip(np.repeat(np.arange(o.N), pat[0]), pat[2], pat[3])])
    y = O.apply_spa(o, pat, xs)
    yd = O.apply(o, x, trace_row=None)
    assert np.abs(y[:-1] - yd.transpose(0, 2, 1).reshape(-1)).max() <= 1e-13 * np.abs(yd).max()
    assert abs(y[-1] - np.trace(x[0])) <= 1e-14


def test_apply_spa_is_the_restricted_dense_matrix():
    # brute force: column j of the sparse apply is the pattern's rows of column
    # j of the dense matrix, for a pattern that drops elements at random
    o = O.Model(O.siam_params(nk=1, ntop=1, V=0.3, U=1.5))
    rng = np.random.default_rng(3)
    mask = rng.random((o.N, o.d, o.d)) < 0.6
    mask[0] |= np.eye(o.d, dtype=bool)
    pat = O.spa_pattern_from_mask(mask)
    a_of = np.repeat(np.arange(o.N), pat[0])
    A = O.dense_matrix(o, trace_row=None)
    n = len(a_of)
    idx = lambda a, i, j: 2 * (a * o.d * o.d + i + j * o.d)   # as_real: Fortran order, (Re, Im)
    for e in range(n):
        for part in (0, 1):
            xs = np.zeros(n, complex)
            xs[e] = 1.0 if part == 0 else 1j
            y = O.apply_spa(o, pat, xs)
            col = A[:, idx(a_of[e], pat[2][e] - 1, pat[3][e] - 1) + part]
            want = np.array([col[idx(a, i - 1, j - 1)] + 1j * col[idx(a, i - 1, j - 1) + 1]
                             for a, i, j in zip(a_of, pat[2], pat[3])])
            assert np.abs(y[:-1] - want).max() <= 1e-12 * max(np.abs(A).max(), 1.0)


def test_calcderiv_spa_is_the_sparse_apply_and_the_dense_derivative():
    """calcderiv_spa_omp at fixed H is the sparse apply without its trace
    element; on the full pattern with a lead shift it is the dense
    derivative at that shift, element for element."""
    import dataclasses
    o = O.Model(O.siam_params(nk=2, ntop=2, V=0.3, U=1.5))
    rng = np.random.default_rng(8)
    mask = rng.random((o.N, o.d, o.d)) < 0.7
    mask[0] |= np.eye(o.d, dtype=bool)
    pat = O.spa_pattern_from_mask(mask)
    xs = rng.standard_normal(len(pat[2])) + 1j * rng.standard_normal(len(pat[2]))
    assert np.array_equal(O.calcderiv_spa(o, pat, xs), O.apply_spa(o, pat, xs)[:-1])
    ot = O.Model(dataclasses.replace(O.siam_params(nk=2, ntop=2, V=0.3, U=1.5), steady=False))
    full = O.spa_pattern_from_mask(np.ones((ot.N, ot.d, ot.d), bool))
    x = O.seeded_ado_vector(ot, 22)
    a_of = np.repeat(np.arange(ot.N), full[0])
    xf = x[a_of, full[2] - 1, full[3] - 1]
    g = O.gdraw_at(ot, O.engy_shift(ot, 0.8, 0, np.array([[2.0, 2.0], [3.0, 3.0]])))
    yd = O.apply(ot, x, trace_row=None, gdraw=g)
    y = O.calcderiv_spa(ot, full, xf, gdraw=g)
    assert np.array_equal(y, yd[a_of, full[2] - 1, full[3] - 1])


def test_spa_pattern_is_the_selection_rule_when_every_lead_couples():
    """prelude_spa's pattern (exp(a H) at tier 0, products with the c
    operators tier by tier) is a numerical construction; with every orbital
    coupled to every lead it must reproduce, element for element, the (N, S_z)
    selection rule the symmetry sectors are built from (physical_projection),
    for a diagonal H (SIAM) and a hopping one (double dot)."""
    for p in (O.siam_params(nk=2, ntop=3, V=0.3, U=1.5), O.dqd_params(nk=2, ntop=2)):
        o = O.Model(p)
        nnz, ind, irow, icol = O.spa_pattern_build(o, seed=3)
        phys = O.physical_projection(o, np.ones((o.N, o.d, o.d), complex), hermitian=False) != 0
        pn, _, pr, pc = O.spa_pattern_from_mask(phys)
        assert np.array_equal(nnz, pn) and np.array_equal(irow, pr) and np.array_equal(icol, pc)
        assert ind[0] == 1 and np.array_equal(np.diff(ind), nnz[:-1])


def test_spa_pattern_holds_the_steady_state_and_ignores_the_mixing():
    """Every element of the exact steady state lies in the pattern (the
    sparse mode loses nothing), including a model whose left lead is
    decoupled (its drawers empty their ADOs, which the steady state leaves at
    zero); the mixing coefficients change no element of the pattern."""
    for p in (O.siam_params(nk=2, ntop=2, V=0.3, U=1.5), O.siam_params(nk=2, ntop=2, V=0.3, U=1.5, gl=0.0)):
        o = O.Model(p)
        pats = [O.spa_pattern_build(o, seed=s) for s in (1, 2)]
        assert all(np.array_equal(a, b) for a, b in zip(*pats))
        nnz, _, irow, icol = pats[0]
        mask = np.zeros((o.N, o.d, o.d), bool)
        mask[np.repeat(np.arange(o.N), nnz), irow - 1, icol - 1] = True
        x, _, _ = O.steady_state(o, tol=1e-13)
        assert not np.any(np.abs(x[~mask]) > 1e-12 * np.abs(x).max())
        if p.linewidth[0][0] == 0.0:
            assert nnz.sum() < np.count_nonzero(O.physical_projection(o, np.ones_like(x), hermitian=False))


def _trun_cases():
    import dataclasses
    return [O.siam_params(nk=2, ntop=1, V=0.5, U=1.0, T=0.3), O.dqd_params(nk=1, ntop=1, T=0.3),
            O.siam_params(nk=1, ntop=2, V=0.4, U=1.5, T=0.3),
            dataclasses.replace(O.siam_params(nk=1, ntop=1, V=0.5, U=1.0, T=0.3), lsdraw=True)]


def test_trun_der_term_is_the_eliminated_next_tier():
    """The terminator restated from the reference's loop (per top ADO and
    drawer, the virtual ADO's lower couplings, the resolvent from the
    Liouvillian eigenbasis) equals the hierarchy one tier higher with that
    tier eliminated by dense solves: x_V = -S_V^-1 [A_{L+1}(x, 0)]_V, then
    [A_{L+1}(x, x_V)] - A_L x on tiers 0..L.  (On the physical subspace:
    the stored hierarchy keeps one ADO of each conjugate pair, and the pair
    rule that lets either stand for the other holds there.)"""
    import dataclasses
    for p in _trun_cases():
        o = O.Model(p)
        o1 = O.Model(dataclasses.replace(p, ntop=p.ntop + 1))
        d, n = o.d, o.N
        assert all(np.array_equal(a, b) for a, b in zip(o.tiers, o1.tiers[:o.L + 1]))
        # a vector of the physical subspace every state and iterate lives in
        # (parity-allowed elements, and the conjugation rule of the paired
        # storage on self-paired ADOs): a random mix of A^k applied to a
        # trace-one tier-0 state
        v = np.zeros((n, d, d), complex)
        v[0] = np.diag(np.arange(1.0, d + 1))
        x = np.zeros_like(v)
        rng = np.random.default_rng(0x7E11)
        for _ in range(o.L + 3):
            v = O.apply(o, v, trace_row=None)
            v /= np.linalg.norm(v)
            x += rng.standard_normal() * v
        X = np.zeros((o1.N, d, d), complex)
        X[:n] = x
        B = O.apply(o1, X, trace_row=None)[n:]
        lsup = np.kron(np.eye(d), o.H) - np.kron(o.H.T, np.eye(d))
        for v, K in enumerate(o1.tiers[o1.L]):
            g = o1.gdraw[K].sum()
            X[n + v] = np.linalg.solve(-1j * lsup + g * np.eye(d * d), -B[v].reshape(-1, order="F")).reshape(
                d, d, order="F")
        T = O.apply(o1, X, trace_row=None)[:n] - O.apply(o, x, trace_row=None)
        Tr = O.trun_der_term(o, x)
        assert np.abs(T[:o.first[o.L]]).max() <= 1e-13 * np.abs(T).max()   # only the top tier changes
        assert np.linalg.norm(Tr - T) <= 1e-12 * np.linalg.norm(T)


def test_trun_der_steady_state_is_the_next_tier():
    """With the terminator the steady state on tiers 0..L is exactly that of
    the hierarchy one tier higher (its top tier's equation is what the
    terminator solves), and differs from the plain tier-L truncation."""
    import dataclasses
    for p in _trun_cases()[:1] + _trun_cases()[3:]:
        o = O.Model(p)
        o1 = O.Model(dataclasses.replace(p, ntop=p.ntop + 1))
        xt, _, _ = O.steady_state(o, tol=1e-13, trun_der=True)
        x1, _, _ = O.steady_state(o1, tol=1e-13)
        xp, _, _ = O.steady_state(o, tol=1e-13)
        assert np.abs(xt - x1[:o.N]).max() <= 1e-11
        assert np.abs(xp[0] - x1[0]).max() >= 1e-5


def test_system_field_limit_is_the_shifted_level():
    """dhs(t) of calcdhs's dedot (1 - exp(-t/aD)) on n_up + n_dn: long after
    the switch-on, hs + dhs is the Hamiltonian of the model whose level sits
    dedot higher (built independently by calchs), and so is its derivative."""
    import dataclasses
    p = dataclasses.replace(O.siam_params(nk=1, ntop=2, V=0.6, U=1.5, T=0.2), steady=False)
    o = O.Model(p)
    p2 = dataclasses.replace(O.siam_params(nk=1, ntop=2, V=0.6, U=1.5, T=0.2, eps=-1.0 + 0.3), steady=False)
    o2 = O.Model(p2)
    n = sum(o.ams[0][s].T @ o.ams[0][s] for s in range(2))
    terms = [(n, 0, (0.3, 2.0))]
    H = O.h_at(o, 1e4, terms)
    assert np.abs(H - o2.H).max() <= 1e-12 * np.abs(o2.H).max()
    assert np.abs(O.h_at(o, 0.0, terms) - o.H).max() == 0.0
    x = O.seeded_ado_vector(o, 23)
    y = O.apply(o, x, trace_row=None, H=H)
    y2 = O.apply(o2, x, trace_row=None)
    assert np.abs(y - y2).max() <= 1e-12 * np.abs(y2).max()
    # window and gate time functions at their edges (calcdhs.f90:112, :267)
    assert O.dhs_coef(3, [1.0, 1.0, 2.0], 1.0) == 0.0 and O.dhs_coef(4, [1.0, 1.0, 2.0], 1.0) == 1.0
    assert O.dhs_coef(3, [1.0, 1.0, 2.0], 2.0) == 0.0 and O.dhs_coef(4, [1.0, 1.0, 2.0], 2.0) == 0.0


# ---- time-dependent lead shift (getengyshift, calcderiv_omp(tt), intstp) ----
def _td_pair():
    import dataclasses
    p = O.siam_params(nk=1, ntop=2, V=0.6, U=1.5, T=0.2)
    return O.Model(dataclasses.replace(p, steady=False)), O.Model(p)


def test_field_limits_are_the_steady_bias():
    # the exponential switch-on reaches the steady bias for t >> tchar, the
    # sine shift equals it at a quarter period and vanishes at half a period
    td, st = _td_pair()
    tc = np.full((2, 2), 3.0)
    assert np.abs(O.gdraw_at(td, O.engy_shift(td, 1e6, 0, tc)) - st.gdraw).max() <= 1e-14 * np.abs(st.gdraw).max()
    assert np.abs(O.gdraw_at(td, O.engy_shift(td, 0.75, 1, tc)) - st.gdraw).max() <= 1e-14 * np.abs(st.gdraw).max()
    assert np.abs(O.gdraw_at(td, O.engy_shift(td, 1.5, 1, tc)) - td.cgamma).max() <= 1e-14 * np.abs(st.gdraw).max()
    assert np.abs(O.engy_shift(td, 5.0, 0, tc, t_off=4.0)).max() == 0.0


def test_rk4_with_field_is_fourth_order():
    # a time-dependent rate must be sampled at the RK4 stage times for the
    # global error to fall 16x per halving of dt
    td, _ = _td_pair()
    x0 = O.physical_projection(td, O.seeded_ado_vector(td, 6))
    field = (1, np.full((2, 2), 2.0), None, 1e99)
    T = 0.8
    ref = O.rk4_field(td, x0, 0.1, T / 256, 256, field)
    e1 = np.abs(O.rk4_field(td, x0, 0.1, T / 16, 16, field) - ref).max()
    e2 = np.abs(O.rk4_field(td, x0, 0.1, T / 32, 32, field) - ref).max()
    assert 12.0 < e1 / e2 < 20.0
