"""The fused TFQMR iteration's algebra, restated in numpy on the CPU.

The device's block-Jacobi steady solve runs TFQMR (Freund 1993, the method of
tfqmr_double/solve_tfqmr.f90:97-179) in four launches per iteration: step 1
with the previous iteration's deferred x and d updates and the v update, the
operator on y2 with step 2, the operator on w, and one fold of all sums, with
sigma = rt.v taken one recurrence step from a directly summed rt.v.  Here
that schedule is restated line by line and checked against the textbook
iteration (two half steps, u, v and d as in Saad, Iterative Methods, alg.
7.8) on a random nonsymmetric system: the same iterates in exact arithmetic,
so the x after every flushed iteration agrees to rounding.  A schedule with a
plausible slip (sigma from the stale v, a missing deferred term) does not.
"""
import numpy as np
import pytest


def system(n=160, seed=3):
    rng = np.random.default_rng(seed)
    A = np.eye(n) + 0.35 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    rt = np.sign(rng.standard_normal(n))
    return A, b, rt


def textbook(A, b, rt, iters):
    """x after each full iteration (two half steps) of Freund's TFQMR from x = 0."""
    x = np.zeros_like(b)
    r = b.copy()
    w = r.copy()
    u1 = r.copy()
    v = A @ u1
    au1 = v.copy()
    d = np.zeros_like(b)
    tau, theta, eta = np.linalg.norm(r), 0.0, 0.0
    rho = rt @ r
    out = []
    for _ in range(iters):
        alpha = rho / (rt @ v)
        u2 = u1 - alpha * v
        au2 = A @ u2
        for u, au in ((u1, au1), (u2, au2)):
            w = w - alpha * au
            d = u + (theta**2 * eta / alpha) * d
            theta = np.linalg.norm(w) / tau
            c = 1.0 / np.sqrt(1.0 + theta**2)
            tau = tau * theta * c
            eta = c * c * alpha
            x = x + eta * d
        rho_new = rt @ w
        beta = rho_new / rho
        rho = rho_new
        u1 = w + beta * u2
        au1 = A @ u1
        v = au1 + beta * (au2 + beta * v)
        out.append(x.copy())
    return out


def fused(A, b, rt, iters, slip=None):
    """The device schedule: k_step1_fused, M^-1 A y2 with step 2, M^-1 A w,
    k_tfq_fold; x flushed (tfq_flush) after each iteration for comparison."""
    x = np.zeros_like(b)
    w = b.copy()                         # k_init_r: w = r, y2 = d = 0
    y2 = np.zeros_like(b)
    d = np.zeros_like(b)
    aw = A @ w                           # k_init_v: v = aw, u2 = 0
    v = aw.copy()
    u2 = np.zeros_like(b)
    tau = np.sqrt(w @ w)
    rho = rt @ w
    alpha = rho / (rt @ v)
    beta = dcoef = eta_a = eta_b = cy = 0.0
    cd = 1.0

    def half(a, tau_, alpha_):
        theta = np.sqrt(a) / tau_
        c = 1.0 / np.sqrt(1.0 + theta * theta)
        eta = c * c * alpha_
        return theta, tau_ * theta * c, eta, theta * theta * eta / alpha_

    out = []
    for _ in range(iters):
        # k_step1_fused
        db = cy * y2 + cd * d
        x = x + eta_a * d + eta_b * db
        v_old = v
        v = aw + beta * (2.0 * u2 + beta * v)
        sv = rt @ (v_old if slip == "stale_v" else v)
        y1 = w + beta * y2
        u1 = aw + beta * u2
        y2 = y1 - alpha * v
        w = w - alpha * u1
        d = y1 + dcoef * db
        a1 = w @ w
        # operator on y2, step 2 in its epilogue
        u2 = A @ y2
        w = w - alpha * u2
        a2, rw, ru = w @ w, rt @ w, rt @ u2
        # operator on w, rt.aw in its epilogue
        aw = A @ w
        ra = rt @ aw
        # k_tfq_fold
        _, tau, eta_a, cd = half(a1, tau, alpha)
        cy = 1.0
        theta, tau, eta_b, _ = half(a2, tau, alpha)
        if slip == "no_second_eta":
            eta_b = 0.0
        beta = rw / rho
        rho = rw
        sigma = ra + beta * (2.0 * ru + beta * sv)
        alpha = rho / sigma
        dcoef = theta * theta * eta_b / alpha
        # tfq_flush (a copy: the iteration itself keeps the deferred form)
        out.append(x + eta_a * d + eta_b * (cy * y2 + cd * d))
    return out


def test_fused_schedule_has_the_textbook_iterates():
    A, b, rt = system()
    ref = textbook(A, b, rt, 24)
    got = fused(A, b, rt, 24)
    xs = np.linalg.solve(A, b)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert np.abs(g - r).max() <= 1e-11 * np.abs(r).max(), k
    # and the iteration converges (the comparison is not of two stalled runs)
    assert np.linalg.norm(ref[-1] - xs) <= 1e-6 * np.linalg.norm(xs)


@pytest.mark.parametrize("slip", ["stale_v", "no_second_eta"])
def test_a_slipped_schedule_is_caught(slip):
    A, b, rt = system()
    ref = textbook(A, b, rt, 12)
    got = fused(A, b, rt, 12, slip=slip)
    assert max(np.abs(g - r).max() / np.abs(r).max() for g, r in zip(got, ref)) > 1e-6
