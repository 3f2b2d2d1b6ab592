"""Time propagation with the reference's outputs ``curr.data`` and ``popu.data``.

main/main.f90:1440-1500 runs ``intstp`` step by step and, after every step,
writes one line per file in format 518, ``(f12.4, 2x, 32(e15.6e3, 1x))``:

* nalf = 1: curr.data ``t/u_t, eshift(1,1), J_up, J_dn, J_up + J_dn,
  J_up - J_dn, occ``; popu.data ``t, eshift(1,1), occ_up, occ_dn, occ``;
* nalf = 2: curr.data ``t/u_t, eshift(1,1), eshift(2,1), J_L, J_R, J_L + J_R,
  occ_up, occ_dn, occ``; popu.data ``t, eshift(1,1), eshift(2,1), occ_up,
  occ_dn, occ``;
* more leads: curr.data ``t/u_t, (eshift(a,1), eshift(a,2)), (J_a,up, J_a,dn),
  occ_up, occ_dn, occ``; popu.data ``t, (eshift(a,1), eshift(a,2)), occ_up,
  occ_dn, occ``.

Currents in nA (td/calcj.f90 with jconst), times in fs, shifts in eV; with
runits = 1 time and current are in units of hbar / Gamma and e Gamma / hbar,
Gamma the sum over leads of the largest linewidth (main/main.f90:1231-1244).
The steps are the device RK4 of td/intstp.f90 (heomq_rk4_at), with the lead
and system fields of the context at each stage time.
"""
from __future__ import annotations

import math

import numpy as np

HBAR = 0.658211928
JCONST = 1.602177e5


def fortran_e(v: float, w: int = 15, d: int = 6, e: int = 3) -> str:
    """Fortran ``Ew.dEe`` output: 0.dddddd mantissa, signed e-digit exponent."""
    v = float(v)
    if v == 0.0 or not math.isfinite(v):
        body = ("0." + "0" * d + "E+" + "0" * e) if v == 0.0 else ("NaN" if v != v else ("Infinity" if v > 0 else "-Infinity"))
        return body.rjust(w)
    x = abs(v)
    exp = math.floor(math.log10(x)) + 1
    mant = round(x / 10.0 ** exp, d)
    if mant >= 1.0:   # rounding carried into the next decade
        exp += 1
        mant = round(x / 10.0 ** exp, d)
    elif mant < 0.1:
        exp -= 1
        mant = round(x / 10.0 ** exp, d)
    s = f"{mant:.{d}f}"   # '0.dddddd'
    body = ("-" if v < 0 else "") + s + ("E+" if exp >= 0 else "E-") + str(abs(exp)).rjust(e, "0")
    return body.rjust(w)


def format_518(t: float, values) -> str:
    """One line of format 518: f12.4, 2x, then e15.6e3 values each followed by 1x."""
    return f"{t:12.4f}" + "  " + "".join(fortran_e(v) + " " for v in values)


def _units(ctx, runits: int):
    if runits != 1:
        return 1.0, 1.0
    p = ctx.model.params
    g = sum(max(abs(w) for w in p.linewidth[a][: p.norbs]) / HBAR for a in range(p.nalf))
    return g, 1.0 / JCONST / g


def _shifts(ctx, t: float):
    p = ctx.model.params
    es = ctx.field_shifts(t)
    if ctx.field_type() < 0 and p is not None and p.steady:
        es = np.array([[p.engyshift[a][s] for s in range(2)] for a in range(p.nalf)])
    return es


def record(ctx, x, t: float, runits: int = 0):
    """(curr.data values, popu.data values) of state x at time t."""
    nalf = ctx.model.nalf
    cur, occ = ctx.observables(x)
    occu, occd = float(occ[:, 0].sum()), float(occ[:, 1].sum())
    oc = occu + occd
    es = _shifts(ctx, t)
    ut, uj = _units(ctx, runits)
    if nalf == 1:
        ju, jd = cur[0, 0] * uj, cur[0, 1] * uj
        cv = [es[0, 0], ju, jd, ju + jd, ju - jd, oc]
        pv = [es[0, 0], occu, occd, oc]
    elif nalf == 2:
        jl, jr = (cur[0, 0] + cur[0, 1]) * uj, (cur[1, 0] + cur[1, 1]) * uj
        cv = [es[0, 0], es[1, 0], jl, jr, jl + jr, occu, occd, oc]
        pv = [es[0, 0], es[1, 0], occu, occd, oc]
    else:
        cv = [v for a in range(nalf) for v in (es[a, 0], es[a, 1])]
        cv += [v * uj for a in range(nalf) for v in (cur[a, 0], cur[a, 1])] + [occu, occd, oc]
        pv = [v for a in range(nalf) for v in (es[a, 0], es[a, 1])] + [occu, occd, oc]
    return t * ut, cv, t, pv


def propagate(ctx, x, t0: float, dt: float, nstep: int, curr="curr.data", popu="popu.data", runits: int = 0,
              every: int = 1):
    """Advance x (device state) by nstep RK4 steps of dt from t0, writing a
    curr.data and popu.data line after every `every` steps; returns the final
    time."""
    t = t0
    with open(curr, "w") as fc, open(popu, "w") as fp:
        for k in range(1, nstep + 1):
            ctx.rk4(x, dt, 1, t0=t)
            t = t0 + k * dt
            if k % every == 0:
                tc, cv, tp, pv = record(ctx, x, t, runits)
                fc.write(format_518(tc, cv) + "\n")
                fp.write(format_518(tp, pv) + "\n")
    return t
