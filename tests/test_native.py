"""CPU tests of the native library: exports, host tables against the oracle,
the reference-array drop-in loader and the error behaviour (no GPU needed)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import heomq
from heomq import _lib
from oracle import heom_oracle as O
from ref_arrays import reference_arrays

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "heomq.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(heomq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    h = ctypes.CDLL(str(_lib.LIB_PATH))
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(h, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_header_constants_match_the_binding():
    # status codes, trace rows, Hamiltonian kinds and stats lengths of heomq.h
    text = (ROOT / "include" / "heomq.h").read_text()
    defs = {k: int(v) for k, v in re.findall(r"#define\s+HEOMQ_([A-Z0-9_]+)\s+(-?\d+)", text)}
    want = dict(OK=_lib.OK, BAD_DIM=_lib.BAD_DIM, UNSUPPORTED=_lib.UNSUPPORTED, HIP_ERROR=_lib.HIP_ERROR,
                NOT_CONVERGED=_lib.NOT_CONVERGED, BREAKDOWN=_lib.BREAKDOWN, STOPPED=_lib.STOPPED,
                BAD_ARG=_lib.BAD_ARG, TRACE_NONE=_lib.TRACE_NONE, TRACE_LAST=_lib.TRACE_LAST,
                TRACE_RDO=_lib.TRACE_RDO, HAM_SIAM=_lib.HAM_SIAM, HAM_DQD=_lib.HAM_DQD,
                STATS_LEN=_lib.STATS_LEN, STATS_MIXED_LEN=_lib.STATS_MIXED_LEN,
                STATS_CHEBY_LEN=_lib.STATS_CHEBY_LEN)
    for k, v in want.items():
        assert defs[k] == v, k


def test_fortran_shim_exports_reference_names():
    path = ROOT / "heom-quick2_amd" / "heomq" / "libheomq_fortran.so"
    if not path.exists():
        pytest.skip("Fortran compiler not available at build time")
    h = ctypes.CDLL(str(path))
    for name in ("calc_ax_tfqmr_omp_", "calc_ax_bicg_omp_", "calc_atrx_bicg_omp_", "calcderiv_omp_",
                 "heomq_fortran_init_", "solve_tfqmr_gpu_", "solve_bicg_gpu_", "calc_ax_jacobi_omp_",
                 "solve_jacobi_gpu_", "calc_ax_tfqmr_spa_omp_", "heomq_register_spa_pattern_",
                 "calcderiv_spa_omp_", "heomq_register_td_spa_arrays_", "heomq_fortran_set_system_field_"):
        assert hasattr(h, name), name


CASES = [
    (heomq.siam(nk=3, ntop=3, V=0.3, U=1.5), O.siam_params(nk=3, ntop=3, V=0.3, U=1.5)),
    (heomq.two_impurity(nk=2, ntop=2), O.dqd_params(nk=2, ntop=2)),
    (heomq.siam(nk=2, ntop=3, lsdraw=True), O.siam_params(nk=2, ntop=3, lsdraw=True)),
    (heomq.siam(nk=1, ntop=4, T=0.2), O.siam_params(nk=1, ntop=4, T=0.2)),
]


@pytest.mark.parametrize("pp,op", CASES)
def test_native_tables_match_oracle(pp, op):
    m = heomq.Model(pp)
    o = O.Model(op)
    assert m.N == o.N and list(m.first) == list(o.first)
    cb, cd, g = m.coefs()
    assert np.abs(cb - o.cb).max() <= 1e-13 * np.abs(o.cb).max()
    assert np.abs(cd - o.cd).max() <= 1e-13 * np.abs(o.cd).max()
    assert np.abs(g - o.gdraw).max() <= 1e-12 * np.abs(o.gdraw).max()
    H, A = m.operators()
    assert np.abs(H - o.H).max() == 0.0
    for mm in range(o.norbs):
        for s in range(2):
            assert np.array_equal(A[mm * 2 + s], o.ams[mm][s])
    for n in range(1, m.L + 1):
        assert np.array_equal(m.tier(n), o.tiers[n])
    lo, up = m.coupling()
    olo = np.concatenate([np.concatenate([o.lower[n][0], o.lower[n][1]], axis=2).reshape(-1, 4)
                          for n in range(1, o.L + 1)])
    assert np.array_equal(lo, olo)
    oup = np.concatenate([np.concatenate([o.upper[n][0], o.upper[n][1]], axis=2).reshape(-1, 4)
                          for n in range(0, o.L)])
    mask = oup[:, 0] >= 0
    assert np.array_equal(up[mask], oup[mask]) and (up[~mask, 0] == -1).all()


def test_native_lscale_coefficients_match_oracle():
    # lscale: cb and cd divided by |cb|^(1/2) of the unscaled cb, the rates unchanged
    import dataclasses
    for pp, op in ((heomq.two_impurity(nk=6, ntop=1, lscale=True), dataclasses.replace(O.dqd_params(nk=6, ntop=1), lscale=True)),
                   (heomq.siam(nk=3, ntop=2, lscale=True), dataclasses.replace(O.siam_params(nk=3, ntop=2), lscale=True))):
        m, o = heomq.Model(pp), O.Model(op)
        cb, cd, g = m.coefs()
        assert np.abs(cb - o.cb).max() <= 1e-12 * np.abs(o.cb).max()
        assert np.abs(cd - o.cd).max() <= 1e-12 * np.abs(o.cd).max()
        assert np.abs(np.abs(cb) - o.dbsqrt).max() <= 1e-12 * o.dbsqrt.max()   # |cb / |cb|^(1/2)| = |cb|^(1/2)
        assert np.abs(g - o.gdraw).max() <= 1e-12 * np.abs(o.gdraw).max()
        with pytest.raises(heomq.HeomqError):   # the representation is part of the shape
            m.set_params(dataclasses.replace(pp, lscale=False))


def test_reference_loader_takes_dbsqrt_and_unique_index_tables():
    # the reference's lscale factors, and a hierarchy without repeated drawers
    # (absent upper couplings: indexcoef = 0)
    import dataclasses
    o = O.Model(dataclasses.replace(O.dqd_params(nk=1, ntop=2), lscale=True))
    arr = reference_arrays(o)
    assert arr["dbsqrt"] is not None and (arr["indexcoef"] == 0).any()
    m = heomq.Model.from_reference(**arr)
    cb, cd, g = m.coefs()
    assert np.abs(cb - o.cb).max() == 0.0 and np.abs(cd - o.cd).max() == 0.0
    arr["dbsqrt"] = -arr["dbsqrt"]
    with pytest.raises(heomq.HeomqError):
        heomq.Model.from_reference(**arr)


def test_pade_native_matches_numpy_eigensolver():
    # the native Sturm-bisection eigenvalues reproduce LAPACK's through cb/cgamma
    for nk in (1, 4, 8, 16, 30):
        m = heomq.Model(heomq.siam(nk=nk, ntop=1))
        o = O.Model(O.siam_params(nk=nk, ntop=1))
        _, _, g = m.coefs()
        assert np.abs(g - o.gdraw).max() <= 1e-12 * np.abs(o.gdraw).max()


@pytest.mark.parametrize("pp,op", [CASES[2], (heomq.two_impurity(nk=1, ntop=2, lsdraw=True),
                                              O.dqd_params(nk=1, ntop=2, lsdraw=True))])
def test_reference_loader_round_trip(pp, op):
    o = O.Model(op)
    m_ref = heomq.Model.from_reference(**reference_arrays(o))
    m = heomq.Model(pp)
    assert m_ref.N == m.N and list(m_ref.first) == list(m.first)
    for n in range(m.L + 1):
        assert np.array_equal(m_ref.tier(n), m.tier(n))
    for a, b in zip(m_ref.operators(), m.operators()):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    lo1, up1 = m_ref.coupling()
    lo2, up2 = m.coupling()
    assert np.array_equal(lo1, lo2)
    assert np.array_equal(up1, up2)
    assert np.abs(m_ref.coefs()[0] - m.coefs()[0]).max() <= 1e-13 * np.abs(m.coefs()[0]).max()
    assert np.abs(m_ref.coefs()[2] - m.coefs()[2]).max() <= 1e-13 * np.abs(m.coefs()[2]).max()


def test_model_knows_its_orbitals_and_leads_without_parameters():
    # a model loaded from the reference's arrays carries no parameters: the
    # sizes of its observables (current (nalf, nspin), occupation (norbs,
    # nspin)) come from the model itself, for one, two and three orbitals
    three = O.Params(norbs=3, nspin=2, nalf=2, nmats=1, bandwidth=(10.0, 10.0),
                     linewidth=((0.1, 0.08, 0.12), (0.1, 0.08, 0.12)), temperature=(0.2, 0.2),
                     engyshift=((0.1, 0.1), (-0.1, -0.1)), ham="siam", hpar=dict(eup=-0.5, edown=-0.5, uu=1.0),
                     ntop=1)
    for op, norbs in ((O.siam_params(nk=2, ntop=2), 1), (O.dqd_params(nk=2, ntop=2), 2), (three, 3)):
        o = O.Model(op)
        m = heomq.Model.from_reference(**reference_arrays(o))
        assert m.params is None and (m.norbs, m.nalf, m.d) == (norbs, 2, 4 ** norbs)
        H, A = m.operators()
        assert A.shape == (2 * norbs, m.d, m.d)
    m = heomq.Model(heomq.siam(nk=2, ntop=2))
    assert (m.norbs, m.nalf) == (1, 2)


def test_reference_loader_rejects_shifted_indextable():
    # indextable without tier 0's empty record (ball2drawer.f90:40-44) shifts
    # every tuple; the loader must refuse it instead of reading past the end
    o = O.Model(O.siam_params(nk=1, ntop=2, lsdraw=True))
    arr = reference_arrays(o)
    arr["indextable"] = np.ascontiguousarray(np.concatenate([arr["indextable"][1:], [10 ** 6]]).astype(np.int32))
    with pytest.raises(heomq.HeomqError) as e:
        heomq.Model.from_reference(**arr)
    assert e.value.code == _lib.BAD_ARG


@pytest.mark.parametrize("short", ["indextable", "indexcoef"])
def test_reference_loader_checks_table_lengths(short):
    # a coupling stream or index table shorter than the positions the other
    # tables name (lall / ntable of the reference's common block info3) is
    # refused before anything past its end is read
    o = O.Model(O.siam_params(nk=1, ntop=2, lsdraw=True))
    arr = reference_arrays(o)
    key = "ntable" if short == "indextable" else "lall"
    arr[key] = np.size(arr[short]) - 1
    with pytest.raises(heomq.HeomqError) as e:
        heomq.Model.from_reference(**arr)
    assert e.value.code == _lib.BAD_ARG
    arr[key] = np.size(arr[short])
    heomq.Model.from_reference(**arr).close()


def test_reference_loader_rejects_bad_dims():
    o = O.Model(O.siam_params(nk=1, ntop=2, lsdraw=True))
    arr = reference_arrays(o)
    arr["nrho"] = 8
    with pytest.raises(heomq.HeomqError) as e:
        heomq.Model.from_reference(**arr)
    assert e.value.code == _lib.BAD_DIM


@pytest.mark.parametrize("bad", [dict(norbs=4), dict(nmats=0), dict(ntop=0)])
def test_bad_params_raise(bad):
    p = heomq.siam()
    for k, v in bad.items():
        setattr(p, k, v)
    with pytest.raises(heomq.HeomqError):
        heomq.Model(p)


def test_set_params_keeps_shape():
    m = heomq.Model(heomq.siam(nk=2, ntop=2))
    m.set_params(heomq.siam(nk=2, ntop=2, V=0.5, eps=-0.2))
    o = O.Model(O.siam_params(nk=2, ntop=2, V=0.5, eps=-0.2))
    assert np.abs(m.coefs()[2] - o.gdraw).max() <= 1e-12 * np.abs(o.gdraw).max()
    with pytest.raises(heomq.HeomqError):
        m.set_params(heomq.siam(nk=3, ntop=2))


def test_cfg_sizes():
    # hierarchy sizes of the benchmark configurations (unique-index, Hermitian-paired)
    assert heomq.Model(heomq.siam()).N == 31297
    assert heomq.Model(heomq.two_impurity()).N == 117321


@pytest.mark.parametrize("pp,op", CASES + [(heomq.siam(nk=3, ntop=3, V=0.3, U=1.5, T=0.2),
                                            O.siam_params(nk=3, ntop=3, V=0.3, U=1.5, T=0.2))])
def test_packed_layout_host(pp, op):
    """The packed layout built on the host: element count equals the oracle's
    selection rules, and the top-tier block tables fit the light kernel's
    257-entry LDS tables even where many ADOs hold no element (d = 4)."""
    lay = heomq.Model(pp).packed_layout()
    o = O.Model(op)
    mask = O.physical_projection(o, np.ones((o.N, o.d, o.d), complex), hermitian=False) != 0
    assert lay["packed"] == int(mask.sum())
    assert lay["empty_ados"] == int((mask.reshape(o.N, -1).sum(1) == 0).sum())
    assert 1 <= lay["max_ados_per_light_block"] <= 257
    assert lay["n_conserved"] == 1 and lay["sz_conserved"] == 1


def test_packed_layout_benchmark_sizes():
    lay = heomq.Model(heomq.two_impurity()).packed_layout()
    assert lay["packed"] == 1797356 and lay["max_light_couplings"] == 3
    lay = heomq.Model(heomq.siam()).packed_layout()
    assert lay["packed"] == 37012 and lay["empty_ados"] > 0 and lay["max_ados_per_light_block"] <= 257


@pytest.mark.parametrize("pp,op", [(heomq.siam(nk=2, ntop=3, V=0.3, U=1.5), O.siam_params(nk=2, ntop=3, V=0.3, U=1.5)),
                                   (heomq.two_impurity(nk=2, ntop=2), O.dqd_params(nk=2, ntop=2)),
                                   (heomq.two_impurity(), None)])
def test_jacobi_eigensystem_host(pp, op):
    """The block-Jacobi eigensystem built on the host: H U = U diag(E), U unitary
    and block-diagonal over the conserved groups, and the default shift equals
    the oracle's (eigenvalues of the d^2 x d^2 commutator matrix)."""
    m = heomq.Model(pp)
    same = op is not None          # the benchmark-size model: only its H is rebuilt in the oracle
    o = O.Model(op if same else O.dqd_params(nk=1, ntop=1))
    info, E, U = m.jacobi_info()
    H = o.H
    assert np.abs(H @ U - U * E[None, :]).max() <= 1e-13 * max(np.abs(H).max(), 1.0)
    assert np.abs(U.conj().T @ U - np.eye(o.d)).max() <= 1e-13
    N, S2 = O.fock_quantum_numbers(o)
    nz = np.abs(U) > 0
    assert not np.any(nz & ((N[:, None] != N[None, :]) | (S2[:, None] != S2[None, :])))
    ej, gap, cbt, nnz = O.jacobi_ejac(o)
    assert abs(info["gap"] - gap) <= 1e-12 * max(gap, 1.0)
    assert int(info["nnzjac"]) == nnz
    if same:
        assert abs(info["ejac"] - ej) <= 1e-12 * ej
        assert abs(info["cb_term"] - cbt) <= 1e-12 * cbt


@pytest.mark.parametrize("pf,of", [
    (lambda: heomq.siam(nk=2, ntop=3, V=0.3, U=1.5), lambda: O.siam_params(nk=2, ntop=3, V=0.3, U=1.5)),
    (lambda: heomq.two_impurity(nk=2, ntop=2), lambda: O.dqd_params(nk=2, ntop=2)),
    (lambda: heomq.siam(nk=2, ntop=3, V=0.3, U=1.5, gl=0.0), lambda: O.siam_params(nk=2, ntop=3, V=0.3, U=1.5, gl=0.0)),
])
def test_spa_pattern_builder_matches_oracle(pf, of):
    # the native prelude_spa restatement against the oracle's (independent
    # mixing coefficients: the pattern must not depend on them)
    nat = heomq.Model(pf()).spa_pattern(seed=5)
    ora = O.spa_pattern_build(O.Model(of()), seed=3)
    for a, b in zip(nat, ora):
        assert np.array_equal(np.asarray(a, np.int64), np.asarray(b, np.int64))


def test_spa_pattern_builder_full_two_impurity_size():
    # the two-impurity workload (d = 16, M = 112, L = 3): the built pattern has
    # exactly the symmetry-packed layout's elements
    m = heomq.Model(heomq.two_impurity())
    nnz, ind, irow, icol = m.spa_pattern()
    assert len(irow) == m.packed_layout()["packed"] == int(nnz.sum())
    assert ind[0] == 1 and np.array_equal(np.diff(ind), nnz[:-1])

