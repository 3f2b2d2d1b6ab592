"""The reference's state and table files, read and written on the host.

A state saved by the reference can seed a solve here, and a state solved here
can be handed back to the reference (its ``initialize2`` / ``initialize2_spa``
read it when a job restarts from a saved state).

Formats (main/main.f90:1085-1120 and :1655-1700 write them;
main/initialize2.f90:43-82 and fermion/initialize2_spa.f90:24-104 read them):

``variables.gr`` / ``variables.st`` / ``variables.sav`` / ``variables_td.sav``
    Fortran unformatted sequential records, each framed by its byte length as
    a 4-byte integer before and after: (norbs, nspin, ncor, ntier, nalf) as
    five int32; nunk as one int64; then one record per ADO holding
    rho(1:nrho, 1:nrho, lni), complex*16 in column-major order.

``rho_spa.gr`` / ``rho_spa.st`` / ``rho_spa.sav`` / ``rho_spa_td.sav``
    ``form='binary'``: the same items with no record framing; for each ADO
    with nnz_spa(lni) > 0, its nnz_spa(lni) complex*16 values in the order of
    the sparse pattern (module sparsemod: nnz_spa, ind_spa, irow_spa,
    icol_spa, 1-based).

``indextable.sav``
    ``form='binary'``: ntable as int64, then indextable(1:ntable) as int32.

``coefindex.data`` and ``auxindex.data`` (main/buildcoefindex.f90:683-709
write them, :44-140 read them back instead of rebuilding)
    ``form='binary'``: (ntier, ncor, norbs, nspin, nalf, numfff, ntier0,
    ndrawer_slow, ncor_slow) as int32; (lad, lad_fast, lset_fast) as 4-byte
    logicals, idegen_fast and ndegen_fast as int32, dratio_fast as real*8;
    lfilter, nfilter_count, nfilter_long; ltrun_der; lsimple, lwalf; lscreen;
    lomit, dfact_omit; lall as int64; then per coupling indexcoef (int64),
    itypecoef, ifactfcoef, ifactrcoef (int8).  auxindex.data: nunk as int64,
    then index_coef_ref(1:nunk) as int64.

``TAPE_<n>.resume`` / ``TAPE.resume`` (td/resumejob.f90:15-103, the
propagators' checkpoints)
    ``form='binary'``: tt as real*8; (ntier, norbs, nspin, ncor, nalf,
    numfff) as int32; nunk as int64; lhb as a 4-byte logical (then four
    int32 of the heat bath when set); then either lunk_spa as int64 and
    rho_spa(1:lunk_spa), or the nunk dense ADO blocks, complex*16.

ADO order is the reference's (tier-major, drawer tuples sorted), the order
of the dense vectors of this package: x[a] is ADO a, its memory the
column-major d x d block.
"""
from __future__ import annotations

from pathlib import Path
from typing import NamedTuple

import numpy as np

_I4 = np.dtype("<i4")
_I8 = np.dtype("<i8")
_C16 = np.dtype("<c16")


class StateHeader(NamedTuple):
    norbs: int
    nspin: int
    ncor: int
    ntier: int     # number of tiers held, tier 0 included (the reference's ntier)
    nalf: int
    nunk: int      # ADOs held


class FileFormatError(ValueError):
    pass


def header_of(model, nunk: int | None = None) -> StateHeader:
    """The header the reference writes for a hierarchy like `model`'s."""
    return StateHeader(int(model.norbs), 2, int(model.ncor), int(model.L) + 1, int(model.nalf),
                       int(model.N if nunk is None else nunk))


def indextable_of(model) -> np.ndarray:
    """The reference's indextable for `model`'s hierarchy: tier 0's empty
    record (a single 0, main/ball2drawer.f90:40-44), then every ADO's sorted
    drawer tuple, 1-based, tier by tier."""
    parts = [np.zeros(1, np.int32)]
    for n in range(1, model.L + 1):
        parts.append((model.tier(n).astype(np.int32) + 1).ravel())
    return np.concatenate(parts)


def _tier_lengths(model) -> np.ndarray:
    return np.diff(np.asarray(model.first[: model.L + 2], np.int64))


# ---------------------------------------------------------------------------
# unformatted sequential records
# ---------------------------------------------------------------------------
def _record(payload: bytes) -> bytes:
    n = len(payload)
    if n >= 2 ** 31:
        raise FileFormatError("record longer than 2 GiB: the reference's 4-byte framing cannot hold it")
    m = np.array([n], _I4).tobytes()
    return m + payload + m


class _Records:
    def __init__(self, data: bytes, path):
        self.data, self.pos, self.path = data, 0, path

    def next(self) -> memoryview:
        d, p = self.data, self.pos
        if p + 4 > len(d):
            raise FileFormatError(f"{self.path}: file ends before a record")
        n = int(np.frombuffer(d, _I4, 1, p)[0])
        if n < 0 or p + 8 + n > len(d):
            raise FileFormatError(f"{self.path}: record of {n} bytes at offset {p} runs past the end")
        tail = int(np.frombuffer(d, _I4, 1, p + 4 + n)[0])
        if tail != n:
            raise FileFormatError(f"{self.path}: record framing {n} / {tail} at offset {p} do not match")
        self.pos = p + 8 + n
        return memoryview(d)[p + 4: p + 4 + n]


def _check_against(h: StateHeader, model, name: str) -> tuple[int, bool]:
    """The checks of main/initialize2.f90:47-67: same norbs, nspin, ncor, nalf,
    more than one tier; returns (ADOs to read, lsmall)."""
    m = header_of(model)
    if (h.norbs, h.nspin, h.ncor, h.nalf) != (m.norbs, m.nspin, m.ncor, m.nalf) or h.ntier <= 1:
        raise FileFormatError(f"error input file <{name}>: header {tuple(h[:5])}, model {tuple(m[:5])}")
    lengths = _tier_lengths(model)
    lnj = int(lengths[: min(h.ntier, m.ntier)].sum())
    small = h.ntier < m.ntier
    if small and lnj != h.nunk:
        raise FileFormatError(f"incompatible <{name}>: {h.ntier} tiers should hold {lnj} ADOs, it holds {h.nunk}")
    if not small and h.nunk < m.nunk:
        raise FileFormatError(f"incompatible <{name}>: {h.nunk} ADOs, the model has {m.nunk}")
    return (h.nunk if small else m.nunk), small


def write_variables(path, x, header: StateHeader) -> None:
    """Write a dense state as ``variables.*`` (main/main.f90:1112-1118).
    x: (nunk, d, d) complex128 with x[a] in column-major memory order (as the
    vectors of this package), a numpy array or a tensor on any device."""
    x = _host_c16(x)
    if x.ndim != 3 or x.shape[1] != x.shape[2] or x.shape[0] != header.nunk:
        raise FileFormatError(f"x must be (nunk={header.nunk}, d, d), got {x.shape}")
    with open(path, "wb") as f:
        f.write(_record(np.array(header[:5], _I4).tobytes()))
        f.write(_record(np.array([header.nunk], _I8).tobytes()))
        for a in range(header.nunk):
            f.write(_record(x[a].tobytes()))


def read_variables(path, model=None, d: int | None = None):
    """Read ``variables.*``; returns (header, x) with x (n, d, d) complex128.
    With `model` the reference's restart rules apply (main/initialize2.f90:
    47-82): the header must match, a file with fewer tiers fills the first
    ADOs and leaves the rest zero, a file with more tiers is cut to the
    model's ADOs; x then has the model's N ADOs."""
    path = Path(path)
    r = _Records(path.read_bytes(), path)
    hdr = np.frombuffer(r.next(), _I4)
    if hdr.size != 5:
        raise FileFormatError(f"{path}: first record holds {hdr.size * 4} bytes, expected 20")
    nu = np.frombuffer(r.next(), _I8)
    if nu.size != 1:
        raise FileFormatError(f"{path}: second record must hold nunk as one 8-byte integer")
    h = StateHeader(*(int(v) for v in hdr), int(nu[0]))
    if model is not None:
        d = model.d
        nread, small = _check_against(h, model, path.name)
        nout = model.N
    else:
        nread = nout = h.nunk
    x = np.zeros((nout, d or 0, d or 0), np.complex128) if d else None
    for a in range(nread):
        rec = np.frombuffer(r.next(), _C16)
        if x is None:
            dd = int(round(np.sqrt(rec.size)))
            if dd * dd != rec.size:
                raise FileFormatError(f"{path}: ADO record of {rec.size} values is not a square block")
            d = dd
            x = np.zeros((nout, d, d), np.complex128)
        if rec.size != d * d:
            raise FileFormatError(f"{path}: ADO {a + 1} holds {rec.size} values, expected {d * d}")
        x[a] = rec.reshape(d, d)
    if x is None:
        x = np.zeros((nout, d or 0, d or 0), np.complex128)
    return h, x


# ---------------------------------------------------------------------------
# form='binary' streams
# ---------------------------------------------------------------------------
def write_rho_spa(path, x_spa, nnz_spa, ind_spa, header: StateHeader) -> None:
    """Write a sparse state as ``rho_spa.*`` (main/main.f90:1093-1104): the
    header, nunk, then each ADO's nnz_spa values (ind_spa: 1-based starts)."""
    xs = _host_c16(x_spa).ravel()
    nnz = np.asarray(nnz_spa, np.int64)
    ind = np.asarray(ind_spa, np.int64)
    if nnz.size != header.nunk or ind.size != header.nunk:
        raise FileFormatError("nnz_spa and ind_spa need nunk entries")
    with open(path, "wb") as f:
        f.write(np.array(header[:5], _I4).tobytes())
        f.write(np.array([header.nunk], _I8).tobytes())
        for a in range(header.nunk):
            if nnz[a] > 0:
                s = int(ind[a]) - 1
                if s < 0 or s + nnz[a] > xs.size:
                    raise FileFormatError(f"ADO {a + 1}: ind_spa / nnz_spa run outside x_spa")
                f.write(xs[s: s + int(nnz[a])].tobytes())


def read_rho_spa(path, nnz_spa, ind_spa, lunk_spa: int, model=None):
    """Read ``rho_spa.*`` with the pattern it was written with
    (fermion/initialize2_spa.f90:80-104); returns (header, x_spa) with x_spa
    (lunk_spa,) complex128.  With `model`: the header checks and the
    fewer-tiers rule of the reference (later ADOs left zero)."""
    path = Path(path)
    data = path.read_bytes()
    if len(data) < 28:
        raise FileFormatError(f"{path}: shorter than its header")
    h = StateHeader(*(int(v) for v in np.frombuffer(data, _I4, 5, 0)), int(np.frombuffer(data, _I8, 1, 20)[0]))
    nnz = np.asarray(nnz_spa, np.int64)
    ind = np.asarray(ind_spa, np.int64)
    nread = h.nunk
    if model is not None:
        nread, _ = _check_against(h, model, path.name)
    if nnz.size < nread or ind.size < nread:
        raise FileFormatError(f"{path}: the pattern covers {nnz.size} ADOs, the file needs {nread}")
    x = np.zeros(int(lunk_spa), np.complex128)
    pos = 28
    for a in range(nread):
        k = int(nnz[a])
        if k <= 0:
            continue
        if pos + 16 * k > len(data):
            raise FileFormatError(f"{path}: ends inside ADO {a + 1}")
        s = int(ind[a]) - 1
        if s < 0 or s + k > x.size:
            raise FileFormatError(f"ADO {a + 1}: ind_spa / nnz_spa run outside lunk_spa")
        x[s: s + k] = np.frombuffer(data, _C16, k, pos)
        pos += 16 * k
    return h, x


def remap_rho_spa(x_spa0, pattern0, pattern1, lunk_spa1: int, d: int) -> np.ndarray:
    """A sparse state moved from the pattern it was saved with to another one
    (fermion/initialize2_spa.f90:42-74, iread_spa = 1): each ADO expanded to
    its d x d block, then the new pattern's elements taken from it.
    pattern = (nnz_spa, ind_spa, irow_spa, icol_spa), 1-based."""
    nnz0, ind0, ir0, ic0 = (np.asarray(v, np.int64) for v in pattern0)
    nnz1, ind1, ir1, ic1 = (np.asarray(v, np.int64) for v in pattern1)
    xs0 = _host_c16(x_spa0).ravel()
    out = np.zeros(int(lunk_spa1), np.complex128)
    for a in range(min(nnz0.size, nnz1.size)):
        blk = np.zeros((d, d), np.complex128)
        if nnz0[a] > 0:
            s0 = int(ind0[a]) - 1
            e = slice(s0, s0 + int(nnz0[a]))
            blk[ir0[e] - 1, ic0[e] - 1] = xs0[e]
        if nnz1[a] > 0:
            s1 = int(ind1[a]) - 1
            e = slice(s1, s1 + int(nnz1[a]))
            out[e] = blk[ir1[e] - 1, ic1[e] - 1]
    return out


def dense_to_spa(x, pattern) -> np.ndarray:
    """The pattern's elements of a dense state (x[a] column-major)."""
    nnz, ind, ir, ic = (np.asarray(v, np.int64) for v in pattern)
    xd = _host_c16(x)
    out = np.zeros(ir.size, np.complex128)
    for a in range(nnz.size):
        if nnz[a] > 0:
            e = slice(int(ind[a]) - 1, int(ind[a]) - 1 + int(nnz[a]))
            out[e] = xd[a].reshape(-1)[(ir[e] - 1) + (ic[e] - 1) * xd.shape[1]]
    return out


def spa_to_dense(x_spa, pattern, n_ado: int, d: int) -> np.ndarray:
    """A dense state (x[a] column-major, zero off the pattern) from a sparse one."""
    nnz, ind, ir, ic = (np.asarray(v, np.int64) for v in pattern)
    xs = _host_c16(x_spa).ravel()
    out = np.zeros((n_ado, d * d), np.complex128)
    for a in range(min(n_ado, nnz.size)):
        if nnz[a] > 0:
            e = slice(int(ind[a]) - 1, int(ind[a]) - 1 + int(nnz[a]))
            out[a, (ir[e] - 1) + (ic[e] - 1) * d] = xs[e]
    return out.reshape(n_ado, d, d)


def write_indextable(path, indextable) -> None:
    """``indextable.sav`` (main/main.f90:1657-1661)."""
    t = np.ascontiguousarray(indextable, _I4)
    with open(path, "wb") as f:
        f.write(np.array([t.size], _I8).tobytes())
        f.write(t.tobytes())


def read_indextable(path) -> np.ndarray:
    data = Path(path).read_bytes()
    if len(data) < 8:
        raise FileFormatError(f"{path}: shorter than its length field")
    n = int(np.frombuffer(data, _I8, 1, 0)[0])
    if n < 0 or len(data) != 8 + 4 * n:
        raise FileFormatError(f"{path}: {len(data)} bytes do not hold {n} int32 entries")
    return np.frombuffer(data, _I4, n, 8).copy()


_COEF_DT = np.dtype([("indexcoef", "<i8"), ("itypecoef", "i1"), ("ifactfcoef", "i1"), ("ifactrcoef", "i1")])


def write_coef_tables(coefindex_path, auxindex_path, tables: dict, header: StateHeader, numfff: int = 0) -> None:
    """coefindex.data and auxindex.data for the default flags (no adiabatic
    modes, filter, terminator, simple or screened hierarchy, omission): the
    reference reads them back when the header matches.  tables: indexcoef,
    itypecoef, ifactfcoef, ifactrcoef (lall each) and index_coef_ref (nunk),
    the module arrays of module/matmod.f90:7-9."""
    ic = np.asarray(tables["indexcoef"], np.int64)
    rec = np.empty(ic.size, _COEF_DT)
    rec["indexcoef"] = ic
    for k in ("itypecoef", "ifactfcoef", "ifactrcoef"):
        v = np.asarray(tables[k])
        if v.size != ic.size:
            raise FileFormatError(f"{k} needs lall = {ic.size} entries")
        rec[k] = v.astype(np.int8)
    ref = np.asarray(tables["index_coef_ref"], np.int64)
    if ref.size != header.nunk:
        raise FileFormatError(f"index_coef_ref needs nunk = {header.nunk} entries")
    f4 = lambda *v: np.array(v, _I4).tobytes()
    with open(coefindex_path, "wb") as f:
        f.write(f4(header.ntier, header.ncor, header.norbs, header.nspin, header.nalf, numfff, header.ntier, 0, 0))
        f.write(f4(0, 0, 0, 0, 0) + np.array([0.0], "<f8").tobytes())   # lad, lad_fast, lset_fast, idegen, ndegen, dratio
        f.write(f4(0, 0, 0))                                          # lfilter, nfilter_count, nfilter_long
        f.write(f4(0))                                                # ltrun_der
        f.write(f4(0, 0))                                             # lsimple, lwalf
        f.write(f4(0))                                                # lscreen
        f.write(f4(0) + np.array([0.0], "<f8").tobytes())            # lomit, dfact_omit
        f.write(np.array([ic.size], _I8).tobytes())
        f.write(rec.tobytes())
    with open(auxindex_path, "wb") as f:
        f.write(np.array([ref.size], _I8).tobytes())
        f.write(ref.tobytes())


def read_coef_tables(coefindex_path, auxindex_path, model=None):
    """Read coefindex.data and auxindex.data; returns (header dict, tables
    dict).  Files written with any of the non-default flags are refused (the
    loader takes the default-flag tables only).  With `model` the header
    checks of main/buildcoefindex.f90:56-100 and the nunk check of auxindex
    apply."""
    data = Path(coefindex_path).read_bytes()
    if len(data) < 36 + 28 + 12 + 4 + 8 + 4 + 12 + 8:
        raise FileFormatError(f"{coefindex_path}: shorter than its header")
    h1 = [int(v) for v in np.frombuffer(data, _I4, 9, 0)]
    hdr = dict(zip(("ntier", "ncor", "norbs", "nspin", "nalf", "numfff", "ntier0", "ndrawer_slow", "ncor_slow"), h1))
    flags = dict(lad=np.frombuffer(data, _I4, 1, 36)[0], lfilter=np.frombuffer(data, _I4, 1, 64)[0],
                 ltrun_der=np.frombuffer(data, _I4, 1, 76)[0], lsimple=np.frombuffer(data, _I4, 1, 80)[0],
                 lwalf=np.frombuffer(data, _I4, 1, 84)[0], lscreen=np.frombuffer(data, _I4, 1, 88)[0],
                 lomit=np.frombuffer(data, _I4, 1, 92)[0])
    on = [k for k, v in flags.items() if v != 0]
    if on:
        raise FileFormatError(f"{coefindex_path}: written with {', '.join(on)}; the default-flag tables only")
    if model is not None:
        m = header_of(model)
        want = dict(ntier=m.ntier, ncor=m.ncor, norbs=m.norbs, nspin=m.nspin, nalf=m.nalf, numfff=0)
        bad = {k: hdr[k] for k in want if hdr[k] != want[k]}
        if bad or hdr["ntier0"] != m.ntier:
            raise FileFormatError(f"{coefindex_path}: header {hdr} does not match the model {want}")
    lall = int(np.frombuffer(data, _I8, 1, 104)[0])
    if len(data) != 112 + 11 * lall:
        raise FileFormatError(f"{coefindex_path}: {len(data) - 112} bytes do not hold {lall} 11-byte entries")
    rec = np.frombuffer(data, _COEF_DT, lall, 112)
    aux = Path(auxindex_path).read_bytes()
    nunk = int(np.frombuffer(aux, _I8, 1, 0)[0])
    if len(aux) != 8 + 8 * nunk:
        raise FileFormatError(f"{auxindex_path}: {len(aux) - 8} bytes do not hold {nunk} int64 entries")
    if model is not None and nunk != model.N:
        raise FileFormatError(f"{auxindex_path}: nunk {nunk}, the model has {model.N}")
    hdr["lall"], hdr["nunk"] = lall, nunk
    tables = dict(indexcoef=rec["indexcoef"].copy(), itypecoef=rec["itypecoef"].copy(),
                  ifactfcoef=rec["ifactfcoef"].copy(), ifactrcoef=rec["ifactrcoef"].copy(),
                  index_coef_ref=np.frombuffer(aux, _I8, nunk, 8).copy())
    return hdr, tables


class ResumeHeader(NamedTuple):
    t: float
    ntier: int
    norbs: int
    nspin: int
    ncor: int
    nalf: int
    numfff: int
    nunk: int
    lhb: bool
    hb: tuple = ()          # (npade_hb, ndrude_hb, nmode_hb, itype_hb) when lhb
    lunk_spa: int = -1      # sparse checkpoints: the COO length; -1 dense


def resume_name(t: float, dt: float, nresume: int, nfbase: int = 0) -> str:
    """The checkpoint file name td/resumejob.f90:81-84 forms at time t."""
    n = nfbase + int((t + 1e-9) / dt) // nresume
    return f"TAPE_{n}.resume"


def write_resume(path, t: float, x, model=None, header: StateHeader | None = None, numfff: int = 0,
                 sparse: bool = False) -> None:
    """A propagation checkpoint as td/resumejob.f90:85-101 writes it: x is the
    dense state (nunk, d, d) or, with sparse=True, rho_spa (lunk_spa,).
    The header comes from `model` (or `header`); lhb is false."""
    h = header_of(model) if header is None else header
    xs = _host_c16(x)
    with open(path, "wb") as f:
        f.write(np.array([t], "<f8").tobytes())
        f.write(np.array([h.ntier, h.norbs, h.nspin, h.ncor, h.nalf, numfff], _I4).tobytes())
        f.write(np.array([h.nunk], _I8).tobytes())
        f.write(np.array([0], _I4).tobytes())   # lhb = .false.
        if sparse:
            v = xs.ravel()
            f.write(np.array([v.size], _I8).tobytes())
            f.write(v.tobytes())
        else:
            if xs.ndim != 3 or xs.shape[0] != h.nunk:
                raise FileFormatError(f"x must be (nunk={h.nunk}, d, d), got {xs.shape}")
            f.write(xs.tobytes())


def read_resume(path, model=None, sparse: bool = False, d: int | None = None, lunk_spa: int | None = None):
    """Read a checkpoint (td/resumejob.f90:17-75); returns (ResumeHeader, x).
    With `model` the checks of :26-47 apply: same ntier, norbs, nspin, ncor,
    nalf and nunk (numfff must be 0 here), and for sparse files lunk_spa
    when given.  Dense files need the block size d (from `model` or `d`)."""
    path = Path(path)
    data = path.read_bytes()
    if len(data) < 8 + 24 + 8 + 4:
        raise FileFormatError(f"{path}: shorter than its header")
    t = float(np.frombuffer(data, "<f8", 1, 0)[0])
    ntier, norbs, nspin, ncor, nalf, numfff = (int(v) for v in np.frombuffer(data, _I4, 6, 8))
    nunk = int(np.frombuffer(data, _I8, 1, 32)[0])
    lhb = int(np.frombuffer(data, _I4, 1, 40)[0]) != 0
    pos, hb = 44, ()
    if lhb:
        hb = tuple(int(v) for v in np.frombuffer(data, _I4, 4, pos))
        pos += 16
    if model is not None:
        m = header_of(model)
        if (ntier, norbs, nspin, ncor, nalf, numfff) != (m.ntier, m.norbs, m.nspin, m.ncor, m.nalf, 0) or lhb:
            raise FileFormatError(f"bad {path.name} to read 1: {(ntier, norbs, nspin, ncor, nalf, numfff)}")
        if nunk != m.nunk:
            raise FileFormatError(f"bad {path.name} to read 2: nunk {nunk}, model {m.nunk}")
        d = model.d
    lsp = -1
    if sparse:
        lsp = int(np.frombuffer(data, _I8, 1, pos)[0])
        pos += 8
        if lunk_spa is not None and lsp != lunk_spa:
            raise FileFormatError(f"bad {path.name} to read 3: lunk_spa {lsp}, expected {lunk_spa}")
        if len(data) != pos + 16 * lsp:
            raise FileFormatError(f"{path}: {len(data) - pos} bytes do not hold {lsp} complex values")
        x = np.frombuffer(data, _C16, lsp, pos).copy()
    else:
        if d is None:
            raise FileFormatError("dense checkpoints need the block size d (or a model)")
        if len(data) != pos + 16 * nunk * d * d:
            raise FileFormatError(f"{path}: {len(data) - pos} bytes do not hold {nunk} blocks of {d}x{d}")
        x = np.frombuffer(data, _C16, nunk * d * d, pos).reshape(nunk, d, d).copy()
    return ResumeHeader(t, ntier, norbs, nspin, ncor, nalf, numfff, nunk, lhb, hb, lsp), x


def _host_c16(x) -> np.ndarray:
    if hasattr(x, "detach"):   # a torch tensor, on any device
        x = x.detach().to("cpu").numpy()
    return np.ascontiguousarray(x, np.complex128)
