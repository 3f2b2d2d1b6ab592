"""The reference's state and table files (heomq.files): byte layout, the
restart rules of main/initialize2.f90:43-82 and fermion/initialize2_spa.f90,
and the layout a Fortran compiler itself gives these records (amdflang, when
the toolchain has it)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import heomq
from heomq import files as F
from oracle import heom_oracle as O
from oracle.ref_tables import reference_arrays


def _state(n, d, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, d, d)) + 1j * rng.normal(size=(n, d, d))


def test_variables_round_trip_and_framing(tmp_path):
    x = _state(5, 4, 1)
    h = F.StateHeader(1, 2, 3, 3, 2, 5)
    p = tmp_path / "variables.st"
    F.write_variables(p, x, h)
    raw = p.read_bytes()
    i4 = np.frombuffer(raw, np.int32)
    # record 1: 20 bytes of five int32, framed; record 2: 8 bytes of int64
    assert list(i4[:7]) == [20, 1, 2, 3, 3, 2, 20]
    assert i4[7] == 8 and np.frombuffer(raw[32:40], np.int64)[0] == 5 and i4[10] == 8
    assert i4[11] == 16 * 16   # first ADO record: 4x4 complex*16
    assert len(raw) == 28 + 16 + 5 * (8 + 256)
    # ADO records hold rho(:,:,a) column-major: the memory of x[a]
    assert np.array_equal(np.frombuffer(raw[48:48 + 256], np.complex128), x[0].ravel())
    h2, x2 = F.read_variables(p)
    assert h2 == h and np.array_equal(x2, x)


def test_variables_rejects_broken_framing(tmp_path):
    p = tmp_path / "variables.sav"
    F.write_variables(p, _state(2, 2, 2), F.StateHeader(1, 2, 3, 2, 2, 2))
    raw = bytearray(p.read_bytes())
    raw[-1] ^= 1   # last trailer
    p.write_bytes(bytes(raw))
    with pytest.raises(F.FileFormatError):
        F.read_variables(p)
    p.write_bytes(bytes(raw[:-20]))
    with pytest.raises(F.FileFormatError):
        F.read_variables(p)


def test_restart_rules_follow_initialize2(tmp_path):
    """Fewer tiers: the first ADOs filled, the rest zero; more tiers: cut to
    the model; a different bath or leads: refused (initialize2.f90:47-82)."""
    small = heomq.Model(heomq.siam(nk=2, ntop=1))
    big = heomq.Model(heomq.siam(nk=2, ntop=2))
    xs = _state(small.N, small.d, 3)
    p = tmp_path / "variables.sav"
    F.write_variables(p, xs, F.header_of(small))
    h, x = F.read_variables(p, model=big)
    assert h.ntier == 2 and x.shape == (big.N, big.d, big.d)
    assert np.array_equal(x[: small.N], xs) and not np.any(x[small.N:])
    # the tiers the small hierarchy holds are the big one's first ADOs
    assert np.array_equal(F.indextable_of(small), F.indextable_of(big)[: F.indextable_of(small).size])
    xb = _state(big.N, big.d, 4)
    F.write_variables(p, xb, F.header_of(big))
    h, x = F.read_variables(p, model=small)
    assert np.array_equal(x, xb[: small.N])
    other = heomq.Model(heomq.siam(nk=3, ntop=2))   # one more Pade term: ncor differs
    with pytest.raises(F.FileFormatError):
        F.read_variables(p, model=other)
    F.write_variables(p, xs[:3], F.StateHeader(*F.header_of(small)[:5], 3))   # ADO count of no tier
    with pytest.raises(F.FileFormatError):
        F.read_variables(p, model=big)


def _pattern(n_ado, d, seed, density=0.5):
    rng = np.random.default_rng(seed)
    nnz, ind, ir, ic = [], [], [], []
    pos = 1
    for a in range(n_ado):
        mask = rng.random((d, d)) < density
        if a == 0:
            mask[:] = True
        j, i = np.nonzero(mask.T)   # (column, row) pairs in column-major element order
        nnz.append(len(i))
        ind.append(pos)
        ir.extend((i + 1).tolist())
        ic.extend((j + 1).tolist())
        pos += len(i)
    return (np.array(nnz, np.int32), np.array(ind, np.int64), np.array(ir, np.int32), np.array(ic, np.int32)), pos - 1


def test_rho_spa_round_trip_remap_and_dense(tmp_path):
    n, d = 7, 4
    pat0, lunk0 = _pattern(n, d, 5)
    pat1, lunk1 = _pattern(n, d, 6, density=0.7)
    x = _state(n, d, 7)
    xs0 = F.dense_to_spa(x, pat0)
    assert xs0.size == lunk0
    # dense -> sparse -> dense keeps exactly the pattern's elements
    back = F.spa_to_dense(xs0, pat0, n, d)
    for a in range(n):
        s = slice(pat0[1][a] - 1, pat0[1][a] - 1 + pat0[0][a])
        m = np.zeros(d * d, bool)
        m[(pat0[2][s] - 1) + (pat0[3][s] - 1) * d] = True
        assert np.array_equal(back[a].ravel()[m], x[a].ravel()[m]) and not np.any(back[a].ravel()[~m])
    h = F.StateHeader(1, 2, 3, 3, 2, n)
    p = tmp_path / "rho_spa.st"
    F.write_rho_spa(p, xs0, pat0[0], pat0[1], h)
    raw = p.read_bytes()
    assert len(raw) == 20 + 8 + 16 * lunk0   # no record framing
    assert list(np.frombuffer(raw[:20], np.int32)) == [1, 2, 3, 3, 2]
    h2, xr = F.read_rho_spa(p, pat0[0], pat0[1], lunk0)
    assert h2 == h and np.array_equal(xr, xs0)
    # a state saved with one pattern, read into another (iread_spa = 1)
    xs1 = F.remap_rho_spa(xr, pat0, pat1, lunk1, d)
    assert np.array_equal(xs1, F.dense_to_spa(F.spa_to_dense(xs0, pat0, n, d), pat1))


def test_indextable_matches_reference_tables(tmp_path):
    m = heomq.Model(heomq.siam(nk=2, ntop=3))
    o = O.Model(O.siam_params(nk=2, ntop=3))
    t = F.indextable_of(m)
    assert np.array_equal(t, reference_arrays(o)["indextable"])
    p = tmp_path / "indextable.sav"
    F.write_indextable(p, t)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:8], np.int64)[0] == t.size and len(raw) == 8 + 4 * t.size
    assert np.array_equal(F.read_indextable(p), t)


FORTRAN = """
program files_check
implicit none
integer :: norbs, nspin, ncor, ntier, nalf, ni, nj
integer*8 :: nunk, lni
complex*16 :: rho(3,3,4), acc
! the reference's writes (main/main.f90), with known values
norbs = 1; nspin = 2; ncor = 5; ntier = 3; nalf = 2; nunk = 4
do lni = 1, nunk
  do nj = 1, 3
    do ni = 1, 3
      rho(ni,nj,lni) = dcmplx(dble(ni + 10*nj), dble(lni) - 0.5d0)
    end do
  end do
end do
open(unit=25, file='variables.st', form='unformatted', status='unknown')
rewind(25)
write(25)norbs, nspin, ncor, ntier, nalf
write(25)nunk
do lni=1,nunk
   write(25)((rho(ni,nj,lni), ni=1,3), nj=1,3)
end do
close(25)
open(unit=51, file='rho_spa.st', form='binary', status='unknown')
rewind(51)
write(51)norbs, nspin, ncor, ntier, nalf
write(51)nunk
write(51)(rho(ni,1,1), ni=1,3)
write(51)(rho(ni,2,3), ni=2,3)
close(51)
! the reference's read (main/initialize2.f90) of a file written by Python
open(unit=15, file='variables.sav', form='unformatted', status='old')
rewind(15)
read(15)norbs, nspin, ncor, ntier, nalf
read(15)nunk
acc = (0.d0, 0.d0)
do lni=1,nunk
   read(15)((rho(ni,nj,lni), ni=1,3), nj=1,3)
   do nj = 1, 3
     do ni = 1, 3
       acc = acc + rho(ni,nj,lni) * dble(ni + 3*(nj-1) + 9*(lni-1))
     end do
   end do
end do
close(15)
write(6,'(5i6,i6)') norbs, nspin, ncor, ntier, nalf, nunk
write(6,'(2es25.16)') dble(acc), dimag(acc)
end program
"""


def _flang():
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang"), shutil.which("gfortran")):
        if c and Path(c).exists():
            return c
    return None


@pytest.mark.skipif(_flang() is None, reason="no Fortran compiler in this toolchain")
def test_layout_matches_a_fortran_compiler(tmp_path):
    """Files written by Fortran's own unformatted / binary I/O read back here,
    and a file written here reads back in Fortran."""
    xs = _state(2, 3, 9)
    F.write_variables(tmp_path / "variables.sav", xs, F.StateHeader(1, 2, 5, 2, 2, 2))
    (tmp_path / "files_check.f90").write_text(FORTRAN)
    subprocess.run([_flang(), "-O0", "files_check.f90", "-o", "files_check"], cwd=tmp_path, check=True,
                   capture_output=True)
    out = subprocess.run(["./files_check"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out[:6]] == [1, 2, 5, 2, 2, 2]
    # weight ni + 3 (nj - 1) + 9 (lni - 1): one plus the element's column-major offset in the file
    acc = np.sum(xs.reshape(2, 9) * (np.arange(18).reshape(2, 9) + 1))
    assert abs(float(out[6]) - acc.real) <= 1e-12 * np.abs(xs).sum() * 18
    assert abs(float(out[7]) - acc.imag) <= 1e-12 * np.abs(xs).sum() * 18
    rho = np.zeros((4, 3, 3), np.complex128)   # rho[lni-1] as x[a]: memory column-major
    for lni in range(1, 5):
        for nj in range(1, 4):
            for ni in range(1, 4):
                rho[lni - 1].reshape(-1)[(ni - 1) + 3 * (nj - 1)] = complex(ni + 10 * nj, lni - 0.5)
    h, x = F.read_variables(tmp_path / "variables.st")
    assert h == F.StateHeader(1, 2, 5, 3, 2, 4) and np.array_equal(x, rho)
    nnz, ind = np.array([3, 0, 2, 0]), np.array([1, 4, 4, 6])
    h, xsp = F.read_rho_spa(tmp_path / "rho_spa.st", nnz, ind, 5)
    assert h.nunk == 4
    assert np.array_equal(xsp[:3], rho[0].reshape(-1)[0:3])
    assert np.array_equal(xsp[3:5], rho[2].reshape(-1)[4:6])


def test_resume_round_trip_and_checks(tmp_path):
    m = heomq.Model(heomq.siam(nk=2, ntop=2))
    x = _state(m.N, m.d, 13)
    p = tmp_path / F.resume_name(2.5, 0.01, 50)
    assert p.name == "TAPE_5.resume"   # int((2.5 + 1e-9) / 0.01) / 50 = 250 / 50
    F.write_resume(p, 2.5, x, model=m)
    raw = p.read_bytes()
    assert len(raw) == 8 + 24 + 8 + 4 + 16 * m.N * m.d * m.d
    h, x2 = F.read_resume(p, model=m)
    assert h.t == 2.5 and h.nunk == m.N and not h.lhb and np.array_equal(x2, x)
    other = heomq.Model(heomq.siam(nk=2, ntop=3))
    with pytest.raises(F.FileFormatError):
        F.read_resume(p, model=other)
    xs = x.reshape(-1)[:37]
    F.write_resume(p, 0.75, xs, model=m, sparse=True)
    h, xr = F.read_resume(p, model=m, sparse=True, lunk_spa=37)
    assert h.lunk_spa == 37 and h.t == 0.75 and np.array_equal(xr, xs)
    with pytest.raises(F.FileFormatError):
        F.read_resume(p, model=m, sparse=True, lunk_spa=36)


FORTRAN_RESUME = """
program resume_check
implicit none
integer :: ntier, norbs, nspin, ncor, nalf, numfff, ni, nj
integer*8 :: nunk, lni
logical :: lhb
real*8 :: tt
complex*16 :: rho(2,2,3), acc
! the checkpoint as td/resumejob.f90 writes it
tt = 1.25d0; ntier = 3; norbs = 1; nspin = 2; ncor = 4; nalf = 2; numfff = 0; nunk = 3; lhb = .false.
do lni = 1, nunk
  do nj = 1, 2
    do ni = 1, 2
      rho(ni,nj,lni) = dcmplx(dble(ni + 2*nj), -dble(lni))
    end do
  end do
end do
open(unit=23, file='TAPE_7.resume', form='binary', status='unknown')
rewind(23)
write(23)tt
write(23)ntier, norbs, nspin, ncor, nalf, numfff
write(23)nunk
write(23)lhb
do lni=1,nunk
  write(23)((rho(ni,nj,lni), ni=1,2), nj=1,2)
end do
close(23)
! and read as it reads TAPE.resume (written here by Python)
open(unit=23, file='TAPE.resume', form='binary', status='old')
read(23)tt
read(23)ntier, norbs, nspin, ncor, nalf, numfff
read(23)nunk
read(23)lhb
acc = (0.d0, 0.d0)
do lni=1,nunk
  read(23)((rho(ni,nj,lni), ni=1,2), nj=1,2)
  do nj = 1, 2
    do ni = 1, 2
      acc = acc + rho(ni,nj,lni) * dble(ni + 2*(nj-1) + 4*(lni-1))
    end do
  end do
end do
close(23)
write(6,'(f10.4,6i4,i6,l2)') tt, ntier, norbs, nspin, ncor, nalf, numfff, nunk, lhb
write(6,'(2es25.16)') dble(acc), dimag(acc)
end program
"""


@pytest.mark.skipif(_flang() is None, reason="no Fortran compiler in this toolchain")
def test_resume_layout_matches_a_fortran_compiler(tmp_path):
    xs = _state(3, 2, 14)
    F.write_resume(tmp_path / "TAPE.resume", 0.5, xs, header=F.StateHeader(1, 2, 4, 3, 2, 3))
    (tmp_path / "resume_check.f90").write_text(FORTRAN_RESUME)
    subprocess.run([_flang(), "-O0", "resume_check.f90", "-o", "resume_check"], cwd=tmp_path, check=True,
                   capture_output=True)
    out = subprocess.run(["./resume_check"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.split()
    assert float(out[0]) == 0.5 and [int(v) for v in out[1:8]] == [3, 1, 2, 4, 2, 0, 3] and out[8] == "F"
    acc = np.sum(xs.reshape(3, 4) * (np.arange(12).reshape(3, 4) + 1))
    tol = 1e-12 * np.abs(xs).sum() * 12
    assert abs(float(out[9]) - acc.real) <= tol and abs(float(out[10]) - acc.imag) <= tol
    h, x = F.read_resume(tmp_path / "TAPE_7.resume", d=2)
    assert (h.t, h.ntier, h.norbs, h.nspin, h.ncor, h.nalf, h.numfff, h.nunk, h.lhb) == (1.25, 3, 1, 2, 4, 2, 0, 3, False)
    want = np.zeros((3, 2, 2), complex)
    for lni in range(1, 4):
        for nj in range(1, 3):
            for ni in range(1, 3):
                want[lni - 1].reshape(-1)[(ni - 1) + 2 * (nj - 1)] = complex(ni + 2 * nj, -lni)
    assert np.array_equal(x, want)


def test_coef_tables_round_trip_into_the_loader(tmp_path):
    """coefindex.data / auxindex.data written from the reference's module
    arrays, read back, and loaded: the same hierarchy and couplings as the
    model built here."""
    # the reference's hierarchy keeps repeated drawers (lsdraw), so every
    # upper-tier coupling of its table has a neighbour
    o = O.Model(O.siam_params(nk=2, ntop=2, lsdraw=True))
    arr = reference_arrays(o)
    m = heomq.Model(heomq.siam(nk=2, ntop=2, lsdraw=True))
    F.write_coef_tables(tmp_path / "coefindex.data", tmp_path / "auxindex.data", arr, F.header_of(m))
    raw = (tmp_path / "coefindex.data").read_bytes()
    lall = len(arr["indexcoef"])
    assert len(raw) == 112 + 11 * lall and np.frombuffer(raw[104:112], np.int64)[0] == lall
    hdr, tab = F.read_coef_tables(tmp_path / "coefindex.data", tmp_path / "auxindex.data", model=m)
    assert hdr["nunk"] == m.N and hdr["ntier"] == m.L + 1
    for k in tab:
        assert np.array_equal(tab[k], np.asarray(arr[k]).astype(tab[k].dtype)), k
    arr2 = dict(arr, **tab)
    m2 = heomq.Model.from_reference(**arr2)
    lo1, up1 = m.coupling()
    lo2, up2 = m2.coupling()
    assert np.array_equal(lo1, lo2) and np.array_equal(up1, up2)
    with pytest.raises(F.FileFormatError):
        F.read_coef_tables(tmp_path / "coefindex.data", tmp_path / "auxindex.data",
                           model=heomq.Model(heomq.siam(nk=3, ntop=2, lsdraw=True)))


FORTRAN_COEF = """
program coef_check
implicit none
integer :: ntier, ncor, norbs, nspin, nalf, numfff, ntier0, ndrawer_slow, ncor_slow
integer :: idegen_fast, ndegen_fast, nfilter_count, nfilter_long
logical :: lad, lad_fast, lset_fast, lfilter, ltrun_der, lsimple, lwalf, lscreen, lomit
real*8 :: dratio_fast, dfact_omit
integer*8 :: lall, lni, nunk
integer*8 :: indexcoef(3), index_coef_ref(2)
integer*1 :: itypecoef(3), ifactfcoef(3), ifactrcoef(3)
ntier = 3; ncor = 4; norbs = 1; nspin = 2; nalf = 2; numfff = 0; ntier0 = 3; ndrawer_slow = 0; ncor_slow = 0
lad = .false.; lad_fast = .false.; lset_fast = .false.; idegen_fast = 0; ndegen_fast = 0; dratio_fast = 0.d0
lfilter = .false.; nfilter_count = 0; nfilter_long = 0; ltrun_der = .false.; lsimple = .false.; lwalf = .false.
lscreen = .false.; lomit = .false.; dfact_omit = 0.d0
lall = 3; nunk = 2
indexcoef = (/ 7_8, 123456789012_8, 1_8 /)
itypecoef = (/ 0_1, 1_1, 0_1 /); ifactfcoef = (/ 1_1, -1_1, 1_1 /); ifactrcoef = (/ -1_1, 1_1, 1_1 /)
index_coef_ref = (/ 1_8, 4_8 /)
! the writes of main/buildcoefindex.f90
open(unit=17, file='coefindex.data', form='binary', status='unknown')
rewind(17)
write(17)ntier, ncor, norbs, nspin, nalf, numfff, ntier0, ndrawer_slow, ncor_slow
write(17)lad, lad_fast, lset_fast, idegen_fast, ndegen_fast, dratio_fast
write(17)lfilter, nfilter_count, nfilter_long
write(17)ltrun_der
write(17)lsimple, lwalf
write(17)lscreen
write(17)lomit, dfact_omit
write(17)lall
do lni=1,lall
   write(17)indexcoef(lni), itypecoef(lni), ifactfcoef(lni), ifactrcoef(lni)
end do
close(17)
open(unit=37, file='auxindex.data', form='binary', status='unknown')
rewind(37)
write(37)nunk
do lni=1,nunk
   write(37)index_coef_ref(lni)
end do
close(37)
end program
"""


@pytest.mark.skipif(_flang() is None, reason="no Fortran compiler in this toolchain")
def test_coef_tables_layout_matches_a_fortran_compiler(tmp_path):
    (tmp_path / "coef_check.f90").write_text(FORTRAN_COEF)
    subprocess.run([_flang(), "-O0", "coef_check.f90", "-o", "coef_check"], cwd=tmp_path, check=True,
                   capture_output=True)
    subprocess.run(["./coef_check"], cwd=tmp_path, check=True, capture_output=True)
    hdr, tab = F.read_coef_tables(tmp_path / "coefindex.data", tmp_path / "auxindex.data")
    assert (hdr["ntier"], hdr["ncor"], hdr["lall"], hdr["nunk"]) == (3, 4, 3, 2)
    assert list(tab["indexcoef"]) == [7, 123456789012, 1]
    assert list(tab["itypecoef"]) == [0, 1, 0] and list(tab["ifactfcoef"]) == [1, -1, 1]
    assert list(tab["ifactrcoef"]) == [-1, 1, 1] and list(tab["index_coef_ref"]) == [1, 4]
    # and the bytes this package writes for the same tables are the compiler's
    F.write_coef_tables(tmp_path / "c2.data", tmp_path / "a2.data", tab, F.StateHeader(1, 2, 4, 3, 2, 2))
    assert (tmp_path / "c2.data").read_bytes() == (tmp_path / "coefindex.data").read_bytes()
    assert (tmp_path / "a2.data").read_bytes() == (tmp_path / "auxindex.data").read_bytes()
