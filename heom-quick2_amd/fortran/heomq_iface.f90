!
! Drop-in Fortran layer over libheomq (ISO_C_BINDING).
!
! The reference calls its Liouvillian apply through fixed subroutine
! signatures from its solvers (tfqmr_double/solve_tfqmr.f90:127,140,202,204,
! 267,269,...; bicg/solvebicg.f90:69,98,119; td/intstp.f90:32-93).  This file
! exports the same names and argument meanings, backed by the MI355X engine:
!
!   calc_ax_tfqmr_omp(dim0, rcgtmp, rcgrhs)   tfqmr_double/calc_ax_tfqmr_omp.f90:1,15-17
!   calc_ax_bicg_omp(dim0, rcgtmp, rcgrhs)    bicg/calc_ax_bicg_omp.f90:1,16-17
!   calc_atrx_bicg_omp(dim0, rcgtmp, rcgrhs)  bicg/calc_atrx_bicg_omp.f90:1,35-36
!   calcderiv_omp(tt)                         td/calcderiv_omp.f90:1-7 (module arrays
!                                             rhotmp -> rhorhs, registered once; the
!                                             lead field of heomq_fortran_set_field at tt)
!   solve_tfqmr_gpu(rho, crit, maxit, info)   solver-level replacement of solve_tfqmr
!   solve_bicg_gpu(rho, crit, maxit, info)    solver-level replacement of solvebicg
!   calc_ax_jacobi_omp(dim0, rcgtmp, rcgrhs)  jacobi/calc_ax_jacobi_omp.f90:1,15-17
!   calcderiv_spa_omp(tt)                     td/calcderiv_spa_omp.f90:1-11 (module
!                                             arrays rhotmp_spa -> rhorhs_spa, registered
!                                             once by heomq_register_td_spa_arrays)
!   calc_ax_tfqmr_spa_omp(ldim, rcgtmp, rcgrhs)
!                                             tfqmr_double/calc_ax_tfqmr_spa_omp.f90:1,16-18
!                                             (pattern of module sparsemod registered
!                                             once by heomq_register_spa_pattern)
!   solve_jacobi_gpu(rho, crit, maxit, info)  solver-level replacement of solve_jacobi
!
! heomq_fortran_init uploads the reference's own module arrays once (after
! buildcoefindex), as heomq_model_from_reference + heomq_ctx_create.  Errors
! print to unit 6 and stop (ERROR STOP, so the exit status is nonzero), as
! the reference's own checks do
! (e.g. calc_ax_tfqmr_omp.f90:44-48).
!
module heomq_iface
  use iso_c_binding
  implicit none
  type(c_ptr), save :: heomq_model_h = c_null_ptr
  type(c_ptr), save :: heomq_ctx_h = c_null_ptr
  type(c_ptr), save :: td_rhotmp = c_null_ptr, td_rhorhs = c_null_ptr
  type(c_ptr), save :: td_rhotmp_spa = c_null_ptr, td_rhorhs_spa = c_null_ptr

  interface
    function c_last_error() bind(C, name='heomq_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    integer(c_int) function c_model_from_reference(nrho, nunk, ntier, norbs, nspin, nalf, ncor, igs, &
        ntable, lall, indextable, nfirst, nlast, ifirst, indexcoef, itypecoef, ifactfcoef, ifactrcoef, index_coef_ref, &
        indexdraw, cb, cd, cgamma, dipm, ilead, mspin, morbs, mpm, eleadinfty, hs, amsall, dbsqrt, out) &
        bind(C, name='heomq_model_from_reference')
      import :: c_int, c_int32_t, c_int64_t, c_int8_t, c_double, c_double_complex, c_ptr
      integer(c_int32_t), value :: nrho, ntier, norbs, nspin, nalf, ncor, igs
      integer(c_int64_t), value :: nunk, ntable, lall
      integer(c_int32_t) :: indextable(*), indexdraw(*), ilead(*), mspin(*), morbs(*), mpm(*)
      integer(c_int64_t) :: nfirst(*), nlast(*), ifirst(*), indexcoef(*), index_coef_ref(*)
      integer(c_int8_t) :: itypecoef(*), ifactfcoef(*), ifactrcoef(*)
      complex(c_double_complex) :: cb(*), cd(*), cgamma(*), hs(*)
      real(c_double) :: dipm(*), eleadinfty(*), amsall(*)
      type(c_ptr), value :: dbsqrt
      type(c_ptr) :: out
    end function

    integer(c_int) function c_ctx_create(m, device, out) bind(C, name='heomq_ctx_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: m
      integer(c_int), value :: device
      type(c_ptr) :: out
    end function

    integer(c_int) function c_calc_ax(ctx, dim0, x, y) bind(C, name='heomq_calc_ax_tfqmr')
      import :: c_int, c_int32_t, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dim0
      complex(c_double_complex) :: x(*), y(*)
    end function

    integer(c_int) function c_calc_ax_bicg(ctx, dim0, x, y) bind(C, name='heomq_calc_ax_bicg')
      import :: c_int, c_int32_t, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dim0
      complex(c_double_complex) :: x(*), y(*)
    end function

    integer(c_int) function c_calc_atrx_bicg(ctx, dim0, x, y) bind(C, name='heomq_calc_atrx_bicg')
      import :: c_int, c_int32_t, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dim0
      complex(c_double_complex) :: x(*), y(*)
    end function

    integer(c_int) function c_solve_bicg_host(ctx, rho, crit, maxit, stats) bind(C, name='heomq_solve_bicg_host')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: rho(*)
      real(c_double), value :: crit
      integer(c_int), value :: maxit
      real(c_double) :: stats(*)
    end function

    integer(c_int) function c_spa_pattern(ctx, lunk, nnz, ind, irow, icol) bind(C, name='heomq_spa_pattern')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: lunk
      integer(c_int32_t) :: nnz(*), irow(*), icol(*)
      integer(c_int64_t) :: ind(*)
    end function

    integer(c_int) function c_calc_ax_spa(ctx, ldim, x, y) bind(C, name='heomq_calc_ax_tfqmr_spa')
      import :: c_int, c_int64_t, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: ldim
      complex(c_double_complex) :: x(*), y(*)
    end function

    integer(c_int) function c_calc_ax_jacobi(ctx, dim0, x, y) bind(C, name='heomq_calc_ax_jacobi')
      import :: c_int, c_int32_t, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dim0
      complex(c_double_complex) :: x(*), y(*)
    end function

    integer(c_int) function c_solve_jacobi_host(ctx, rho, crit, maxit, stats) bind(C, name='heomq_solve_jacobi_host')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: rho(*)
      real(c_double), value :: crit
      integer(c_int), value :: maxit
      real(c_double) :: stats(*)
    end function

    integer(c_int) function c_calcderiv_tt(ctx, tt, x, y) bind(C, name='heomq_calcderiv_tt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, y
      real(c_double), value :: tt
    end function

    integer(c_int) function c_set_system_field(ctx, nterm, ops, kinds, pars) bind(C, name='heomq_ctx_set_system_field')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nterm
      complex(c_double), intent(in) :: ops(*)
      integer(c_int32_t), intent(in) :: kinds(*)
      real(c_double), intent(in) :: pars(*)
    end function

    integer(c_int) function c_calcderiv_spa_tt(ctx, tt, x, y) bind(C, name='heomq_calcderiv_spa_tt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, y
      real(c_double), value :: tt
    end function

    integer(c_int) function c_set_field(ctx, ftype, tchar, dc, toff) bind(C, name='heomq_ctx_set_field')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ftype
      real(c_double) :: tchar(*), dc(*)
      real(c_double), value :: toff
    end function

    integer(c_int) function c_calcderiv(ctx, x, y) bind(C, name='heomq_calcderiv')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, x, y
    end function

    integer(c_int) function c_solve_host(ctx, rho, crit, maxit, stats) bind(C, name='heomq_solve_steady_host')
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: rho(*)
      real(c_double), value :: crit
      integer(c_int), value :: maxit
      real(c_double) :: stats(*)
    end function
  end interface

contains

  subroutine heomq_fail(where, rc)
    character(len=*), intent(in) :: where
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: k
    type(c_ptr) :: p
    p = c_last_error()
    call c_f_pointer(p, msg, [512])
    text = ' '
    do k = 1, 512
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    write(6,*)
    write(6,*) ' error! ', where, ' status ', rc, ' : ', trim(text)
    error stop 1
  end subroutine heomq_fail

end module heomq_iface

! One-time upload of the reference's module arrays (module/matmod.f90); with
! lscale /= 0, dbsqrt(norbs,nspin,ncor,nalf,nsgn) of main/evaluatepara.f90:2931-2960
! (cb and cd then the scaled ones), else dbsqrt is not read.
subroutine heomq_fortran_init(nrho, nunk, ntier, norbs, nspin, nalf, ncor, igroundsteady, &
    ntable, lall, indextable, nfirst, nlast, ifirst, indexcoef, itypecoef, ifactfcoef, ifactrcoef, index_coef_ref, &
    indexdraw, cb, cd, cgamma, dipm, ilead, mspin, morbs, mpm, eleadinfty, hs, amsall, lscale, dbsqrt, device)
  use heomq_iface
  implicit none
  integer(c_int32_t), intent(in) :: nrho, ntier, norbs, nspin, nalf, ncor, igroundsteady, device
  integer(c_int64_t), intent(in) :: nunk, ntable, lall
  integer(c_int32_t), intent(in) :: indextable(*), indexdraw(*), ilead(*), mspin(*), morbs(*), mpm(*)
  integer(c_int64_t), intent(in) :: nfirst(*), nlast(*), ifirst(*), indexcoef(*), index_coef_ref(*)
  integer(c_int8_t), intent(in) :: itypecoef(*), ifactfcoef(*), ifactrcoef(*)
  complex(c_double_complex), intent(in) :: cb(*), cd(*), cgamma(*), hs(*)
  real(c_double), intent(in) :: dipm(*), eleadinfty(*), amsall(*)
  integer(c_int32_t), intent(in) :: lscale
  real(c_double), intent(in), target :: dbsqrt(*)
  integer(c_int) :: rc
  type(c_ptr) :: pdb
  pdb = c_null_ptr
  if (lscale /= 0) pdb = c_loc(dbsqrt)
  rc = c_model_from_reference(nrho, nunk, ntier, norbs, nspin, nalf, ncor, igroundsteady, &
       ntable, lall, indextable, nfirst, nlast, ifirst, indexcoef, itypecoef, ifactfcoef, ifactrcoef, index_coef_ref, &
       indexdraw, cb, cd, cgamma, dipm, ilead, mspin, morbs, mpm, eleadinfty, hs, amsall, pdb, heomq_model_h)
  if (rc /= 0) call heomq_fail('heomq_fortran_init (tables)', rc)
  rc = c_ctx_create(heomq_model_h, int(device, c_int), heomq_ctx_h)
  if (rc /= 0) call heomq_fail('heomq_fortran_init (device)', rc)
end subroutine heomq_fortran_init

! tfqmr_double/calc_ax_tfqmr_omp.f90:1,15-17 -- y = A x with the trace row
subroutine calc_ax_tfqmr_omp(dim0, rcgtmp, rcgrhs)
  use heomq_iface
  implicit none
  integer, intent(in) :: dim0
  complex*16, intent(in) :: rcgtmp(*)
  complex*16, intent(out) :: rcgrhs(*)
  integer(c_int) :: rc
  rc = c_calc_ax(heomq_ctx_h, int(dim0, c_int32_t), rcgtmp, rcgrhs)
  if (rc /= 0) call heomq_fail('calc_ax_tfqmr_omp', rc)
end subroutine calc_ax_tfqmr_omp

! bicg/calc_ax_bicg_omp.f90:1,16-17 -- the BiCG-convention apply; rcgtmp's
! rho0 is rebuilt from its lower triangle in place, as the reference does
subroutine calc_ax_bicg_omp(dim0, rcgtmp, rcgrhs)
  use heomq_iface
  implicit none
  integer, intent(in) :: dim0
  complex*16, intent(inout) :: rcgtmp(*)
  complex*16, intent(inout) :: rcgrhs(*)
  integer(c_int) :: rc
  rc = c_calc_ax_bicg(heomq_ctx_h, int(dim0, c_int32_t), rcgtmp, rcgrhs)
  if (rc /= 0) call heomq_fail('calc_ax_bicg_omp', rc)
end subroutine calc_ax_bicg_omp

! bicg/calc_atrx_bicg_omp.f90:1,35-36 -- the real transpose of the BiCG-
! convention apply; rcgtmp's rho0 is reduced in place to its lower triangle
! with a real diagonal (calc_atrx_bicg_omp.f90:103-110)
subroutine calc_atrx_bicg_omp(dim0, rcgtmp, rcgrhs)
  use heomq_iface
  implicit none
  integer, intent(in) :: dim0
  complex*16, intent(inout) :: rcgtmp(*)
  complex*16, intent(inout) :: rcgrhs(*)
  integer(c_int) :: rc
  rc = c_calc_atrx_bicg(heomq_ctx_h, int(dim0, c_int32_t), rcgtmp, rcgrhs)
  if (rc /= 0) call heomq_fail('calc_atrx_bicg_omp', rc)
end subroutine calc_atrx_bicg_omp

! td/calcderiv_omp.f90 reads module array rhotmp and writes rhorhs; the
! caller registers their addresses once with heomq_register_td_arrays.
subroutine heomq_register_td_arrays(rhotmp, rhorhs)
  use heomq_iface
  implicit none
  complex*16, intent(in), target :: rhotmp(*)
  complex*16, intent(in), target :: rhorhs(*)
  td_rhotmp = c_loc(rhotmp)
  td_rhorhs = c_loc(rhorhs)
end subroutine heomq_register_td_arrays

subroutine calcderiv_omp(tt)
  use heomq_iface
  implicit none
  real*8, intent(in) :: tt
  integer(c_int) :: rc
  if (.not. c_associated(td_rhotmp)) then
    write(6,*) ' error! calcderiv_omp: rhotmp/rhorhs not registered '
    error stop 1
  end if
  rc = c_calcderiv_tt(heomq_ctx_h, real(tt, c_double), td_rhotmp, td_rhorhs)
  if (rc /= 0) call heomq_fail('calcderiv_omp', rc)
end subroutine calcderiv_omp

! The time-dependent lead shift of main/getengyshift.f90 (module variables
! fieldtype, tchar(nalf,nspin) in fs, amps_dc(nalf,nspin) in 1/fs as
! evaluatepara.f90:2372 stores it, t_off); the amplitudes are the engyshift
! the context was built with.  Call once after heomq_fortran_init for a
! time-dependent run (igroundsteady = 0).
subroutine heomq_fortran_set_field(fieldtype, nalf, nspin, tchar, amps_dc, t_off)
  use heomq_iface
  implicit none
  integer, intent(in) :: fieldtype, nalf, nspin
  real*8, intent(in) :: tchar(nalf, nspin), amps_dc(nalf, nspin), t_off
  real(c_double) :: tc(nalf * nspin), dc(nalf * nspin)
  integer :: a, s
  integer(c_int) :: rc
  do a = 1, nalf
    do s = 1, nspin
      tc((a - 1) * nspin + s) = tchar(a, s)
      dc((a - 1) * nspin + s) = amps_dc(a, s) * 0.658211928d0   ! back to eV (include/sizes:23)
    end do
  end do
  rc = c_set_field(heomq_ctx_h, int(fieldtype, c_int), tc, dc, real(t_off, c_double))
  if (rc /= 0) call heomq_fail('heomq_fortran_set_field', rc)
end subroutine heomq_fortran_set_field

! Solver-level drop-in for tfqmr_double/solve_tfqmr.f90: rho (nrho,nrho,nunk)
! holds the initial guess on entry and the steady state on exit; crit is the
! relative residual target; info(1) = status (0 ok, 4 not converged, 8
! breakdown as DUTFX), info(2) = iterations.
subroutine solve_tfqmr_gpu(rho, crit, maxit, info)
  use heomq_iface
  implicit none
  complex*16, intent(inout) :: rho(*)
  real*8, intent(in) :: crit
  integer, intent(in) :: maxit
  integer, intent(out) :: info(2)
  real(c_double) :: stats(8)
  integer(c_int) :: rc
  rc = c_solve_host(heomq_ctx_h, rho, real(crit, c_double), int(maxit, c_int), stats)
  info(1) = int(rc)
  info(2) = int(stats(1))
  if (rc /= 0 .and. rc /= 4) call heomq_fail('solve_tfqmr_gpu', rc)
end subroutine solve_tfqmr_gpu

! Solver-level drop-in for bicg/solvebicg.f90: rho holds the initial guess on
! entry and the steady state on exit (rho0 Hermitian); crit is the absolute
! ||r|| criterion of solvebicg.f90:124-126; info(1) = status, info(2) = iterations.
subroutine solve_bicg_gpu(rho, crit, maxit, info)
  use heomq_iface
  implicit none
  complex*16, intent(inout) :: rho(*)
  real*8, intent(in) :: crit
  integer, intent(in) :: maxit
  integer, intent(out) :: info(2)
  real(c_double) :: stats(8)
  integer(c_int) :: rc
  rc = c_solve_bicg_host(heomq_ctx_h, rho, real(crit, c_double), int(maxit, c_int), stats)
  info(1) = int(rc)
  info(2) = int(stats(1))
  if (rc /= 0 .and. rc /= 4) call heomq_fail('solve_bicg_gpu', rc)
end subroutine solve_bicg_gpu

! jacobi/calc_ax_jacobi_omp.f90:1,15-17 -- one block-Jacobi sweep: rcgrhs is
! the new iterate from rcgtmp (default shift of prelude_jacobi.f90:84-113)
subroutine calc_ax_jacobi_omp(dim0, rcgtmp, rcgrhs)
  use heomq_iface
  implicit none
  integer, intent(in) :: dim0
  complex*16, intent(in) :: rcgtmp(*)
  complex*16, intent(out) :: rcgrhs(*)
  integer(c_int) :: rc
  rc = c_calc_ax_jacobi(heomq_ctx_h, int(dim0, c_int32_t), rcgtmp, rcgrhs)
  if (rc /= 0) call heomq_fail('calc_ax_jacobi_omp', rc)
end subroutine calc_ax_jacobi_omp

! solve_jacobi on the device (jacobi/solve_jacobi.f90): rho (the full
! hierarchy, reference layout) holds the start on entry -- its rho0 made
! Hermitian from the lower triangle and required to have unit trace -- and
! the last iterate on return.  info(1) = status (0 converged, 4 not
! converged), info(2) = sweeps.
subroutine solve_jacobi_gpu(rho, crit, maxit, info)
  use heomq_iface
  implicit none
  complex*16, intent(inout) :: rho(*)
  real*8, intent(in) :: crit
  integer, intent(in) :: maxit
  integer, intent(out) :: info(2)
  real(c_double) :: stats(8)
  integer(c_int) :: rc
  rc = c_solve_jacobi_host(heomq_ctx_h, rho, real(crit, c_double), int(maxit, c_int), stats)
  info(1) = int(rc)
  info(2) = int(stats(1))
  if (rc /= 0 .and. rc /= 4) call heomq_fail('solve_jacobi_gpu', rc)
end subroutine solve_jacobi_gpu

! The sparse pattern of module sparsemod (fermion/prelude_spa.f90:532-749):
! lunk_spa elements; nnz_spa and ind_spa per ADO; irow_spa / icol_spa
! (module/sparsemod.f90:28-31).  Registered once, after prelude_spa.
subroutine heomq_register_spa_pattern(lunk_spa, nnz_spa, ind_spa, irow_spa, icol_spa)
  use heomq_iface
  implicit none
  integer*8, intent(in) :: lunk_spa
  integer, intent(in) :: nnz_spa(*), irow_spa(*), icol_spa(*)
  integer*8, intent(in) :: ind_spa(*)
  integer(c_int) :: rc
  rc = c_spa_pattern(heomq_ctx_h, int(lunk_spa, c_int64_t), nnz_spa, ind_spa, irow_spa, icol_spa)
  if (rc /= 0) call heomq_fail('heomq_register_spa_pattern', rc)
end subroutine heomq_register_spa_pattern

! The time-dependent system Hamiltonian of main/calcdhs.f90 for
! calcderiv_omp(tt): dhs(t) = sum_k f_k(t) ops(:,:,k), ops Hermitian in eV,
! kinds(k) and pars(4,k) the time function of each term (heomq.h lists them);
! nterm = 0 removes it.
subroutine heomq_fortran_set_system_field(nterm, ops, kinds, pars)
  use heomq_iface
  implicit none
  integer, intent(in) :: nterm
  complex*16, intent(in) :: ops(*)
  integer, intent(in) :: kinds(*)
  real*8, intent(in) :: pars(*)
  integer(c_int) :: rc
  rc = c_set_system_field(heomq_ctx_h, int(nterm, c_int), ops, kinds, pars)
  if (rc /= 0) call heomq_fail('heomq_fortran_set_system_field', rc)
end subroutine heomq_fortran_set_system_field

! td/calcderiv_spa_omp.f90 reads module array rhotmp_spa and writes
! rhorhs_spa (module sparsemod); the caller registers their addresses once.
subroutine heomq_register_td_spa_arrays(rhotmp_spa, rhorhs_spa)
  use heomq_iface
  implicit none
  complex*16, intent(in), target :: rhotmp_spa(*)
  complex*16, intent(in), target :: rhorhs_spa(*)
  td_rhotmp_spa = c_loc(rhotmp_spa)
  td_rhorhs_spa = c_loc(rhorhs_spa)
end subroutine heomq_register_td_spa_arrays

! td/calcderiv_spa_omp.f90:1-11 -- rhorhs_spa = P L(tt) P^T rhotmp_spa over
! the registered sparse pattern (H fixed, the lead shift at tt)
subroutine calcderiv_spa_omp(tt)
  use heomq_iface
  implicit none
  real*8, intent(in) :: tt
  integer(c_int) :: rc
  if (.not. c_associated(td_rhotmp_spa)) then
    write(6,*) ' error! calcderiv_spa_omp: rhotmp_spa/rhorhs_spa not registered '
    error stop 1
  end if
  rc = c_calcderiv_spa_tt(heomq_ctx_h, real(tt, c_double), td_rhotmp_spa, td_rhorhs_spa)
  if (rc /= 0) call heomq_fail('calcderiv_spa_omp', rc)
end subroutine calcderiv_spa_omp

! tfqmr_double/calc_ax_tfqmr_spa_omp.f90:1,16-18 -- the apply on COO storage:
! rcgrhs(1:lunk_spa) = P A P^T rcgtmp, rcgrhs(lunk_spa + 1) = tr rho0;
! ldim /= lunk_spa is fatal as in :61-65
subroutine calc_ax_tfqmr_spa_omp(ldim, rcgtmp, rcgrhs)
  use heomq_iface
  implicit none
  integer*8, intent(in) :: ldim
  complex*16, intent(in) :: rcgtmp(*)
  complex*16, intent(out) :: rcgrhs(*)
  integer(c_int) :: rc
  rc = c_calc_ax_spa(heomq_ctx_h, int(ldim, c_int64_t), rcgtmp, rcgrhs)
  if (rc /= 0) call heomq_fail('calc_ax_tfqmr_spa_omp', rc)
end subroutine calc_ax_tfqmr_spa_omp
