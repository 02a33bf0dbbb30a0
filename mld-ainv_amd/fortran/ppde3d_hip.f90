! Sample driver in the style of the reference's tests/pdegen/ppde3d.f90:
! generate the 3-D convection-diffusion matrix as the local block of a
! psb_dspmat_type with its descriptor, build an AINV / INVK / INVT
! preconditioner through the solver objects' build(a, desc_a, info), solve
! with a Krylov method on the GPU with the right-hand side and solution in
! device storage (psb_d_vect_type with the mld_d_vect_hip_type mold), and
! print the same quantities (build time, solve time, iterations, error
! estimate, preconditioner nonzeros).
!
! stdin deck, one value per line (text after the value is ignored):
!   method (BICGSTAB|CG|CGS|BICG|RGMRES|BICGSTABL), preconditioner
!   (AINV|INVK|INVT|ML), grid size N, eps, itmax, sub_fillin, inv_fillin,
!   sub_iluthrs, inv_thresh, ainv_alg, factor storage (F64|F32|BF16) and,
!   optionally, irst (restart of RGMRES / l of BiCGStab(l), as the
!   reference decks' IRST line, tests/pdegen/runs/mld3d.inp:10); for ML
!   then, optionally, the multilevel fields of tests/pdegen/runs/ppde.inp:
!   smoother factors (AINV|INVK|INVT), smoother sweeps, levels,
!   aggregation kind (SMOOTHED|NONSMOOTHED), correction (MULT|ADD), side
!   (PRE|POST|TWOSIDE), coarsest-level factors, coarsest-level sweeps,
!   aggregation threshold
program ppde3d_hip
  use iso_c_binding
  use psb_base_mod
  use mld_ainv_hip_cbind, only: ainv_gen_pde3d, last_error
  use mld_d_ainv_hip_mod
  implicit none
  character(len=20) :: method, ptype, alg, fprec
  character(len=20) :: mlsub, mlkind, mltype, mlpos, mlcoarse
  integer :: mlsweeps, mlnlev, mlcsweeps
  real(c_double) :: mlthr
  integer :: idim, itmax, fill1, fill2, info, iter, i, irst, ios
  real(c_double) :: eps, thr1, thr2, err
  integer(c_int32_t) :: n
  integer(c_int64_t) :: nnz
  integer(c_int64_t), allocatable, target :: irp(:)
  integer(c_int32_t), allocatable, target :: ja(:)
  real(c_double), allocatable, target :: val(:)
  real(c_double), allocatable :: bh(:), xh(:), ones(:)
  class(mld_d_base_ainv_solver_type), allocatable :: prec
  type(psb_dspmat_type) :: a
  type(psb_desc_type) :: desc_a
  type(psb_d_vect_type) :: b, x
  type(mld_d_vect_hip_type) :: vmold
  integer(8) :: t0, t1, t2, rate

  read(*, *) method
  read(*, *) ptype
  read(*, *) idim
  read(*, *) eps
  read(*, *) itmax
  read(*, *) fill1
  read(*, *) fill2
  read(*, *) thr1
  read(*, *) thr2
  read(*, *) alg
  read(*, *) fprec
  read(*, *, iostat=ios) irst
  if (ios /= 0) irst = 0
  ! multilevel fields (defaults: the ppde.inp deck, coarse sweeps aside)
  mlsub = 'AINV'; mlsweeps = 1; mlnlev = 3; mlkind = 'SMOOTHED'
  mltype = 'MULT'; mlpos = 'POST'; mlcoarse = 'AINV'; mlcsweeps = 15
  mlthr = 0.1d0
  if (ios == 0) read(*, *, iostat=ios) mlsub
  if (ios == 0) read(*, *, iostat=ios) mlsweeps
  if (ios == 0) read(*, *, iostat=ios) mlnlev
  if (ios == 0) read(*, *, iostat=ios) mlkind
  if (ios == 0) read(*, *, iostat=ios) mltype
  if (ios == 0) read(*, *, iostat=ios) mlpos
  if (ios == 0) read(*, *, iostat=ios) mlcoarse
  if (ios == 0) read(*, *, iostat=ios) mlcsweeps
  if (ios == 0) read(*, *, iostat=ios) mlthr

  call mld_ainv_init(0, info)
  if (info /= 0) call fail('init')
  info = ainv_gen_pde3d(int(idim, c_int32_t), 0_c_int, n, nnz, c_null_ptr, &
       & c_null_ptr, c_null_ptr)
  if (info /= 0) call fail('generate')
  allocate(irp(n + 1), ja(nnz), val(nnz))
  info = ainv_gen_pde3d(int(idim, c_int32_t), 0_c_int, n, nnz, c_loc(irp), &
       & c_loc(ja), c_loc(val))
  if (info /= 0) call fail('generate')
  ! the library generates 0-based CSR; the Fortran side works 1-based
  irp = irp + 1
  ja = ja + 1
  call a%csr_bld(n, n, irp, ja, val, info)
  if (info /= 0) call fail('matrix')
  desc_a%n_row = n
  desc_a%n_col = n

  select case (trim(ptype))
  case ('INVK')
    allocate(mld_d_invk_solver_type :: prec)
  case ('INVT')
    allocate(mld_d_invt_solver_type :: prec)
  case ('ML')
    allocate(mld_d_ml_prec_type :: prec)
  case default
    allocate(mld_d_ainv_solver_type :: prec)
  end select
  call prec%set('SUB_FILLIN', fill1, info)
  call prec%set('INV_FILLIN', fill2, info)
  call prec%set('SUB_ILUTHRS', thr1, info)
  call prec%set('INV_THRESH', thr2, info)
  if (trim(ptype) == 'AINV' .or. trim(ptype) == 'ML') &
       & call prec%set('AINV_ALG', trim(alg), info)
  if (trim(ptype) == 'ML') then
    call prec%set('SUB_SOLVE', trim(mlsub), info)
    call prec%set('SMOOTHER_SWEEPS', mlsweeps, info)
    call prec%set('N_PREC_LEVS', mlnlev, info)
    call prec%set('AGGR_KIND', trim(mlkind), info)
    call prec%set('ML_TYPE', trim(mltype), info)
    call prec%set('SMOOTHER_POS', trim(mlpos), info)
    call prec%set('COARSE_SUBSOLVE', trim(mlcoarse), info)
    call prec%set('COARSE_SWEEPS', mlcsweeps, info)
    call prec%set('AGGR_THRESH', mlthr, info)
    if (info /= 0) call fail('multilevel settings')
  end if
  call prec%set('FACTOR_PREC', trim(fprec), info)
  call prec%check(info)

  call system_clock(t0, rate)
  call prec%build(a, desc_a, info, vmold=vmold)
  if (info /= 0) call fail('build')
  call system_clock(t1)

  ! b = A * 1 (the df_sample.f90 convention when no rhs is given), b and x
  ! in device storage
  allocate(ones(n), bh(n), xh(n))
  ones = 1.0d0
  bh = 0.0d0
  do i = 1, n
    bh(i) = sum(val(irp(i):irp(i + 1) - 1) * ones(ja(irp(i):irp(i + 1) - 1)))
  end do
  xh = 0.0d0
  call b%bld(bh, info, mold=vmold)
  if (info == 0) call x%bld(xh, info, mold=vmold)
  if (info /= 0) call fail('vectors')
  call mld_krylov(trim(method), a, prec, b, x, eps, desc_a, info, &
       & itmax=itmax, iter=iter, err=err, irst=irst, istop=2)
  if (info /= 0) call fail('krylov')
  call system_clock(t2)
  xh = x%get_vect()

  call prec%descr(info)
  write(*, '(a,i0)') 'Matrix order                 : ', n
  write(*, '(a,es12.5)') 'Preconditioner time (s)      : ', &
       & real(t1 - t0, 8) / real(rate, 8)
  write(*, '(a,es12.5)') 'Time to solve system (s)     : ', &
       & real(t2 - t1, 8) / real(rate, 8)
  write(*, '(a,i0)') 'Number of iterations         : ', iter
  write(*, '(a,es12.5)') 'Convergence indicator on exit: ', err
  write(*, '(a,i0)') 'Total nonzeros for PREC      : ', prec%get_nzeros()
  write(*, '(a,es12.5)') 'Max error |x - 1|            : ', maxval(abs(xh - 1.0d0))
  call b%free(info)
  call x%free(info)
  call prec%free(info)

contains

  subroutine fail(what)
    character(len=*), intent(in) :: what
    write(*, '(a,a,a,i0,a,a)') 'ERROR in ', what, ': info=', info, ' ', &
         & trim(last_error())
    stop 1
  end subroutine

end program ppde3d_hip
