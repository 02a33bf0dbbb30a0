! Exercises the solver objects through the reference's interfaces on one
! pde3d grid (stdin: kind AINV|INVK|INVT, grid size N, factor storage
! F64|F32|BF16):
!   * build(a, desc_a, info, vmold=...) on a psb_dspmat_type block;
!   * apply_a on host arrays and apply_v on psb_d_vect_type in device
!     storage ('N' and 'T', alpha / beta), which must agree bit for bit
!     (the same kernels), written to stdout for the caller's oracle check;
!   * apply_v refusing a wv of fewer than 2 vectors (the reference's
!     "invalid wv size");
!   * clone (a new object of the same type, same results) and
!     clone_settings, cseti / csetr / csetc with idx, descr with coarse,
!     get_wrksz, sizeof, get_nzeros, update (the shifted pivots),
!     stringval / algname.
program apply_check_hip
  use iso_c_binding
  use psb_base_mod
  use mld_d_base_solver_mod
  use mld_ainv_hip_cbind, only: ainv_gen_pde3d, last_error
  use mld_d_ainv_hip_mod
  implicit none
  character(len=20) :: ptype, fprec
  integer :: idim, info, i
  integer(c_int32_t) :: n
  integer(c_int64_t) :: nnz
  integer(c_int64_t), allocatable, target :: irp(:)
  integer(c_int32_t), allocatable, target :: ja(:)
  real(c_double), allocatable, target :: val(:)
  class(mld_d_base_ainv_solver_type), allocatable :: sv
  class(mld_d_base_solver_type), allocatable :: cl
  class(mld_d_base_ainv_solver_type), allocatable :: other
  type(mld_d_invk_solver_type) :: invk
  type(mld_d_ainv_solver_type) :: names
  type(psb_dspmat_type) :: a
  type(psb_desc_type) :: desc
  type(psb_d_vect_type) :: xv, yv, wv(2), wv1(1)
  type(mld_d_vect_hip_type) :: vmold
  real(c_double), allocatable :: x(:), y0(:), ya(:), yd(:), yc(:), work(:)
  real(c_double), allocatable :: yt(:), ytd(:), shift(:)
  real(c_double), parameter :: alpha = 0.75d0, beta = -0.5d0

  read(*, *) ptype
  read(*, *) idim
  read(*, *) fprec
  call mld_ainv_init(0, info)
  if (info /= 0) call fail('init')
  info = ainv_gen_pde3d(int(idim, c_int32_t), 0_c_int, n, nnz, c_null_ptr, &
       & c_null_ptr, c_null_ptr)
  allocate(irp(n + 1), ja(nnz), val(nnz))
  info = ainv_gen_pde3d(int(idim, c_int32_t), 0_c_int, n, nnz, c_loc(irp), &
       & c_loc(ja), c_loc(val))
  if (info /= 0) call fail('generate')
  irp = irp + 1
  ja = ja + 1
  call a%csr_bld(n, n, irp, ja, val, info)
  desc%n_row = n
  desc%n_col = n

  select case (trim(ptype))
  case ('INVK')
    allocate(mld_d_invk_solver_type :: sv)
    call sv%set('SUB_FILLIN', 1, info, idx=1)
    call sv%set('INV_FILLIN', 1, info, idx=1)
  case ('INVT')
    allocate(mld_d_invt_solver_type :: sv)
    call sv%set('SUB_FILLIN', 4, info)
    call sv%set('INV_FILLIN', 4, info)
    call sv%set('SUB_ILUTHRS', 1d-2, info, idx=2)
    call sv%set('INV_THRESH', 1d-2, info)
  case default
    allocate(mld_d_ainv_solver_type :: sv)
    call sv%set('AINV_ALG', 'LLK', info, idx=1)
    call sv%set('SUB_FILLIN', 8, info)
    call sv%set('SUB_ILUTHRS', 0.1d0, info)
  end select
  if (info /= 0) call fail('set')
  call sv%set('FACTOR_PREC', trim(fprec), info)
  call sv%check(info)
  call sv%build(a, desc, info, vmold=vmold)
  if (info /= 0) call fail('build')

  allocate(x(n), y0(n), work(5 * n))
  do i = 1, n
    x(i) = sin(real(i, c_double))
    y0(i) = cos(real(i, c_double))
  end do

  ! host arrays
  ya = y0
  call sv%apply(alpha, x, beta, ya, desc, 'N', work, info)
  if (info /= 0) call fail('apply_a N')
  yt = y0
  call sv%apply(alpha, x, beta, yt, desc, 'T', work, info)
  if (info /= 0) call fail('apply_a T')
  ! device vectors
  call xv%bld(x, info, mold=vmold)
  call yv%bld(y0, info, mold=vmold)
  call wv(1)%bld(y0, info, mold=vmold)
  call wv(2)%bld(y0, info, mold=vmold)
  call sv%apply(alpha, xv, beta, yv, desc, 'N', work, wv, info)
  if (info /= 0) call fail('apply_v N')
  yd = yv%get_vect()
  call yv%set_vect(y0)
  call sv%apply(alpha, xv, beta, yv, desc, 'T', work, wv, info)
  if (info /= 0) call fail('apply_v T')
  ytd = yv%get_vect()
  write(*, '(a,es12.5)') 'apply_a vs apply_v N : ', maxval(abs(ya - yd))
  write(*, '(a,es12.5)') 'apply_a vs apply_v T : ', maxval(abs(yt - ytd))
  call wv1(1)%bld(y0, info)
  call sv%apply(alpha, xv, beta, yv, desc, 'N', work, wv1, info)
  write(*, '(a,i0)') 'apply_v with 1 wv info : ', info

  ! clone: same type, same results
  call sv%clone(cl, info)
  if (info /= 0) call fail('clone')
  yc = y0
  call cl%apply(alpha, x, beta, yc, desc, 'N', work, info)
  write(*, '(a,es12.5)') 'clone apply diff     : ', maxval(abs(yc - ya))
  select type (cl)
  type is (mld_d_ainv_solver_type)
    write(*, '(a)') 'clone type           : AINV'
  type is (mld_d_invk_solver_type)
    write(*, '(a)') 'clone type           : INVK'
  type is (mld_d_invt_solver_type)
    write(*, '(a)') 'clone type           : INVT'
  class default
    write(*, '(a)') 'clone type           : other'
  end select
  call cl%free(info)
  ! settings into an object of the same type; another type is refused, as
  ! the reference's clone_settings refuses it
  allocate(other, mold=sv)
  call sv%clone_settings(other, info)
  write(*, '(a,i0)') 'clone_settings info  : ', info
  call other%free(info)
  if (same_type_as(sv, invk)) then
    write(*, '(a,i0)') 'clone_settings other : ', 1
  else
    call sv%clone_settings(invk, info)
    write(*, '(a,i0)') 'clone_settings other : ', merge(1, 0, info /= 0)
  end if
  call invk%free(info)

  call sv%descr(info, iout=6, coarse=.true.)
  write(*, '(a,i0)') 'get_wrksz            : ', sv%get_wrksz()
  write(*, '(a,i0)') 'get_nzeros           : ', sv%get_nzeros()
  write(*, '(a,i0)') 'sizeof positive      : ', merge(1, 0, sv%sizeof() > 0)
  write(*, '(a,i0)') 'stringval SYM-LLK    : ', names%stringval('sym-llk')
  write(*, '(a,a)') 'algname 1            : ', trim(names%algname(1))

  ! host-array apply results for the caller's oracle check
  write(*, '(a)') 'Y_N'
  do i = 1, n
    write(*, '(es25.17)') ya(i)
  end do
  write(*, '(a)') 'Y_T'
  do i = 1, n
    write(*, '(es25.17)') yt(i)
  end do

  ! update: pivots of A + diag(shift), then the N apply again
  allocate(shift(n))
  shift = 0.5d0
  call sv%update(shift, desc, info)
  if (info /= 0) call fail('update')
  ya = y0
  call sv%apply(alpha, x, beta, ya, desc, 'N', work, info)
  write(*, '(a)') 'Y_U'
  do i = 1, n
    write(*, '(es25.17)') ya(i)
  end do

  call xv%free(info)
  call yv%free(info)
  call sv%free(info)

contains

  subroutine fail(what)
    character(len=*), intent(in) :: what
    write(*, '(a,a,a,i0,a,a)') 'ERROR in ', what, ': info=', info, ' ', &
         & trim(last_error())
    stop 1
  end subroutine

end program apply_check_hip
