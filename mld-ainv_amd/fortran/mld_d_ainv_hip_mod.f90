! Fortran solver objects over the HIP library, with the procedure names of
! the reference plugin's types:
!   mld_d_base_ainv_solver_type  (src/mld_d_base_ainv_mod.F90:45-68)
!   mld_d_ainv_solver_type       (src/mld_d_ainv_solver.F90:46-67)
!   mld_d_invk_solver_type       (src/mld_d_invk_solver.f90:47-57)
!   mld_d_invt_solver_type       (src/mld_d_invt_solver.f90:46-58)
! and mld_d_ml_prec_type, the multilevel preconditioner of MLD2P4
! (prec%init('ML'), configured as tests/pdegen/runs/ppde.inp:10-29 does) with
! these factors as smoother and coarsest-level solver, set through the same
! set / build / apply procedures (ML_TYPE, SMOOTHER_POS, N_PREC_LEVS, ...).
! Each procedure is one call through iso_c_binding; info follows the
! reference convention (0 = success).  The local matrix block is passed as
! 1-based CSR arrays (what psb_dspmat_type holds for a CSR local block);
! vectors are host arrays.  MLD2P4 itself is not linked here: the types
! extend the minimal base in mld_d_base_solver_stub.f90, which a build
! against a real MLD2P4 replaces by mld_d_base_solver_mod.
module mld_d_ainv_hip_mod
  use iso_c_binding
  use mld_ainv_hip_cbind
  use mld_d_base_solver_stub
  implicit none
  private

  public :: mld_d_base_ainv_solver_type, mld_d_ainv_solver_type, &
       & mld_d_invk_solver_type, mld_d_invt_solver_type, &
       & mld_d_ml_prec_type, mld_krylov, &
       & mld_ainv_init

  type, extends(mld_d_base_solver_type) :: mld_d_base_ainv_solver_type
    type(c_ptr) :: h = c_null_ptr
    integer(c_int32_t) :: n = 0
  contains
    procedure, pass(sv) :: kind_id => base_kind_id
    procedure, pass(sv) :: ensure => base_ensure
    procedure, pass(sv) :: build => base_build
    procedure, pass(sv) :: apply => base_apply
    procedure, pass(sv) :: cseti => base_cseti
    procedure, pass(sv) :: csetr => base_csetr
    procedure, pass(sv) :: csetc => base_csetc
    generic, public :: set => cseti, csetr, csetc
    procedure, pass(sv) :: check => base_check
    procedure, pass(sv) :: descr => base_descr
    procedure, pass(sv) :: default => base_default
    procedure, pass(sv) :: free => base_free
    procedure, pass(sv) :: clone => base_clone
    procedure, pass(sv) :: clone_settings => base_clone_settings
    procedure, pass(sv) :: clear_data => base_clear_data
    procedure, pass(sv) :: cnv => base_cnv
    procedure, pass(sv) :: dump => base_dump
    procedure, pass(sv) :: update_a => base_update_a
    procedure, pass(sv) :: sizeof => base_sizeof
    procedure, pass(sv) :: get_nzeros => base_get_nzeros
    procedure, nopass :: get_wrksz => base_get_wrksz
  end type mld_d_base_ainv_solver_type

  type, extends(mld_d_base_ainv_solver_type) :: mld_d_ainv_solver_type
  contains
    procedure, pass(sv) :: kind_id => ainv_kind_id
  end type
  type, extends(mld_d_base_ainv_solver_type) :: mld_d_invk_solver_type
  contains
    procedure, pass(sv) :: kind_id => invk_kind_id
  end type
  type, extends(mld_d_base_ainv_solver_type) :: mld_d_invt_solver_type
  contains
    procedure, pass(sv) :: kind_id => invt_kind_id
  end type
  type, extends(mld_d_base_ainv_solver_type) :: mld_d_ml_prec_type
  contains
    procedure, pass(sv) :: kind_id => ml_kind_id
  end type

contains

  subroutine mld_ainv_init(device, info)
    integer, intent(in) :: device
    integer, intent(out) :: info
    info = ainv_init(int(device, c_int))
  end subroutine

  integer function base_kind_id(sv)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    base_kind_id = ainv_kind_ainv
  end function
  integer function ainv_kind_id(sv)
    class(mld_d_ainv_solver_type), intent(in) :: sv
    ainv_kind_id = ainv_kind_ainv
  end function
  integer function invk_kind_id(sv)
    class(mld_d_invk_solver_type), intent(in) :: sv
    invk_kind_id = ainv_kind_invk
  end function
  integer function invt_kind_id(sv)
    class(mld_d_invt_solver_type), intent(in) :: sv
    invt_kind_id = ainv_kind_invt
  end function
  integer function ml_kind_id(sv)
    class(mld_d_ml_prec_type), intent(in) :: sv
    ml_kind_id = ainv_kind_ml
  end function

  ! creates the C object on first use (applies the defaults)
  subroutine base_ensure(sv, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer, intent(out) :: info
    info = 0
    if (.not. c_associated(sv%h)) info = ainv_create(int(sv%kind_id(), c_int), sv%h)
  end subroutine

  ! sv%build on the local block, 1-based CSR (mld_d_*_solver_bld)
  subroutine base_build(sv, n, irp, ja, val, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer, intent(in) :: n
    integer(c_int64_t), intent(in) :: irp(:)
    integer(c_int32_t), intent(in) :: ja(:)
    real(c_double), intent(in) :: val(:)
    integer, intent(out) :: info
    call sv%ensure(info)
    if (info /= 0) return
    info = ainv_build_csr(sv%h, int(n, c_int32_t), irp, ja, val, 1_c_int)
    if (info == 0) sv%n = n
  end subroutine

  ! sv%apply: y = alpha * prec(x) + beta * y (mld_d_base_ainv_solver_apply)
  subroutine base_apply(alpha, sv, x, beta, y, trans, info)
    real(c_double), intent(in) :: alpha, beta
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    real(c_double), intent(in) :: x(:)
    real(c_double), intent(inout) :: y(:)
    character(len=1), intent(in) :: trans
    integer, intent(out) :: info
    real(c_double), allocatable :: xc(:)
    if (.not. c_associated(sv%h)) then
      info = 6
      return
    end if
    xc = x
    info = ainv_apply(sv%h, alpha, xc, beta, y, trans, 1_c_int, 1_c_int64_t, &
         & 1_c_int64_t, ainv_host_ptrs, c_null_ptr)
  end subroutine

  subroutine base_cseti(sv, what, val, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    integer, intent(in) :: val
    integer, intent(out) :: info
    call sv%ensure(info)
    if (info == 0) info = ainv_cseti(sv%h, cstr(what), int(val, c_int))
  end subroutine

  subroutine base_csetr(sv, what, val, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    real(c_double), intent(in) :: val
    integer, intent(out) :: info
    call sv%ensure(info)
    if (info == 0) info = ainv_csetr(sv%h, cstr(what), val)
  end subroutine

  subroutine base_csetc(sv, what, val, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what, val
    integer, intent(out) :: info
    call sv%ensure(info)
    if (info == 0) info = ainv_csetc(sv%h, cstr(what), cstr(val))
  end subroutine

  subroutine base_check(sv, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer, intent(out) :: info
    call sv%ensure(info)
    if (info == 0) info = ainv_check(sv%h)
  end subroutine

  subroutine base_descr(sv, info, iout)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer, intent(out) :: info
    integer, intent(in), optional :: iout
    character(kind=c_char) :: buf(2048)
    character(len=2048) :: text
    integer :: i, u
    u = 6
    if (present(iout)) u = iout
    call sv%ensure(info)
    if (info /= 0) return
    info = ainv_descr(sv%h, buf, int(size(buf), c_size_t))
    text = ' '
    do i = 1, size(buf)
      if (buf(i) == c_null_char) exit
      text(i:i) = buf(i)
    end do
    write(u, '(a)') trim(text)
  end subroutine

  subroutine base_default(sv)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer :: info
    call sv%ensure(info)
    if (info == 0) info = ainv_default(sv%h)
  end subroutine

  subroutine base_free(sv, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer, intent(out) :: info
    info = 0
    if (c_associated(sv%h)) info = ainv_free(sv%h)
    sv%h = c_null_ptr
    sv%n = 0
  end subroutine

  subroutine base_clone(sv, svout, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    class(mld_d_base_ainv_solver_type), intent(inout) :: svout
    integer, intent(out) :: info
    call svout%free(info)
    call sv%ensure(info)
    if (info == 0) info = ainv_clone(sv%h, svout%h)
    svout%n = sv%n
  end subroutine

  subroutine base_clone_settings(sv, svout, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    class(mld_d_base_ainv_solver_type), intent(inout) :: svout
    integer, intent(out) :: info
    call sv%ensure(info)
    if (info == 0) call svout%ensure(info)
    if (info == 0) info = ainv_clone_settings(sv%h, svout%h)
  end subroutine

  subroutine base_clear_data(sv, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer, intent(out) :: info
    info = 0
    if (c_associated(sv%h)) info = ainv_clear_data(sv%h)
    sv%n = 0
  end subroutine

  ! sv%dump (mld_d_base_ainv_solver_dmp.f90:56-106): W, d and Z as
  ! MatrixMarket files <prefix>_p000_l<level>_{wmat,diag,zmat}.mtx when
  ! solver is present and true
  subroutine base_dump(sv, level, info, prefix, head, solver)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    integer, intent(in) :: level
    integer, intent(out) :: info
    character(len=*), intent(in), optional :: prefix, head
    logical, intent(in), optional :: solver
    character(len=80) :: prefix_
    character(len=120) :: head_
    integer(c_int) :: solver_
    prefix_ = 'dump_ainv_d'
    if (present(prefix)) prefix_ = prefix
    head_ = ' '
    if (present(head)) head_ = head
    solver_ = 0
    if (present(solver)) then
      if (solver) solver_ = 1
    end if
    info = 0
    if (c_associated(sv%h)) &
         & info = ainv_dump(sv%h, cstr(prefix_), int(level, c_int), solver_, &
         & cstr(head_))
  end subroutine

  ! sv%update_a (mld_d_base_ainv_update_a.f90:53-79): pivots of the
  ! shifted operator A + diag(x) with the same W and Z
  subroutine base_update_a(sv, x, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    real(c_double), intent(in) :: x(:)
    integer, intent(out) :: info
    real(c_double), allocatable :: xc(:)
    if (.not. c_associated(sv%h)) then
      info = ainv_err_state
      return
    end if
    xc = x
    info = ainv_update_a(sv%h, xc, int(size(x), c_int64_t), ainv_host_ptrs)
  end subroutine

  ! re-mold the resident factors: 'F64', 'F32' or 'BF16'
  subroutine base_cnv(sv, info, prec)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer, intent(out) :: info
    character(len=*), intent(in) :: prec
    integer(c_int) :: p
    select case (prec)
    case ('F32', 'f32')
      p = ainv_prec_f32
    case ('BF16', 'bf16')
      p = ainv_prec_bf16
    case default
      p = ainv_prec_f64
    end select
    call sv%ensure(info)
    if (info == 0) info = ainv_cnv(sv%h, p)
  end subroutine

  function base_sizeof(sv) result(val)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    integer(c_int64_t) :: val
    integer :: info
    val = 0
    if (c_associated(sv%h)) info = ainv_sizeof(sv%h, val)
  end function

  function base_get_nzeros(sv) result(val)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    integer(c_int64_t) :: val
    integer :: info
    val = 0
    if (c_associated(sv%h)) info = ainv_get_nzeros(sv%h, val)
  end function

  function base_get_wrksz() result(val)
    integer :: val
    val = 2
  end function

  ! psb_krylov(method, a, prec, b, x, eps, ...) with istop = 2 on the local
  ! block (1-based CSR) and host vectors; irst: restart of RGMRES, l of
  ! BICGSTABL (absent: 10 and 2)
  subroutine mld_krylov(method, n, irp, ja, val, prec, b, x, eps, itmax, &
       & iter, err, info, irst)
    character(len=*), intent(in) :: method
    integer, intent(in) :: n, itmax
    integer, intent(in), optional :: irst
    integer(c_int64_t), intent(in) :: irp(:)
    integer(c_int32_t), intent(in) :: ja(:)
    real(c_double), intent(in) :: val(:), b(:), eps
    class(mld_d_base_ainv_solver_type), intent(inout) :: prec
    real(c_double), intent(inout) :: x(:)
    integer, intent(out) :: iter, info
    real(c_double), intent(out) :: err
    type(c_ptr) :: a
    integer(c_int32_t) :: it(1)
    real(c_double) :: er(1), bc(size(b))
    integer :: i2
    integer(c_int) :: nl
    nl = 0
    if (present(irst)) nl = irst
    a = c_null_ptr
    info = ainv_mat_create(int(n, c_int32_t), irp, ja, val, 1_c_int, a)
    if (info /= 0) return
    bc = b
    info = ainv_krylov(cstr(method), a, prec%h, bc, x, 1_c_int, eps, &
         & int(itmax, c_int), 2_c_int, nl, it, er, ainv_host_ptrs, c_null_ptr)
    iter = it(1)
    err = er(1)
    i2 = ainv_mat_free(a)
  end subroutine

end module mld_d_ainv_hip_mod
