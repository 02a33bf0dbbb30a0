! Fortran solver objects over the HIP library, with the types and
! type-bound procedures of the reference plugin:
!   mld_d_base_ainv_solver_type  (src/mld_d_base_ainv_mod.F90:45-68)
!   mld_d_ainv_solver_type       (src/mld_d_ainv_solver.F90:46-67)
!   mld_d_invk_solver_type       (src/mld_d_invk_solver.f90:47-57)
!   mld_d_invt_solver_type       (src/mld_d_invt_solver.f90:46-58)
! and mld_d_ml_prec_type, the multilevel preconditioner of MLD2P4
! (prec%init('ML'), configured as tests/pdegen/runs/ppde.inp:10-29 does) with
! these factors as smoother and coarsest-level solver.
!
! The types extend mld_d_base_solver_type and override its procedures with
! the reference's argument lists: build(a, desc_a, sv, info, b, amold,
! vmold, imold), apply_v(alpha, sv, x, beta, y, desc_data, trans, work, wv,
! info, init, initu), apply_a(alpha, sv, x, beta, y, desc_data, trans, work,
! info, init, initu), cseti / csetr / csetc(sv, what, val, info, idx),
! descr(sv, info, iout, coarse), clone(sv, svout, info), ... .  Here they
! compile against the stand-ins of psb_base_mod and mld_d_base_solver_mod
! (psb_base_standin.f90, mld_d_base_solver_standin.f90); against installed
! PSBLAS / MLD2P4 the same source uses theirs.
!
! Each procedure is one call through iso_c_binding (info 0 = success, the
! reference convention).  The factors live on the GPU.  Vectors whose
! storage class is mld_d_vect_hip_type (device memory; pass it as vmold to
! build, or as the mold of psb_d_vect_type%bld) are applied in place on the
! device; vectors with host storage are staged through device memory.
module mld_d_ainv_hip_mod
  use iso_c_binding
  use psb_base_mod
  use mld_d_base_solver_mod
  use mld_ainv_hip_cbind
  implicit none
  private

  public :: mld_d_base_ainv_solver_type, mld_d_ainv_solver_type, &
       & mld_d_invk_solver_type, mld_d_invt_solver_type, &
       & mld_d_ml_prec_type, mld_d_vect_hip_type, mld_krylov, &
       & mld_ainv_init

  ! ---- vector storage in device memory ------------------------------------
  type, extends(psb_d_base_vect_type) :: mld_d_vect_hip_type
    type(c_ptr) :: d = c_null_ptr
    integer(psb_ipk_) :: nr = 0
  contains
    procedure, pass(x) :: get_nrows => hv_get_nrows
    procedure, pass(x) :: get_vect => hv_get_vect
    procedure, pass(x) :: set_vect => hv_set_vect
    procedure, pass(x) :: alloc => hv_alloc
    procedure, pass(x) :: dev_ptr => hv_dev_ptr
    procedure, pass(x) :: free => hv_free
  end type mld_d_vect_hip_type

  ! ---- solver objects --------------------------------------------------------
  type, extends(mld_d_base_solver_type) :: mld_d_base_ainv_solver_type
    type(c_ptr) :: h = c_null_ptr
    integer(psb_ipk_) :: n = 0
  contains
    procedure, pass(sv) :: kind_id => base_kind_id
    procedure, pass(sv) :: ensure => base_ensure
    procedure, pass(sv) :: build => base_build
    procedure, pass(sv) :: apply_v => base_apply_vect
    procedure, pass(sv) :: apply_a => base_apply
    procedure, pass(sv) :: cseti => base_cseti
    procedure, pass(sv) :: csetr => base_csetr
    procedure, pass(sv) :: csetc => base_csetc
    procedure, pass(sv) :: check => base_check
    procedure, pass(sv) :: descr => base_descr
    procedure, pass(sv) :: default => base_default
    procedure, pass(sv) :: free => base_free
    procedure, pass(sv) :: clone => base_clone
    procedure, pass(sv) :: clone_settings => base_clone_settings
    procedure, pass(sv) :: clear_data => base_clear_data
    procedure, pass(sv) :: cnv => base_cnv
    procedure, pass(sv) :: dump => base_dump
    procedure, pass(sv) :: sizeof => base_sizeof
    procedure, pass(sv) :: get_nzeros => base_get_nzeros
    procedure, nopass :: get_wrksz => base_get_wrksz
    procedure, pass(sv) :: update_a => base_update_a
    generic, public :: update => update_a
    procedure, pass(sv) :: set_prec => base_set_prec
  end type mld_d_base_ainv_solver_type

  type, extends(mld_d_base_ainv_solver_type) :: mld_d_ainv_solver_type
  contains
    procedure, pass(sv) :: kind_id => ainv_kind_id
    procedure, nopass :: stringval => d_ainv_stringval
    procedure, nopass :: algname => d_ainv_algname
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
    integer(psb_ipk_), intent(in) :: device
    integer(psb_ipk_), intent(out) :: info
    info = ainv_init(int(device, c_int))
  end subroutine

  ! ---- mld_d_vect_hip_type ----------------------------------------------------
  function hv_get_nrows(x) result(res)
    class(mld_d_vect_hip_type), intent(in) :: x
    integer(psb_ipk_) :: res
    res = x%nr
  end function

  function hv_get_vect(x) result(res)
    class(mld_d_vect_hip_type), intent(inout) :: x
    real(psb_dpk_), allocatable, target :: res(:)
    integer(c_int) :: i2
    allocate(res(x%nr), source=dzero)
    if (x%nr > 0) i2 = ainv_memcpy(c_loc(res), x%d, &
         & int(x%nr, c_size_t) * 8_c_size_t, ainv_d2h)
  end function

  subroutine hv_set_vect(x, val)
    class(mld_d_vect_hip_type), intent(inout) :: x
    real(psb_dpk_), intent(in) :: val(:)
    real(psb_dpk_), allocatable, target :: vc(:)
    integer(c_int) :: i2
    integer(psb_ipk_) :: info
    if (size(val) /= x%nr) call x%alloc(int(size(val), psb_ipk_), info)
    vc = val
    if (x%nr > 0) i2 = ainv_memcpy(x%d, c_loc(vc), &
         & int(x%nr, c_size_t) * 8_c_size_t, ainv_h2d)
  end subroutine

  subroutine hv_alloc(x, n, info)
    class(mld_d_vect_hip_type), intent(inout) :: x
    integer(psb_ipk_), intent(in) :: n
    integer(psb_ipk_), intent(out) :: info
    real(psb_dpk_), allocatable, target :: z(:)
    call x%free(info)
    x%nr = n
    info = ainv_dev_alloc(int(max(n, 1), c_size_t) * 8_c_size_t, x%d)
    if (info /= 0) return
    allocate(z(n))
    z = dzero
    if (n > 0) info = ainv_memcpy(x%d, c_loc(z), int(n, c_size_t) * 8_c_size_t, &
         & ainv_h2d)
  end subroutine

  function hv_dev_ptr(x) result(p)
    class(mld_d_vect_hip_type), intent(in) :: x
    type(c_ptr) :: p
    p = x%d
  end function

  subroutine hv_free(x, info)
    class(mld_d_vect_hip_type), intent(inout) :: x
    integer(psb_ipk_), intent(out) :: info
    info = 0
    if (c_associated(x%d)) info = ainv_dev_free(x%d)
    x%d = c_null_ptr
    x%nr = 0
  end subroutine

  ! ---- kinds --------------------------------------------------------------------
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
    integer(psb_ipk_), intent(out) :: info
    info = 0
    if (.not. c_associated(sv%h)) &
         & info = ainv_create(int(sv%kind_id(), c_int), sv%h)
  end subroutine

  ! sv%build (mld_d_ainv_solver_bld.f90:35-94): the factors of the local
  ! block -- the first desc_a%get_local_rows() rows of a (all of a when the
  ! descriptor is empty), columns outside the block dropped -- or of b
  ! when present.  amold / vmold / imold: the reference's re-mold hooks;
  ! the factors are laid out for the GPU in any case.
  subroutine base_build(a, desc_a, sv, info, b, amold, vmold, imold)
    type(psb_dspmat_type), intent(in), target :: a
    type(psb_desc_type), intent(inout) :: desc_a
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    type(psb_dspmat_type), intent(in), target, optional :: b
    class(psb_d_base_sparse_mat), intent(in), optional :: amold
    class(psb_d_base_vect_type), intent(in), optional :: vmold
    class(psb_i_base_vect_type), intent(in), optional :: imold
    type(psb_dspmat_type), pointer :: m
    integer(psb_ipk_) :: n
    m => a
    if (present(b)) m => b
    if (.not. allocated(m%a)) then
      info = 1
      return
    end if
    n = desc_a%get_local_rows()
    if (n <= 0) n = m%get_nrows()
    call sv%ensure(info)
    if (info /= 0) return
    info = ainv_build_csr(sv%h, int(n, c_int32_t), m%a%irp, m%a%ja, &
         & m%a%val, 1_c_int)
    if (info == 0) sv%n = n
  end subroutine

  ! sv%apply_v (mld_d_base_ainv_solver_apply_vect.f90:35-166):
  ! y = alpha * Z (d o (W x)) + beta * y ('N'), the factors' transposes
  ! for 'T' / 'C'; work and wv are the caller's scratch in the reference
  ! (wv needs 2 vectors); init / initu are ignored, as there.  Device
  ! vectors are used in place; host vectors are staged.
  subroutine base_apply_vect(alpha, sv, x, beta, y, desc_data, trans, work, &
       & wv, info, init, initu)
    type(psb_desc_type), intent(in) :: desc_data
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_d_vect_type), intent(inout) :: y
    real(psb_dpk_), intent(in) :: alpha, beta
    character(len=1), intent(in) :: trans
    real(psb_dpk_), target, intent(inout) :: work(:)
    type(psb_d_vect_type), intent(inout) :: wv(:)
    integer(psb_ipk_), intent(out) :: info
    character, intent(in), optional :: init
    type(psb_d_vect_type), intent(inout), optional :: initu
    real(psb_dpk_), allocatable, target :: xh(:), yh(:)
    type(c_ptr) :: px, py
    if (.not. c_associated(sv%h)) then
      info = ainv_err_state
      return
    end if
    if (size(wv) < 2) then  ! the reference's "invalid wv size"
      info = 4
      return
    end if
    if (.not. allocated(x%v) .or. .not. allocated(y%v)) then
      info = 1
      return
    end if
    px = x%v%dev_ptr()
    py = y%v%dev_ptr()
    if (c_associated(px) .and. c_associated(py)) then
      info = ainv_apply(sv%h, alpha, px, beta, py, trans, 1_c_int, &
           & 1_c_int64_t, 1_c_int64_t, 0_c_int, c_null_ptr)
      if (info == 0) info = ainv_device_sync()
    else
      xh = x%get_vect()
      yh = y%get_vect()
      info = ainv_apply(sv%h, alpha, c_loc(xh), beta, c_loc(yh), trans, &
           & 1_c_int, 1_c_int64_t, 1_c_int64_t, ainv_host_ptrs, c_null_ptr)
      if (info == 0) call y%set_vect(yh)
    end if
  end subroutine

  ! sv%apply_a (mld_d_base_ainv_solver_apply.f90:35-152) on host arrays
  subroutine base_apply(alpha, sv, x, beta, y, desc_data, trans, work, info, &
       & init, initu)
    type(psb_desc_type), intent(in) :: desc_data
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    real(psb_dpk_), intent(inout) :: x(:)
    real(psb_dpk_), intent(inout) :: y(:)
    real(psb_dpk_), intent(in) :: alpha, beta
    character(len=1), intent(in) :: trans
    real(psb_dpk_), target, intent(inout) :: work(:)
    integer(psb_ipk_), intent(out) :: info
    character, intent(in), optional :: init
    real(psb_dpk_), intent(inout), optional :: initu(:)
    real(psb_dpk_), allocatable, target :: xc(:), yc(:)
    if (.not. c_associated(sv%h)) then
      info = ainv_err_state
      return
    end if
    ! contiguous host copies (x, y may be array sections); staged to the
    ! device by the library
    xc = x
    yc = y
    info = ainv_apply(sv%h, alpha, c_loc(xc), beta, c_loc(yc), trans, &
         & 1_c_int, 1_c_int64_t, 1_c_int64_t, ainv_host_ptrs, c_null_ptr)
    if (info == 0) y = yc
  end subroutine

  ! sv%cseti / csetr / csetc (mld_d_ainv_solver_cseti.f90:55-62, ...):
  ! idx (a level index in MLD2P4) is not used by these solvers
  subroutine base_cseti(sv, what, val, info, idx)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    integer(psb_ipk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: idx
    call sv%ensure(info)
    if (info == 0) info = ainv_cseti(sv%h, cstr(what), int(val, c_int))
  end subroutine

  subroutine base_csetr(sv, what, val, info, idx)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    real(psb_dpk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: idx
    call sv%ensure(info)
    if (info == 0) info = ainv_csetr(sv%h, cstr(what), val)
  end subroutine

  subroutine base_csetc(sv, what, val, info, idx)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    character(len=*), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: idx
    call sv%ensure(info)
    if (info == 0) info = ainv_csetc(sv%h, cstr(what), cstr(val))
  end subroutine

  ! FACTOR_PREC by name ('F64' | 'F32' | 'BF16'): the re-mold of the
  ! factors this build offers in place of a sparse storage mold
  subroutine base_set_prec(sv, prec, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: prec
    integer(psb_ipk_), intent(out) :: info
    call sv%csetc('FACTOR_PREC', prec, info)
  end subroutine

  subroutine base_check(sv, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    call sv%ensure(info)
    if (info == 0) info = ainv_check(sv%h)
  end subroutine

  ! sv%descr (mld_d_ainv_solver_descr.f90): on iout (default 6); with
  ! coarse = .true. under a "coarse solver" heading
  subroutine base_descr(sv, info, iout, coarse)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iout
    logical, intent(in), optional :: coarse
    character(kind=c_char) :: buf(2048)
    character(len=2048) :: text
    integer :: i, u
    u = 6
    if (present(iout)) u = iout
    info = 0
    if (.not. c_associated(sv%h)) then
      write(u, '(a)') '  (not configured)'
      return
    end if
    info = ainv_descr(sv%h, buf, int(size(buf), c_size_t))
    text = ' '
    do i = 1, size(buf)
      if (buf(i) == c_null_char) exit
      text(i:i) = buf(i)
    end do
    if (present(coarse)) then
      if (coarse) write(u, '(a)') '  Coarse solver:'
    end if
    write(u, '(a)') trim(text)
  end subroutine

  subroutine base_default(sv)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer(psb_ipk_) :: info
    call sv%ensure(info)
    if (info == 0) info = ainv_default(sv%h)
  end subroutine

  subroutine base_free(sv, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    info = 0
    if (c_associated(sv%h)) info = ainv_free(sv%h)
    sv%h = c_null_ptr
    sv%n = 0
  end subroutine

  ! sv%clone (mld_d_ainv_solver_clone.f90): svout becomes an object of the
  ! same type holding a copy of the settings and factors
  subroutine base_clone(sv, svout, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    class(mld_d_base_solver_type), allocatable, intent(inout) :: svout
    integer(psb_ipk_), intent(out) :: info
    if (allocated(svout)) then
      call svout%free(info)
      deallocate(svout)
    end if
    allocate(svout, mold=sv)
    call sv%ensure(info)
    if (info /= 0) return
    select type (so => svout)
    class is (mld_d_base_ainv_solver_type)
      so%h = c_null_ptr
      info = ainv_clone(sv%h, so%h)
      so%n = sv%n
    class default
      info = 4
    end select
  end subroutine

  subroutine base_clone_settings(sv, svout, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    class(mld_d_base_solver_type), intent(inout) :: svout
    integer(psb_ipk_), intent(out) :: info
    call sv%ensure(info)
    if (info /= 0) return
    select type (so => svout)
    class is (mld_d_base_ainv_solver_type)
      call so%ensure(info)
      if (info == 0) info = ainv_clone_settings(sv%h, so%h)
    class default
      info = 1
    end select
  end subroutine

  subroutine base_clear_data(sv, info)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    info = 0
    if (c_associated(sv%h)) info = ainv_clear_data(sv%h)
    sv%n = 0
  end subroutine

  ! sv%cnv (mld_d_base_ainv_solver_cnv.f90:56-60): the factors are laid out
  ! for the GPU at build; a storage mold changes nothing more here (the
  ! factor precision is set_prec / FACTOR_PREC)
  subroutine base_cnv(sv, info, amold, vmold, imold)
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    class(psb_d_base_sparse_mat), intent(in), optional :: amold
    class(psb_d_base_vect_type), intent(in), optional :: vmold
    class(psb_i_base_vect_type), intent(in), optional :: imold
    info = 0
    if (.not. c_associated(sv%h)) info = ainv_err_state
  end subroutine

  ! sv%dump (mld_d_base_ainv_solver_dmp.f90:56-106): W, d and Z as
  ! MatrixMarket files <prefix>_p000_l<level>_{wmat,diag,zmat}.mtx when
  ! solver is present and true (one process: global_num changes nothing)
  subroutine base_dump(sv, desc, level, info, prefix, head, solver, &
       & global_num)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    type(psb_desc_type), intent(in) :: desc
    integer(psb_ipk_), intent(in) :: level
    integer(psb_ipk_), intent(out) :: info
    character(len=*), intent(in), optional :: prefix, head
    logical, optional, intent(in) :: solver, global_num
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
  subroutine base_update_a(sv, x, desc_data, info)
    type(psb_desc_type), intent(in) :: desc_data
    class(mld_d_base_ainv_solver_type), intent(inout) :: sv
    real(psb_dpk_), intent(in) :: x(:)
    integer(psb_ipk_), intent(out) :: info
    real(psb_dpk_), allocatable, target :: xc(:)
    if (.not. c_associated(sv%h)) then
      info = ainv_err_state
      return
    end if
    xc = x
    info = ainv_update_a(sv%h, c_loc(xc), int(size(x), c_int64_t), &
         & ainv_host_ptrs)
  end subroutine

  function base_sizeof(sv) result(val)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    integer(psb_epk_) :: val
    integer :: info
    val = 0
    if (c_associated(sv%h)) info = ainv_sizeof(sv%h, val)
  end function

  function base_get_nzeros(sv) result(val)
    class(mld_d_base_ainv_solver_type), intent(in) :: sv
    integer(psb_epk_) :: val
    integer :: info
    val = 0
    if (c_associated(sv%h)) info = ainv_get_nzeros(sv%h, val)
  end function

  ! two work vectors (mld_d_base_ainv_mod.F90:245-254)
  function base_get_wrksz() result(val)
    integer(psb_ipk_) :: val
    val = 2
  end function

  ! mld_d_ainv_solver_type%stringval / algname (mld_d_ainv_solver.F90:
  ! 222-290): the algorithm ids of the reference's names, and back
  function d_ainv_stringval(string) result(val)
    character(len=*), intent(in) :: string
    integer(psb_ipk_) :: val
    character(len=len(string)) :: up
    integer :: i
    up = adjustl(string)
    do i = 1, len(up)
      if (up(i:i) >= 'a' .and. up(i:i) <= 'z') &
           & up(i:i) = achar(iachar(up(i:i)) - 32)
    end do
    select case (trim(up))
    case ('LLK')
      val = ainv_alg_llk
    case ('STAB-LLK')
      val = ainv_alg_s_ft_llk
    case ('SYM-LLK')
      val = ainv_alg_s_llk
    case ('MLK')
      val = ainv_alg_mlk
    case ('LLK-NOTH')
      val = ainv_alg_llk_noth
    case default
      val = -1
    end select
  end function

  function d_ainv_algname(ialg) result(val)
    integer(psb_ipk_), intent(in) :: ialg
    character(len=40) :: val
    select case (ialg)
    case (ainv_alg_mlk)
      val = 'Left-looking, list merge'
    case (ainv_alg_llk)
      val = 'Left-looking'
    case (ainv_alg_s_llk)
      val = 'Symmetric Left-looking'
    case (ainv_alg_s_ft_llk)
      val = 'Stabilized Left-looking'
    case (ainv_alg_llk_noth)
      val = 'Left-looking, no alpha threshold'
    case default
      val = 'Unknown alg variant'
    end select
  end function

  ! psb_krylov(method, a, prec, b, x, eps, desc_a, info, itmax, iter, err,
  ! itrace, irst, istop) on the local block with istop = 2 (the argument
  ! list of tests/pdegen/ppde3d.f90:280-281).  irst: restart of RGMRES, l
  ! of BICGSTABL (absent: 10 and 2).  b and x in device storage
  ! (mld_d_vect_hip_type) are used in place.
  subroutine mld_krylov(method, a, prec, b, x, eps, desc_a, info, itmax, &
       & iter, err, itrace, irst, istop)
    character(len=*), intent(in) :: method
    type(psb_dspmat_type), intent(in) :: a
    class(mld_d_base_ainv_solver_type), intent(inout) :: prec
    type(psb_d_vect_type), intent(inout) :: b, x
    real(psb_dpk_), intent(in) :: eps
    type(psb_desc_type), intent(in) :: desc_a
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: itmax, itrace, irst, istop
    integer(psb_ipk_), intent(out), optional :: iter
    real(psb_dpk_), intent(out), optional :: err
    type(c_ptr) :: am, pb, px
    integer(c_int32_t) :: it(1)
    real(c_double) :: er(1)
    real(psb_dpk_), allocatable, target :: bh(:), xh(:)
    integer(c_int) :: nl, im, ist, flags
    integer :: i2
    integer(psb_ipk_) :: n
    nl = 0
    if (present(irst)) nl = irst
    im = 1000
    if (present(itmax)) im = itmax
    ist = 2
    if (present(istop)) ist = istop
    n = desc_a%get_local_rows()
    if (n <= 0) n = a%get_nrows()
    am = c_null_ptr
    info = ainv_mat_create(int(n, c_int32_t), a%a%irp, a%a%ja, a%a%val, &
         & 1_c_int, am)
    if (info /= 0) return
    pb = b%v%dev_ptr()
    px = x%v%dev_ptr()
    if (c_associated(pb) .and. c_associated(px)) then
      flags = 0
    else
      bh = b%get_vect()
      xh = x%get_vect()
      pb = c_loc(bh)
      px = c_loc(xh)
      flags = ainv_host_ptrs
    end if
    info = ainv_krylov(cstr(method), am, prec%h, pb, px, 1_c_int, eps, im, &
         & ist, nl, it, er, flags, c_null_ptr)
    if (flags /= 0 .and. info == 0) call x%set_vect(xh)
    if (present(iter)) iter = it(1)
    if (present(err)) err = er(1)
    i2 = ainv_mat_free(am)
  end subroutine

end module mld_d_ainv_hip_mod
