! Stand-in for MLD2P4's mld_d_base_solver_mod (MLD2P4 >= 2.1.1 is not
! installed here): the base solver type with the type-bound procedures the
! reference plugin's solvers override, each with the argument list (names,
! kinds and intents) of the reference's overriding procedures:
!   build    src/mld_d_ainv_solver.F90:95-111  (mld_d_ainv_solver_bld)
!   apply_a  src/mld_d_base_ainv_mod.F90:97-113
!   apply_v  src/mld_d_base_ainv_mod.F90:115-131
!   cseti / csetc / csetr  src/mld_d_ainv_solver.F90:128-169 (with idx)
!   descr    src/mld_d_ainv_solver.F90:171-184 (with coarse)
!   check, clone, clone_settings, free, clear_data, cnv, dump, sizeof,
!   get_nzeros, get_wrksz, default
! so that a type which overrides them here overrides MLD2P4's base solver
! the same way.  The base procedures report "not implemented" (info 1).
module mld_d_base_solver_mod
  use psb_base_mod
  implicit none

  type :: mld_d_base_solver_type
  contains
    procedure, pass(sv) :: check => d_base_solver_check
    procedure, pass(sv) :: build => d_base_solver_bld
    procedure, pass(sv) :: cnv => d_base_solver_cnv
    procedure, pass(sv) :: clone => d_base_solver_clone
    procedure, pass(sv) :: clone_settings => d_base_solver_clone_settings
    procedure, pass(sv) :: apply_v => d_base_solver_apply_vect
    procedure, pass(sv) :: apply_a => d_base_solver_apply
    generic, public :: apply => apply_a, apply_v
    procedure, pass(sv) :: free => d_base_solver_free
    procedure, pass(sv) :: clear_data => d_base_solver_clear_data
    procedure, pass(sv) :: cseti => d_base_solver_cseti
    procedure, pass(sv) :: csetc => d_base_solver_csetc
    procedure, pass(sv) :: csetr => d_base_solver_csetr
    generic, public :: set => cseti, csetc, csetr
    procedure, pass(sv) :: default => d_base_solver_default
    procedure, pass(sv) :: descr => d_base_solver_descr
    procedure, pass(sv) :: sizeof => d_base_solver_sizeof
    procedure, pass(sv) :: get_nzeros => d_base_solver_get_nzeros
    procedure, nopass :: get_wrksz => d_base_solver_get_wrksz
    procedure, pass(sv) :: dump => d_base_solver_dmp
  end type mld_d_base_solver_type

  integer(psb_ipk_), parameter, private :: not_impl = 1

contains

  subroutine d_base_solver_check(sv, info)
    class(mld_d_base_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
  end subroutine

  subroutine d_base_solver_bld(a, desc_a, sv, info, b, amold, vmold, imold)
    type(psb_dspmat_type), intent(in), target :: a
    type(psb_desc_type), intent(inout) :: desc_a
    class(mld_d_base_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    type(psb_dspmat_type), intent(in), target, optional :: b
    class(psb_d_base_sparse_mat), intent(in), optional :: amold
    class(psb_d_base_vect_type), intent(in), optional :: vmold
    class(psb_i_base_vect_type), intent(in), optional :: imold
    info = not_impl
  end subroutine

  subroutine d_base_solver_cnv(sv, info, amold, vmold, imold)
    class(mld_d_base_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    class(psb_d_base_sparse_mat), intent(in), optional :: amold
    class(psb_d_base_vect_type), intent(in), optional :: vmold
    class(psb_i_base_vect_type), intent(in), optional :: imold
    info = psb_success_
  end subroutine

  subroutine d_base_solver_clone(sv, svout, info)
    class(mld_d_base_solver_type), intent(inout) :: sv
    class(mld_d_base_solver_type), allocatable, intent(inout) :: svout
    integer(psb_ipk_), intent(out) :: info
    info = not_impl
  end subroutine

  subroutine d_base_solver_clone_settings(sv, svout, info)
    class(mld_d_base_solver_type), intent(inout) :: sv
    class(mld_d_base_solver_type), intent(inout) :: svout
    integer(psb_ipk_), intent(out) :: info
    info = not_impl
  end subroutine

  subroutine d_base_solver_apply_vect(alpha, sv, x, beta, y, desc_data, &
       & trans, work, wv, info, init, initu)
    type(psb_desc_type), intent(in) :: desc_data
    class(mld_d_base_solver_type), intent(inout) :: sv
    type(psb_d_vect_type), intent(inout) :: x
    type(psb_d_vect_type), intent(inout) :: y
    real(psb_dpk_), intent(in) :: alpha, beta
    character(len=1), intent(in) :: trans
    real(psb_dpk_), target, intent(inout) :: work(:)
    type(psb_d_vect_type), intent(inout) :: wv(:)
    integer(psb_ipk_), intent(out) :: info
    character, intent(in), optional :: init
    type(psb_d_vect_type), intent(inout), optional :: initu
    info = not_impl
  end subroutine

  subroutine d_base_solver_apply(alpha, sv, x, beta, y, desc_data, trans, &
       & work, info, init, initu)
    type(psb_desc_type), intent(in) :: desc_data
    class(mld_d_base_solver_type), intent(inout) :: sv
    real(psb_dpk_), intent(inout) :: x(:)
    real(psb_dpk_), intent(inout) :: y(:)
    real(psb_dpk_), intent(in) :: alpha, beta
    character(len=1), intent(in) :: trans
    real(psb_dpk_), target, intent(inout) :: work(:)
    integer(psb_ipk_), intent(out) :: info
    character, intent(in), optional :: init
    real(psb_dpk_), intent(inout), optional :: initu(:)
    info = not_impl
  end subroutine

  subroutine d_base_solver_free(sv, info)
    class(mld_d_base_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
  end subroutine

  subroutine d_base_solver_clear_data(sv, info)
    class(mld_d_base_solver_type), intent(inout) :: sv
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
  end subroutine

  subroutine d_base_solver_cseti(sv, what, val, info, idx)
    class(mld_d_base_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    integer(psb_ipk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: idx
    info = psb_success_  ! unknown keys are ignored, as MLD2P4 does
  end subroutine

  subroutine d_base_solver_csetc(sv, what, val, info, idx)
    class(mld_d_base_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    character(len=*), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: idx
    info = psb_success_
  end subroutine

  subroutine d_base_solver_csetr(sv, what, val, info, idx)
    class(mld_d_base_solver_type), intent(inout) :: sv
    character(len=*), intent(in) :: what
    real(psb_dpk_), intent(in) :: val
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: idx
    info = psb_success_
  end subroutine

  subroutine d_base_solver_default(sv)
    class(mld_d_base_solver_type), intent(inout) :: sv
  end subroutine

  subroutine d_base_solver_descr(sv, info, iout, coarse)
    class(mld_d_base_solver_type), intent(in) :: sv
    integer(psb_ipk_), intent(out) :: info
    integer(psb_ipk_), intent(in), optional :: iout
    logical, intent(in), optional :: coarse
    info = psb_success_
  end subroutine

  function d_base_solver_sizeof(sv) result(val)
    class(mld_d_base_solver_type), intent(in) :: sv
    integer(psb_epk_) :: val
    val = 0
  end function

  function d_base_solver_get_nzeros(sv) result(val)
    class(mld_d_base_solver_type), intent(in) :: sv
    integer(psb_epk_) :: val
    val = 0
  end function

  function d_base_solver_get_wrksz() result(val)
    integer(psb_ipk_) :: val
    val = 0
  end function

  subroutine d_base_solver_dmp(sv, desc, level, info, prefix, head, solver, &
       & global_num)
    class(mld_d_base_solver_type), intent(in) :: sv
    type(psb_desc_type), intent(in) :: desc
    integer(psb_ipk_), intent(in) :: level
    integer(psb_ipk_), intent(out) :: info
    character(len=*), intent(in), optional :: prefix, head
    logical, optional, intent(in) :: solver, global_num
    info = psb_success_
  end subroutine

end module mld_d_base_solver_mod
