! Stand-in for the parts of PSBLAS's psb_base_mod that the solver objects'
! interfaces name (PSBLAS >= 3.4 is not installed here).  It keeps PSBLAS's
! names and kinds so that mld_d_ainv_hip_mod compiles unchanged against the
! real module:
!   psb_ipk_ / psb_lpk_ / psb_epk_ / psb_dpk_   integer and real kinds
!   psb_dspmat_type        the local matrix block, a CSR storage class
!                          (psb_d_base_sparse_mat; PSBLAS's CSR class is
!                          psb_d_csr_sparse_mat)
!   psb_desc_type          the communication descriptor, here only the
!                          local row / column counts one process needs
!   psb_d_vect_type        a vector whose storage is a class extending
!                          psb_d_base_vect_type (host array here; device
!                          storage classes extend it, as PSBLAS's GPU
!                          vector classes do)
!   psb_i_base_vect_type   integer vector storage (a mold argument only)
module psb_base_mod
  use iso_c_binding
  implicit none

  integer, parameter :: psb_ipk_ = c_int32_t
  integer, parameter :: psb_lpk_ = c_int64_t
  integer, parameter :: psb_epk_ = c_int64_t
  integer, parameter :: psb_dpk_ = c_double
  integer(psb_ipk_), parameter :: psb_success_ = 0
  real(psb_dpk_), parameter :: dzero = 0.0_psb_dpk_, done = 1.0_psb_dpk_

  ! ---- sparse matrix: a 1-based CSR block --------------------------------
  type :: psb_d_base_sparse_mat
    integer(psb_ipk_) :: m = 0, n = 0
    integer(psb_lpk_), allocatable :: irp(:)
    integer(psb_ipk_), allocatable :: ja(:)
    real(psb_dpk_), allocatable :: val(:)
  end type psb_d_base_sparse_mat

  type :: psb_dspmat_type
    class(psb_d_base_sparse_mat), allocatable :: a
  contains
    procedure, pass(a) :: get_nrows => dspmat_get_nrows
    procedure, pass(a) :: get_ncols => dspmat_get_ncols
    procedure, pass(a) :: get_nzeros => dspmat_get_nzeros
    procedure, pass(a) :: csr_bld => dspmat_csr_bld
    procedure, pass(a) :: free => dspmat_free
  end type psb_dspmat_type

  ! ---- descriptor of the local block -------------------------------------
  type :: psb_desc_type
    integer(psb_ipk_) :: n_row = 0, n_col = 0
  contains
    procedure, pass(desc) :: get_local_rows => desc_get_local_rows
    procedure, pass(desc) :: get_local_cols => desc_get_local_cols
  end type psb_desc_type

  ! ---- vectors ------------------------------------------------------------
  type :: psb_d_base_vect_type
    real(psb_dpk_), allocatable :: v(:)
  contains
    procedure, pass(x) :: get_nrows => dbv_get_nrows
    procedure, pass(x) :: get_vect => dbv_get_vect
    procedure, pass(x) :: set_vect => dbv_set_vect
    procedure, pass(x) :: alloc => dbv_alloc
    procedure, pass(x) :: dev_ptr => dbv_dev_ptr
    procedure, pass(x) :: free => dbv_free
  end type psb_d_base_vect_type

  type :: psb_i_base_vect_type
    integer(psb_ipk_), allocatable :: v(:)
  end type psb_i_base_vect_type

  type :: psb_d_vect_type
    class(psb_d_base_vect_type), allocatable :: v
  contains
    procedure, pass(x) :: bld => dv_bld
    procedure, pass(x) :: get_nrows => dv_get_nrows
    procedure, pass(x) :: get_vect => dv_get_vect
    procedure, pass(x) :: set_vect => dv_set_vect
    procedure, pass(x) :: free => dv_free
  end type psb_d_vect_type

contains

  function dspmat_get_nrows(a) result(res)
    class(psb_dspmat_type), intent(in) :: a
    integer(psb_ipk_) :: res
    res = 0
    if (allocated(a%a)) res = a%a%m
  end function

  function dspmat_get_ncols(a) result(res)
    class(psb_dspmat_type), intent(in) :: a
    integer(psb_ipk_) :: res
    res = 0
    if (allocated(a%a)) res = a%a%n
  end function

  function dspmat_get_nzeros(a) result(res)
    class(psb_dspmat_type), intent(in) :: a
    integer(psb_epk_) :: res
    res = 0
    if (allocated(a%a)) then
      if (allocated(a%a%irp)) res = a%a%irp(a%a%m + 1) - 1
    end if
  end function

  ! the local block from 1-based CSR arrays
  subroutine dspmat_csr_bld(a, m, n, irp, ja, val, info)
    class(psb_dspmat_type), intent(inout) :: a
    integer(psb_ipk_), intent(in) :: m, n
    integer(psb_lpk_), intent(in) :: irp(:)
    integer(psb_ipk_), intent(in) :: ja(:)
    real(psb_dpk_), intent(in) :: val(:)
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    if (size(irp) < m + 1) then
      info = 1
      return
    end if
    if (allocated(a%a)) deallocate(a%a)
    allocate(psb_d_base_sparse_mat :: a%a)
    a%a%m = m
    a%a%n = n
    a%a%irp = irp(1:m + 1)
    a%a%ja = ja(1:irp(m + 1) - 1)
    a%a%val = val(1:irp(m + 1) - 1)
  end subroutine

  subroutine dspmat_free(a)
    class(psb_dspmat_type), intent(inout) :: a
    if (allocated(a%a)) deallocate(a%a)
  end subroutine

  function desc_get_local_rows(desc) result(res)
    class(psb_desc_type), intent(in) :: desc
    integer(psb_ipk_) :: res
    res = desc%n_row
  end function

  function desc_get_local_cols(desc) result(res)
    class(psb_desc_type), intent(in) :: desc
    integer(psb_ipk_) :: res
    res = desc%n_col
  end function

  function dbv_get_nrows(x) result(res)
    class(psb_d_base_vect_type), intent(in) :: x
    integer(psb_ipk_) :: res
    res = 0
    if (allocated(x%v)) res = size(x%v)
  end function

  function dbv_get_vect(x) result(res)
    class(psb_d_base_vect_type), intent(inout) :: x
    real(psb_dpk_), allocatable :: res(:)
    res = x%v
  end function

  subroutine dbv_set_vect(x, val)
    class(psb_d_base_vect_type), intent(inout) :: x
    real(psb_dpk_), intent(in) :: val(:)
    x%v = val
  end subroutine

  subroutine dbv_alloc(x, n, info)
    class(psb_d_base_vect_type), intent(inout) :: x
    integer(psb_ipk_), intent(in) :: n
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    if (allocated(x%v)) deallocate(x%v)
    allocate(x%v(n), stat=info)
    if (info == 0) x%v = dzero
  end subroutine

  ! device address of the storage (none: host storage)
  function dbv_dev_ptr(x) result(p)
    class(psb_d_base_vect_type), intent(in) :: x
    type(c_ptr) :: p
    p = c_null_ptr
  end function

  subroutine dbv_free(x, info)
    class(psb_d_base_vect_type), intent(inout) :: x
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    if (allocated(x%v)) deallocate(x%v)
  end subroutine

  ! x from a host array, its storage a copy of mold's class if present
  subroutine dv_bld(x, val, info, mold)
    class(psb_d_vect_type), intent(inout) :: x
    real(psb_dpk_), intent(in) :: val(:)
    integer(psb_ipk_), intent(out) :: info
    class(psb_d_base_vect_type), intent(in), optional :: mold
    integer(psb_ipk_) :: i2
    if (allocated(x%v)) then
      call x%v%free(i2)
      deallocate(x%v)
    end if
    if (present(mold)) then
      allocate(x%v, mold=mold)
    else
      allocate(psb_d_base_vect_type :: x%v)
    end if
    call x%v%alloc(int(size(val), psb_ipk_), info)
    if (info == psb_success_) call x%v%set_vect(val)
  end subroutine

  function dv_get_nrows(x) result(res)
    class(psb_d_vect_type), intent(in) :: x
    integer(psb_ipk_) :: res
    res = 0
    if (allocated(x%v)) res = x%v%get_nrows()
  end function

  function dv_get_vect(x) result(res)
    class(psb_d_vect_type), intent(inout) :: x
    real(psb_dpk_), allocatable :: res(:)
    if (allocated(x%v)) then
      res = x%v%get_vect()
    else
      allocate(res(0))
    end if
  end function

  subroutine dv_set_vect(x, val)
    class(psb_d_vect_type), intent(inout) :: x
    real(psb_dpk_), intent(in) :: val(:)
    if (allocated(x%v)) call x%v%set_vect(val)
  end subroutine

  subroutine dv_free(x, info)
    class(psb_d_vect_type), intent(inout) :: x
    integer(psb_ipk_), intent(out) :: info
    info = psb_success_
    if (allocated(x%v)) then
      call x%v%free(info)
      deallocate(x%v)
    end if
  end subroutine

end module psb_base_mod
