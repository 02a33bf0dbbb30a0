! iso_c_binding interfaces to libmldainv_hip.so (include/mld_ainv_hip.h).
! Thin layer: one interface per C entry point, no logic.
module mld_ainv_hip_cbind
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ainv_success = 0
  integer(c_int), parameter :: ainv_kind_ainv = 1, ainv_kind_invk = 2, &
       & ainv_kind_invt = 3, ainv_kind_ml = 4
  integer(c_int), parameter :: ainv_alg_llk = 1, ainv_alg_s_llk = 2, &
       & ainv_alg_s_ft_llk = 3, ainv_alg_llk_noth = 4, ainv_alg_mlk = 5
  integer(c_int), parameter :: ainv_prec_f64 = 0, ainv_prec_f32 = 1, &
       & ainv_prec_bf16 = 2
  integer(c_int), parameter :: ainv_host_ptrs = 1
  integer(c_int), parameter :: ainv_h2d = 1, ainv_d2h = 2, ainv_d2d = 3
  integer(c_int), parameter :: ainv_err_state = 6

  interface
    function ainv_init(device) bind(c, name='ainv_init')
      import :: c_int
      integer(c_int), value :: device
      integer(c_int) :: ainv_init
    end function
    function ainv_last_error() bind(c, name='ainv_last_error')
      import :: c_ptr
      type(c_ptr) :: ainv_last_error
    end function
    function ainv_create(kind, h) bind(c, name='ainv_create')
      import :: c_int, c_ptr
      integer(c_int), value :: kind
      type(c_ptr) :: h
      integer(c_int) :: ainv_create
    end function
    function ainv_free(h) bind(c, name='ainv_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: ainv_free
    end function
    function ainv_default(h) bind(c, name='ainv_default')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: ainv_default
    end function
    function ainv_cseti(h, what, val) bind(c, name='ainv_cseti')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char) :: what(*)
      integer(c_int), value :: val
      integer(c_int) :: ainv_cseti
    end function
    function ainv_csetr(h, what, val) bind(c, name='ainv_csetr')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      character(kind=c_char) :: what(*)
      real(c_double), value :: val
      integer(c_int) :: ainv_csetr
    end function
    function ainv_csetc(h, what, val) bind(c, name='ainv_csetc')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char) :: what(*), val(*)
      integer(c_int) :: ainv_csetc
    end function
    function ainv_check(h) bind(c, name='ainv_check')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: ainv_check
    end function
    function ainv_descr(h, buf, len) bind(c, name='ainv_descr')
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: h
      character(kind=c_char) :: buf(*)
      integer(c_size_t), value :: len
      integer(c_int) :: ainv_descr
    end function
    function ainv_build_csr(h, n, irp, ja, val, base) &
         & bind(c, name='ainv_build_csr')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
      integer(c_int64_t) :: irp(*)
      integer(c_int32_t) :: ja(*)
      real(c_double) :: val(*)
      integer(c_int), value :: base
      integer(c_int) :: ainv_build_csr
    end function
    function ainv_cnv(h, prec) bind(c, name='ainv_cnv')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: prec
      integer(c_int) :: ainv_cnv
    end function
    ! x, y: device addresses, or host ones (c_loc) with ainv_host_ptrs
    function ainv_apply(h, alpha, x, beta, y, trans, k, ldx, ldy, flags, &
         & stream) bind(c, name='ainv_apply')
      import :: c_int, c_ptr, c_double, c_char, c_int64_t
      type(c_ptr), value :: h
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: x, y
      character(kind=c_char), value :: trans
      integer(c_int), value :: k, flags
      integer(c_int64_t), value :: ldx, ldy
      type(c_ptr), value :: stream
      integer(c_int) :: ainv_apply
    end function
    function ainv_get_nzeros(h, nz) bind(c, name='ainv_get_nzeros')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t) :: nz
      integer(c_int) :: ainv_get_nzeros
    end function
    function ainv_sizeof(h, nb) bind(c, name='ainv_sizeof')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t) :: nb
      integer(c_int) :: ainv_sizeof
    end function
    function ainv_get_wrksz(h, w) bind(c, name='ainv_get_wrksz')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: w
      integer(c_int) :: ainv_get_wrksz
    end function
    function ainv_clone(h, out) bind(c, name='ainv_clone')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr) :: out
      integer(c_int) :: ainv_clone
    end function
    function ainv_clone_settings(h, out) bind(c, name='ainv_clone_settings')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, out
      integer(c_int) :: ainv_clone_settings
    end function
    function ainv_clear_data(h) bind(c, name='ainv_clear_data')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: ainv_clear_data
    end function
    function ainv_update_a(h, x, nx, flags) bind(c, name='ainv_update_a')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      type(c_ptr), value :: x
      integer(c_int64_t), value :: nx
      integer(c_int), value :: flags
      integer(c_int) :: ainv_update_a
    end function
    function ainv_dump(h, prefix, level, solver, head) &
         & bind(c, name='ainv_dump')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char) :: prefix(*), head(*)
      integer(c_int), value :: level, solver
      integer(c_int) :: ainv_dump
    end function
    function ainv_mat_create(n, irp, ja, val, base, a) &
         & bind(c, name='ainv_mat_create')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: n
      integer(c_int64_t) :: irp(*)
      integer(c_int32_t) :: ja(*)
      real(c_double) :: val(*)
      integer(c_int), value :: base
      type(c_ptr) :: a
      integer(c_int) :: ainv_mat_create
    end function
    function ainv_mat_free(a) bind(c, name='ainv_mat_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: a
      integer(c_int) :: ainv_mat_free
    end function
    function ainv_krylov(method, a, prec, b, x, k, eps, itmax, istop, irst, &
         & iter, err, flags, stream) bind(c, name='ainv_krylov')
      import :: c_int, c_ptr, c_char, c_double, c_int32_t
      character(kind=c_char) :: method(*)
      type(c_ptr), value :: a, prec
      type(c_ptr), value :: b, x
      integer(c_int), value :: k, itmax, istop, irst, flags
      real(c_double), value :: eps
      integer(c_int32_t) :: iter(*)
      real(c_double) :: err(*)
      type(c_ptr), value :: stream
      integer(c_int) :: ainv_krylov
    end function
    ! device memory for vectors kept on the GPU (no counterpart in the
    ! reference, where PSBLAS vectors own their storage)
    function ainv_dev_alloc(bytes, p) bind(c, name='ainv_dev_alloc')
      import :: c_int, c_ptr, c_size_t
      integer(c_size_t), value :: bytes
      type(c_ptr) :: p
      integer(c_int) :: ainv_dev_alloc
    end function
    function ainv_dev_free(p) bind(c, name='ainv_dev_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: p
      integer(c_int) :: ainv_dev_free
    end function
    function ainv_memcpy(dst, src, bytes, kind) bind(c, name='ainv_memcpy')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int), value :: kind
      integer(c_int) :: ainv_memcpy
    end function
    function ainv_device_sync() bind(c, name='ainv_device_sync')
      import :: c_int
      integer(c_int) :: ainv_device_sync
    end function
    ! irp / ja / val: c_loc of the arrays, or c_null_ptr to query sizes
    function ainv_gen_pde3d(nn, hconv, n, nnz, irp, ja, val) &
         & bind(c, name='ainv_gen_pde3d')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      integer(c_int32_t), value :: nn
      integer(c_int), value :: hconv
      integer(c_int32_t) :: n
      integer(c_int64_t) :: nnz
      type(c_ptr), value :: irp, ja, val
      integer(c_int) :: ainv_gen_pde3d
    end function
  end interface

contains

  ! Fortran string -> NUL-terminated C string
  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char) :: c(len_trim(s) + 1)
    integer :: i
    do i = 1, len_trim(s)
      c(i) = s(i:i)
    end do
    c(len_trim(s) + 1) = c_null_char
  end function cstr

  ! last error message of the library
  function last_error() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(ainv_last_error(), p, [256])
    do i = 1, 256
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function last_error

end module mld_ainv_hip_cbind
