! fsim_capi.F90 -- ISO_C_BINDING declarations of libfsim_hip.so (include/fsim.h).
!
! The derived types mirror the C structs field by field (bind(C) layout); the
! interfaces are the C entry points.  The Fortran modules props, dev_mirror,
! circ_lib, sfc_lib, ptcl_lib and lapl_lib call only these.
module fsim_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FS_NI = 0, FS_NE = 1, FS_NT = 2, FS_NM = 3, FS_PH = 4
  integer(c_int), parameter :: FS_SIG = 9
  integer(c_int), parameter :: FS_REC_LEN = 10

  type, bind(C) :: fs_grid_desc
    integer(c_int32_t) :: nx, ny, grid_x, grid_y
    real(c_double) :: ratio_x, ratio_y, l, w, ew, gap
    integer(c_int32_t) :: cyl, rwall
  end type

  type, bind(C) :: fs_params
    real(c_double) :: p_torr, gam, vset, res_ohm
    integer(c_int32_t) :: rf, circuit
    real(c_double) :: t_rise, t_width, t_fall, dt0, lapl_rtol
    integer(c_int32_t) :: lapl_maxit, lapl_chunk, debug_keep_k, lapl_solver, use_graphs, lapl_warm
    integer(c_int32_t) :: stage_tile
  end type

  type, bind(C) :: fs_step_info
    real(c_double) :: err_cur, dt_stage, dt_next, t_m, Vd_pl, Ida, Idc, Vsrc
    integer(c_int32_t) :: attempts, status
  end type

  type, bind(C) :: fs_solve_stats
    integer(c_int32_t) :: iters, lerr
    real(c_double) :: resid, bnorm, seconds
  end type

  interface
    integer(c_int) function fs_create(g, p, h) bind(C, name="fs_create")
      import :: c_int, c_ptr, fs_grid_desc, fs_params
      type(fs_grid_desc), intent(in) :: g
      type(fs_params), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function fs_destroy(h) bind(C, name="fs_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fs_get_grid(h, x, y, dx, dlx, dy, dly, r, tx, ty) &
        bind(C, name="fs_get_grid")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, x, y, dx, dlx, dy, dly, r, tx, ty
    end function
    integer(c_int) function fs_set_field(h, field, level, host, ld) bind(C, name="fs_set_field")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h, host
      integer(c_int32_t), value :: field, level, ld
    end function
    integer(c_int) function fs_get_field(h, field, level, host, ld) bind(C, name="fs_get_field")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h, host
      integer(c_int32_t), value :: field, level, ld
    end function
    integer(c_int) function fs_init_uniform(h) bind(C, name="fs_init_uniform")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fs_init_gaussian(h, seed, peak, sigma, noise) &
        bind(C, name="fs_init_gaussian")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), value :: seed
      real(c_double), value :: peak, sigma, noise
    end function
    integer(c_int) function fs_circ_init(h, vset) bind(C, name="fs_circ_init")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: vset
    end function
    integer(c_int) function fs_get_scalars(h, out) bind(C, name="fs_get_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(11)
    end function
    integer(c_int) function fs_set_scalars(h, vals) bind(C, name="fs_set_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: vals(11)
    end function
    integer(c_int) function fs_ptcl_step(h, info) bind(C, name="fs_ptcl_step")
      import :: c_int, c_ptr, fs_step_info
      type(c_ptr), value :: h
      type(fs_step_info), intent(out) :: info
    end function
    integer(c_int) function fs_lapl_solve(h, lerr, st) bind(C, name="fs_lapl_solve")
      import :: c_int, c_int32_t, c_ptr, fs_solve_stats
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: lerr
      type(fs_solve_stats), intent(out) :: st
    end function
    integer(c_int) function fs_accept(h) bind(C, name="fs_accept")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fs_reject(h) bind(C, name="fs_reject")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fs_constants(p, gam, out) bind(C, name="fs_constants")
      import :: c_int, c_double
      real(c_double), value :: p, gam
      real(c_double), intent(out) :: out(40)
    end function
    type(c_ptr) function fs_last_error() bind(C, name="fs_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! Stop with the library's message on a fatal status (the reference stops too).
  subroutine fs_check(code, what)
    integer(c_int), intent(in) :: code
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (code >= 0) return
    call c_f_pointer(fs_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(a, a, i0, a, 512a)') what, ': status ', code, ': ', msg(1:n)
    stop 1
  end subroutine
end module fsim_capi
