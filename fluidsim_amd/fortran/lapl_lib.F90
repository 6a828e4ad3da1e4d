! lapl_lib.F90 (host) -- semi-implicit Poisson module over libfsim_hip.
!
! Keeps the interface of the reference module lapl_lib (lapl_lib.F90:14 ph,
! :22 lapl_init, :57 lapl_solve).  PETSc is replaced by the library's
! matrix-free PCG; the potential ph is a host mirror filled by lapl_sync.
module lapl_lib
  use iso_c_binding
  use fsim_capi
  use props
  implicit none

  real(8), allocatable, target :: ph(:,:)
  type(fs_solve_stats) :: last_solve

contains

  subroutine lapl_init(g)
    type(grid), intent(in) :: g
    allocate(ph(g%bx+2, g%by+2))
    ph = 0
  end subroutine

  ! *** lapl_solve(g, ne, ni, nt, sig, lerr) (lapl_lib.F90:57-121) ***
  ! The level-1 densities are device-resident; the arguments are kept for the
  ! interface.  lerr = 1 when the solve did not converge (ph unchanged).
  subroutine lapl_solve(g, ne, ni, nt, sig, lerr)
    type(grid), intent(in) :: g
    real(8), intent(in) :: ne(:,:), ni(:,:), nt(:,:), sig(:)
    integer, intent(out) :: lerr
    integer(c_int32_t) :: l
    integer(c_int) :: rc
    rc = fs_lapl_solve(g%h, l, last_solve)
    call fs_check(rc, 'lapl_solve')
    lerr = l
    if (size(ne) + size(ni) + size(nt) + size(sig) < 0) write(*,*) 'unreachable'
  end subroutine

  subroutine lapl_sync(g)
    type(grid), intent(in) :: g
    call fs_check(fs_get_field(g%h, FS_PH, 1, c_loc(ph), g%bx + 2), 'get ph')
  end subroutine
end module lapl_lib
