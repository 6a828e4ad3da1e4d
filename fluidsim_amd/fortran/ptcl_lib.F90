! ptcl_lib.F90 (host) -- species module over libfsim_hip.
!
! Keeps the interface of the reference module ptcl_lib (ptcl_lib.F90:7-10
! module arrays, :15 ptcl_init, :32 ptcl_step).  The densities live on the
! device; the module arrays ni, ne, nt, nm (bx+2, by+2, 3) are host mirrors
! filled by ptcl_sync (used for output only).
module ptcl_lib
  use iso_c_binding
  use fsim_capi
  use props
  use circ_lib
  implicit none

  real(8), allocatable, target :: ni(:,:,:), ne(:,:,:), nt(:,:,:), nm(:,:,:)
  type(fs_step_info) :: last_step

contains

  ! *** ptcl_init (ptcl_lib.F90:15-30); ic = 1 selects the Gaussian blob ***
  subroutine ptcl_init(g, ic, seed)
    type(grid), intent(inout) :: g
    integer, intent(in) :: ic
    integer(c_int64_t), intent(in) :: seed
    allocate(ni(g%bx+2, g%by+2, 3), ne(g%bx+2, g%by+2, 3), &
             nt(g%bx+2, g%by+2, 3), nm(g%bx+2, g%by+2, 3))
    if (ic == 1) then
      ! 1e8 cm^-3 peak, sigma 0.1 cm, 1% noise
      call fs_check(fs_init_gaussian(g%h, seed, 1d14 * x0**3, 1d-3 / x0, 1d-2), &
                    'fs_init_gaussian')
    else
      call fs_check(fs_init_uniform(g%h), 'fs_init_uniform')
    end if
    call ptcl_sync(g)
  end subroutine

  ! *** ptcl_step(g, ph, Em) (ptcl_lib.F90:32-97): five RK-Merson stages,
  ! circuit, controller and restarts, all on the device.  ph and Em are kept
  ! for the interface: the potential is device-resident and Em = 0.
  subroutine ptcl_step(g, ph, Em)
    type(grid), intent(inout) :: g
    real(8), intent(inout) :: ph(:,:), Em(:)
    call fs_check(fs_ptcl_step(g%h, last_step), 'ptcl_step')
    call circ_sync(g)
    if (size(ph) < 0 .or. size(Em) < 0) write(*,*) 'unreachable'
  end subroutine

  subroutine ptcl_sync(g)
    type(grid), intent(in) :: g
    integer :: lev
    integer(c_int32_t) :: ld
    ld = g%bx + 2
    do lev = 1, 3
      call fs_check(fs_get_field(g%h, FS_NI, lev, c_loc(ni(1,1,lev)), ld), 'get ni')
      call fs_check(fs_get_field(g%h, FS_NE, lev, c_loc(ne(1,1,lev)), ld), 'get ne')
      call fs_check(fs_get_field(g%h, FS_NT, lev, c_loc(nt(1,1,lev)), ld), 'get nt')
      call fs_check(fs_get_field(g%h, FS_NM, lev, c_loc(nm(1,1,lev)), ld), 'get nm')
    end do
  end subroutine
end module ptcl_lib
