! circ_lib.F90 (host) -- external circuit state over libfsim_hip.
!
! Keeps the public names of the reference module (circ_lib.F90:7-14): Vd_pl,
! Vd_mi, Vd_or, Ida, Idc, Res and circ_init.  circ_step itself runs on the
! device inside ptcl_step (one small kernel per stage); the host mirrors are
! refreshed by circ_sync.
module circ_lib
  use iso_c_binding
  use fsim_capi
  use props
  implicit none

  real(8) :: Vd_pl, Vd_mi, Vd_or, Ida, Idc, Vsrc, Res

contains

  ! *** External Circuit Initialization (circ_lib.F90:213-223) ***
  subroutine circ_init(g, Vset)
    type(grid), intent(inout) :: g
    real(8), intent(in) :: Vset
    real(c_double) :: s(11)
    call fs_check(fs_get_scalars(g%h, s), 'fs_get_scalars')
    s(5) = Vset; s(6) = Vset; s(7) = Vset; s(10) = Vset
    call fs_check(fs_set_scalars(g%h, s), 'fs_set_scalars')
    Res = res_ohm * e / (ph0 * t0)  ! main.F90:286
    call circ_sync(g)
  end subroutine

  subroutine circ_sync(g)
    type(grid), intent(inout) :: g
    real(c_double) :: s(11)
    call fs_check(fs_get_scalars(g%h, s), 'fs_get_scalars')
    g%t = s(1); g%dt = s(2); t_m = s(4)
    Vd_pl = s(5); Vd_mi = s(6); Vd_or = s(7); Ida = s(8); Idc = s(9); Vsrc = s(10)
  end subroutine
end module circ_lib
