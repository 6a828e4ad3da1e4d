! circ_lib.F90 (host) -- external circuit state over libfsim_hip.
!
! Keeps the public names of the reference module (circ_lib.F90:7-14): Vd_pl,
! Vd_mi, Vd_or, Ida, Idc, Vsrc, Vmax, Res, Cap and circ_init(Vset).  circ_step
! itself runs on the device inside ptcl_step; the host values are exchanged
! with the device by ptcl_lib at each module call.
module circ_lib
  use iso_c_binding
  use fsim_capi
  use props
  implicit none

  real(8) :: Vd_pl = 0, Vd_mi = 0, Vd_or = 0, Ida = 0, Idc = 0, Vsrc = 0, Vmax = 0
  real(8) :: Res = 0, Cap = 0

contains

  ! *** External Circuit Initialization (circ_lib.F90:213-223) ***
  subroutine circ_init(Vset)
    real(8), intent(in) :: Vset
    Vmax  = Vset
    Vsrc  = Vset
    Vd_pl = Vset
    Vd_mi = Vset
    Vd_or = Vset
    Cap = 1d-11 * ph0 / e
    Res = res_ohm * e / (ph0 * t0)  ! main.F90:286
    call fs_check(fs_circ_init(dev_h, Vset), 'circ_init')
  end subroutine
end module circ_lib
