! sfc_lib.F90 (host) -- dielectric wall surface charge over libfsim_hip.
!
! Keeps the public names of the reference module (sfc_lib.F90:7-15): the
! levels sig_pl, sig_mi, sig_or (bx+2) and sfc_init(g).  sfc_step runs on the
! device inside ptcl_step, fused into the stage kernel's top row; the host
! levels are exchanged with the device by ptcl_lib at each module call when
! the wall is dielectric (rwall).
module sfc_lib
  use iso_c_binding
  use props
  use dev_mirror
  implicit none

  real(8), allocatable, target :: sig_pl(:), sig_mi(:), sig_or(:)

contains

  ! *** Surface Charge Initialization (sfc_lib.F90:82-90) ***
  subroutine sfc_init(g)
    type(grid), intent(inout) :: g
    allocate(sig_pl(g%bx+2), sig_mi(g%bx+2), sig_or(g%bx+2))
    sig_pl = 0d0
    sig_mi = 0d0
    sig_or = 0d0
  end subroutine

  ! host levels <- device (rwall only: without the wall the charge is zero)
  subroutine sfc_sync(g)
    type(grid), intent(in) :: g
    if (.not. rwall) return
    call pull1(g%h, 1_c_int32_t, sig_pl, sh_sig(:,1))
    call pull1(g%h, 2_c_int32_t, sig_mi, sh_sig(:,2))
    call pull1(g%h, 3_c_int32_t, sig_or, sh_sig(:,3))
  end subroutine

  subroutine sfc_push(g)
    type(grid), intent(in) :: g
    if (.not. rwall) return
    call push1(g%h, 1_c_int32_t, sig_pl, sh_sig(:,1))
    call push1(g%h, 2_c_int32_t, sig_mi, sh_sig(:,2))
    call push1(g%h, 3_c_int32_t, sig_or, sh_sig(:,3))
  end subroutine
end module sfc_lib
