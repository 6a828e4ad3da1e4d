! dev_mirror.F90 (host) -- keeps the device state and the host module arrays
! in step at the module calls.
!
! The reference's main loop edits its module arrays between the module calls:
! the accept / reject block rotates the density levels, the circuit voltages
! and the wall charge and changes g%t and g%dt (main.F90:83-121).  The device
! owns the state, so each call first compares every host array it reads with
! a shadow copy of what the device holds and uploads the arrays that differ,
! and afterwards downloads what it wrote (refreshing the shadows).  A caller
! that edits its host arrays therefore drives the device exactly as the
! reference's arrays drive the CPU code, at the cost of the copies.
!
! With resident = .true. the driver keeps the state on the device (accept
! and reject by fs_accept / fs_reject), no array crosses at the calls, and
! the host arrays are refreshed only by the explicit *_sync calls.
module dev_mirror
  use iso_c_binding
  use, intrinsic :: ieee_arithmetic
  use fsim_capi
  implicit none

  logical :: resident = .false.

  ! what the device holds: densities (bx+2, by+2, level, species ni/ne/nt/nm),
  ! potential, wall charge (bx+2, sig_pl/sig_mi/sig_or), scalars
  real(8), allocatable :: sh_n(:,:,:,:), sh_ph(:,:), sh_sig(:,:)
  real(8) :: sh_sc(11)

contains

  ! Shadows start as NaN: no host value equals them, so the first call
  ! uploads every array.
  subroutine mirror_init(bx, by)
    integer, intent(in) :: bx, by
    real(8) :: nan
    nan = ieee_value(nan, ieee_quiet_nan)
    if (allocated(sh_n)) deallocate(sh_n, sh_ph, sh_sig)
    allocate(sh_n(bx+2, by+2, 3, 4), sh_ph(bx+2, by+2), sh_sig(bx+2, 3))
    sh_n = nan
    sh_ph = nan
    sh_sig = nan
    sh_sc = nan
  end subroutine

  ! upload a (bx+2, by+2) field level when it differs from the shadow
  subroutine push2(h, field, level, a, sh)
    type(c_ptr), intent(in) :: h
    integer(c_int32_t), intent(in) :: field, level
    real(8), intent(in), target, contiguous :: a(:,:)
    real(8), intent(inout) :: sh(:,:)
    if (all(a == sh)) return
    call fs_check(fs_set_field(h, field, level, c_loc(a), int(size(a, 1), c_int32_t)), &
                  'upload field')
    sh = a
  end subroutine

  subroutine pull2(h, field, level, a, sh)
    type(c_ptr), intent(in) :: h
    integer(c_int32_t), intent(in) :: field, level
    real(8), intent(inout), target, contiguous :: a(:,:)
    real(8), intent(inout) :: sh(:,:)
    call fs_check(fs_get_field(h, field, level, c_loc(a), int(size(a, 1), c_int32_t)), &
                  'download field')
    sh = a
  end subroutine

  ! the same for one wall-charge level (bx+2)
  subroutine push1(h, level, a, sh)
    type(c_ptr), intent(in) :: h
    integer(c_int32_t), intent(in) :: level
    real(8), intent(in), target, contiguous :: a(:)
    real(8), intent(inout) :: sh(:)
    if (all(a == sh)) return
    call fs_check(fs_set_field(h, FS_SIG, level, c_loc(a), 0_c_int32_t), 'upload sig')
    sh = a
  end subroutine

  subroutine pull1(h, level, a, sh)
    type(c_ptr), intent(in) :: h
    integer(c_int32_t), intent(in) :: level
    real(8), intent(inout), target, contiguous :: a(:)
    real(8), intent(inout) :: sh(:)
    call fs_check(fs_get_field(h, FS_SIG, level, c_loc(a), 0_c_int32_t), 'download sig')
    sh = a
  end subroutine
end module dev_mirror
