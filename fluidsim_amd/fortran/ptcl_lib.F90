! ptcl_lib.F90 (host) -- species module over libfsim_hip.
!
! Keeps the interface of the reference module ptcl_lib (ptcl_lib.F90:7-10
! module arrays ni, ne, nt, nm (bx+2, by+2, 3) and err_prev, :15 ptcl_init(g),
! :32 ptcl_step(g, ph, Em)).  The five RK-Merson stages, the circuit, the wall
! charge, the controller and its restarts run on the device; the module
! arrays, ph, the wall charge and the scalars are exchanged with the device at
! each call (dev_mirror), so a caller that edits them between calls, as the
! reference's main loop does, is honoured.
module ptcl_lib
  use iso_c_binding
  use fsim_capi
  use dev_mirror
  use props
  use circ_lib
  use sfc_lib
  implicit none

  real(8), allocatable, target :: ni(:,:,:), ne(:,:,:), nt(:,:,:), nm(:,:,:)
  real(8) :: err_prev = 1
  ! initial condition: 0 the reference's uniform n_init (ptcl_lib.F90:22-29),
  ! 1 a Gaussian blob with seeded noise
  integer :: ic = 0
  integer(c_int64_t) :: seed = 0
  type(fs_step_info) :: last_step

  integer(c_int32_t), parameter, private :: SPECIES(4) = [FS_NI, FS_NE, FS_NT, FS_NM]

contains

  ! *** ptcl_init (ptcl_lib.F90:15-30) ***
  subroutine ptcl_init(g)
    type(grid), intent(inout) :: g
    allocate(ni(g%bx+2, g%by+2, 3), ne(g%bx+2, g%by+2, 3), &
             nt(g%bx+2, g%by+2, 3), nm(g%bx+2, g%by+2, 3))
    if (ic == 1) then
      ! 1e8 cm^-3 peak, sigma 0.1 cm, 1% noise
      call fs_check(fs_init_gaussian(g%h, seed, 1d14 * x0**3, 1d-3 / x0, 1d-2), &
                    'fs_init_gaussian')
      call ptcl_sync(g)
    else
      ni = n_init
      ne = n_init
      nt = ne * Tg * kb / e / ph0
      nm = n_init
      call ptcl_push(g)
    end if
  end subroutine

  ! *** ptcl_step(g, ph, Em) (ptcl_lib.F90:32-97) ***
  ! Em is the microwave field of em_lib, which this build does not carry: it
  ! must be zero.
  subroutine ptcl_step(g, ph, Em)
    type(grid), intent(inout) :: g
    real(8), intent(inout), target, contiguous :: ph(:,:)
    real(8), intent(inout) :: Em(:)
    if (any(Em /= 0d0)) then
      write(*,*) 'Error: ptcl_step with Em /= 0: the microwave field is not built. Stop.'
      stop 4
    end if
    if (size(ph, 1) /= g%bx + 2 .or. size(ph, 2) /= g%by + 2) then
      write(*,*) 'Error: ptcl_step: ph is not (bx+2, by+2). Stop.'
      stop 4
    end if
    if (.not. resident) then
      call ptcl_push(g)
      call push2(g%h, FS_PH, 1_c_int32_t, ph, sh_ph)
      call sfc_push(g)
    end if
    call scalars_push(g)
    call fs_check(fs_ptcl_step(g%h, last_step), 'ptcl_step')
    call scalars_pull(g)
    if (.not. resident) then
      call ptcl_sync(g)
      call pull2(g%h, FS_PH, 1_c_int32_t, ph, sh_ph)
      call sfc_sync(g)
    end if
  end subroutine

  ! host arrays -> device, the levels that differ from what it holds
  subroutine ptcl_push(g)
    type(grid), intent(in) :: g
    integer :: lev
    do lev = 1, 3
      call push2(g%h, SPECIES(1), int(lev, c_int32_t), ni(:,:,lev), sh_n(:,:,lev,1))
      call push2(g%h, SPECIES(2), int(lev, c_int32_t), ne(:,:,lev), sh_n(:,:,lev,2))
      call push2(g%h, SPECIES(3), int(lev, c_int32_t), nt(:,:,lev), sh_n(:,:,lev,3))
      call push2(g%h, SPECIES(4), int(lev, c_int32_t), nm(:,:,lev), sh_n(:,:,lev,4))
    end do
  end subroutine

  ! device -> host arrays
  subroutine ptcl_sync(g)
    type(grid), intent(in) :: g
    integer :: lev
    do lev = 1, 3
      call pull2(g%h, SPECIES(1), int(lev, c_int32_t), ni(:,:,lev), sh_n(:,:,lev,1))
      call pull2(g%h, SPECIES(2), int(lev, c_int32_t), ne(:,:,lev), sh_n(:,:,lev,2))
      call pull2(g%h, SPECIES(3), int(lev, c_int32_t), nt(:,:,lev), sh_n(:,:,lev,3))
      call pull2(g%h, SPECIES(4), int(lev, c_int32_t), nm(:,:,lev), sh_n(:,:,lev,4))
    end do
  end subroutine

  ! time, circuit and controller scalars: host -> device when changed
  subroutine scalars_push(g)
    type(grid), intent(in) :: g
    real(c_double) :: s(11)
    ! before the first exchange the shadow is unset: take the device's values
    ! (its step count ts has no host variable)
    if (any(sh_sc /= sh_sc)) call fs_check(fs_get_scalars(g%h, sh_sc), 'fs_get_scalars')
    s = [g%t, g%dt, err_prev, t_m, Vd_pl, Vd_mi, Vd_or, Ida, Idc, Vsrc, sh_sc(11)]
    if (all(s == sh_sc)) return
    call fs_check(fs_set_scalars(g%h, s), 'fs_set_scalars')
    sh_sc = s
  end subroutine

  subroutine scalars_pull(g)
    type(grid), intent(inout) :: g
    real(c_double) :: s(11)
    call fs_check(fs_get_scalars(g%h, s), 'fs_get_scalars')
    g%t = s(1); g%dt = s(2); err_prev = s(3); t_m = s(4)
    Vd_pl = s(5); Vd_mi = s(6); Vd_or = s(7); Ida = s(8); Idc = s(9); Vsrc = s(10)
    sh_sc = s
  end subroutine
end module ptcl_lib
