! lapl_lib.F90 (host) -- semi-implicit Poisson module over libfsim_hip.
!
! Keeps the interface of the reference module lapl_lib (lapl_lib.F90:14 ph,
! :22 lapl_init(g), :57 lapl_solve(g, ne, ni, nt, sig, lerr)).  PETSc is
! replaced by the library's matrix-free PCG.  The densities, the wall charge
! and ph given to lapl_solve are uploaded where they differ from what the
! device holds (dev_mirror) and the corrected ph is downloaded.
module lapl_lib
  use iso_c_binding
  use fsim_capi
  use dev_mirror
  use props
  use ptcl_lib, only: scalars_push, scalars_pull
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
  ! ne, ni, nt are the level-1 densities (bx+2, by+2), sig the wall charge
  ! sig_pl (bx+2, used with rwall).  lerr = 1 when the solve did not converge
  ! (ph unchanged).
  subroutine lapl_solve(g, ne, ni, nt, sig, lerr)
    type(grid), intent(inout) :: g
    real(8), intent(in), target, contiguous :: ne(:,:), ni(:,:), nt(:,:), sig(:)
    integer, intent(out) :: lerr
    integer(c_int32_t) :: l
    if (size(ne, 1) /= g%bx + 2 .or. size(ne, 2) /= g%by + 2 .or. &
        any(shape(ni) /= shape(ne)) .or. any(shape(nt) /= shape(ne)) .or. &
        size(sig) /= g%bx + 2) then
      write(*,*) 'Error: lapl_solve: densities not (bx+2, by+2) or sig not bx+2. Stop.'
      stop 4
    end if
    if (.not. resident) then
      call push2(g%h, FS_NE, 1_c_int32_t, ne, sh_n(:,:,1,2))
      call push2(g%h, FS_NI, 1_c_int32_t, ni, sh_n(:,:,1,1))
      call push2(g%h, FS_NT, 1_c_int32_t, nt, sh_n(:,:,1,3))
      call push2(g%h, FS_PH, 1_c_int32_t, ph, sh_ph)
      if (rwall) call push1(g%h, 1_c_int32_t, sig, sh_sig(:,1))
    end if
    call scalars_push(g)
    call fs_check(fs_lapl_solve(g%h, l, last_solve), 'lapl_solve')
    lerr = l
    call scalars_pull(g)
    if (.not. resident) call pull2(g%h, FS_PH, 1_c_int32_t, ph, sh_ph)
  end subroutine

  subroutine lapl_sync(g)
    type(grid), intent(in) :: g
    call pull2(g%h, FS_PH, 1_c_int32_t, ph, sh_ph)
  end subroutine
end module lapl_lib
