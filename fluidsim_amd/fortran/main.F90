! main.F90 (host) -- driver of the MI355X plasma-transport path.
!
! Follows the reference driver (main.F90): defaults (main.F90:27-44), flag
! pairs (read_in, main.F90:172-295), the main loop with the two Poisson-only
! warm-up passes and accept / reject (main.F90:65-121), the progress line and
! the output files (writeOut, main.F90:297-333) in the reference's raw FP64
! format, so plot.py / save.py read them unchanged.  No MPI, no PETSc: one
! process drives one GPU through libfsim_hip.
program main
  use iso_c_binding
  use fsim_capi
  use props
  use circ_lib
  use ptcl_lib
  use lapl_lib
  implicit none

  type(grid), target :: g
  integer :: ts = 0, nx, ny, dof, lerr, max_steps, ic, iters_sum, nsave
  integer(c_int64_t) :: seed
  real(8) :: l, w, ew, gap, vl, dt, t_fin, t_sv, t_sv0, EmFlx = 0
  real(8), allocatable, target :: Em(:), sig(:)
  integer(8) :: c0, c1, crate, cpr
  character(256) :: path
  logical :: verbose

  ! defaults (main.F90:28-44)
  nx = 40; ny = 8; dof = 1
  p = 10d0
  call load_constants()
  l = 1e-2 / x0
  w = 1d-2 / x0
  ew = 2e-4 / x0
  gap = 1e-3 / x0
  dt = 5.0e-6
  t_fin = 100
  t_sv = 1e-4
  t_sv0 = 1e-4
  vl = 0
  res_ohm = 1e4
  max_steps = huge(1)
  ic = 0
  seed = 0
  path = ''
  verbose = .false.

  call read_in

  call g_init(g, nx, ny, 1, 1, dof, l, w, ew, gap, trim(path), vl, dt)
  call lapl_init(g)
  call ptcl_init(g, ic, seed)
  call circ_init(g, vl)
  allocate(Em(g%by+2), sig(g%bx+2))
  Em = 0
  sig = 0
  if (len_trim(path) > 0) call write_mesh

  g%t = 0
  g%dt = dt
  iters_sum = 0
  nsave = 0
  call system_clock(c0, crate)
  cpr = c0
  do
    if (g%t >= t_fin .or. ts >= max_steps) exit

    t_m = 1e9
    if (ts > 1) call ptcl_step(g, ph, Em)

    call lapl_solve(g, ne(:,:,1), ni(:,:,1), nt(:,:,1), sig, lerr)
    iters_sum = iters_sum + last_solve%iters

    ! accept (main.F90:83-101) / reject (main.F90:103-121)
    if (lerr == 0) then
      call fs_check(fs_accept(g%h), 'accept')
      ts = ts + 1
    else
      call fs_check(fs_reject(g%h), 'reject')
    end if
    call circ_sync(g)

    call system_clock(c1)
    if (verbose .or. real(c1 - cpr, 8) / crate >= 5d0) then
      write(*,11) ts, g%t, g%dt, last_solve%iters, last_solve%resid, Vd_pl * ph0
      cpr = c1
    end if

    if (len_trim(path) > 0 .and. t_sv <= g%t) then
      call writeOut
      t_sv = t_sv + t_sv0
      if (g%t > 2d-1) then
        t_sv0 = t_sv0 * 1.01
        t_sv0 = min(t_sv0, 2d-2)
      end if
    end if
  end do
  if (len_trim(path) > 0) call writeOut

  call system_clock(c1)
  write(*,12) ts, g%t, real(c1 - c0, 8) / crate, iters_sum
  call fs_check(fs_destroy(g%h), 'fs_destroy')

11 format('Timestep:', i7, '  Time:', es11.4, '  dT:', es10.3, '  pcg:', i6, &
          '  res:', es9.2, '  Vd:', f8.2)
12 format('Finished ', i0, ' steps, t = ', es12.5, ' us in ', f9.3, ' s (', i0, &
          ' PCG iterations)')

contains

  ! Flag pairs as the reference's read_in (main.F90:172-295), plus the
  ! runtime physics the reference fixes at compile time.
  subroutine read_in
    integer :: i, narg
    character(256) :: arg, val
    narg = command_argument_count()
    if (mod(narg, 2) /= 0) then
      write(*,*) 'Usage: fsim_main <flag> <value> ...'
      write(*,*) '  -nx -ny -l -w -ew -t -dt -v -r -rf -unif -cyl (reference flags)'
      write(*,*) '  -p -gam -grid -ratio -circuit -ic -seed -steps -rtol -out -verbose'
      stop
    end if
    do i = 1, narg / 2
      call get_command_argument(2 * i - 1, arg)
      call get_command_argument(2 * i, val)
      select case (arg)
        case ('-nx'); read(val,*) nx
        case ('-ny'); read(val,*) ny
        case ('-l'); read(val,*) l; l = l / x0
        case ('-w'); read(val,*) w; w = w / x0
        case ('-ew'); read(val,*) ew; ew = ew / x0
        case ('-t'); read(val,*) t_fin
        case ('-dt'); read(val,*) dt
        case ('-v'); read(val,*) vl; vl = vl / ph0
        case ('-r'); read(val,*) res_ohm
        case ('-rf'); read(val,*) rf
        case ('-unif'); read(val,*) unif
        case ('-cyl'); read(val,*) cyl
        case ('-p'); read(val,*) p
        case ('-gam'); read(val,*) gam
        case ('-grid'); read(val,*) grid_x
        case ('-ratio'); read(val,*) ratio
        case ('-circuit'); read(val,*) circuit
        case ('-ic'); read(val,*) ic
        case ('-seed'); read(val,*) seed
        case ('-steps'); read(val,*) max_steps
        case ('-rtol'); read(val,*) lapl_rtol
        case ('-solver'); read(val,*) lapl_solver
        case ('-out'); path = val
        case ('-verbose'); read(val,*) verbose
        case ('-trise'); read(val,*) t_rise
        case ('-twidth'); read(val,*) t_width
        case ('-tfall'); read(val,*) t_fall
      end select
    end do
    if (ny < 2) then
      write(*,*) 'Error: this build is two-dimensional (ny >= 2). Stop.'
      stop 3
    end if
  end subroutine

  subroutine append(name, a)
    character(*), intent(in) :: name
    real(8), intent(in) :: a(:)
    integer :: u
    open(newunit=u, file=trim(path)//name, access='stream', form='unformatted', &
         status='unknown', position='append')
    write(u) a
    close(u)
  end subroutine

  subroutine write_mesh
    integer :: u
    character(16) :: names(12)
    integer :: k
    names = [character(16) :: 'meshx.dat', 'meshy.dat', 'time.dat', 'dt.dat', 'vd.dat', &
             'ida.dat', 'idc.dat', 'flx.dat', 'f1.dat', 'f2.dat', 'f3.dat', 'f4.dat']
    do k = 1, 12
      open(newunit=u, file=trim(path)//trim(names(k)), access='stream', &
           form='unformatted', status='replace')
      close(u)
    end do
    open(newunit=u, file=trim(path)//'f5.dat', access='stream', form='unformatted', &
         status='replace')
    close(u)
    call append('meshx.dat', g%x(2:g%bx+1) * x0)
    call append('meshy.dat', g%y(2:g%by+1) * x0)
  end subroutine

  ! writeOut (main.F90:297-333): interior fields, reference units and order
  subroutine writeOut
    call ptcl_sync(g)
    call lapl_sync(g)
    call append('f1.dat', reshape(ph(2:g%bx+1, 2:g%by+1) * ph0, [g%bx * g%by]))
    call append('f2.dat', reshape(ne(2:g%bx+1, 2:g%by+1, 1) / x0**3, [g%bx * g%by]))
    call append('f3.dat', reshape(ni(2:g%bx+1, 2:g%by+1, 1) / x0**3, [g%bx * g%by]))
    call append('f4.dat', reshape(nt(2:g%bx+1, 2:g%by+1, 1) / x0**3 * ph0 / 1.5, &
                                  [g%bx * g%by]))
    call append('f5.dat', reshape(nm(2:g%bx+1, 2:g%by+1, 1) / x0**3, [g%bx * g%by]))
    call append('time.dat', [g%t])
    call append('vd.dat', [Vd_pl * ph0])
    call append('ida.dat', [Ida * e / t0])
    call append('idc.dat', [Idc * e / t0])
    call append('dt.dat', [g%dt])
    call append('flx.dat', [EmFlx])
    nsave = nsave + 1
  end subroutine
end program main
