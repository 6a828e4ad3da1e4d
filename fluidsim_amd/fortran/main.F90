! main.F90 (host) -- driver of the MI355X plasma-transport path.
!
! Follows the reference driver (main.F90): defaults (main.F90:27-44), flag
! pairs (read_in, main.F90:172-295), the module calls in the reference's
! order and with its arguments (g_init, lapl_init, ptcl_init, circ_init,
! sfc_init, main.F90:51-55; ptcl_step and lapl_solve, main.F90:75-80), the
! accept / reject block on the module arrays (main.F90:83-121), the progress
! block (main.F90:123-139) and the output files (writeOut, main.F90:297-333)
! in the reference's raw FP64 format, so plot.py / save.py read them
! unchanged.  No MPI, no PETSc: one process drives one GPU through
! libfsim_hip.
!
! -resident T keeps the state on the device instead: accept and reject run
! there (fs_accept / fs_reject) and the module arrays are refreshed only for
! the output, which removes the host copies from every pass.
program main
  use iso_c_binding
  use fsim_capi
  use dev_mirror
  use props
  use circ_lib
  use sfc_lib
  use ptcl_lib
  use lapl_lib
  implicit none

  type(grid), target :: g
  integer :: ts = 0, ts1, ts2, nx, ny, px, py, dof, lerr, max_steps, iters_sum
  real(8) :: l, w, ew, gap, vl, dt, t_fin, t_pr, t_sv, t_sv0, time1, time2, sim_start, &
             EmFlx = 0
  real(8), allocatable, target :: Em(:)
  character(256) :: path
  logical :: verbose

  call cpu_time(sim_start)
  call cpu_time(time1)
  ts1 = 0

  ! defaults (main.F90:28-44), the one-dimensional ny = 1 included
  nx = 40; ny = 1; px = 1; py = 1; dof = 1
  call load_constants()
  l = 1e-2 / x0
  w = 1d-2 / x0
  ew = 2e-4 / x0
  gap = 1e-3 / x0
  dt = 5.0e-6
  t_fin = 100
  t_pr = 0d0
  t_sv = 1e-4
  t_sv0 = 1e-4
  vl = 0 / ph0
  res_ohm = 1e4
  max_steps = huge(1)
  path = ''
  verbose = .false.

  call read_in

  ! Initialize grid and arrays (main.F90:51-55)
  call g_init(g, nx, ny, px, py, dof, l, w, ew, gap, trim(path))
  call lapl_init(g)
  call ptcl_init(g)
  call circ_init(vl)
  call sfc_init(g)
  allocate(Em(g%by+2))
  Em = 0

  g%t = 0
  g%dt = dt
  iters_sum = 0

  do
    if (g%t >= t_fin .or. ts >= max_steps) exit

    ! Solve ne system
    t_m = 1e9
    if (ts > 1) call ptcl_step(g, ph, Em)

    ! Solve ph system
    call lapl_solve(g, ne(:,:,1), ni(:,:,1), nt(:,:,1), sig_pl, lerr)
    iters_sum = iters_sum + last_solve%iters

    if (resident) then
      ! accept / reject on the device
      if (lerr == 0) then
        call fs_check(fs_accept(g%h), 'accept')
        ts = ts + 1
      else
        call fs_check(fs_reject(g%h), 'reject')
      end if
      call scalars_pull(g)
    ! Accept step (main.F90:83-101)
    else if (lerr == 0) then
      ts = ts + 1
      g%t = g%t + g%dt

      ni(:,:,3) = ni(:,:,2)
      ne(:,:,3) = ne(:,:,2)
      nm(:,:,3) = nm(:,:,2)
      nt(:,:,3) = nt(:,:,2)

      ni(:,:,2) = ni(:,:,1)
      ne(:,:,2) = ne(:,:,1)
      nm(:,:,2) = nm(:,:,1)
      nt(:,:,2) = nt(:,:,1)

      Vd_or = Vd_mi
      Vd_mi = Vd_pl

      sig_or = sig_mi
      sig_mi = sig_pl
    ! Reject step (main.F90:103-121)
    else
      g%dt = g%dt / 2d0

      ni(:,:,1) = ni(:,:,3)
      ne(:,:,1) = ne(:,:,3)
      nm(:,:,1) = nm(:,:,3)
      nt(:,:,1) = nt(:,:,3)

      ni(:,:,2) = ni(:,:,3)
      ne(:,:,2) = ne(:,:,3)
      nm(:,:,2) = nm(:,:,3)
      nt(:,:,2) = nt(:,:,3)

      Vd_pl = Vd_or
      Vd_mi = Vd_or

      sig_pl = sig_or
      sig_mi = sig_or
    end if

    ! Print out some information (main.F90:123-139)
    if (t_pr >= 5d0 .or. verbose) then
      call cpu_time(time2)
      ts2 = ts
      if (resident) call ptcl_sync(g)
      write(*,*)
      write(*,11) ts, g%t, (time2 - time1) / g%dt / float(max(ts2 - ts1, 1)) / 60.
      write(*,12) g%dt, t_m, sqrt(abs(EmFlx)) * ph0 / x0
      write(*,15) Vd_pl * ph0, Ida * e / t0, Idc * e / t0
      write(*,16) sqrt(maxval(ne) * wp02) / t0 / 2e9 / pi, maxval(ne) / x0**3
      write(*,17) last_solve%iters, last_solve%resid
      t_pr = 0
      time1 = time2
      ts1 = ts
    else if (mod(ts, 100) == 0) then
      call cpu_time(time2)
      t_pr = time2 - time1
    end if

    ! Save data
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

  call cpu_time(time1)
  write(*,*)
  write(*,9) int(time1 - sim_start) / 3600, mod(int(time1 - sim_start) / 60, 60)
  write(*,18) ts, g%t, iters_sum
  write(*,*)
  call fs_check(fs_destroy(g%h), 'fs_destroy')

11 format('Timestep:', i7, '  Time:', es9.2, '  time/us:', f6.2, ' min')
12 format('   dT:  ', es9.2, '   tm:', es9.2,  '   Em:', es9.2)
15 format('   Vd:', f7.2, '     Ida:', es9.2, '       Idc:', es9.2)
16 format('   wp:', f7.3, '      ne:', es9.2)
17 format('   pcg:', i6, '      res:', es9.2)
9  format('Simulation finished in ', i0, ' hr ', i0, ' min')
18 format('Finished ', i0, ' steps, t = ', es12.5, ' (', i0, ' PCG iterations)')

contains

  ! Flag pairs as the reference's read_in (main.F90:172-295), plus the
  ! runtime physics the reference fixes at compile time.  The output path is
  ! given by -out (the reference derives it from p, V, nx, ny); without it
  ! nothing is written.
  subroutine read_in
    integer :: i, narg
    character(256) :: arg, val
    narg = command_argument_count()
    if (mod(narg, 2) /= 0) then
      write(*,*)
      write(*,*) 'Usage:   fsim_main <options>'
      write(*,*) 'Options: -nx <nx>, -ny <ny>, -px <px>, -py <py>'
      write(*,*) '         -l <l>, -w <w>, -ew <ew>, -t <t>, -dt <dt>, -v <V>, -r <ohm>'
      write(*,*) '         -unif <T/F>, -cyl <T/F>, -wall <T/F>, -n0 <cm^-3>, -rf <0..4>'
      write(*,*) '         -p -gam -grid -ratio -circuit -ic -seed -steps -rtol -solver'
      write(*,*) '         -warm -trise -twidth -tfall -out <dir/> -verbose -resident'
      write(*,*)
      stop
    end if
    do i = 1, narg / 2
      call get_command_argument(2 * i - 1, arg)
      call get_command_argument(2 * i, val)
      select case (arg)
        case ('-nx'); read(val,*) nx
        case ('-ny'); read(val,*) ny
        case ('-px'); read(val,*) px
        case ('-py'); read(val,*) py
        case ('-l'); read(val,*) l; l = l / x0
        case ('-w'); read(val,*) w; w = w / x0
        case ('-ew'); read(val,*) ew; ew = ew / x0
        case ('-t'); read(val,*) t_fin
        case ('-dt'); read(val,*) dt
        case ('-v'); read(val,*) vl; vl = vl / ph0
        case ('-r'); read(val,*) res_ohm
        case ('-unif'); read(val,*) unif
        case ('-cyl'); read(val,*) cyl
        case ('-wall'); read(val,*) rwall
        case ('-n0'); read(val,*) n_init; n_init = n_init * x0**3
        case ('-rf'); read(val,*) rf
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
        case ('-warm'); read(val,*) lapl_warm
        case ('-out'); path = val
        case ('-verbose'); read(val,*) verbose
        case ('-resident'); read(val,*) resident
        case ('-trise'); read(val,*) t_rise
        case ('-twidth'); read(val,*) t_width
        case ('-tfall'); read(val,*) t_fall
        case default
          write(*,*) 'Error: unknown option ', trim(arg), '. Stop.'
          stop 2
      end select
    end do
    if (px * py /= 1) then
      write(*,*) 'Error: px * py must equal the number of processes (1). Stop.'
      stop 2
    end if
    if (ny < 1 .or. nx < 3) then
      write(*,*) 'Error: need nx >= 3 and ny >= 1. Stop.'
      stop 3
    end if
  end subroutine

  ! writeOut (main.F90:297-333): interior fields, reference units and order
  subroutine writeOut
    if (resident) then
      call ptcl_sync(g)
      call lapl_sync(g)
    end if
    call savedat(trim(path)//'f1.dat', ph * ph0)
    call savedat(trim(path)//'f2.dat', ne(:,:,1) / x0**3)
    call savedat(trim(path)//'f3.dat', ni(:,:,1) / x0**3)
    call savedat(trim(path)//'f4.dat', nt(:,:,1) / x0**3 * ph0 / 1.5)
    call savedat(trim(path)//'f5.dat', nm(:,:,1) / x0**3)
    call saveval(trim(path)//'time.dat', g%t)
    call saveval(trim(path)//'vd.dat', Vd_pl * ph0)
    call saveval(trim(path)//'ida.dat', Ida * e / t0)
    call saveval(trim(path)//'idc.dat', Idc * e / t0)
    call saveval(trim(path)//'dt.dat', g%dt)
    call saveval(trim(path)//'flx.dat', EmFlx)
  end subroutine
end program main
