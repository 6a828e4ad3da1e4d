! props.F90 (host) -- grid type, constants, g_init and savedat over libfsim_hip.
!
! Keeps the interface of the reference module `props` (props.F90:16-21 type
! grid, props.F90:52 g_init(g, nx, ny, px, py, dof, l, w, ew, gap, path),
! props.F90:389 savedat) without MPI: one process, one GPU, one handle.  The
! grid metrics are computed by the library (fs_get_grid) so that host and
! device agree bit for bit.  The run parameters that the reference fixes at
! compile time or reads in read_in are module variables here, set by the
! driver before g_init.
module props
  use iso_c_binding
  use fsim_capi
  use dev_mirror
  implicit none

  real(8) :: t_m = 1e9

  type :: grid
    integer :: nx, ny, bx, by, dof
    integer, allocatable :: type_x(:,:), type_y(:,:)
    real(8) :: l, w, ew, gap, dt, t
    real(8), allocatable :: dx(:), dlx(:), dy(:), dly(:), r(:), x(:), y(:)
    type(c_ptr) :: h = c_null_ptr      ! the device state of this discharge
  end type

  ! constants (props.F90:24-44), filled from the library's table; n_init is
  ! taken from the table once, so that the driver may override it (-n0)
  real(8) :: pi, eps0, e, kb, me, x0, ph0, t0, Tg, ninf, n_zero, n_init, wp02
  logical, private :: consts_loaded = .false.

  ! run parameters that the reference fixes at compile time or reads in read_in
  real(8) :: p = 10d0, gam = 0.1d0, res_ohm = 1e4, ratio = 1.05d0
  real(8) :: t_rise = 0.01d0, t_width = 0.1d0, t_fall = 0.01d0
  real(8) :: lapl_rtol = 1d-10, dt0 = 5d-6
  integer :: rf = 0, circuit = 0, grid_x = 0, grid_y = 0, lapl_maxit = 20000, lapl_solver = 1
  integer :: lapl_warm = 0
  logical :: unif = .True., cyl = .False., rwall = .False.

  ! the handle of the last g_init, for the module calls without a grid
  ! argument (circ_init(Vset))
  type(c_ptr) :: dev_h = c_null_ptr

contains

  subroutine load_constants()
    real(c_double) :: tab(40)
    call fs_check(fs_constants(p, gam, tab), 'fs_constants')
    pi = tab(1); eps0 = tab(2); e = tab(4); kb = tab(5); me = tab(6); x0 = tab(7)
    ph0 = tab(8); t0 = tab(9); Tg = tab(12); ninf = tab(14); n_zero = tab(15)
    if (.not. consts_loaded) n_init = tab(16)
    wp02 = t0**2 * e**2 / (eps0 * x0**3 * me)   ! props.F90:36
    consts_loaded = .true.
  end subroutine

  ! *** Initialize grid and device state (props.F90:52-260) ***
  ! One rank drives one GPU, so px = py = 1.  With a path, the mesh files
  ! meshx.dat / meshy.dat are written and the output series created empty, as
  ! the reference's g_init does.
  subroutine g_init(g, nx, ny, px, py, dof, l, w, ew, gap, path)
    type(grid), intent(inout), target :: g
    integer, intent(in) :: nx, ny, px, py, dof
    real(8), intent(in) :: l, w, ew, gap
    character(*), intent(in) :: path
    type(fs_grid_desc) :: gd
    type(fs_params) :: pr
    integer :: rc, k, u
    character(8) :: names(13)

    if (px * py /= 1) then
      write(*,*) 'Error: one process drives one GPU (px = py = 1). Stop.'
      stop 2
    end if
    call load_constants()
    if (ny == 1) then  ! props.F90:64-65
      cyl = .False.
      rwall = .False.
    end if
    g%nx = nx; g%ny = ny; g%bx = nx; g%by = ny; g%dof = dof
    g%l = l; g%w = w; g%ew = ew; g%gap = gap

    gd%nx = nx; gd%ny = ny
    gd%grid_x = grid_x; gd%grid_y = grid_y
    if (.not. unif .and. grid_x == 0) gd%grid_x = 1
    if (.not. unif .and. grid_y == 0) gd%grid_y = 1
    gd%ratio_x = ratio; gd%ratio_y = 1d0
    gd%l = l; gd%w = w; gd%ew = ew; gd%gap = gap
    gd%cyl = merge(1, 0, cyl); gd%rwall = merge(1, 0, rwall)

    ! the source amplitude comes with circ_init(Vset), the time step from the
    ! driver's g%dt at the first module call
    pr%p_torr = p; pr%gam = gam; pr%vset = 0d0; pr%res_ohm = res_ohm
    pr%rf = rf; pr%circuit = circuit
    pr%t_rise = t_rise; pr%t_width = t_width; pr%t_fall = t_fall
    pr%dt0 = dt0; pr%lapl_rtol = lapl_rtol; pr%lapl_maxit = lapl_maxit
    pr%lapl_chunk = 32; pr%debug_keep_k = 0; pr%lapl_solver = lapl_solver
    pr%use_graphs = 1; pr%lapl_warm = lapl_warm; pr%stage_tile = 0

    rc = fs_create(gd, pr, g%h)
    call fs_check(rc, 'fs_create')
    dev_h = g%h
    call mirror_init(nx, ny)

    allocate(g%x(nx+2), g%y(ny+2), g%dx(nx+1), g%dlx(nx), g%dy(ny+1), g%dly(ny), &
             g%r(ny+2), g%type_x(nx,ny), g%type_y(nx,ny))
    rc = fs_get_grid(g%h, c_loc(g%x), c_loc(g%y), c_loc(g%dx), c_loc(g%dlx), &
                     c_loc(g%dy), c_loc(g%dly), c_loc(g%r), c_loc(g%type_x), &
                     c_loc(g%type_y))
    call fs_check(rc, 'fs_get_grid')
    if (.not. unif) write(*,31) g%l / float(g%nx+1) / minval(g%dx), maxval(g%dx) / minval(g%dx)
31  format('Non-uniform Grid, dx_u / dx_min:', f6.2, '   dx_max / dx_min:', f6.2)

    if (len(path) > 0) then
      names = [character(8) :: 'time', 'dt', 'vd', 'ida', 'idc', 'flx', 'f1', 'f2', 'f3', &
               'f4', 'f5', 'meshx', 'meshy']
      do k = 1, 13
        open(newunit=u, file=path//trim(names(k))//'.dat', access='stream', &
             form='unformatted', status='replace')
        if (k == 12) write(u) g%x(2:g%bx+1) * x0
        if (k == 13 .and. g%ny > 1) write(u) g%y(2:g%by+1) * x0  ! props.F90:262-264
        close(u)
      end do
    end if
  end subroutine

  ! *** Save Data (props.F90:389-399): append the interior of a (bx+2, by+2)
  ! field to a raw FP64 file ***
  subroutine savedat(path, dat)
    character(*), intent(in) :: path
    real(8), intent(in) :: dat(:,:)
    integer :: u
    open(newunit=u, file=path, access='stream', form='unformatted', status='unknown', &
         position='append')
    write(u) dat(2:size(dat,1)-1, 2:size(dat,2)-1)
    close(u)
  end subroutine

  ! append scalars to a raw FP64 series
  subroutine saveval(path, v)
    character(*), intent(in) :: path
    real(8), intent(in) :: v
    integer :: u
    open(newunit=u, file=path, access='stream', form='unformatted', status='unknown', &
         position='append')
    write(u) v
    close(u)
  end subroutine
end module props
