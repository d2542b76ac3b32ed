! mhd_driver.f90 -- Fortran host program for the MI355X step.
!
! Keeps the time loop, guards and outputs of the reference's program main
! (t3dsdsc.f:26-207) and calls the GPU step through the C ABI:
!   setdt  -> mhd_compute_dt      stepon -> mhd_step
!   diagn  -> mhd_diagnostics     recrd  -> mhd_get_state + restart record
! Inputs in the run directory (written by dasha_amd.rundir):
!   mhd.nml      namelist /run/ (grid sizes, physics flags and constants,
!                nstop, nst, ncase, dtime, dnstep)
!   grid.bin     xx, yy, zz, dx, dy, dz; with gridgen = .true. the program
!                builds the mesh itself as gridpntnew2 does (t3dsdsc.f:4380-4690,
!                switches uniformx/y/z, halfx/y/z and the extents of the namelist;
!                non-periodic y) and writes gridx.dat and grid.dat like the reference
!   equil.bin    rhoi, bxi, byi, bzi, pri
!   mhd<nst><000> restart record with the initial state (readin format)
! Magnetosphere mode (magneto = .true.; the reference's own case): with
! initia = .true. the program builds the initial state itself as initia does
! (t3dsdsc.f:771-814): the point dipole, or with mxdat the field of
! inp/binx000, biny000, binz000 (mhd_io), and writes it as mhdinit in the
! restart format; equil.bin and the restart file are not read then.  With
! mxdatfc the inflow face data come from fcinp/binx000 .. binz000.
! Outputs: fort.20 (dt, time, nstep per step, t3dsdsc.f:180-181),
! data/stepnm (nstep, time, dt, t3dsdsc.f:157-169), current.dat (diagn,
! t3dsdsc.f:1228-1230), restart files mhd<nst><000> (recrd), the plot files
! data/m2d<nst>.dat (recrd2: the y-midplane) and data/m3d<nst> (recrd3: the
! volume) of rho, v = m/rho, B, p and J, and at the end of the run the final
! state in the restart format as mhdfinal.  Exit status: 0 done, 2 dt
! collapse (t3dsdsc.f:142-154), 3 current blow-up (t3dsdsc.f:1232), 1 any
! other error.
! With nadj > 0 in the namelist the program instead differentiates nadj
! steps from the initial state: it records the dt sequence of a plain run
! (setdt per step), restores the initial state and calls mhd_adjoint_run
! (objective adjobj: 0 magnetic energy, 1 reconnected flux), writing
! adjoint.bin = (value, d value/d eta0, d value/d x(mx+4,my,mz+4,8)).
program mhd_driver
  use iso_c_binding
  use mhd_gpu
  use mhd_io
  implicit none
  integer :: mx, my, mz, nstop, nst, ncase, dnstep, nend, nint, nadj, adjobj
  logical :: periody, hall, whistler
  logical :: magneto, initia, mxdat, mxdatfc, inflow, radial_hall, partial_e
  real(c_double) :: di, eta0, fmu0, frho0, fpr0, gamma, cfl, dtime, rcr
  ! gridpntnew2's switches and extents (mhd03.for)
  logical :: gridgen, uniformx, uniformy, uniformz, halfx, halfy, halfz
  real(c_double) :: xmin0, xmin, xmax, ymin, ymax, zmin0, zmin, zmax, dymin, dymax, yhr, dzmin, dzmax, zhr
  namelist /run/ mx, my, mz, periody, hall, whistler, di, eta0, fmu0, frho0, fpr0, gamma, cfl, &
       nstop, nst, ncase, dtime, dnstep, nend, nint, nadj, adjobj, &
       magneto, initia, mxdat, mxdatfc, inflow, radial_hall, partial_e, rcr, &
       gridgen, uniformx, uniformy, uniformz, halfx, halfy, halfz, &
       xmin0, xmin, xmax, ymin, ymax, zmin0, zmin, zmax, dymin, dymax, yhr, dzmin, dzmax, zhr
  type(mhd_config) :: cfg
  type(mhd_diag) :: dg
  type(c_ptr) :: h
  real(c_double), allocatable :: xx(:), yy(:), zz(:), dx(:), dy(:), dz(:)
  real(c_double), allocatable :: rhoi(:,:,:), bxi(:,:,:), byi(:,:,:), bzi(:,:,:), pri(:,:,:)
  real(c_double), allocatable :: x(:,:,:,:)
  real(c_double) :: time, dt, dt0, timein
  integer :: nstep, nstin, rc, u, ios, ncase_in, nst_in
  ! step at which output file nst was written (common /run/ nstp, mhd01.for:28)
  integer :: nstp(0:10000)
  logical :: exist

  periody = .true.; hall = .true.; whistler = .true.
  di = 0; eta0 = 0; fmu0 = 0; frho0 = 0; fpr0 = 0; gamma = 5.d0/3.d0; cfl = 0.4d0
  nstop = 10; nst = 1; ncase = 1; dtime = 1.d0; dnstep = 10; nend = 5011; nint = 1
  nadj = 0; adjobj = 0
  magneto = .false.; initia = .false.; mxdat = .false.; mxdatfc = .false.
  inflow = .false.; radial_hall = .false.; partial_e = .false.; rcr = 1.d0
  nstp = 0
  ! mhd03.for's mesh (default real literals)
  gridgen = .false.
  uniformx = .true.; uniformy = .true.; uniformz = .true.
  halfx = .false.; halfy = .false.; halfz = .false.
  xmin0 = -2.1; xmin = -2.1; xmax = 5.65; ymin = 0.; ymax = 2.5; zmin0 = -2.5; zmin = 0.; zmax = 2.5
  dymin = 0.01; dymax = 1.0; yhr = 0.0; dzmin = 0.01; dzmax = 1.0; zhr = 0.0
  open(newunit=u, file='mhd.nml', status='old', action='read', iostat=ios)
  if (ios /= 0) call fail('cannot open mhd.nml', 1)
  read(u, nml=run)
  close(u)

  allocate(xx(mx+4), yy(my), zz(mz+4), dx(mx+4), dy(my), dz(mz+4))
  allocate(rhoi(mx+4,my,mz+4), bxi(mx+4,my,mz+4), byi(mx+4,my,mz+4), bzi(mx+4,my,mz+4), &
       pri(mx+4,my,mz+4), x(mx+4,my,mz+4,8))
  if (gridgen) then
     call build_mesh()
  else
     open(newunit=u, file='grid.bin', form='unformatted', status='old', action='read')
     read(u) xx
     read(u) yy
     read(u) zz
     read(u) dx
     read(u) dy
     read(u) dz
     close(u)
  end if
  if (initia) then
     call build_initia()
  else
     open(newunit=u, file='equil.bin', form='unformatted', status='old', action='read')
     read(u) rhoi
     read(u) bxi
     read(u) byi
     read(u) bzi
     read(u) pri
     close(u)
     ! readin (t3dsdsc.f:4822-4828)
     open(newunit=u, file='mhd'//cn(nst)//cn(0), form='unformatted', status='old', action='read')
     read(u) ncase_in, nstep, time, nst_in
     read(u) x
     close(u)
  end if

  cfg%mx = mx; cfg%my = my; cfg%mz = mz
  cfg%periodic_y = merge(1, 0, periody); cfg%hall = merge(1, 0, hall)
  cfg%whistler_dt = merge(1, 0, whistler)
  cfg%di = di; cfg%eta0 = eta0; cfg%fmu0 = fmu0; cfg%frho0 = frho0; cfg%fpr0 = fpr0
  cfg%gamma = gamma; cfg%cfl = cfl
  if (magneto) then
     cfg%magnetosphere = 1; cfg%rcr = rcr
     cfg%inflow = merge(1, 0, inflow); cfg%radial_hall = merge(1, 0, radial_hall)
     cfg%partial_e = merge(1, 0, partial_e)
  end if
  rc = mhd_create(cfg, xx, yy, zz, dx, dy, dz, h)
  if (rc /= MHD_OK) call fail('mhd_create failed', 1)
  call check(mhd_set_equilibrium(h, rhoi, bxi, byi, bzi, pri))
  if (magneto .and. mxdatfc) call set_face()
  call check(mhd_set_state(h, x))
  ! initia's frozen residual (t3dsdsc.f:816-834)
  call check(mhd_init_residual(h))
  call check(mhd_compute_dt(h, dt))
  dt0 = dt
  timein = time
  nstin = nst
  if (nadj > 0) then
     call adjoint_run()
     call mhd_destroy(h)
     stop
  end if
  inquire(file='data/.', exist=exist)
  if (.not. exist) call execute_command_line('mkdir -p data')

  do
     ! setdt (t3dsdsc.f:139-140)
     call check(mhd_compute_dt(h, dt))
     write(*,*) 'main', nstep, time, dt
     if (dt < 0.1d0*dt0) then
        ! t3dsdsc.f:142-154: midplane snapshot and stop
        nst = nst + nint
        call recrd2()
        call mhd_destroy(h)
        stop 2
     end if
     open(newunit=u, file='data/stepnm', position='append', action='write')
     write(u,*) nstep, time, dt
     close(u)
     call check(mhd_step(h, time, dt))
     ! diagn every 10 steps inside stepon (t3dsdsc.f:1174-1182)
     if (mod(nstep, 10) == 0) then
        rc = mhd_diagnostics(h, dg)
        open(newunit=u, file='current.dat', position='append', action='write')
        write(u,'(7(1x,e13.4))') dg%cxmax, dg%cxmin, dg%cymax, dg%cymin, dg%czmax, dg%czmin, time
        close(u)
        if (rc == MHD_ERR_BLOWUP) then
           call mhd_destroy(h)
           stop 3
        end if
        if (rc /= MHD_OK) call fail(trim(mhd_last_error(h)), 1)
     end if
     nstep = nstep + 1
     time = time + dt
     write(20,'(2(1x,f10.5),1x,i5)') dt, time, nstep
     ! output cadence (t3dsdsc.f:185-199)
     if (abs(time - timein - (nst - nstin + 1)*dtime) <= dt .and. (nstep - nstp(nst)) >= dnstep) then
        nstp(nst) = nstep
        nst = nst + nint
        if (mod(nst, 3) == 1) call recrd()
        call recrd2()
        call recrd3()
     end if
     if (nstep > nstop) exit
     if (nst >= nend) exit
  end do
  ! final state, in the restart format, under its own name
  call check(mhd_get_state(h, x))
  open(newunit=u, file='mhdfinal', form='unformatted', status='replace', action='write')
  write(u) ncase, nstep, time, nst
  write(u) x
  close(u)
  call mhd_destroy(h)

contains

  subroutine recrd()
    ! t3dsdsc.f:4840-4861
    integer :: v
    call check(mhd_get_state(h, x))
    open(newunit=v, file='mhd'//cn(nst)//cn(0), form='unformatted', status='replace', action='write')
    write(v) ncase, nstep, time, nst
    write(v) x
    close(v)
  end subroutine recrd

  ! the eleven plot fields of physical point (jx, jy, jz) (1-based physical
  ! indices): rho, vx, vy, vz, bx, by, bz, p, jx, jy, jz with v = m / rho
  ! (t3dsdsc.f:5072-5076, 5331-5336)
  subroutine plot_fields(cur, jx, jy, jz, f)
    real(c_double), intent(in) :: cur(:,:,:,:)
    integer, intent(in) :: jx, jy, jz
    real(c_double), intent(out) :: f(11)
    integer :: i
    do i = 1, 8
       f(i) = x(jx+2, jy, jz+2, i)
    end do
    do i = 2, 4
       f(i) = f(i) / f(1)
    end do
    do i = 1, 3
       f(8+i) = cur(jx+2, jy, jz+2, i)
    end do
  end subroutine plot_fields

  subroutine recrd2()
    ! t3dsdsc.f:4965-5164: the plane jy = my/2+1
    real(c_double), allocatable :: cur(:,:,:,:), bl(:,:,:)
    real(c_double) :: f(11)
    integer :: v, jx, jz, i, nyh
    allocate(cur(mx+4,my,mz+4,3), bl(mx,mz,11))
    call check(mhd_get_state(h, x))
    call check(mhd_get_current(h, cur))
    nyh = my/2 + 1
    do jz = 1, mz
       do jx = 1, mx
          call plot_fields(cur, jx, nyh, jz, f)
          bl(jx,jz,:) = f
       end do
    end do
    call write_m2d('data/m2d'//cn(nst)//'.dat', time, mx, mz, xx(3:mx+2), zz(3:mz+2), bl)
  end subroutine recrd2

  subroutine recrd3()
    ! t3dsdsc.f:5167-5412: the volume, every y row
    real(c_double), allocatable :: cur(:,:,:,:), bl(:,:,:,:)
    real(c_double) :: f(11)
    integer :: v, jx, jy, jz, i
    allocate(cur(mx+4,my,mz+4,3), bl(mx,my,mz,11))
    call check(mhd_get_state(h, x))
    call check(mhd_get_current(h, cur))
    do jz = 1, mz
       do jy = 1, my
          do jx = 1, mx
             call plot_fields(cur, jx, jy, jz, f)
             bl(jx,jy,jz,:) = f
          end do
       end do
    end do
    call write_m3d('data/m3d'//cn(nst), time, mx, my, mz, xx(3:mx+2), yy, zz(3:mz+2), bl)
  end subroutine recrd3

  subroutine adjoint_run()
    ! gradients of the objective after nadj steps (frozen dt sequence)
    real(c_double), allocatable :: dts(:), x0(:,:,:,:), gx(:,:,:,:)
    real(c_double) :: val, geta, ms(2), t
    integer :: n, v
    allocate(dts(nadj), x0(mx+4,my,mz+4,8), gx(mx+4,my,mz+4,8))
    call check(mhd_get_state(h, x0))
    t = time
    do n = 1, nadj
       call check(mhd_compute_dt(h, dts(n)))
       call check(mhd_step(h, t, dts(n)))
       t = t + dts(n)
    end do
    call check(mhd_set_state(h, x0))
    call check(mhd_init_residual(h))
    call check(mhd_adjoint_run(h, nadj, dts, adjobj, 0, MHD_ADJ_RESIDUAL, val, gx, geta, ms))
    write(*,*) 'adjoint', nadj, val, geta, ms
    open(newunit=v, file='adjoint.bin', form='unformatted', status='replace', action='write')
    write(v) val, geta
    write(v) gx
    close(v)
  end subroutine adjoint_run

  subroutine build_initia()
    ! initia (t3dsdsc.f:612-814) of the magnetosphere; the state is also
    ! written as mhdinit (restart format) for comparisons
    real(c_double), allocatable :: bin(:,:,:,:)
    integer :: v, ios, c
    character(len=1), parameter :: comp(3) = ['x', 'y', 'z']
    if (.not. magneto) call fail('initia = .true. builds the magnetosphere case only', 1)
    allocate(bin(mx+4,my,mz+4,3))
    bin = 0.d0
    if (mxdat) then
       do c = 1, 3
          call read_inp_field('inp/bin'//comp(c)//cn(0), mx, my, mz, bin(:,:,:,c), ios)
          if (ios /= 0) call fail('cannot read inp/bin'//comp(c)//cn(0), 1)
       end do
    end if
    call magnetosphere_initia(mx, my, mz, xx, yy, zz, mxdat, bin, x, rhoi, bxi, byi, bzi, pri)
    ncase_in = ncase; nstep = 0; time = 0.d0; nst_in = nst
    open(newunit=v, file='mhdinit', form='unformatted', status='replace', action='write')
    write(v) ncase, nstep, time, nst
    write(v) x
    close(v)
  end subroutine build_initia

  subroutine build_mesh()
    ! gridpntnew2 (t3dsdsc.f:4380-4690) and input's placement of the global
    ! mesh in the one domain's padded arrays (t3dsdsc.f:296-305: physical
    ! points at jx = 3..mx+2, ghosts 0; y as it is), with the reference's
    ! gridx.dat (t3dsdsc.f:4453-4458) and grid.dat (t3dsdsc.f:287-293)
    real(c_double) :: xxt(mx), dxt(mx), zzt(mz), dzt(mz)
    integer :: v, ios, j
    if (periody) call fail('gridgen builds the reference''s non-periodic mesh; periody needs grid.bin', 1)
    call mesh_x(mx, xmin0, xmin, xmax, uniformx, halfx, xxt, dxt, ios)
    if (ios /= 0) call fail('the graded x mesh of gridpntnew2 has 1441 points', 1)
    call mesh_yz(my, ymin, ymax, -ymax, uniformy, halfy, dymin, dymax, yhr, yy, dy)
    call mesh_yz(mz, zmin, zmax, zmin0, uniformz, halfz, dzmin, dzmax, zhr, zzt, dzt)
    xx = 0.d0; dx = 0.d0; zz = 0.d0; dz = 0.d0
    xx(3:mx+2) = xxt; dx(3:mx+2) = dxt
    zz(3:mz+2) = zzt; dz(3:mz+2) = dzt
    open(newunit=v, file='gridx.dat', status='replace', form='formatted')
    write(v, '(2(i5,2(1x,f13.5)))') (j, xxt(j), dxt(j), j = 1, mx)
    close(v)
    open(newunit=v, file='grid.dat', status='replace', form='formatted')
    write(v, '(6(1x,f13.5))') (xxt(j), dxt(j), j = 1, mx), (zzt(j), dzt(j), j = 1, mz), (yy(j), dy(j), j = 1, my)
    close(v)
  end subroutine build_mesh

  subroutine set_face()
    ! fcinp/bin* (t3dsdsc.f:728-768) -> bxinfc, byinfc, bzinfc
    real(c_double), allocatable :: f(:,:,:)
    integer :: ios, c
    character(len=1), parameter :: comp(3) = ['x', 'y', 'z']
    allocate(f(my, mz+4, 3))
    do c = 1, 3
       call read_fcinp_field('fcinp/bin'//comp(c)//cn(0), my, mz, f(:,:,c), ios)
       if (ios /= 0) call fail('cannot read fcinp/bin'//comp(c)//cn(0), 1)
    end do
    call check(mhd_set_inflow_face(h, f(:,:,1), f(:,:,2), f(:,:,3)))
  end subroutine set_face

  subroutine check(r)
    integer(c_int), intent(in) :: r
    if (r /= MHD_OK) call fail(trim(mhd_last_error(h)), 1)
  end subroutine check

  subroutine fail(msg, code)
    character(len=*), intent(in) :: msg
    integer, intent(in) :: code
    write(0,*) 'mhd_driver: ', msg, ' (status', code, ')'
    stop 1
  end subroutine fail

  character(len=3) function cn(n)
    ! t3dsdsc.f:5416-5436
    integer, intent(in) :: n
    write(cn, '(i3.3)') n
  end function cn

end program mhd_driver
