! mhd_io.f90 -- host-side run-directory I/O and set-up of the reference's
! magnetosphere case, for the Fortran host program (no GPU calls here).
!
!   read_inp_field     inp/binx|biny|binz<rank>: the magnetic field of initia's
!                      mxdat branch (t3dsdsc.f:674-726), list-directed values,
!                      z fastest, then x, then y, over the physical points of
!                      one domain (rank 000: jx = 3..mx+2, jz = 3..mz+2)
!   read_fcinp_field   fcinp/binx|biny|binz<rank>: the inflow face data of the
!                      mxdatfc branch (t3dsdsc.f:728-768), z fastest, then y,
!                      scaled by -50 into bxinfc(my, mz+4)
!   magnetosphere_initia  initia's state (t3dsdsc.f:771-814): density and
!                      pressure shells about r = 1.04, v = 0, the point dipole
!                      14 (-3xz, -3yz, x^2+y^2-2z^2) / r^5 (zero for r < 0.1) or
!                      the inp field times -25/3; the equilibrium is that state
!   write_m2d, write_m3d  recrd2 / recrd3 plot files (t3dsdsc.f:5116-5159,
!                      5365-5408)
!   mesh_x, mesh_yz    gridpntnew2's meshes (t3dsdsc.f:4380-4690): uniform
!                      axes, the graded dayside x mesh, graded y / z axes
!                      with a fine band, mirrored about their origin
module mhd_io
  use iso_c_binding, only: c_double
  implicit none

contains

  subroutine read_inp_field(path, mx, my, mz, a, ios)
    character(len=*), intent(in) :: path
    integer, intent(in) :: mx, my, mz
    real(c_double), intent(out) :: a(mx+4, my, mz+4)
    integer, intent(out) :: ios
    real(c_double), allocatable :: bo(:)
    integer :: u, jx, jy, jz
    allocate(bo(mx*my*mz))
    a = 0.d0
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) return
    read(u, *, iostat=ios) bo
    close(u)
    if (ios /= 0) return
    ! bxin(jx,jy,jz) = bxo(n_x*n_z*(jy-1) + n_z*(jx-3) + jz-2), n_x = mx, n_z = mz
    do jy = 1, my
       do jx = 3, mx + 2
          do jz = 3, mz + 2
             a(jx, jy, jz) = bo(mx*mz*(jy-1) + mz*(jx-3) + jz-2)
          end do
       end do
    end do
  end subroutine read_inp_field

  subroutine read_fcinp_field(path, my, mz, a, ios)
    character(len=*), intent(in) :: path
    integer, intent(in) :: my, mz
    real(c_double), intent(out) :: a(my, mz+4)
    integer, intent(out) :: ios
    real(c_double), allocatable :: bo(:)
    integer :: u, jy, jz
    allocate(bo(my*mz))
    a = 0.d0
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) return
    read(u, *, iostat=ios) bo
    close(u)
    if (ios /= 0) return
    ! bxinfc(jy,jz) = bxofc(n_z*(jy-1) + jz-2) * (-50.), n_z = mz
    do jy = 1, my
       do jz = 3, mz + 2
          a(jy, jz) = bo(mz*(jy-1) + jz-2) * (-50.)
       end do
    end do
  end subroutine read_fcinp_field

  ! initia's state on the physical points (ghost columns stay 0); with
  ! mxdat the field is bin(:,:,:,1:3) * (-25./3.) (single-precision literals
  ! as the reference writes them)
  subroutine magnetosphere_initia(mx, my, mz, xx, yy, zz, mxdat, bin, x, rhoi, bxi, byi, bzi, pri)
    integer, intent(in) :: mx, my, mz
    real(c_double), intent(in) :: xx(mx+4), yy(my), zz(mz+4)
    logical, intent(in) :: mxdat
    real(c_double), intent(in) :: bin(mx+4, my, mz+4, 3)
    real(c_double), intent(out) :: x(mx+4, my, mz+4, 8)
    real(c_double), intent(out), dimension(mx+4, my, mz+4) :: rhoi, bxi, byi, bzi, pri
    real(c_double) :: rr0, rr1, r2, r4, r5
    integer :: jx, jy, jz
    x = 0.d0
    do jz = 3, mz + 2
       do jy = 1, my
          do jx = 3, mx + 2
             rr0 = sqrt(xx(jx)**2 + yy(jy)**2 + zz(jz)**2)
             rr1 = sqrt((xx(jx) - 0.0)**2 + yy(jy)**2 + zz(jz)**2)
             x(jx,jy,jz,1) = 1.d-3*exp((-(rr0-1.04)*(rr0-1.04)/(2.*0.15*0.15))) + 0.1*1.d-3
             if (mxdat) then
                x(jx,jy,jz,5) = bin(jx,jy,jz,1)*(-25./3.)
                x(jx,jy,jz,6) = bin(jx,jy,jz,2)*(-25./3.)
                x(jx,jy,jz,7) = bin(jx,jy,jz,3)*(-25./3.)
             else if (rr1 >= 0.1) then
                r2 = rr0*rr0
                r4 = r2*r2
                r5 = rr0*r4
                x(jx,jy,jz,5) = 14.0*(-3.)*xx(jx)*zz(jz)/r5
                x(jx,jy,jz,6) = 14.0*(-3.)*yy(jy)*zz(jz)/r5
                x(jx,jy,jz,7) = 14.0*(xx(jx)**2+yy(jy)**2-2.*zz(jz)**2)/r5
             end if
             x(jx,jy,jz,8) = 0.2*1.d-3*exp((-(rr0-1.04)*(rr0-1.04)/(2.*0.15*0.15))) + 0.1*0.2*1.d-3
          end do
       end do
    end do
    rhoi = x(:,:,:,1)
    bxi = x(:,:,:,5)
    byi = x(:,:,:,6)
    bzi = x(:,:,:,7)
    pri = x(:,:,:,8)
  end subroutine magnetosphere_initia

  ! recrd2 (t3dsdsc.f:5116-5159): the plane of bl(mx, mz, 11)
  subroutine write_m2d(path, time, mx, mz, xxt, zzt, bl)
    character(len=*), intent(in) :: path
    real(c_double), intent(in) :: time
    integer, intent(in) :: mx, mz
    real(c_double), intent(in) :: xxt(mx), zzt(mz), bl(mx, mz, 11)
    integer :: v, jx, jz, i
    open(newunit=v, file=path, status='replace', form='formatted')
    write(v,*) 2, mx, mz
    write(v,*) time
    write(v,'(5(1x,e14.5E3))') xxt, zzt
    do i = 1, 11
       write(v,*) i
       write(v,'(5(1x,e14.5E3))') ((bl(jx,jz,i), jx=1,mx), jz=1,mz)
    end do
    close(v)
  end subroutine write_m2d

  ! recrd3 (t3dsdsc.f:5365-5408): the volume bl(mx, my, mz, 11)
  subroutine write_m3d(path, time, mx, my, mz, xxt, yy, zzt, bl)
    character(len=*), intent(in) :: path
    real(c_double), intent(in) :: time
    integer, intent(in) :: mx, my, mz
    real(c_double), intent(in) :: xxt(mx), yy(my), zzt(mz), bl(mx, my, mz, 11)
    integer :: v, jx, jy, jz, i
    open(newunit=v, file=path, status='replace', form='formatted')
    write(v,*) 3, mx, my, mz
    write(v,*) time
    write(v,'(11(1x,e10.4))') xxt, yy, zzt
    do i = 1, 11
       write(v,*) i
       write(v,'(11(1x,e10.4))') (((bl(jx,jy,jz,i), jx=1,mx), jy=1,my), jz=1,mz)
    end do
    close(v)
  end subroutine write_m3d

  ! ---------------------------------------------------------------------
  ! gridpntnew2 (t3dsdsc.f:4380-4690): the global mesh of each axis, as the
  ! points p(1..n) and the spacing array d(1..n).  Default-real literals of
  ! the reference stay real(4) constants (0.33 enters as 0.330000013...,
  ! 0.33 * count is a real(4) product) and its integer assignments truncate,
  ! so the points round as the reference's.  The meshes assume n odd (the
  ! gap counts are 2 ((n-1)/2)); an even n leaves the last point as the
  ! reference leaves it (0, spacing 0) unless the half-axis mirror sets it.
  ! ---------------------------------------------------------------------

  ! x: uniform from xmin0 (xmin with halfx) in steps (xmax - xmin) / gaps,
  ! or the dayside mesh of three tanh-graded segments from -30 to 80 (400,
  ! 2 x 300 and 440 gaps: n = 1441 only; ios = 1 otherwise)
  subroutine mesh_x(n, xmin0, xmin, xmax, uniform, half, p, d, ios)
    integer, intent(in) :: n
    real(c_double), intent(in) :: xmin0, xmin, xmax
    logical, intent(in) :: uniform, half
    real(c_double), intent(out) :: p(n), d(n)
    integer, intent(out) :: ios
    integer, parameter :: n1 = 400, n2 = 300, n3 = 440
    real(c_double) :: step, lo, shift, g
    integer :: ngap, j
    p = 0.d0
    d = 0.d0
    ios = 0
    ngap = 2 * ((n - 1) / 2)
    if (uniform) then
       step = (xmax - xmin) / real(ngap, kind(1.0))
       lo = xmin0
       if (half) lo = xmin
       do j = 1, ngap + 1
          d(j) = step
          p(j) = lo + (j - 1) * step
       end do
       return
    end if
    if (n /= n1 + 2 * n2 + n3 + 1) then
       ios = 1
       return
    end if
    ! segment 1 (x in -30..-10 before the final stretch): gaps falling
    ! towards the nose, stored in reverse
    do j = 1, n1
       d(n1 - j + 1) = seg_gap(j, n1, 20.d0, 12.d0)
    end do
    ! segment 2: the same profile on both sides of x = 0, continued from
    ! the end of segment 1
    shift = d(n1) - seg_gap(1, n2, 10.d0, 8.d0)
    do j = 1, n2
       g = shift + seg_gap(j, n2, 10.d0, 8.d0)
       d(n1 + j) = g
       d(n1 + 2 * n2 - j + 1) = g
    end do
    ! segment 3: the tail, continued from the end of segment 2
    shift = d(n1 + 2 * n2) - seg_gap(1, n3, 70.d0, 15.d0)
    do j = 1, n3
       d(n1 + 2 * n2 + j) = shift + seg_gap(j, n3, 70.d0, 15.d0)
    end do
    ! points from -30, then every gap widened alike so that the last point is 80
    p(1) = -30.
    do j = 1, n - 1
       p(j + 1) = p(j) + d(j)
    end do
    shift = (80. - p(n)) / real(n - 1, kind(1.0))
    do j = 1, n - 1
       d(j) = d(j) + shift
       p(j + 1) = p(j) + d(j)
    end do
  end subroutine mesh_x

  ! gap j of a dayside segment of m gaps over length len, graded by tanh
  ! from dmin = 0.02 with the given width (t3dsdsc.f:4403-4447)
  real(c_double) function seg_gap(j, m, len, width)
    integer, intent(in) :: j, m
    real(c_double), intent(in) :: len, width
    real(c_double) :: dmin, dmax
    dmin = 0.02
    dmax = (len - m * dmin) / (0.33 * m)
    seg_gap = 0.5 * (dmax + dmin) + 0.5 * (dmax - dmin) * tanh((width * j - width * m * 2 / 3.) / real(m, kind(1.0)))
  end function seg_gap

  ! y and z: n points over lo..hi (half) or full..hi with the mirror image
  ! of the half axis (full = -ymax for y, zmin0 for z).  Uniform: constant
  ! steps.  Graded: dmin in the band lo..hr (graded there from dmax when
  ! dmax /= dmin), then a tanh rise over the rest, every gap widened alike
  ! to end at hi; without half the axis is mirrored about lo.
  subroutine mesh_yz(n, lo, hi, full, uniform, half, dmin, dmax_in, hr, p, d)
    integer, intent(in) :: n
    real(c_double), intent(in) :: lo, hi, full, dmin, dmax_in, hr
    logical, intent(in) :: uniform, half
    real(c_double), intent(out) :: p(n), d(n)
    real(c_double), parameter :: width = 10.
    real(c_double) :: dmax, step, q, qm, q1, qm1, base, shift
    real(c_double), allocatable :: hp(:), hd(:)
    integer :: npt, ngap, nhr, nr, j
    p = 0.d0
    d = 0.d0
    npt = 2 * ((n - 1) / 2) + 1      ! points of the full axis
    ngap = npt - 1                   ! gaps of the graded (or half) axis
    if (.not. half) ngap = (n - 1) / 2
    p(1) = lo
    if (uniform) then
       step = (hi - lo) / real(ngap, kind(1.0))
       base = full
       if (half) base = lo
       do j = 1, npt
          d(j) = step
          p(j) = base + (j - 1) * step
       end do
       if (.not. half) return
    else
       dmax = dmax_in
       if (dmax == dmin) then
          nhr = int((hr - lo) / dmin + 1)
          nr = ngap + 1 - nhr
          dmax = (2. * (hi - hr) - nr * dmin) / real(nr, kind(1.0))
          q = 0.5 * (dmax + dmin)
          qm = 0.5 * (dmax - dmin)
          do j = 1, nhr - 1
             d(j) = q + qm * tanh((-width * nr / 2.) / nr)
             p(j + 1) = p(j) + d(j)
          end do
       else
          nhr = int(2 * (hr - lo) / (dmax + dmin) + 1)
          nr = ngap + 1 - nhr
          q1 = 0.5 * (dmax + dmin)
          qm1 = 0.5 * (dmax - dmin)
          dmax = (2. * (hi - hr) - nr * dmin) / real(nr, kind(1.0))
          q = 0.5 * (dmax + dmin)
          qm = 0.5 * (dmax - dmin)
          ! joins the band's profile to the rise at the band's edge
          shift = q + qm * tanh((-width * nr / 2.) / nr) - (q1 - qm1 * tanh((width * nhr / 2.) / nhr))
          do j = 1, nhr - 1
             d(j) = shift + q1 - qm1 * tanh((width * (j + 1) - width * nhr / 3.) / nhr)
             p(j + 1) = p(j) + d(j)
          end do
       end if
       do j = nhr, ngap
          d(j) = q + qm * tanh((width * (j - nhr) - width * nr * 2 / 3.) / nr)
          p(j + 1) = p(j) + d(j)
       end do
       shift = (hi - p(ngap + 1)) / real(ngap, kind(1.0))
       do j = 1, ngap
          d(j) = d(j) + shift
          p(j + 1) = p(j) + d(j)
       end do
       d(ngap + 1) = d(ngap)
    end if
    allocate(hp(2 * ngap + 1), hd(2 * ngap + 1))
    if (half) then
       ! an even n: the npt computed points reversed and negated, the last
       ! point the image of point ngap
       if (n /= npt) then
          do j = 1, npt
             hp(j) = -p(npt - j + 1)
             hd(j) = d(npt - j + 1)
          end do
          p(1:npt) = hp(1:npt)
          p(n) = -p(ngap)
          d(1:npt) = hd(1:npt)
          d(n) = d(ngap)
       end if
    else
       ! the mirror image below lo, then the axis; the centre point is the
       ! image of the first (written last, as the reference's loop does)
       do j = 1, ngap + 1
          hp(j + ngap) = p(j)
          hd(j + ngap) = d(j)
       end do
       do j = 1, ngap + 1
          hp(j) = -p(ngap - j + 2)
          hd(j) = d(ngap - j + 2)
       end do
       p(1:2 * ngap + 1) = hp
       d(1:2 * ngap + 1) = hd
    end if
  end subroutine mesh_yz

end module mhd_io
