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

end module mhd_io
