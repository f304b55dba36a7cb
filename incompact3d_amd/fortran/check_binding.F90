! Fortran caller of libx3d.so through x3d_binding and x3d_drop_in, written
! as the reference calls its hot path: procedure pointers with the abstract
! interfaces DERIVATIVE_X / _Y / _YY / _Z of module_param.f90:135-168 bound
! to the drop-in subroutines (schemes.f90:31-66), the staggered subroutines
! with the argument lists navier.f90 uses, poisson(rhs), and the module
! coefficient arrays the reference passes (ffx, fsx, fwx, ...).
!
!   check_binding cpu            schemes() coefficients on the host (no GPU)
!                                and the status of calls made without a
!                                context, one "name index value" per line
!   check_binding gpu <file>     the input field, then derivatives, staggered
!                                operators, Poisson and 3-D FFT on the
!                                32 x 33 x 16 box with walls in y, written to
!                                <file> (stream access, Fortran order)
program check_binding
  use iso_c_binding
  use x3d_binding
  use x3d_drop_in
  implicit none

  abstract interface
     subroutine DERIVATIVE_X(t,u,r,s,ff,fs,fw,nx,ny,nz,npaire,lind)
       import :: mytype
       integer :: nx,ny,nz,npaire
       real(mytype), dimension(nx,ny,nz) :: t,u,r
       real(mytype), dimension(ny,nz):: s
       real(mytype), dimension(nx):: ff,fs,fw
       real(mytype) :: lind
     end subroutine DERIVATIVE_X
     subroutine DERIVATIVE_Y(t,u,r,s,ff,fs,fw,pp,nx,ny,nz,npaire,lind)
       import :: mytype
       integer :: nx,ny,nz,npaire
       real(mytype), dimension(nx,ny,nz) :: t,u,r
       real(mytype), dimension(nx,nz):: s
       real(mytype), dimension(ny):: ff,fs,fw,pp
       real(mytype) :: lind
     end subroutine DERIVATIVE_Y
     subroutine DERIVATIVE_YY(t,u,r,s,ff,fs,fw,nx,ny,nz,npaire,lind)
       import :: mytype
       integer :: nx,ny,nz,npaire
       real(mytype), dimension(nx,ny,nz) :: t,u,r
       real(mytype), dimension(nx,nz):: s
       real(mytype), dimension(ny):: ff,fs,fw
       real(mytype) :: lind
     end subroutine DERIVATIVE_YY
     subroutine DERIVATIVE_Z(t,u,r,s,ff,fs,fw,nx,ny,nz,npaire,lind)
       import :: mytype
       integer :: nx,ny,nz,npaire
       real(mytype), dimension(nx,ny,nz) :: t,u,r
       real(mytype), dimension(nx,ny):: s
       real(mytype), dimension(nz):: ff,fs,fw
       real(mytype) :: lind
     end subroutine DERIVATIVE_Z
  end interface

  procedure(DERIVATIVE_X), pointer :: derx, derxx
  procedure(DERIVATIVE_Y), pointer :: dery
  procedure(DERIVATIVE_YY), pointer :: deryy
  procedure(DERIVATIVE_Z), pointer :: derz, derzz

  character(len=16) :: mode
  character(len=256) :: path

  call get_command_argument(1, mode)
  if (trim(mode) == 'cpu') then
     call run_cpu()
  else if (trim(mode) == 'gpu') then
     call get_command_argument(2, path)
     call run_gpu(trim(path))
  else
     write(*,*) 'usage: check_binding cpu | gpu <file>'
     error stop 2
  end if

contains

  function params(nx, ny, nz, ncly) result(p)
    integer, intent(in) :: nx, ny, nz, ncly
    type(x3d_params) :: p
    real(c_double), parameter :: twopi = 6.283185307179586476925286766559_c_double
    p%nx = nx; p%ny = ny; p%nz = nz
    p%xlx = twopi; p%yly = twopi; p%zlz = twopi
    p%nclx1 = 0; p%nclxn = 0; p%ncly1 = ncly; p%nclyn = ncly; p%nclz1 = 0; p%nclzn = 0
    p%ifirstder = 4; p%isecondder = 4; p%ipinter = 3
    p%re = 1600.0_c_double; p%dt = 0.005_c_double
    p%itimescheme = 3; p%istret = 0
    p%nu0nu = 4.0_c_double; p%cnu = 0.44_c_double; p%beta = 0.259_c_double
  end function

  subroutine run_cpu()
    character(len=8), parameter :: names(10) = [character(len=8) :: 'ffx', 'fsx', 'fwx', 'sfx', 'ffy', &
         'ffyp', 'sfyp', 'cfy6', 'cifi6y', 'cwip6y']
    real(c_double) :: out(64), t(4), u(4), r(4), s(4), ff(4), lind
    integer(c_int) :: n, i, k, ierr, nx, ny, nz, npaire
    type(x3d_params) :: p
    p = params(32, 33, 16, 2)
    do k = 1, size(names)
       call x3d_check(x3d_schemes_host(p, x3d_cstr(names(k)), out, 64_c_int, n))
       do i = 1, n
          write(*, '(a, 1x, i0, 1x, es25.17e3)') trim(names(k)), i - 1, out(i)
       end do
    end do
    ! entries called without x3d_init report X3D_ERR_STATE instead of failing
    nx = 4; ny = 1; nz = 1; npaire = 0; lind = 0
    u = 1; ff = 0
    ierr = x3d_host_derx_00(t, u, r, s, ff, ff, ff, nx, ny, nz, npaire, lind)
    write(*, '(a, 1x, i0)') 'status_host_derx_00', ierr
    ierr = x3d_host_poisson(u, nx, ny, nz)
    write(*, '(a, 1x, i0)') 'status_host_poisson', ierr
    write(*, '(2a)') 'message ', x3d_error_message()
  end subroutine

  subroutine run_gpu(path)
    character(len=*), intent(in) :: path
    integer, parameter :: nx = 32, ny = 33, nz = 16, nym = ny - 1
    real(mytype), dimension(nx,ny,nz) :: u, t, r
    real(mytype), dimension(nx,nym,nz) :: pp, tp, rp
    real(mytype), dimension(ny,nz) :: sx
    real(mytype), dimension(nx,nz) :: sy
    real(mytype), dimension(nx,ny) :: sz
    real(mytype), dimension(nx) :: ffx, fsx, fwx, sfx, ssx, swx
    real(mytype), dimension(ny) :: ffy, fsy, fwy, sfy, ssy, swy, ppy
    real(mytype), dimension(nz) :: ffz, fsz, fwz, sfz, ssz, swz
    real(mytype), dimension(nym) :: cfy6, csy6, cwy6, ppyi
    real(mytype), dimension(ny) :: cfip6y, csip6y, cwip6y
    complex(mytype), dimension(nx,nym,nz/2+1) :: spec
    real(mytype) :: lind
    type(x3d_params) :: p
    integer :: i, j, k, unit
    real(c_double) :: x, y, z, h
    p = params(nx, ny, nz, 2)
    call x3d_check(x3d_init(p))
    call coef('ffx', ffx); call coef('fsx', fsx); call coef('fwx', fwx)
    call coef('sfx', sfx); call coef('ssx', ssx); call coef('swx', swx)
    call coef('ffy', ffy); call coef('fsy', fsy); call coef('fwy', fwy)
    call coef('sfy', sfy); call coef('ssy', ssy); call coef('swy', swy)
    call coef('ffz', ffz); call coef('fsz', fsz); call coef('fwz', fwz)
    call coef('sfz', sfz); call coef('ssz', ssz); call coef('swz', swz)
    call coef('cfy6', cfy6); call coef('csy6', csy6); call coef('cwy6', cwy6)
    call coef('cfip6y', cfip6y); call coef('csip6y', csip6y); call coef('cwip6y', cwip6y)
    ppy = 1; ppyi = 1; lind = 0
    ! a smooth O(1) field that is not a mode of any operator
    h = 6.283185307179586476925286766559_c_double
    do k = 1, nz
       do j = 1, ny
          do i = 1, nx
             x = (i - 1) * h / nx; y = (j - 1) * h / (ny - 1); z = (k - 1) * h / nz
             u(i,j,k) = sin(x + 0.3_c_double) * cos(0.7_c_double * y) * cos(2 * z) + 0.02_c_double * y * y &
                  + 0.1_c_double * sin(3 * x + 2 * z)
          end do
       end do
    end do
    ! schemes.f90:31-66 for periodic x/z and walls in y
    derx => derx_00_x3d; derxx => derxx_00_x3d
    dery => dery_22_x3d; deryy => deryy_22_x3d
    derz => derz_00_x3d; derzz => derzz_00_x3d
    open(newunit=unit, file=path, access='stream', form='unformatted', status='replace')
    write(unit) u
    call derx(t, u, r, sx, ffx, fsx, fwx, nx, ny, nz, 0, lind); write(unit) t
    call dery(t, u, r, sy, ffy, fsy, fwy, ppy, nx, ny, nz, 0, lind); write(unit) t
    call derz(t, u, r, sz, ffz, fsz, fwz, nx, ny, nz, 0, lind); write(unit) t
    call derxx(t, u, r, sx, sfx, ssx, swx, nx, ny, nz, 0, lind); write(unit) t
    call deryy(t, u, r, sy, sfy, ssy, swy, nx, ny, nz, 0, lind); write(unit) t
    call derzz(t, u, r, sz, sfz, ssz, swz, nx, ny, nz, 0, lind); write(unit) t
    ! divergence-style calls (navier.f90:321-327, 413-418)
    call deryvp_x3d(tp, u, rp, sy, cfy6, csy6, cwy6, ppyi, nx, ny, nym, nz, 0); write(unit) tp
    pp = tp
    call derypv_x3d(t, pp, r, sy, cfip6y, csip6y, cwip6y, cfy6, csy6, cwy6, ppy, nx, nym, ny, nz, 1)
    write(unit) t
    ! poisson(rhs) on the zero-mean pressure-grid field, then the 3-D FFT
    pp = pp - sum(pp) / size(pp)
    call poisson_x3d(pp); write(unit) pp
    call x3d_check(x3d_host_fft3d_r2c(tp, spec, nx, nym, nz)); write(unit) spec
    close(unit)
    call x3d_check(x3d_finalize())
    write(*, '(a)') 'check_binding gpu: done'
  end subroutine

  subroutine coef(name, a)
    character(len=*), intent(in) :: name
    real(mytype), intent(out) :: a(:)
    real(c_double) :: buf(4096)
    integer(c_int) :: n
    call x3d_check(x3d_get_coefficients(x3d_cstr(name), buf, 4096_c_int, n))
    if (n /= size(a)) then
       write(*, '(3a, i0, a, i0)') 'coefficient ', name, ': length ', n, ' expected ', size(a)
       error stop 3
    end if
    a = real(buf(1:n), mytype)
  end subroutine
end program check_binding
