! ISO_C_BINDING interfaces of the C-ABI in include/rast3dmhd.h.
! The derived type MHD_DESC is interoperable with the C struct mhd_desc
! (same members, same order).
module rast3dmhd_c
  use iso_c_binding
  implicit none

  ! mhd_desc%precision
  integer(c_int), parameter :: MHD_FP64 = 0, MHD_MIXED = 1

  type, bind(c) :: mhd_desc
    integer(c_int) :: npx, npy, npz
    integer(c_int) :: lmag, lrot
    integer(c_int) :: izcon, itcon, ibcon
    integer(c_int) :: precision
    real(c_double) :: hx, h2x, hy, h2y, hz, h2z, c13, c23, c43
    real(c_double) :: gam1, gam2, gam3, zeta1, zeta2
    real(c_double) :: cv, ocv, ore, re, repr, theta, grav, sf, gamma
    real(c_double) :: omx, omz
    real(c_double) :: orm, rm, obeta
    real(c_double) :: tb
    real(c_double) :: tp
    real(c_double) :: dzf
    integer(c_int) :: id, lrem, lshr, npey, npez
    real(c_double) :: dh, dv, tu, dztu, dztb, rlax, tstart, toff
    real(c_double) :: xp, yp, zp, tc, qfh, hh, zmax
  end type mhd_desc

  interface
    integer(c_int) function mhd_create(desc, h) bind(c, name='mhd_create')
      import :: c_int, c_ptr, mhd_desc
      type(mhd_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function mhd_set_metrics(h, dxxdx, d2xxdx2, ddx, dyydy, d2yydy2, ddy, &
                                            dzzdz, d2zzdz2, ddz, rkapa, dkapa) &
                                            bind(c, name='mhd_set_metrics')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: dxxdx(*), d2xxdx2(*), ddx(*), dyydy(*), d2yydy2(*), ddy(*)
      real(c_double), intent(in) :: dzzdz(*), d2zzdz2(*), ddz(*), rkapa(*), dkapa(*)
    end function

    integer(c_int) function mhd_set_coordinates(h, exx, wyy, zee) bind(c, name='mhd_set_coordinates')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: exx(*), wyy(*), zee(*)
    end function

    integer(c_int) function mhd_set_background(h, ro0, tt0) bind(c, name='mhd_set_background')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ro0(*), tt0(*)
    end function

    integer(c_int) function mhd_rebase(h) bind(c, name='mhd_rebase')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function mhd_set_rebase_interval(h, nsteps) bind(c, name='mhd_set_rebase_interval')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function mhd_upload(h, ru, rv, rw, ro, tt, bx, by, bz) bind(c, name='mhd_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ru(*), rv(*), rw(*), ro(*), tt(*), bx(*), by(*), bz(*)
    end function

    integer(c_int) function mhd_download(h, ru, rv, rw, ro, tt, bx, by, bz) bind(c, name='mhd_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: ru(*), rv(*), rw(*), ro(*), tt(*), bx(*), by(*), bz(*)
    end function

    integer(c_int) function mhd_step(h, dt, umach) bind(c, name='mhd_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: dt, umach
    end function

    integer(c_int) function mhd_set_fixed_dt(h, dt) bind(c, name='mhd_set_fixed_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
    end function

    integer(c_int) function mhd_set_clock(h, nit, timc, timi) bind(c, name='mhd_set_clock')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nit
      real(c_double), value :: timc, timi
    end function

    type(c_ptr) function mhd_last_error(h) bind(c, name='mhd_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function

    subroutine mhd_destroy(h) bind(c, name='mhd_destroy')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine
  end interface

contains

  ! Prints the library's last error for handle h (null: the last error of
  ! a call without a handle) with the failing call's name.
  subroutine mhd_report(h, what, rc)
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    character(len=512) :: line
    p = mhd_last_error(h)
    line = ' '
    if (c_associated(p)) then
      call c_f_pointer(p, msg, [512])
      do n = 1, 512
        if (msg(n) == c_null_char) exit
        line(n:n) = msg(n)
      end do
    end if
    write(6, '(A,A,I3,A,A)') what, ': error ', rc, ': ', trim(line)
  end subroutine mhd_report

  ! mhd_report, then stop, as the reference does on invalid settings
  ! (WRITE(6,*) ... ; STOP) in a single-process run.
  subroutine mhd_fail(h, what, rc)
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    integer(c_int), intent(in) :: rc
    call mhd_report(h, what, rc)
    stop
  end subroutine mhd_fail
end module rast3dmhd_c
