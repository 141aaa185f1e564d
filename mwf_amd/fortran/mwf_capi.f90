!
! mwf_capi.f90 -- ISO_C_BINDING interface to libmwf.so (include/mwf.h).
!
! Every function returns 0 or a negative MWF_E* code; mwf_check turns a
! failure into  stop 'mwf: <message>', the reference's error convention
! (program/io.f90:135,177,195).
!
module mwf_capi
   use iso_c_binding
   implicit none

   integer(c_int32_t), parameter :: MWF_FP64 = 0, MWF_MIXED = 1

   ! mwf_params: the compile-time constants of program/parameters.f90 (:7-46)
   type, bind(C) :: mwf_params
      integer(c_int32_t) :: k0, mm, nn, dragall, precision, reserved
      real(c_double)     :: alpha, gamma, re, dt, drag, vdrag, hypo, theta
   end type mwf_params

   interface
      integer(c_int) function mwf_create(p, h) bind(C, name='mwf_create')
         import :: c_int, c_ptr, mwf_params
         type(mwf_params), intent(in) :: p
         type(c_ptr), intent(out) :: h
      end function mwf_create

      integer(c_int) function mwf_destroy(h) bind(C, name='mwf_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function mwf_destroy

      integer(c_int) function mwf_set_state(h, re, im) bind(C, name='mwf_set_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: re(*), im(*)
      end function mwf_set_state

      integer(c_int) function mwf_get_state(h, re, im) bind(C, name='mwf_get_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: re(*), im(*)
      end function mwf_get_state

      integer(c_int) function mwf_put_state(h, re, im) bind(C, name='mwf_put_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: re(*), im(*)
      end function mwf_put_state

      integer(c_int) function mwf_spec2phys(h, re, im, phys) bind(C, name='mwf_spec2phys')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: re(*), im(*)
         real(c_double), intent(out) :: phys(*)
      end function mwf_spec2phys

      integer(c_int) function mwf_phys2spec(h, phys, re, im) bind(C, name='mwf_phys2spec')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: phys(*)
         real(c_double), intent(out) :: re(*), im(*)
      end function mwf_phys2spec

      integer(c_int) function mwf_step(h, nsteps) bind(C, name='mwf_step')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: nsteps
      end function mwf_step

      integer(c_int) function mwf_energy(h, e) bind(C, name='mwf_energy')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: e
      end function mwf_energy

      integer(c_int) function mwf_nonlinear(h, re, im) bind(C, name='mwf_nonlinear')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: re(*), im(*)
      end function mwf_nonlinear

      integer(c_int) function mwf_mpt2phys(h, re, im, phys) bind(C, name='mwf_mpt2phys')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: re(*), im(*)
         real(c_double), intent(out) :: phys(*)
      end function mwf_mpt2phys

      integer(c_int) function mwf_sync(h) bind(C, name='mwf_sync')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function mwf_sync

      type(c_ptr) function mwf_last_error() bind(C, name='mwf_last_error')
         import :: c_ptr
      end function mwf_last_error

      integer(c_int) function mwf_ic_gaussian(p, seed, scl, k0, re, im) bind(C, name='mwf_ic_gaussian')
         import :: c_int, c_double, c_int64_t, mwf_params
         type(mwf_params), intent(in) :: p
         integer(c_int64_t), value :: seed
         real(c_double), value :: scl, k0
         real(c_double), intent(out) :: re(*), im(*)
      end function mwf_ic_gaussian

      integer(c_int) function mwf_get_onl(h, re, im, first_step) bind(C, name='mwf_get_onl')
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: h
         real(c_double), intent(out) :: re(*), im(*)
         integer(c_int32_t), intent(out) :: first_step
      end function mwf_get_onl

      integer(c_int) function mwf_set_onl(h, re, im) bind(C, name='mwf_set_onl')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: re(*), im(*)
      end function mwf_set_onl

      integer(c_int) function mwf_spectrum_max(h, m_max, n_max) bind(C, name='mwf_spectrum_max')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: m_max(*), n_max(*)
      end function mwf_spectrum_max

      integer(c_int) function mwf_vel_history(h, nh, y, ans) bind(C, name='mwf_vel_history')
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: nh
         real(c_double), value :: y
         real(c_double), intent(out) :: ans(*)
      end function mwf_vel_history

      integer(c_int) function mwf_autotune(mode) bind(C, name='mwf_autotune')
         import :: c_int, c_int32_t
         integer(c_int32_t), value :: mode
      end function mwf_autotune

      integer(c_int) function mwf_turb_sample(h, dm, dn, out) bind(C, name='mwf_turb_sample')
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: dm, dn
         real(c_double), intent(out) :: out(*)
      end function mwf_turb_sample
   end interface

contains

   subroutine mwf_check(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      character(256) :: text
      integer :: i
      if (rc == 0) return
      text = ' '
      call c_f_pointer(mwf_last_error(), msg, [256])
      do i = 1, 256
         if (msg(i) == c_null_char) exit
         text(i:i) = msg(i)
      end do
      print *, 'mwf: ', what, ': ', trim(text)
      stop 'mwf: GPU call failed'
   end subroutine mwf_check

end module mwf_capi
