!*************************************************************************
! transform.mwf.f90 -- GPU variant of module transform
! (program/transform.fftw5.f90).
!
! Same module name and public procedures as the reference's FFTW variant,
! selected the same way (Makefile:6-7,68-70: TRANSFORM picks the file), so
! that the kept main.f90 (use transform, call tra_precompute, main.f90:12,44)
! and the utilities that call tra_precompute / tra_spec2phys (utils/*.f90)
! link without FFTW.  The time step itself does not come through here: the
! GPU velocity variant runs its transforms inside the step kernels.
!
!   tra_precompute   creates the device plans (a libmwf handle of this size)
!   tra_spec2phys    type(spec) -> type(phys), 3/2-dealiased inverse
!   tra_phys2spec    type(phys) -> type(spec), scaled forward, truncated
!
! The arrays keep the reference's layouts (program/variables.f90:8-20) and
! are passed through as they are.
!*************************************************************************
module transform
   use iso_c_binding
   use parameters
   use variables
   use mwf_capi
   implicit none
   save

   type(c_ptr), private :: tra_h = c_null_ptr

contains

   ! program/transform.fftw5.f90:33-59 (FFTW plans); here the device plans
   ! and twiddle tables of one FP64 handle of this size
   subroutine tra_precompute()
      type(mwf_params) :: p
      p%k0 = i_K0
      p%mm = i_MM
      p%nn = i_NN
      p%dragall = merge(1, 0, s_dragall)
      p%precision = MWF_FP64
      p%reserved = 0
      p%alpha = d_alpha
      p%gamma = d_gamma
      p%re = d_Re
      p%dt = d_dt
      p%drag = d_drag
      p%vdrag = d_vdrag
      p%hypo = d_HYPO
      p%theta = d_theta
      if (c_associated(tra_h)) call mwf_check(mwf_destroy(tra_h), 'mwf_destroy')
      call mwf_check(mwf_create(p, tra_h), 'mwf_create')
   end subroutine tra_precompute

   ! program/transform.fftw5.f90:64-87
   subroutine tra_spec2phys(s, p)
      type (spec), intent(in)  :: s
      type (phys), intent(out) :: p
      call mwf_check(mwf_spec2phys(tra_h, s%Re, s%Im, p%Re), 'mwf_spec2phys')
   end subroutine tra_spec2phys

   ! program/transform.fftw5.f90:122-145
   subroutine tra_phys2spec(p, s)
      type (phys), intent(in)  :: p
      type (spec), intent(out) :: s
      call mwf_check(mwf_phys2spec(tra_h, p%Re, s%Re, s%Im), 'mwf_phys2spec')
   end subroutine tra_phys2spec

end module transform
