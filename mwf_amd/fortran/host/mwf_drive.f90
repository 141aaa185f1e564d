!
! mwf_drive.f90 -- driver with the time loop of program/main.f90:20-27 over
! the GPU velocity variant: diagnostics before each step (energy every
! i_save_rate2 steps to vel_energy.dat as '(2e20.12)', program/io.f90:460),
! then vel_TS, then t += dt.  The initial state comes from the seeded
! Gaussian generator of libmwf (mwf_ic_gaussian); the final state is written
! to final_state.bin (Re then Im, raw doubles) after vel_sync_host, and the
! device diagnostics of the final state to diag.bin (raw doubles): energy,
! spectrum maxima m_, n_, the history at y = 0.25, the turb sample grid, then
! the AB2 history vel_onl (Re, Im) and nst.  Last, the transform variant:
! tra_spec2phys of the spec s(k,m,n) = (sin(k + 3m + 7n), cos(2k - m + 5n))
! and tra_phys2spec of that phys, to transform.bin (phys, then spec Re, Im).
!   usage: mwf_drive <nsteps> <seed> <rms> <k0>
!
program mwf_drive
   use iso_c_binding
   use parameters
   use variables
   use velocity
   use transform
   use mwf_capi
   implicit none
   type(phys) :: vp, tp
   type(spec) :: ts, ts2
   integer :: k, m, n
   type(mwf_params) :: p
   double precision :: E, scl, k0
   double precision :: m_(0:i_MM-1), n_(0:i_NN-1), H(4, i_H), T(3, 0:i_SN-1, 0:i_SMp-1)
   integer :: nsteps, io_KE, uo
   integer(c_int64_t) :: seed
   character(64) :: arg

   call get_command_argument(1, arg); read(arg, *) nsteps
   call get_command_argument(2, arg); read(arg, *) seed
   call get_command_argument(3, arg); read(arg, *) scl
   call get_command_argument(4, arg); read(arg, *) k0

   tim_t = 0d0
   tim_step = 0
   call tra_precompute()
   call vel_precompute()
   p%k0 = i_K0; p%mm = i_MM; p%nn = i_NN; p%dragall = 1; p%precision = MWF_FP64; p%reserved = 0
   p%alpha = d_alpha; p%gamma = d_gamma; p%re = d_Re; p%dt = d_dt
   p%drag = d_drag; p%vdrag = d_vdrag; p%hypo = d_HYPO; p%theta = d_theta
   call mwf_check(mwf_ic_gaussian(p, seed, scl, k0, vel_c%Re, vel_c%Im), 'mwf_ic_gaussian')

   open(newunit=io_KE, status='replace', file='vel_energy.dat')
   do while (tim_step < nsteps)
      call vel_imposesym()
      if (modulo(tim_step, i_save_rate2) == 0) then
         call vel_mpt2phys(vel_c, vp)
         call vel_energy(vp, E)
         write(io_KE, '(2e20.12)') tim_t, E
      end if
      call vel_TS()
      tim_t = tim_t + d_dt
      tim_step = tim_step + 1
   end do
   close(io_KE)
   call vel_sync_host()
   open(newunit=uo, file='final_state.bin', access='stream', form='unformatted', status='replace')
   write(uo) vel_c%Re, vel_c%Im
   close(uo)
   call vel_energy_dev(E)
   call vel_spec_max(m_, n_)
   call vel_history_dev(0.25d0, H)
   call vel_turb_sample_dev(T)
   call vel_sync_onl_host()
   open(newunit=uo, file='diag.bin', access='stream', form='unformatted', status='replace')
   write(uo) E, m_, n_, H, T, vel_onl%Re, vel_onl%Im, merge(1d0, 0d0, nst)
   close(uo)
   do n = 0, i_Np - 1
      do m = 0, i_M1
         do k = 1, i_KK
            ts%Re(k, m, n) = dsin(dble(k + 3 * m + 7 * n))
            ts%Im(k, m, n) = dcos(dble(2 * k - m + 5 * n))
         end do
      end do
   end do
   call tra_spec2phys(ts, tp)
   call tra_phys2spec(tp, ts2)
   open(newunit=uo, file='transform.bin', access='stream', form='unformatted', status='replace')
   write(uo) tp%Re, ts2%Re, ts2%Im
   close(uo)
end program mwf_drive
