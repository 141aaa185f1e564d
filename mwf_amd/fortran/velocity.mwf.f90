!*************************************************************************
! velocity.mwf.f90 -- GPU variant of module velocity (program/velocity.f90).
!
! Same module name and public interface as the reference's velocity module,
! selected like the reference's other variants (Makefile:6-7,68-80): build
! this file instead of program/velocity.f90 and link libmwf.so.  The per-step
! work (vel_nonlinear, the AB2 + implicit update, the transforms) runs on the
! GPU through mwf_capi; the host keeps vel_c for the kept io.f90 / turb.f90.
!
! Host/device coherence of vel_c (the reference's io.f90 reads and writes it
! directly):
!  - vel_precompute marks the host copy as newer, so the first vel_TS uploads
!    what io_load_state put there (program/io.f90:153) and restarts AB2 with
!    forward Euler (nst, program/velocity.f90:24-27,48).
!  - vel_TS downloads vel_c before the step at which io_write2files saves the
!    state (tim_step multiple of i_save_rate1, program/io.f90:89-92).
!  - vel_mpt2phys(vel_c, p) refreshes vel_c from the device first
!    (program/io.f90:98).
!  - vel_sync_host() downloads on demand: call it before cleanup saves the
!    final state (program/main.f90:161), since a CPU-hour stop is not known
!    in advance.
!*************************************************************************
module velocity
   use iso_c_binding
   use mpif
   use parameters
   use variables
   use modes
   use mwf_capi
   implicit none
   save

   type (phys) :: vel_p
   type (mpt), target :: vel_c
   type (mpt)  :: vel_nl, vel_onl
   logical :: nst

   type(c_ptr), private :: mwf_h = c_null_ptr
   logical, private :: host_newer = .true.   ! host vel_c changed since the last upload
   logical, private :: dev_newer = .false.   ! device state is ahead of host vel_c
   logical, private :: restart_ab2 = .true.  ! the next upload starts AB2 afresh (nst)

contains

   subroutine vel_precompute()
      type(mwf_params) :: p
      call var_mpt_init(vel_c)
      call var_mpt_init(vel_nl)
      call var_mpt_init(vel_onl)
      nst = .true.
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
      if (c_associated(mwf_h)) call mwf_check(mwf_destroy(mwf_h), 'mwf_destroy')
      ! timed kernel plans, the part FFTW_PATIENT planning plays in
      ! tra_precompute (program/transform.fftw5.f90:33-59)
      call mwf_check(mwf_autotune(1_c_int32_t), 'mwf_autotune')
      call mwf_check(mwf_create(p, mwf_h), 'mwf_create')
      host_newer = .true.
      dev_newer = .false.
      restart_ab2 = .true.
   end subroutine vel_precompute

   ! upload host vel_c: the first upload after vel_precompute is the loaded
   ! state and starts AB2 with forward Euler (nst, program/velocity.f90:24-27,
   ! 48); later ones (vel_imposesym's projections) replace vel_c and keep the
   ! AB2 history, as the reference's in-place projections do
   subroutine upload_state()
      if (restart_ab2) then
         call mwf_check(mwf_set_state(mwf_h, vel_c%Re, vel_c%Im), 'mwf_set_state')
         restart_ab2 = .false.
      else
         call mwf_check(mwf_put_state(mwf_h, vel_c%Re, vel_c%Im), 'mwf_put_state')
      end if
      host_newer = .false.
   end subroutine upload_state

   ! one step of program/velocity.f90:21-41 on the GPU
   subroutine vel_TS()
      if (host_newer) call upload_state()
      call mwf_check(mwf_step(mwf_h, 1_c_int32_t), 'mwf_step')
      dev_newer = .true.
      nst = .false.
      if (modulo(tim_step + 1, i_save_rate1) == 0) call vel_sync_host()
   end subroutine vel_TS

   subroutine vel_sync_host()
      if (.not. dev_newer) return
      call mwf_check(mwf_get_state(mwf_h, vel_c%Re, vel_c%Im), 'mwf_get_state')
      dev_newer = .false.
   end subroutine vel_sync_host

   ! the device holds the newest state: upload vel_c if the host changed it
   subroutine vel_sync_device()
      if (host_newer) call upload_state()
   end subroutine vel_sync_device

   ! AB2 history for a warm restart (program/velocity.f90:16,24-27): fills
   ! vel_onl and nst from the device, for io_save_state to write beside vel_c
   subroutine vel_sync_onl_host()
      integer(c_int32_t) :: first
      call vel_sync_device()
      call mwf_check(mwf_get_onl(mwf_h, vel_onl%Re, vel_onl%Im, first), 'mwf_get_onl')
      nst = first /= 0
   end subroutine vel_sync_onl_host

   ! after io_load_state has filled vel_c and vel_onl from a warm-restart file:
   ! continue AB2 with that history instead of a forward-Euler first step
   subroutine vel_load_onl()
      call mwf_check(mwf_set_state(mwf_h, vel_c%Re, vel_c%Im), 'mwf_set_state')
      call mwf_check(mwf_set_onl(mwf_h, vel_onl%Re, vel_onl%Im), 'mwf_set_onl')
      host_newer = .false.
      dev_newer = .false.
      restart_ab2 = .false.
      nst = .false.
   end subroutine vel_load_onl

   ! vel_mpt2phys(vel_c) + vel_energy (program/io.f90:98-99) reduced on the
   ! device: only the scalar crosses PCIe
   subroutine vel_energy_dev(e)
      double precision, intent(out) :: e
      call vel_sync_device()
      call mwf_check(mwf_energy(mwf_h, e), 'mwf_energy')
   end subroutine vel_energy_dev

   ! io_save_Uspec's spectrum maxima (program/io.f90:345-395) from the device
   ! state: m_(0:i_MM-1), n_(0:i_NN-1)
   subroutine vel_spec_max(m_, n_)
      double precision, intent(out) :: m_(0:i_MM-1), n_(0:i_NN-1)
      call vel_sync_device()
      call mwf_check(mwf_spectrum_max(mwf_h, m_, n_), 'mwf_spectrum_max')
   end subroutine vel_spec_max

   ! vel_history(vel_mpt2phys(vel_c), y, ans) without moving the physical
   ! field to the host (program/io.f90:472-483)
   subroutine vel_history_dev(y, ans)
      double precision, intent(in) :: y
      double precision, intent(out) :: ans(4, i_H)
      call vel_sync_device()
      call mwf_check(mwf_vel_history(mwf_h, int(i_H, c_int32_t), y, ans), 'mwf_vel_history')
   end subroutine vel_history_dev

   ! turb_sample(vel_mpt2phys(vel_c)) (program/turb.f90:37-47): eposC on the
   ! (i_SN, i_SMp) sample grid into t, the newest slice of turb's tdata
   subroutine vel_turb_sample_dev(t)
      double precision, intent(out) :: t(3, 0:i_SN-1, 0:i_SMp-1)
      call vel_sync_device()
      call mwf_check(mwf_turb_sample(mwf_h, int(i_DM, c_int32_t), int(i_DN, c_int32_t), t), 'mwf_turb_sample')
   end subroutine vel_turb_sample_dev

   ! symmetry projections act on the host copy (program/velocity.f90:68-74);
   ! the next upload replaces vel_c on the device and keeps the AB2 history
   subroutine vel_imposesym()
      if (.not. (s_reflect .or. s_uvreflect .or. s_wreflect)) return
      call vel_sync_host()
      if (s_reflect)   call var_reflect(vel_c)
      if (s_uvreflect) call var_uvreflect(vel_c)
      if (s_wreflect)  call var_wreflect(vel_c)
      host_newer = .true.
   end subroutine vel_imposesym

   ! program/velocity.f90:84-90
   subroutine vel_mpt2phys(u, p)
      type(mpt), intent(in), target :: u
      type(phys), intent(out) :: p
      if (c_associated(c_loc(u), c_loc(vel_c))) call vel_sync_host()
      call mwf_check(mwf_mpt2phys(mwf_h, u%Re, u%Im, p%Re), 'mwf_mpt2phys')
   end subroutine vel_mpt2phys

   ! perturbation energy density at one point: mean over y of |u|^2
   ! (program/velocity.f90:115-126)
   function epos(u) result(e)
      double precision :: e
      double precision, intent(in) :: u(i_KK)
      double precision :: t(i_K0)
      t = nluw(u(1:i_K0), u(1:i_K0))
      e = t(1)
      t = nlvv(u(i_K0+1:2*i_K0-1))
      e = e + t(1)
      t = nluw(u(2*i_K0:i_KK), u(2*i_K0:i_KK))
      e = e + t(1)
   end function epos

   ! the three component energies (program/velocity.f90:128-142)
   function eposC(u) result(e)
      double precision :: e(3)
      double precision, intent(in) :: u(i_KK)
      double precision :: t(i_K0)
      t = nluw(u(1:i_K0), u(1:i_K0))
      e(1) = t(1)
      t = nlvv(u(i_K0+1:2*i_K0-1))
      e(2) = t(1)
      t = nluw(u(2*i_K0:i_KK), u(2*i_K0:i_KK))
      e(3) = t(1)
   end function eposC

   ! mean of epos over the physical grid (program/velocity.f90:144-162)
   subroutine vel_energy(a, e)
      type (phys), intent(in) :: a
      double precision, intent(out) :: e
      integer :: m, n
      e = 0d0
      do m = 0, i_3M - 1
         do n = 0, i_3N - 1
            e = e + epos(a%Re(:, n, m))
         end do
      end do
      e = e / (i_3M * i_3N)
   end subroutine vel_energy

   ! velocity profiles at one height for the history file (program/velocity.f90:164-180)
   subroutine vel_history(V, y, ans)
      type(phys), intent(in) :: V
      double precision, intent(out) :: ans(4, i_H)
      double precision, intent(in) :: y
      integer :: j, ji
      if (mpi_rnk /= 0) return
      do j = 1, i_H
         ji = int((j - 1) * i_3N / i_H) + 1
         ans(1, j) = dble(ji) * d_Lz / i_3N
         ans(2, j) = velU(V%Re(:, ji, 0), y)
         ans(3, j) = velV(V%Re(:, ji, 0), y)
         ans(4, j) = velW(V%Re(:, ji, 0), y)
      end do
   end subroutine vel_history

end module velocity
