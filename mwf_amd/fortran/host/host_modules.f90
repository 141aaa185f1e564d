!
! host_modules.f90 -- minimal host modules with the names the velocity
! variant uses, for running it where the reference tree is absent.  They
! carry the parameters of the 64x64-mode M4 case (Lx = Lz = 8 pi, Re = 200,
! dt = 0.02) and the reference's type layouts (program/variables.f90:8-20).
! In the reference tree the real parameters.f90, variables.f90, mpi.f90 and
! modes.M4.f90 take their place.
!
module mpif
   implicit none
   integer :: mpi_rnk = 0, mpi_sze = 1
end module mpif

module parameters
   implicit none
   save
   double precision            :: d_Re = 200d0
   integer,          parameter :: i_MM = 32, i_NN = 32, i_K0 = 4
   double precision, parameter :: d_PI = 3.1415926535897931d0
   double precision, parameter :: d_Lx = 8d0 * d_PI, d_Lz = 8d0 * d_PI
   double precision, parameter :: d_alpha = 2d0 * d_PI / d_Lx, d_gamma = 2d0 * d_PI / d_Lz
   logical,          parameter :: s_reflect = .false., s_uvreflect = .false., s_wreflect = .false.
   integer,          parameter :: i_save_rate1 = 1000, i_save_rate2 = 1
   double precision, parameter :: d_thdeg = 0d0, d_dt = 0.02d0
   double precision, parameter :: d_HYPO = 0d0, d_drag = 1d-2, d_vdrag = 0d0
   logical,          parameter :: s_dragall = .true.
   integer,          parameter :: i_H = 32
   integer,          parameter :: i_K = 2 * i_K0 - 1, i_KK = 3 * i_K0 - 1, i_M = 2 * (i_MM - 1) + 1
   double precision, parameter :: d_beta = d_PI / 2d0, d_theta = d_thdeg / 180d0 * d_PI
   integer,          parameter :: i_3M = 3 * i_MM, i_3N = 3 * i_NN
   integer,          parameter :: i_Mp = i_3M, i_Np = i_NN, i_M1 = i_M - 1, i_K1 = i_K0 - 1
   integer,          parameter :: i_DM = 6, i_DN = 6, i_MT = 10
   integer,          parameter :: i_SM = i_3M / i_DM, i_SN = i_3N / i_DN, i_SMp = i_SM
   double precision :: tim_t
   integer          :: tim_step
end module parameters

module variables
   use parameters
   implicit none
   type phys
      double precision :: Re(i_KK, 0:i_3N-1, 0:i_Mp-1)
   end type phys
   type mpt
      double precision :: Re(i_K, 0:i_M1, 0:i_Np-1)
      double precision :: Im(i_K, 0:i_M1, 0:i_Np-1)
   end type mpt
   type spec
      double precision :: Re(i_KK, 0:i_M1, 0:i_Np-1)
      double precision :: Im(i_KK, 0:i_M1, 0:i_Np-1)
   end type spec
contains
   subroutine var_mpt_init(a)
      type(mpt), intent(out) :: a
      a%Re = 0d0
      a%Im = 0d0
   end subroutine var_mpt_init
   subroutine var_reflect(c)
      type(mpt), intent(inout) :: c
      stop 'host_modules: symmetry projections are not part of this harness'
   end subroutine var_reflect
   subroutine var_uvreflect(c)
      type(mpt), intent(inout) :: c
      stop 'host_modules: symmetry projections are not part of this harness'
   end subroutine var_uvreflect
   subroutine var_wreflect(c)
      type(mpt), intent(inout) :: c
      stop 'host_modules: symmetry projections are not part of this harness'
   end subroutine var_wreflect
end module variables

! Wall-normal products needed by epos: the y-mean of u*w and of v*v for the
! bases 1, sin(b y), cos(2 b y), ... and cos(b y), sin(2 b y), ...
module modes
   use parameters
   implicit none
contains
   function nluw(u, w) result(ans)
      double precision :: ans(0:i_K1)
      double precision, intent(in) :: u(0:i_K1), w(0:i_K1)
      integer :: k
      ans = 0d0
      ans(0) = u(0) * w(0)
      do k = 1, i_K1
         ans(0) = ans(0) + (u(k) * w(k)) / 2
      end do
   end function nluw
   function nlvv(v) result(ans)
      double precision :: ans(0:i_K1)
      double precision, intent(in) :: v(1:i_K1)
      integer :: k
      ans = 0d0
      do k = 1, i_K1
         ans(0) = ans(0) + v(k)**2 / 2
      end do
   end function nlvv
   double precision function velU(V, y)
      double precision, intent(in) :: V(i_KK), y
      integer :: k
      velU = V(1)
      do k = 1, i_K1
         velU = velU + merge(dcos(k * d_beta * y), dsin(k * d_beta * y), mod(k, 2) == 0) * V(1 + k)
      end do
   end function velU
   double precision function velW(V, y)
      double precision, intent(in) :: V(i_KK), y
      integer :: k
      velW = V(2 * i_K0)
      do k = 1, i_K1
         velW = velW + merge(dcos(k * d_beta * y), dsin(k * d_beta * y), mod(k, 2) == 0) * V(2 * i_K0 + k)
      end do
   end function velW
   double precision function velV(V, y)
      double precision, intent(in) :: V(i_KK), y
      integer :: k
      velV = 0d0
      do k = 1, i_K1
         velV = velV + merge(dcos(k * d_beta * y), dsin(k * d_beta * y), mod(k, 2) == 1) * V(i_K0 + k)
      end do
   end function velV
end module modes
