!
! ref_kat.f90 -- known-answer driver for the reference's compute modules.
!
! TEST INFRASTRUCTURE ONLY.  Linked against the reference's own mpi.f90,
! parameters.f90, modes.M4.f90 and variables.f90, compiled where they lie
! (oracle/Makefile, target "ref").  Reads inputs from kat_in.bin and writes
! the reference's outputs to kat_out.bin, both raw little-endian doubles.
!
!  kat_in.bin : A%Re, A%Im (mpt), S%Re, S%Im (spec), V, VX, VZ (KK x NV)
!  kat_out.bin: header(12), lhs%Re, rhs%Re, mpt2spec(A) Re/Im,
!               spec_grad(S) ux Re/Im, uz Re/Im, spec2mpt(S) Re/Im,
!               then per vector: udotgradu, udotgradu2, nluw, nlvv, nluv,
!               nluyv, nlvvy
!
program ref_kat
   use mpif
   use parameters
   use variables
   use modes
   implicit none
   integer, parameter :: NV = 64
   type(mpt)  :: A, B, L, R
   type(spec) :: S, T, UX, UZ
   double precision :: V(i_KK, NV), VX(i_KK, NV), VZ(i_KK, NV)
   double precision :: hdr(12)
   integer :: j, ui, uo

   call mpi_precompute()
   call par_precompute()
   call var_precompute()

   open(newunit=ui, file='kat_in.bin', access='stream', form='unformatted', status='old')
   read(ui) A%Re, A%Im, S%Re, S%Im, V, VX, VZ
   close(ui)

   call var_LHSRHS(L, R)
   call var_mpt2spec(A, T)
   call var_spec_grad(S, UX, UZ)
   call var_spec2mpt(S, B)

   hdr = 0d0
   hdr(1) = i_K0
   hdr(2) = i_MM
   hdr(3) = i_NN
   hdr(4) = d_alpha
   hdr(5) = d_gamma
   hdr(6) = d_dt
   hdr(7) = d_Re
   hdr(8) = d_drag
   hdr(9) = d_vdrag
   hdr(10) = d_HYPO
   hdr(11) = d_beta
   if (s_dragall) hdr(12) = 1d0

   open(newunit=uo, file='kat_out.bin', access='stream', form='unformatted', status='replace')
   write(uo) hdr
   write(uo) L%Re, R%Re
   write(uo) T%Re, T%Im
   write(uo) UX%Re, UX%Im, UZ%Re, UZ%Im
   write(uo) B%Re, B%Im
   do j = 1, NV
      write(uo) udotgradu(V(:,j), VX(:,j), VZ(:,j))
      write(uo) udotgradu2(V(:,j), VX(:,j), VZ(:,j))
      write(uo) nluw(V(1:i_K0,j), VX(1:i_K0,j))
      write(uo) nlvv(V(i_K0+1:2*i_K0-1,j))
      write(uo) nluv(V(1:i_K0,j), VX(i_K0+1:2*i_K0-1,j))
      write(uo) nluyv(V(1:i_K0,j), V(i_K0+1:2*i_K0-1,j))
      write(uo) nlvvy(V(i_K0+1:2*i_K0-1,j))
   end do
   close(uo)
end program ref_kat
