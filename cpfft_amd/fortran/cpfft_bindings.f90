! cpfft_bindings.f90 -- ISO_C_BINDING interfaces to libcpfft_amd.so
! (include/cpfft.h).  A Fortran caller of the reference's operators
! (G_K_dF, fftPcg, drive_eps_sig, tangent_homo, update) uses these in place
! of the reference routines; device fields are opaque type(c_ptr) values.
module cpfft_bindings
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CPFFT_FN = 0, CPFFT_FN1 = 1, CPFFT_PN = 2, &
       CPFFT_PN1 = 3, CPFFT_DFM = 4, CPFFT_B = 5, CPFFT_K4 = 6, &
       CPFFT_STATE_N = 7, CPFFT_STATE_NP1 = 8, CPFFT_WORK = 9, CPFFT_EPS = 10

  interface
     function cpfft_field(ctx, f) bind(C, name="cpfft_field") result(p)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: f
       type(c_ptr) :: p
     end function cpfft_field

     ! saved locals of FFT_nr3 (FFT_nr3.f:23-34), kept per context
     function cpfft_driver_state(ctx) bind(C, name="cpfft_driver_state") result(p)
       import :: c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr) :: p
     end function cpfft_driver_state

     ! G_K_dF (G_K_dF.f:11-19)
     function cpfft_gkdf(ctx, F, GKF, flgK) bind(C, name="cpfft_gkdf") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, F, GKF
       integer(c_int), value :: flgK
       integer(c_int) :: rc
     end function cpfft_gkdf

     ! fftPcg (FFT_nr3.f:214)
     function cpfft_pcg(ctx, b, x, tol, iters, relres) bind(C, name="cpfft_pcg") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, b, x
       real(c_double), value :: tol
       integer(c_int) :: iters
       real(c_double) :: relres
       integer(c_int) :: rc
     end function cpfft_pcg

     ! drive_eps_sig (drive_eps_sig.f:16)
     function cpfft_drive_eps_sig(ctx, step, iiter) bind(C, name="cpfft_drive_eps_sig") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: step, iiter
       integer(c_int) :: rc
     end function cpfft_drive_eps_sig

     ! tangent_homo (tangent_homo.f:11)
     function cpfft_tangent_homo(ctx, tol, C_homo) bind(C, name="cpfft_tangent_homo") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: tol
       real(c_double) :: C_homo(81)
       integer(c_int) :: rc
     end function cpfft_tangent_homo

     ! warm-start training pairs (no counterpart in the reference)
     function cpfft_mlp_log(ctx, step, stride) bind(C, name="cpfft_mlp_log") result(rc)
       import :: c_ptr, c_int, c_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: step
       integer(c_long), value :: stride
       integer(c_int) :: rc
     end function cpfft_mlp_log

     ! update (update.f:16) + FFT_nr3.f:173-175
     function cpfft_update(ctx) bind(C, name="cpfft_update") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function cpfft_update

     function cpfft_dnrm2(ctx, x, nrm) bind(C, name="cpfft_dnrm2") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, x
       real(c_double) :: nrm
       integer(c_int) :: rc
     end function cpfft_dnrm2

     function cpfft_daxpy(ctx, a, x, y) bind(C, name="cpfft_daxpy") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, x, y
       real(c_double), value :: a
       integer(c_int) :: rc
     end function cpfft_daxpy

     function cpfft_dscal(ctx, a, x) bind(C, name="cpfft_dscal") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, x
       real(c_double), value :: a
       integer(c_int) :: rc
     end function cpfft_dscal

     function cpfft_fill9(ctx, x, v) bind(C, name="cpfft_fill9") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, x
       real(c_double) :: v(9)
       integer(c_int) :: rc
     end function cpfft_fill9

     function cpfft_mean9(ctx, x, out) bind(C, name="cpfft_mean9") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, x
       real(c_double) :: out(9)
       integer(c_int) :: rc
     end function cpfft_mean9

     function cpfft_f2disp(ctx, F, lengths, u) bind(C, name="cpfft_f2disp") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx, F
       real(c_double) :: lengths(3), u(*)
       integer(c_int) :: rc
     end function cpfft_f2disp

     function cpfft_counters(ctx, out) bind(C, name="cpfft_counters") result(rc)
       import :: c_ptr, c_int, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long) :: out(8)
       integer(c_int) :: rc
     end function cpfft_counters

     function cpfft_sync(ctx) bind(C, name="cpfft_sync") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function cpfft_sync
  end interface
end module cpfft_bindings
