! fft_nr3.f90 -- the Fortran driver of the load-step loop, running on the
! MI355X operators of libcpfft_amd.so.
!
! It keeps the control flow of the reference's `compute` command
! (FFT_finite_3d.f:138-147: drive_eps_sig(1,0) then FFT_nr3) and of
! FFT_nr3 (FFT_nr3.f:14-200): mixed boundary conditions through
! NBC_update, the initial-residual solve with the old tangent, the Newton
! loop on the compatible correction dFm with at least one iteration, the
! post-convergence stress update, the P_bar check, tangent_homo and the
! state roll-over.  Fields never leave the device; the host keeps only the
! 9-vectors and the 9x9 system of NBC_update.  The procedures are recursive
! (locals on the stack) and the carried state lives in the context, so one
! thread per context may run cpfft_compute at the same time
! (batch_runner.cpp).
module cpfft_driver
  use iso_c_binding
  use cpfft_bindings
  implicit none
  private
  public :: cpfft_compute

contains

  ! NBC_update (FFT_nr3.f:375-433): rows of prescribed P take the
  ! homogenised tangent, rows of prescribed F the identity; solved by
  ! Gaussian elimination with partial pivoting (dgesv).
  recursive subroutine NBC_update(C_homo, DbarF, P_bar, PBC, isNBC, info)
    real(c_double), intent(in)    :: C_homo(9,9), P_bar(9), PBC(9)
    logical, intent(in)           :: isNBC(9)
    real(c_double), intent(inout) :: DbarF(9)
    integer, intent(out)          :: info
    real(c_double) :: AAA(9,9), bbb(9), t, l
    integer :: i, j, k, p

    do i = 1, 9
      if (isNBC(i)) then
        AAA(i,1:9) = C_homo(1:9,i)
        bbb(i) = PBC(i) - P_bar(i)
      else
        AAA(i,1:9) = 0.0d0
        AAA(i,i) = 1.0d0
        bbb(i) = DbarF(i)
      end if
    end do
    info = 0
    do j = 1, 9
      p = j
      do i = j + 1, 9
        if (abs(AAA(i,j)) > abs(AAA(p,j))) p = i
      end do
      if (AAA(p,j) == 0.0d0) then
        info = j
        return
      end if
      if (p /= j) then
        do k = 1, 9
          t = AAA(j,k); AAA(j,k) = AAA(p,k); AAA(p,k) = t
        end do
        t = bbb(j); bbb(j) = bbb(p); bbb(p) = t
      end if
      do i = j + 1, 9
        l = AAA(i,j) / AAA(j,j)
        do k = j + 1, 9
          AAA(i,k) = AAA(i,k) - l * AAA(j,k)
        end do
        bbb(i) = bbb(i) - l * bbb(j)
      end do
    end do
    do i = 9, 1, -1
      t = bbb(i)
      do k = i + 1, 9
        t = t - AAA(i,k) * bbb(k)
      end do
      bbb(i) = t / AAA(i,i)
    end do
    DbarF(1:9) = bbb(1:9)
  end subroutine NBC_update

  recursive real(c_double) function wall()
    integer(8) :: cnt, rate
    call system_clock(cnt, rate)
    wall = dble(cnt) / dble(rate)
  end function wall

  ! compute + FFT_nr3 for load steps first_step .. first_step+nstep-1.
  ! first_step = 1 starts a run (elastic call and initial tangent); a later
  ! first_step continues the previous run on the same context.
  ! BC_all(9,nstep) holds those steps' boundary values as inlod fills them
  ! (inlod.f:57-63);
  ! returns 0 or the negative code of the reference's die_abort paths:
  ! -2 CG not converged (FFT_nr3.f:335-339), -3 NBC_update singular
  ! (:418-421), -4 Newton maxIter (:116-119), -5 mixed-BC maxIter
  ! (:153-156), other negatives from the operators.
  ! counters(:,step) = Newton iterations, CG iterations, mixed-BC loops,
  ! drive_eps_sig calls; times = pcg, sig-eps, total wall seconds (thyme
  ! categories 1 and 2, thyme.f:15-47).
  ! Steps up to mlp_log_last log warm-start training pairs (every
  ! mlp_log_stride-th voxel) of their converged update before the roll-over;
  ! this has no counterpart in the reference and 0 turns it off.
  recursive integer(c_int) function cpfft_compute(ctx, first_step, nstep, BC_all, isNBC_c, tolPCG, tolNR, &
       maxIter, verbose, Pbar_out, counters, times, mlp_log_last, mlp_log_stride) &
       bind(C, name="cpfft_compute")
    type(c_ptr), value     :: ctx
    integer(c_int), value  :: first_step, nstep, maxIter, verbose, mlp_log_last, mlp_log_stride
    real(c_double), value  :: tolPCG, tolNR
    real(c_double)         :: BC_all(9, nstep), Pbar_out(9, nstep), times(3)
    integer(c_int)         :: isNBC_c(9), counters(4, nstep)

    type(c_ptr) :: Fn1, Pn1, dFm, b
    real(c_double) :: resPbar(3), DbarF(9), barF(9)
    real(c_double) :: FBC(9), PBC(9), Fnorm, resfft, relres, dn, t0, t1
    integer(c_int) :: rc, its
    integer(c_long_long) :: cnt0(8), cnt1(8)
    integer :: step, istep, ii, iiter_EBC, iiter_NBC, info, nnewton, ncalls, nloops
    logical :: isNBC(9), existNBC
    ! state carried from one load step to the next (FFT_nr3.f:23-34 keeps
    ! these as saved locals of the routine); here it lives in the context,
    ! so the routine holds no state of its own and several contexts can be
    ! driven at once from different threads
    real(c_double), pointer :: st(:), barF_t(:), P_bar(:), C_homo(:)

    call c_f_pointer(cpfft_driver_state(ctx), st, [99])
    barF_t => st(1:9)
    P_bar => st(10:18)
    C_homo => st(19:99)
    cpfft_compute = 0
    times = 0.0d0
    counters = 0
    Pbar_out = 0.0d0
    isNBC = isNBC_c /= 0
    existNBC = any(isNBC)
    ! fftPcg aborts on a tolerance outside (eps, 1) (FFT_nr3.f:248-251);
    ! checked here once, so that a continued run whose steps never reach
    ! a CG solve with a nonzero right-hand side still reports it
    if (tolPCG <= epsilon(1.0d0) .or. tolPCG >= 1.0d0) then
      cpfft_compute = -1
      return
    end if
    Fn1 = cpfft_field(ctx, CPFFT_FN1)
    Pn1 = cpfft_field(ctx, CPFFT_PN1)
    dFm = cpfft_field(ctx, CPFFT_DFM)
    b = cpfft_field(ctx, CPFFT_B)
    times(3) = wall()

    if (first_step <= 1) then
      barF_t = [1.0d0, 0.0d0, 0.0d0, 0.0d0, 1.0d0, 0.0d0, 0.0d0, 0.0d0, 1.0d0]
      P_bar = 0.0d0
      ! compute (FFT_finite_3d.f:145): elastic tangent at F = I
      t0 = wall()
      rc = cpfft_drive_eps_sig(ctx, 1, 0)
      if (rc < 0) then
        cpfft_compute = rc
        return
      end if
      times(2) = times(2) + (wall() - t0)

      ! initial homogenised tangent (FFT_nr3.f:47)
      t0 = wall()
      rc = cpfft_tangent_homo(ctx, tolPCG, C_homo)
      times(1) = times(1) + (wall() - t0)
      if (rc < 0) then
        cpfft_compute = rc
        return
      end if
    end if

    do istep = 1, nstep
      step = max(first_step, 1) + istep - 1
      if (verbose /= 0) write(*, 1000) step
      if (verbose /= 0) flush(6)
      rc = cpfft_counters(ctx, cnt0)
      nnewton = 0
      ncalls = 0
      nloops = 0
      PBC = 0.0d0
      FBC = 0.0d0
      DbarF = 0.0d0
      do ii = 1, 9
        if (isNBC(ii)) then
          PBC(ii) = BC_all(ii, istep)
        else
          FBC(ii) = BC_all(ii, istep)
          DbarF(ii) = FBC(ii) - barF_t(ii)
        end if
      end do
      if (existNBC) then
        call NBC_update(C_homo, DbarF, P_bar, PBC, isNBC, info)
        if (info /= 0) then
          cpfft_compute = -3
          return
        end if
      end if
      barF = barF_t + DbarF
      iiter_NBC = 0

      do while (.true.)
        ! initial residual: distribute DbarF over the grid (FFT_nr3.f:85-96)
        rc = cpfft_fill9(ctx, dFm, DbarF)
        rc = cpfft_dnrm2(ctx, Fn1, Fnorm)
        rc = cpfft_daxpy(ctx, 1.0d0, dFm, Fn1)
        t0 = wall()
        rc = cpfft_gkdf(ctx, dFm, b, 1)
        rc = cpfft_dscal(ctx, -1.0d0, b)
        rc = cpfft_pcg(ctx, b, dFm, tolPCG, its, relres)
        times(1) = times(1) + (wall() - t0)
        if (rc < 0) then        ! rc = 1 (zero right-hand side, dFm = 0) goes on
          cpfft_compute = rc
          return
        end if
        rc = cpfft_daxpy(ctx, 1.0d0, dFm, Fn1)
        rc = cpfft_dnrm2(ctx, dFm, dn)
        resfft = dn / Fnorm
        if (verbose /= 0) write(*, 1003) resfft
        if (verbose /= 0) flush(6)
        iiter_EBC = 0
        resfft = 1.0d0          ! at least one iteration (FFT_nr3.f:103)

        do while (resfft > tolNR)
          t0 = wall()
          rc = cpfft_drive_eps_sig(ctx, step, iiter_EBC)
          t1 = wall()
          times(2) = times(2) + (t1 - t0)
          ncalls = ncalls + 1
          if (rc < 0) then
            cpfft_compute = rc
            return
          end if
          rc = cpfft_gkdf(ctx, Pn1, b, 0)
          rc = cpfft_dscal(ctx, -1.0d0, b)
          rc = cpfft_pcg(ctx, b, dFm, tolPCG, its, relres)
          times(1) = times(1) + (wall() - t1)
          if (rc < 0) then
            cpfft_compute = rc
            return
          end if
          rc = cpfft_daxpy(ctx, 1.0d0, dFm, Fn1)
          rc = cpfft_dnrm2(ctx, dFm, dn)
          resfft = dn / Fnorm
          if (verbose /= 0) write(*, 1001) iiter_EBC, resfft
          if (verbose /= 0) flush(6)
          if (iiter_EBC == maxIter) then
            cpfft_compute = -4
            return
          end if
          iiter_EBC = iiter_EBC + 1
          nnewton = nnewton + 1
        end do

        ! post-convergence update (FFT_nr3.f:123)
        t0 = wall()
        rc = cpfft_drive_eps_sig(ctx, step, iiter_EBC)
        times(2) = times(2) + (wall() - t0)
        ncalls = ncalls + 1
        if (rc < 0) then
          cpfft_compute = rc
          return
        end if

        ! P_bar check (FFT_nr3.f:127-149)
        rc = cpfft_mean9(ctx, Pn1, P_bar)
        resPbar = 0.0d0
        do ii = 1, 9
          resPbar(2) = resPbar(2) + P_bar(ii) * P_bar(ii)
          if (.not. isNBC(ii)) cycle
          resPbar(1) = resPbar(1) + (P_bar(ii) - PBC(ii)) * (P_bar(ii) - PBC(ii))
        end do
        if (resPbar(2) < 1.0d-8) then
          resPbar(3) = sqrt(resPbar(1))
        else
          resPbar(3) = sqrt(resPbar(1) / resPbar(2))
        end if
        nloops = nloops + 1
        if (verbose /= 0) write(*, 1002) iiter_NBC, resPbar(3)
        if (verbose /= 0) flush(6)
        if (resPbar(3) <= tolNR) exit
        if (iiter_NBC > maxIter) then
          cpfft_compute = -5
          return
        end if
        ! new homogenised tangent and macroscopic correction (FFT_nr3.f:162-167)
        t0 = wall()
        rc = cpfft_tangent_homo(ctx, tolPCG, C_homo)
        times(1) = times(1) + (wall() - t0)
        if (rc < 0) then
          cpfft_compute = rc
          return
        end if
        DbarF = 0.0d0
        call NBC_update(C_homo, DbarF, P_bar, PBC, isNBC, info)
        if (info /= 0) then
          cpfft_compute = -3
          return
        end if
        barF = barF + DbarF
        iiter_NBC = iiter_NBC + 1
      end do

      if (step <= mlp_log_last) then
        rc = cpfft_mlp_log(ctx, step, int(max(mlp_log_stride, 1), c_long))
        if (rc < 0) then
          cpfft_compute = rc
          return
        end if
      end if

      ! roll-over (FFT_nr3.f:173-179)
      barF_t(1:9) = barF(1:9)
      rc = cpfft_update(ctx)
      Pbar_out(:, istep) = P_bar
      rc = cpfft_counters(ctx, cnt1)
      counters(1, istep) = nnewton
      counters(2, istep) = int(cnt1(2) - cnt0(2))
      counters(3, istep) = nloops
      counters(4, istep) = ncalls
    end do
    rc = cpfft_sync(ctx)
    times(3) = wall() - times(3)

1000 format(/,' load step ',i0)
1001 format('   Newton ',i3,'  |dF|/|F| = ',es11.4)
1002 format('   mixed-BC loop ',i3,'  |Pbar - P_BC| = ',es11.4)
1003 format('   after the macroscopic increment  |dF|/|F| = ',es11.4)
  end function cpfft_compute

end module cpfft_driver
