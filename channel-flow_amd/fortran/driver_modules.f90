!-------------------------------------------------------------------!
! The reference's modules statistics (statistics.f90:22-171),       !
! monitor (monitor.f90:20-265) and finalization                     !
! (finalization.f90:19-33), with their subroutine names, so that    !
! the reference's main program (main.f90:51-87) runs over the HIP   !
! library unchanged.  Reductions run on the device.                 !
!-------------------------------------------------------------------!
Module statistics
  Use global
  Use input_output, Only : output_statistics
  Implicit None
Contains

  ! every nstats steps and at the first step: the plane statistics, u_tau,
  ! Re_tau and the bulk flows on the device, the .stats.txt file, then the
  ! reference's NaN check (Stop 'Error: NaNs!', statistics.f90:165-167)
  Subroutine compute_statistics
    Real(c_double), Allocatable :: out(:)
    Integer(c_int) :: rc
    pressure_computed = .False.
    If ( Mod(istep,nstats) == 0 .Or. istep == 1 ) Then
       pressure_computed = .True.
       Allocate(out(7*ny+4))
       rc = cf_statistics(ctx, out)
       Call sync_scalars_from_device
       Umean  = out(1:ny);        Vmean  = out(ny+1:2*ny);   Wmean  = out(2*ny+1:3*ny)
       U2mean = out(3*ny+1:4*ny); V2mean = out(4*ny+1:5*ny); W2mean = out(5*ny+1:6*ny)
       UVmean = out(6*ny+1:7*ny)
       P2mean = UVmean/Real((nxg-2)*(nzg-2), 8)
       utau = out(7*ny+1); Retau = out(7*ny+2); Qflow_x = out(7*ny+3); Qflow_y = out(7*ny+4)
       Call output_statistics
       Call cf_check(rc, 'compute_statistics')
    End If
  End Subroutine compute_statistics

End Module statistics

Module monitor
  Use global
  Use projection, Only : check_divergence
  Use wallmodel, Only : n_nan_eqwm, n_iter_eqwm
  Implicit None
Contains

  ! every nmonitor steps: means and maxima of U, V, W and nu_t over the
  ! interior, the bulk flows, forcing, wall-model counters, the signed
  ! maximum divergence and the wall time since the last report
  Subroutine output_monitor
    Real(Int64) :: max_divergence, rate
    Integer(Int64) :: cnt, crate
    If ( Mod(istep,nmonitor) /= 0 ) Return
    Call sync_to_host
    Call cf_check(cf_get_nu_t(ctx, nu_t), 'cf_get_nu_t')
    Call sync_scalars_from_device
    Call check_divergence(max_divergence)
    Call system_clock(cnt, crate)
    time2 = cnt
    rate = Real(Max(crate, 1_Int64), 8)
    Write(*,*) 'step number :', istep
    Write(*,*) 'time        :', t
    Write(*,*) 'time step   :', dt
    Write(*,*) ' '
    Write(*,*) 'Retau:      :', Retau
    Write(*,*) ' '
    Write(*,*) 'Maximum U   :', Maxval(U(2:nx-1,2:nyg-1,2:nzg-1))
    Write(*,*) 'Maximum V   :', Maxval(V(2:nxg-1,2:ny-1,2:nzg-1))
    Write(*,*) 'Maximum W   :', Maxval(W(2:nxg-1,2:nyg-1,2:nz-1))
    Write(*,*) 'Mean U      :', Sum(U(2:nx-1,2:nyg-1,2:nzg-1))/Real((nx-1)*(ny-1)*(nz-1), 8)
    Write(*,*) 'Mean V      :', Sum(V(2:nxg-1,2:ny-1,2:nzg-1))/Real((nx-1)*(ny-1)*(nz-1), 8)
    Write(*,*) 'Mean W      :', Sum(W(2:nxg-1,2:nyg-1,2:nz-1))/Real((nx-1)*(ny-1)*(nz-1), 8)
    If ( LES_model > 0 ) Then
       Write(*,*) ' '
       Write(*,*) 'Maximum nu_t: ', Maxval(nu_t(2:nxg-1,2:nyg-1,2:nzg-1))
       Write(*,*) 'Average nu_t: ', Sum(nu_t(2:nxg-1,2:nyg-1,2:nzg-1))/Real(nxg*nyg*nzg, 8)
    End If
    Write(*,*) ' '
    Write(*,*) 'Mean mass flow in x         :', Qflow_x
    Write(*,*) 'Mean mass flow in y         :', Qflow_y
    Write(*,*) 'Mean pressure gradient in x :', dPdx
    Write(*,*) 'Mean pressure gradient in y :', dPdy
    Write(*,*) 'Spanwise rotation Omega_z   :', Omega_z
    Write(*,*) ' '
    If ( iwall_model == 3 ) Then
       Write(*,*) 'Wall model     : Equilibrium wall model (Larsson)'
       Write(*,*) '      jref     : ', jref_eqwm
       Write(*,*) '      NaN eqwm : ', n_nan_eqwm
       Write(*,*) '      Max iter : ', n_iter_eqwm
    Else
       Write(*,*) 'No wall model'
    End If
    Write(*,*) ' '
    Write(*,*) 'Maximum divergence          :', max_divergence
    Write(*,*) 'Elapsed time (s)            :', Real(time2-time1, 8)/rate
    Write(*,*) '------------------------------------------------------'
    Call system_clock(cnt)
    time1 = cnt
  End Subroutine output_monitor

  ! the run's parameters (monitor.f90:155-265), one rank, the device build
  Subroutine summary
    Call system_clock(time1)
    Write(*,*) '------------------------------------------------------------'
    Write(*,*) '              Summary of initial parameters                 '
    Write(*,*) ' '
    Write(*,*) 'Number of processors:', 1
    Write(*,*) 'Device precision    :', Merge('mixed', 'fp64 ', precision == 1)
    Write(*,*) ' '
    Write(*,*) 'filein:  ', Trim(filein)
    Write(*,*) 'fileout: ', Trim(fileout)
    Write(*,*) ' '
    Write(*,*) 'nu     :', nu
    Write(*,*) 'CFL    :', CFL
    Write(*,*) ' '
    If ( x_mass_cte == 1 ) Then
       Write(*,*) 'Constant mass flow in x'
    Elseif ( x_mass_cte == 2 ) Then
       Write(*,*) 'Constant centerline U'
    Else
       Write(*,*) 'dPdx    :', dPdx
    End If
    If ( y_mass_cte == 1 ) Write(*,*) 'Constant mass flow in y'
    Write(*,*) 'dPdz    :', dPdz
    Write(*,*) 'Omega_z :', Omega_z
    Write(*,*) ' '
    Write(*,*) 'nsteps   :', nsteps
    Write(*,*) 'nsave    :', nsave
    Write(*,*) 'nstats   :', nstats
    Write(*,*) 'nmonitor :', nmonitor
    Write(*,*) ' '
    Write(*,*) 'Lx,Ly,Lz:'
    Write(*,*) Lx, Ly, Lz
    Write(*,*) ' '
    Write(*,*) 'nx,nxg,nxm :', nx, nxg, nx-1
    Write(*,*) 'ny,nyg,nym :', ny, nyg, ny-1
    Write(*,*) 'nz,nzg,nzm :', nz, nzg, nz-1
    Write(*,*) ' '
    Write(*,*) 'xg(1),xg(end) :', xg(1), xg(nxg)
    Write(*,*) 'x (1),x (end) :', x(1), x(nx)
    Write(*,*) 'yg(1),yg(end) :', yg(1), yg(nyg)
    Write(*,*) 'y (1),y (end) :', y(1), y(ny)
    Write(*,*) 'zg(1),xz(end) :', zg(1), zg(nzg)
    Write(*,*) 'z (1),z (end) :', z(1), z(nz)
    Write(*,*) ' '
    If ( LES_model == 1 ) Then
       Write(*,*) 'LES model : constant-coefficient Smagorinsky'
    Elseif ( LES_model == 2 ) Then
       Write(*,*) 'LES model : Dynamic Smagorinsky model (xz averaged)'
    Else
       Write(*,*) 'No LES model'
    End If
    Write(*,*) ' '
    If ( Dirichlet_nu_t == 1 ) Then
       Write(*,*) 'Dirichlet BC for nu_t'
    Elseif ( Dirichlet_nu_t == 0 ) Then
       Write(*,*) 'Neumann BC for nu_t'
    Else
       Stop 'Error: unknown nu BC'
    End If
    Write(*,*) ' '
    If ( iwall_model == 3 ) Then
       Write(*,*) 'Wall model     : Equilibrium wall model'
    Else
       Write(*,*) 'No wall model'
    End If
    Write(*,*) ' '
    Write(*,*) '------------------------------------------------------------'
  End Subroutine summary

End Module monitor

Module finalization
  Use global
  Implicit None
Contains
  ! releases the device context (the reference destroys its FFTW plans and
  ! finalises MPI here)
  Subroutine finalize
    If ( c_associated(ctx) ) Call cf_check(cf_destroy(ctx), 'cf_destroy')
    ctx = c_null_ptr
    Write(*,*) 'Done!'
  End Subroutine finalize
End Module finalization
