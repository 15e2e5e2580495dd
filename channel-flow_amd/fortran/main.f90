!--------------------------------------------------------------!
! Driver with the call sequence of main.f90:61-87; steps that   !
! need no host work run back to back on the device.            !
!--------------------------------------------------------------!
Program channel_FD
  Use global
  Use input_output
  Use initialization
  Use time_integration
  Use projection
  Implicit None
  Real(Int64) :: divmax
  Integer(Int32) :: chunk

  Call initialize
  Write(*,'(A,3I6,A,I2)') ' channel_FD (HIP): grid ', nx, ny, nz, '  precision ', precision
  istep = 0
  Do While ( istep < nsteps )
     ! advance to the next step that needs the host (stats, monitor, save)
     chunk = 1
     Do While ( istep + chunk < nsteps .And. Mod(istep+chunk,nstats) /= 0 .And. &
          Mod(istep+chunk,nmonitor) /= 0 .And. Mod(istep+chunk,nsave) /= 0 .And. istep+chunk /= 1 )
        chunk = chunk + 1
     End Do
     Call advance_steps(chunk)
     istep = istep + chunk
     Call compute_statistics
     If ( Mod(istep,nmonitor) == 0 ) Then
        Call check_divergence(divmax)
        Write(*,'(A,I8,A,ES14.6,A,ES12.4,A,ES14.6,A,ES10.3)') ' step ', istep, ' t ', t, ' dt ', dt, &
             ' dPdx ', dPdx, ' div ', divmax
     End If
     Call output_data(.False.)
  End Do
  Call output_data(.True.)
  Call cf_check(cf_destroy(ctx), 'cf_destroy')
End Program channel_FD
