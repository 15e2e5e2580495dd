!--------------------------------------------------------------!
! Driver with the call sequence of main.f90:61-87; runs of      !
! steps that need no host work (statistics, monitor, snapshot)  !
! go to the device as one batch of graph-replayed steps.  The   !
! reference's own main.f90 also builds against these modules.  !
!--------------------------------------------------------------!
Program channel_FD
  Use global
  Use input_output
  Use initialization
  Use time_integration
  Use monitor
  Use statistics
  Use finalization
  Implicit None
  Integer(Int32) :: chunk

  Call initialize
  Call summary
  istep = 0
  Do While ( istep < nsteps )
     ! advance to the next step that needs the host
     chunk = 1
     Do While ( istep + chunk < nsteps .And. Mod(istep+chunk,nstats) /= 0 .And. &
          Mod(istep+chunk,nmonitor) /= 0 .And. Mod(istep+chunk,nsave) /= 0 .And. istep+chunk /= 1 )
        chunk = chunk + 1
     End Do
     Call advance_steps(chunk)
     istep = istep + chunk
     Call compute_statistics
     Call output_monitor
     Call output_data
  End Do
  Call finalize
End Program channel_FD
