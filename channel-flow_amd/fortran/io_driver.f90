!--------------------------------------------------------------!
! Input file, initial field, snapshots, statistics and the      !
! driver loop.  The input file is the reference's positional    !
! format (input_output.f90:21-81) with an optional trailing     !
! block for the settings the reference hard-codes.              !
!--------------------------------------------------------------!
Module input_output
  Use global
  Implicit None
Contains

  Subroutine read_input_parameters
    Character(200) :: dummy_line
    Real(Int64)    :: Rossby_plus, utau_
    Integer        :: ios
    Read(*,*) dummy_line
    Read(*,*) nx_global, ny_global, nz_global
    Read(*,*) dummy_line
    Read(*,*) CFL
    Read(*,*) dummy_line
    Read(*,*) Retau_int, nu
    Read(*,*) dummy_line
    Read(*,*) dPdx, dPdz, x_mass_cte, y_mass_cte
    Read(*,*) dummy_line
    Read(*,*) LES_model
    Read(*,*) dummy_line
    Read(*,*) iwall_model
    Read(*,*) dummy_line
    Read(*,*) alpha_mean_x, alpha_mean_y, alpha_mean_z, alpha_std, freq_mult
    Read(*,*) dummy_line
    Read(*,*) nsteps, nsave, nstats, nmonitor
    Read(*,*) dummy_line
    Read(*,*) filein
    Read(*,*) dummy_line
    Read(*,*) fileout
    Read(*,*) dummy_line
    Read(*,*) nstep_init
    Read(*,*) dummy_line
    Read(*,*) random_init
    Read(*,*) dummy_line
    Read(*,*) Rossby_plus
    Read(*,*) dummy_line
    Read(*,*) Dirichlet_nu_t
    Read(*,*) dummy_line
    Read(*,*) jref_eqwm, est_Tij
    Read(*,*) dummy_line
    Read(*,*) fil_size_2
    Read(*,*) dummy_line
    Read(*,*) dummy_line
    Read(*,*) gwm_cu, gwm_c_test, gwm_diag
    Read(*,*) dummy_line
    Read(*,*) gwm_ai
    ! optional: domain and constants the reference hard-codes
    Lx = 2d0*pi; Ly = 2d0; Lz = pi
    Read(*,*,iostat=ios) dummy_line
    If ( ios == 0 ) Read(*,*,iostat=ios) Lx, Ly, Lz, stretch_y, seed, amp_ic
    If ( ios == 0 ) Read(*,*,iostat=ios) dummy_line
    If ( ios == 0 ) Read(*,*,iostat=ios) precision, Cs, eqwm_ncv, eqwm_stretch, eqwm_tol, eqwm_max_iter, use_graph
    utau_    = Sqrt(Abs(dPdx))
    Omega_z  = Rossby_plus*utau_/2d0
    dPdx_ref = dPdx
  End Subroutine read_input_parameters

  ! face grids of the configured box: x, z uniform over the periodic
  ! length, y tanh-stretched on [0, Ly]
  Subroutine make_grid
    Integer :: i
    Real(Int64) :: eta
    Do i = 1, nx
       x(i) = Real(i-1,8)*(Lx/Real(nx-2,8))
    End Do
    Do i = 1, nz
       z(i) = Real(i-1,8)*(Lz/Real(nz-2,8))
    End Do
    Do i = 1, ny
       eta = -1d0 + 2d0*Real(i-1,8)/Real(ny-1,8)
       If ( stretch_y > 0d0 ) Then
          y(i) = (1d0 + Tanh(stretch_y*eta)/Tanh(stretch_y))*(Ly/2d0)
       Else
          y(i) = (1d0 + eta)*(Ly/2d0)
       End If
    End Do
  End Subroutine make_grid

  ! Snapshot reader: the reference's stream layout (input_output.f90:231-384)
  Subroutine read_input_data(have_trailer)
    Logical, Intent(Out) :: have_trailer
    Integer(Int32) :: n1, nn(3)
    Integer :: ios
    Real(Int64) :: tr(3)
    Open(1, file=filein, access='stream', form='unformatted', action='read')
    Read(1) n1; If ( n1 /= nx ) Stop 'nx_f/=nx'
    Read(1) x
    Read(1) n1; If ( n1 /= ny ) Stop 'ny_f/=ny'
    Read(1) y
    Read(1) n1; If ( n1 /= nz ) Stop 'nz_f/=nz'
    Read(1) z
    Read(1) n1; Read(1) xm
    Read(1) n1; Read(1) ym
    Read(1) n1; Read(1) zm
    Read(1) nn; If ( nn(1)/=nx .Or. nn(2)/=nyg .Or. nn(3)/=nzg ) Stop 'Error! wrong size in input file (U)'
    Read(1) U
    Read(1) nn; If ( nn(1)/=nxg .Or. nn(2)/=ny .Or. nn(3)/=nzg ) Stop 'Error! wrong size in input file (V)'
    Read(1) V
    Read(1) nn; If ( nn(1)/=nxg .Or. nn(2)/=nyg .Or. nn(3)/=nz ) Stop 'Error! wrong size in input file (W)'
    Read(1) W
    ! optional P block and restart trailer written by output_data
    have_trailer = .False.
    Read(1, iostat=ios) nn
    If ( ios == 0 ) Read(1, iostat=ios) P
    If ( ios == 0 ) Read(1, iostat=ios) tr
    If ( ios == 0 ) Then
       t = tr(1); dPdx = tr(2); Qflow_x_0 = tr(3)
       have_trailer = .True.
    End If
    Close(1)
  End Subroutine read_input_data

  ! Snapshot writer: same layout, all z planes of every field (the
  ! reference drops the last plane on one rank), then P, and a trailer
  ! with t, dPdx and Qflow_x_0 for restart
  Subroutine output_data(final)
    Logical, Intent(In) :: final
    Character(200) :: fname
    Character(8)   :: ext
    If ( .Not. final .And. Mod(istep,nsave) /= 0 ) Return
    Call sync_to_host
    Call cf_check(cf_get_pressure(ctx, P), 'cf_get_pressure')
    Write(ext,'(I8)') istep + nstep_init
    fname = Trim(Adjustl(fileout))//'.'//Trim(Adjustl(ext))
    Open(1, file=fname, access='stream', form='unformatted', action='write', status='replace')
    Write(1) Shape(x), x
    Write(1) Shape(y), y
    Write(1) Shape(z), z
    Write(1) Shape(xm), xm
    Write(1) Shape(ym), ym
    Write(1) Shape(zm), zm
    Write(1) nx, nyg, nzg
    Write(1) U
    Write(1) nxg, ny, nzg
    Write(1) V
    Write(1) nxg, nyg, nz
    Write(1) W
    Write(1) nxg, nyg, nzg
    Write(1) P
    Write(1) t, dPdx, Qflow_x_0
    Close(1)
  End Subroutine output_data

  ! Plane statistics (statistics.f90:43-155) and the .stats.txt file
  ! (input_output.f90:549-570)
  Subroutine compute_statistics
    Real(c_double), Allocatable :: out(:)
    Character(200) :: fname
    Character(8)   :: ext
    Integer :: jj
    If ( Mod(istep,nstats) /= 0 .And. istep /= 1 ) Return
    Allocate(out(7*ny))
    Call cf_check(cf_statistics(ctx, out), 'cf_statistics')
    Umean  = out(1:ny);        Vmean  = out(ny+1:2*ny);   Wmean  = out(2*ny+1:3*ny)
    U2mean = out(3*ny+1:4*ny); V2mean = out(4*ny+1:5*ny); W2mean = out(5*ny+1:6*ny)
    UVmean = out(6*ny+1:7*ny)
    Write(ext,'(I8)') istep + nstep_init
    fname = Trim(Adjustl(fileout))//'.'//Trim(Adjustl(ext))//'.stats.txt'
    Open(3, file=fname, form='formatted', action='write')
    Write(3,'(A,4F15.8,4I8)') '%', t, Retau, utau, nu, nx_global, ny_global, nz_global, istep
    Do jj = 1, ny
       Write(3,'(7F15.8)') y(jj), Umean(jj), U2mean(jj), V2mean(jj), W2mean(jj), UVmean(jj), 0d0
    End Do
    Close(3)
  End Subroutine compute_statistics

End Module input_output

Module initialization
  Use global
  Use input_output
  Implicit None
Contains
  ! seeded counter-based field: the same SplitMix64 keys as the Python
  ! host (channel_flow.cases.counter_uniform)
  Function smix(x0) Result(z)
    Integer(Int64), Intent(In) :: x0
    Integer(Int64) :: z
    z = x0 + Int(Z'9E3779B97F4A7C15', Int64)
    z = Ieor(z, Shiftr(z, 30)) * Int(Z'BF58476D1CE4E5B9', Int64)
    z = Ieor(z, Shiftr(z, 27)) * Int(Z'94D049BB133111EB', Int64)
    z = Ieor(z, Shiftr(z, 31))
  End Function smix

  Function urand(comp, i, j, k) Result(r)
    Integer, Intent(In) :: comp, i, j, k
    Real(Int64) :: r
    Integer(Int64) :: key, h
    key = Ior(Ior(Ior(Shiftl(Int(comp,Int64),56), Shiftl(Int(k,Int64),40)), Shiftl(Int(j,Int64),20)), Int(i,Int64))
    h = smix(Ieor(key, smix(Int(seed,Int64))))
    r = Real(Shiftr(h, 11), Int64) * (1d0/9007199254740992d0)
  End Function urand

  Subroutine init_flow
    Integer :: i, j, k
    Do k = 1, nzg
       Do j = 1, nyg
          Do i = 1, nx
             U(i,j,k) = 1.5d0*(1d0 - (yg(j)-1d0)**2) + amp_ic*(2d0*urand(0,i,j,k) - 1d0)
          End Do
       End Do
    End Do
    Do k = 1, nzg
       Do j = 1, ny
          Do i = 1, nxg
             V(i,j,k) = amp_ic*(2d0*urand(1,i,j,k) - 1d0)
          End Do
       End Do
    End Do
    Do k = 1, nz
       Do j = 1, nyg
          Do i = 1, nxg
             W(i,j,k) = amp_ic*(2d0*urand(2,i,j,k) - 1d0)
          End Do
       End Do
    End Do
  End Subroutine init_flow

  Subroutine centres(f, m, g)
    Real(Int64), Intent(In) :: f(:)
    Real(Int64), Intent(Out) :: m(:), g(:)
    Integer :: n
    n = Size(f)
    m = 0.5d0*(f(1:n-1) + f(2:n))
    g(2:n) = m
    g(1) = m(1) - 2d0*(m(1) - f(1))
    g(n+1) = m(n-1) + 2d0*(f(n) - m(n-1))
  End Subroutine centres

  Subroutine initialize
    Type(cf_params_t) :: prm
    Integer :: j
    Real(Int64) :: best
    Logical :: restart
    Call read_input_parameters
    t = 0d0
    nx = nx_global; ny = ny_global; nz = nz_global
    nxg = nx + 1; nyg = ny + 1; nzg = nz + 1
    Allocate(x(nx), y(ny), z(nz), xm(nx-1), ym(ny-1), zm(nz-1), xg(nxg), yg(nyg), zg(nzg))
    Allocate(U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz), P(nxg,nyg,nzg), nu_t(nxg,nyg,nzg))
    Allocate(Umean(ny), Vmean(ny), Wmean(ny), U2mean(ny), V2mean(ny), W2mean(ny), UVmean(ny))
    restart = .False.
    If ( random_init == 1 ) Then
       Call make_grid
    Else
       Call read_input_data(restart)
    End If
    Call centres(x, xm, xg)
    Call centres(y, ym, yg)
    Call centres(z, zm, zg)
    If ( iwall_model == 3 .And. jref_eqwm < 2 ) Then
       ! nearest centre to 0.1 above the lower wall
       best = 1d300
       Do j = 2, nyg/2
          If ( Abs(yg(j) - y(1) - 0.1d0) < best ) Then
             best = Abs(yg(j) - y(1) - 0.1d0); jref_eqwm = j
          End If
       End Do
    End If
    If ( jref_eqwm < 2 ) jref_eqwm = 2
    prm%nx = nx; prm%ny = ny; prm%nz = nz; prm%precision = precision
    prm%cfl = CFL; prm%nu = nu; prm%dPdx = dPdx; prm%dPdz = dPdz; prm%omega_z = Omega_z
    prm%x_mass_cte = x_mass_cte; prm%y_mass_cte = y_mass_cte
    prm%les_model = LES_model; prm%iwall_model = iwall_model
    prm%dirichlet_nu_t = Dirichlet_nu_t; prm%jref_eqwm = jref_eqwm
    prm%cs = Cs; prm%eqwm_ncv = eqwm_ncv; prm%eqwm_max_iter = eqwm_max_iter
    prm%eqwm_stretch = eqwm_stretch; prm%eqwm_tol = eqwm_tol; prm%tau_mean_eqwm = dPdx_ref
    prm%wall_surrogate = 0; prm%device = 0
    Call cf_check(cf_create(prm, x, y, z, ctx), 'cf_create')
    If ( random_init == 1 ) Then
       Call init_flow
       Call sync_to_device
       Call cf_check(cf_apply_boundary_conditions(ctx), 'apply_boundary_conditions')
       Call cf_check(cf_projection_step(ctx, 0), 'compute_projection_step')
       Call cf_check(cf_apply_boundary_conditions(ctx), 'apply_boundary_conditions')
    Else
       Call sync_to_device
    End If
    If ( restart ) Then
       ! continue exactly: time, carried forcing and target bulk flow
       Call sync_scalars_from_device
       Call read_restart_scalars
       Call sync_scalars_to_device
    Else
       Call cf_check(cf_init_mass_flow(ctx), 'cf_init_mass_flow')
       Call sync_scalars_from_device
    End If
  End Subroutine initialize
  Subroutine read_restart_scalars
    Logical :: h
    Call read_input_data(h)
  End Subroutine read_restart_scalars
End Module initialization

