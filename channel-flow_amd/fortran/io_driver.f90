!--------------------------------------------------------------!
! Input file, initial field, snapshots and the statistics file  !
! (modules input_output and initialization of the reference).   !
! The input file is the reference's positional format           !
! (input_output.f90:21-81) with an optional trailing block for  !
! the settings the reference hard-codes.                        !
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

  ! Snapshot reader: the reference's stream layout (input_output.f90:231-384):
  ! grids, then U, V, W.  The reference writer follows them with P and nu_t
  ! (input_output.f90:487-544); both are read when present (P becomes the
  ! module array, nu_t is recomputed by the first substep anyway).  Restart
  ! state that the reference format has no room for (t, dt, the carried
  ! forcing, the target bulk flows, tau_eqwm_initial) lives in the side file
  ! <filein>.restart, which a reference reader never opens.
  Subroutine read_input_data(restart)
    Logical, Intent(Out) :: restart
    Integer(Int32) :: n1, nn(3)
    Integer :: ios
    Open(1, file=filein, access='stream', form='unformatted', action='read')
    Read(1) n1; If ( n1 /= nx ) Stop 'nx_f/=nx'
    Read(1) x
    Read(1) n1; If ( n1 /= ny ) Stop 'ny_f/=ny'
    Read(1) y
    Read(1) n1; If ( n1 /= nz ) Stop 'nz_f/=nz'
    Read(1) z
    Read(1) n1; If ( n1 /= nx-1 ) Stop 'nxm_f/=nxm'
    Read(1) xm
    Read(1) n1; If ( n1 /= ny-1 ) Stop 'nym_f/=nym'
    Read(1) ym
    Read(1) n1; If ( n1 /= nz-1 ) Stop 'nzm_f/=nzm'
    Read(1) zm
    Read(1) nn; If ( nn(1)/=nx .Or. nn(2)/=nyg .Or. nn(3)/=nzg ) Stop 'Error! wrong size in input file (U)'
    Read(1) U
    Read(1) nn; If ( nn(1)/=nxg .Or. nn(2)/=ny .Or. nn(3)/=nzg ) Stop 'Error! wrong size in input file (V)'
    Read(1) V
    Read(1) nn; If ( nn(1)/=nxg .Or. nn(2)/=nyg .Or. nn(3)/=nz ) Stop 'Error! wrong size in input file (W)'
    Read(1) W
    ! optional P and nu_t records, recognised by their extents
    Read(1, iostat=ios) nn
    If ( ios == 0 .And. nn(1) == nxg .And. nn(2) == nyg .And. nn(3) == nzg ) Then
       Read(1, iostat=ios) P
       If ( ios == 0 ) Read(1, iostat=ios) nn
       If ( ios == 0 .And. nn(1) == nxg .And. nn(2) == nyg .And. nn(3) == nzg ) Read(1, iostat=ios) nu_t
    End If
    Close(1)
    Inquire(file=Trim(filein)//'.restart', exist=restart)
  End Subroutine read_input_data

  ! restart side file of a snapshot: the 7 device scalars
  ! (t, dt, dPdx, dPdz, Qflow_x_0, dPdy, Qflow_y_0), then the extents and
  ! values of tau_eqwm_initial (0 0 0 without the wall model)
  Subroutine read_restart_file(fname)
    Character(*), Intent(In) :: fname
    Real(Int64) :: s(7)
    Integer(Int32) :: nn(3)
    Integer :: ios
    Open(2, file=fname, access='stream', form='unformatted', action='read')
    Read(2) s
    t = s(1); dt = s(2); dPdx = s(3); dPdz = s(4); Qflow_x_0 = s(5); dPdy = s(6); Qflow_y_0 = s(7)
    Read(2, iostat=ios) nn
    If ( ios == 0 .And. nn(1) == nx .And. nn(2) == 2 .And. nn(3) == nzg ) Read(2) tau_eqwm_initial
    Close(2)
  End Subroutine read_restart_file

  ! Snapshot writer (input_output.f90:393-544) every nsave steps: grids, U,
  ! V, W, P, nu_t in the reference's record sequence, all z planes of each
  ! field (the reference's one-rank writer drops the last plane).  P is the
  ! pressure the projection saved when compute_statistics ran this step,
  ! the actual pressure (compute_pressure) otherwise, as the reference
  ! chooses it.  The restart side file goes next to it.
  Subroutine output_data
    Character(200) :: fname
    Character(8)   :: ext
    Real(Int64), Allocatable :: ax(:,:,:), az(:,:,:)
    If ( Mod(istep,nsave) /= 0 ) Return
    Call sync_to_host
    If ( pressure_computed ) Then
       Call cf_check(cf_get_pressure(ctx, P), 'cf_get_pressure')
    Else
       Call cf_check(cf_compute_pressure(ctx, P), 'compute_pressure')
    End If
    Call cf_check(cf_get_nu_t(ctx, nu_t), 'cf_get_nu_t')
    Write(ext,'(I8)') istep + nstep_init
    fname = Trim(Adjustl(fileout))//'.'//Trim(Adjustl(ext))
    Write(*,*) 'writting ', Trim(Adjustl(fname))
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
    Write(1) nxg, nyg, nzg
    Write(1) nu_t
    Close(1)
    Call sync_scalars_from_device
    Open(2, file=Trim(fname)//'.restart', access='stream', form='unformatted', action='write', status='replace')
    Write(2) t, dt, dPdx, dPdz, Qflow_x_0, dPdy, Qflow_y_0
    If ( iwall_model == 3 ) Then
       Allocate(ax(nx,2,nzg), az(nxg,2,nz))
       Call cf_check(cf_get_wall(ctx, ax, az, tau_eqwm_initial), 'cf_get_wall')
       Write(2) nx, 2, nzg
       Write(2) tau_eqwm_initial
    Else
       Write(2) 0, 0, 0
    End If
    Close(2)
  End Subroutine output_data

  ! .stats.txt of the plane statistics (input_output.f90:549-570); the last
  ! column is the reference's P2mean, which statistics.f90:117-118 assigns
  ! from the normalised UVmean divided once more by (nxg-2)(nzg-2)
  Subroutine output_statistics
    Character(200) :: fname
    Character(8)   :: ext
    Integer :: jj
    Write(ext,'(I8)') istep + nstep_init
    fname = Trim(Adjustl(fileout))//'.'//Trim(Adjustl(ext))//'.stats.txt'
    Open(3, file=fname, form='formatted', action='write')
    Write(3,'(A,4F15.8,4I8)') '%', t, Retau, utau, nu, nx_global, ny_global, nz_global, istep
    Do jj = 1, ny
       Write(3,'(8F15.8)') y(jj), Umean(jj), U2mean(jj), V2mean(jj), W2mean(jj), UVmean(jj), P2mean(jj)
    End Do
    Close(3)
  End Subroutine output_statistics

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
    Real(Int64) :: best, q, qref
    Real(c_double) :: qd
    Logical :: restart
    Call read_input_parameters
    t = 0d0
    nx = nx_global; ny = ny_global; nz = nz_global
    nxg = nx + 1; nyg = ny + 1; nzg = nz + 1
    Allocate(x(nx), y(ny), z(nz), xm(nx-1), ym(ny-1), zm(nz-1), xg(nxg), yg(nyg), zg(nzg))
    Allocate(U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz), P(nxg,nyg,nzg), nu_t(nxg,nyg,nzg))
    Allocate(Umean(ny), Vmean(ny), Wmean(ny), U2mean(ny), V2mean(ny), W2mean(ny), UVmean(ny), P2mean(ny))
    Allocate(tau_eqwm_initial(nx,2,nzg))
    P = 0d0; nu_t = 0d0; tau_eqwm_initial = dPdx_ref
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
    ! initialization.f90:453-455: the centreline target comes from freq_mult
    Ucenterline_x = 0d0
    If ( x_mass_cte == 2 .And. Retau_int /= 0 ) Ucenterline_x = freq_mult
    prm%nx = nx; prm%ny = ny; prm%nz = nz; prm%precision = precision
    prm%cfl = CFL; prm%nu = nu; prm%dPdx = dPdx; prm%dPdz = dPdz; prm%omega_z = Omega_z
    prm%x_mass_cte = x_mass_cte; prm%y_mass_cte = y_mass_cte
    prm%les_model = LES_model; prm%iwall_model = iwall_model
    prm%dirichlet_nu_t = Dirichlet_nu_t; prm%jref_eqwm = jref_eqwm
    prm%cs = Cs; prm%eqwm_ncv = eqwm_ncv; prm%eqwm_max_iter = eqwm_max_iter
    prm%eqwm_stretch = eqwm_stretch; prm%eqwm_tol = eqwm_tol; prm%tau_mean_eqwm = dPdx_ref
    prm%wall_surrogate = 0; prm%device = 0; prm%u_centerline = Ucenterline_x
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
       ! continue exactly: time, carried forcing, target bulk flows and the
       ! wall model's previous stress
       Call sync_scalars_from_device
       Call read_restart_file(Trim(filein)//'.restart')
       Call sync_scalars_to_device
       If ( iwall_model == 3 ) Call cf_check(cf_set_tau_init(ctx, tau_eqwm_initial), 'cf_set_tau_init')
    Else
       ! initialization.f90:431-463: the bulk flow of the initial field,
       ! rescaled to a tabulated bulk flow when Retau_int asks for it
       If ( Abs(Retau_int) > 0 .And. x_mass_cte == 1 ) Then
          Call cf_check(cf_mean_mass_flow_u(ctx, qd), 'compute_mean_mass_flow_U')
          q = qd
          Write(*,*) 'warning: initial mass flow corrected'
          Write(*,*) '   flag:', Retau_int
          Select Case ( Retau_int )
          Case ( 4200 ); qref = 0.873791265477842d0
          Case ( 2000 ); qref = 0.898245304195718d0
          Case ( 950 );  qref = 0.899952805947840d0
          Case ( 550 );  qref = 0.899885379198784d0
          Case ( 395 );  qref = 0.871974131836648d0
          Case Default
             If ( Retau_int < 0 ) Then
                qref = -Retau_int*q
             Else
                Stop 'Error: Unknown mass flow correction'
             End If
          End Select
          Call sync_to_host
          U = qref/q*U
          Call sync_to_device
       End If
       Call cf_check(cf_init_mass_flow(ctx), 'cf_init_mass_flow')
       Call sync_scalars_from_device
    End If
  End Subroutine initialize
End Module initialization

