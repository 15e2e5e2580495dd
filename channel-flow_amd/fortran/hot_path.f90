!-------------------------------------------------------------------!
! The reference's hot-path modules with their subroutine names and  !
! dummy arguments, each body a call into the HIP library.           !
!   equations.f90:24,186,351   boundary_conditions.f90:20           !
!   projection.f90:20,234      mass_flow.f90:25,61,103,147,192      !
!   pressure.f90:21                                                 !
!   subgrid.f90:23             wallmodel.f90:24   eqwm.f90:90       !
!   time_integration.f90:89,179                                     !
!-------------------------------------------------------------------!
Module equations
  Use global
  Implicit None
Contains
  ! the three right-hand sides share one device evaluation
  Subroutine rhs_all(U_, V_, W_, ru, rv, rw)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Dimension(nxg,ny,nzg), Intent(In) :: V_
    Real(Int64), Dimension(nxg,nyg,nz), Intent(In) :: W_
    Real(Int64), Dimension(2:nx-1,2:nyg-1,2:nzg-1), Intent(Out) :: ru
    Real(Int64), Dimension(2:nxg-1,2:ny-1,2:nzg-1), Intent(Out) :: rv
    Real(Int64), Dimension(2:nxg-1,2:nyg-1,2:nz-1), Intent(Out) :: rw
    Call sync_scalars_to_device
    Call cf_check(cf_compute_rhs(ctx, U_, V_, W_, ru, rv, rw), 'compute_rhs')
  End Subroutine rhs_all

  Subroutine compute_rhs_u(U_, V_, W_, rhs_u)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Dimension(nxg,ny,nzg), Intent(In) :: V_
    Real(Int64), Dimension(nxg,nyg,nz), Intent(In) :: W_
    Real(Int64), Dimension(2:nx-1,2:nyg-1,2:nzg-1), Intent(Out) :: rhs_u
    Real(Int64), Allocatable :: rv(:,:,:), rw(:,:,:)
    Allocate(rv(2:nxg-1,2:ny-1,2:nzg-1), rw(2:nxg-1,2:nyg-1,2:nz-1))
    Call rhs_all(U_, V_, W_, rhs_u, rv, rw)
  End Subroutine compute_rhs_u

  Subroutine compute_rhs_v(U_, V_, W_, rhs_v)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Dimension(nxg,ny,nzg), Intent(In) :: V_
    Real(Int64), Dimension(nxg,nyg,nz), Intent(In) :: W_
    Real(Int64), Dimension(2:nxg-1,2:ny-1,2:nzg-1), Intent(Out) :: rhs_v
    Real(Int64), Allocatable :: ru(:,:,:), rw(:,:,:)
    Allocate(ru(2:nx-1,2:nyg-1,2:nzg-1), rw(2:nxg-1,2:nyg-1,2:nz-1))
    Call rhs_all(U_, V_, W_, ru, rhs_v, rw)
  End Subroutine compute_rhs_v

  Subroutine compute_rhs_w(U_, V_, W_, rhs_w)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Dimension(nxg,ny,nzg), Intent(In) :: V_
    Real(Int64), Dimension(nxg,nyg,nz), Intent(In) :: W_
    Real(Int64), Dimension(2:nxg-1,2:nyg-1,2:nz-1), Intent(Out) :: rhs_w
    Real(Int64), Allocatable :: ru(:,:,:), rv(:,:,:)
    Allocate(ru(2:nx-1,2:nyg-1,2:nzg-1), rv(2:nxg-1,2:ny-1,2:nzg-1))
    Call rhs_all(U_, V_, W_, ru, rv, rhs_w)
  End Subroutine compute_rhs_w
End Module equations

Module boundary_conditions
  Use global
  Implicit None
Contains
  ! acts on the device state (the module arrays after sync_to_device)
  Subroutine apply_boundary_conditions
    Call cf_check(cf_apply_boundary_conditions(ctx), 'apply_boundary_conditions')
  End Subroutine apply_boundary_conditions
End Module boundary_conditions

Module projection
  Use global
  Implicit None
Contains
  Subroutine compute_projection_step
    Call cf_check(cf_projection_step(ctx, rk_step), 'compute_projection_step')
  End Subroutine compute_projection_step

  ! signed maximum as the reference reports it (projection.f90:241-247)
  Subroutine check_divergence(max_divergence)
    Real(Int64), Intent(Out) :: max_divergence
    Real(c_double) :: ms, ma
    Call cf_check(cf_check_divergence(ctx, ms, ma), 'check_divergence')
    max_divergence = ms
  End Subroutine check_divergence
End Module projection

Module pressure
  Use global
  Implicit None
Contains
  ! the actual pressure of the current device state into the module array P
  Subroutine compute_pressure
    Call cf_check(cf_compute_pressure(ctx, P), 'compute_pressure')
  End Subroutine compute_pressure
End Module pressure

Module mass_flow
  Use global
  Implicit None
Contains
  ! of the device state, which mirrors the module array U
  Subroutine compute_mean_mass_flow_U(U_, Qflow)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Intent(Out) :: Qflow
    Real(c_double) :: q
    Call cf_check(cf_mean_mass_flow_u(ctx, q), 'compute_mean_mass_flow_U')
    Qflow = q
  End Subroutine compute_mean_mass_flow_U

  Subroutine compute_mean_mass_flow_V(V_, Qflow)
    Real(Int64), Dimension(nxg,ny,nzg), Intent(In) :: V_
    Real(Int64), Intent(Out) :: Qflow
    Real(c_double) :: q
    Call cf_check(cf_mean_mass_flow_v(ctx, q), 'compute_mean_mass_flow_V')
    Qflow = q
  End Subroutine compute_mean_mass_flow_V

  Subroutine compute_dPx_for_constant_mass_flow(U_, dPdx_)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Intent(Out) :: dPdx_
    Real(Int64) :: q
    Call compute_mean_mass_flow_U(U_, q)
    dPdx_ = Qflow_x_0 - q
  End Subroutine compute_dPx_for_constant_mass_flow

  Subroutine compute_dPy_for_constant_mass_flow(V_, dPdy_)
    Real(Int64), Dimension(nxg,ny,nzg), Intent(In) :: V_
    Real(Int64), Intent(Out) :: dPdy_
    Real(Int64) :: q
    Call compute_mean_mass_flow_V(V_, q)
    dPdy_ = Qflow_y_0 - q
  End Subroutine compute_dPy_for_constant_mass_flow

  Subroutine compute_dPx_for_constant_Ucenterline(U_, dPdx_)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Intent(Out) :: dPdx_
    Real(c_double) :: q
    Call cf_check(cf_mean_centerline_u(ctx, q), 'compute_dPx_for_constant_Ucenterline')
    dPdx_ = Ucenterline_x - q
  End Subroutine compute_dPx_for_constant_Ucenterline
End Module mass_flow

Module subgrid
  Use global
  Implicit None
Contains
  Subroutine compute_eddy_viscosity(U_, V_, W_, avg_nu_t_, nu_t_)
    Real(Int64), Dimension(nx,nyg,nzg), Intent(In) :: U_
    Real(Int64), Dimension(nxg,ny,nzg), Intent(In) :: V_
    Real(Int64), Dimension(nxg,nyg,nz), Intent(In) :: W_
    Real(Int64), Dimension(1,nyg,1), Intent(Out) :: avg_nu_t_
    Real(Int64), Dimension(nxg,nyg,nzg), Intent(Out) :: nu_t_
    Call cf_check(cf_compute_eddy_viscosity(ctx), 'compute_eddy_viscosity')
    Call cf_check(cf_get_nu_t(ctx, nu_t_), 'cf_get_nu_t')
    Call cf_check(cf_get_avg_nu_t(ctx, avg_nu_t_), 'cf_get_avg_nu_t')
  End Subroutine compute_eddy_viscosity
End Module subgrid

Module EQWM
  Use iso_fortran_env, Only : Int32, Int64
  Use cf_hip
  Implicit None
  Private
  Public :: solve_u_eqwm
Contains
  ! eqwm.f90:90-97 signature; rho_eq must be 1 (as every reference call)
  Subroutine solve_u_eqwm(umag_les, mu_lam, rho_eq, h_wm, ncv, stretch_factor, &
       tau_wall, iter, tau_wall_in, tt)
    Real   (Int64), Intent(In)  :: umag_les, mu_lam, rho_eq, h_wm
    Real   (Int64), Intent(In)  :: stretch_factor, tau_wall_in, tt
    Real   (Int64), Intent(Out) :: tau_wall
    Integer(Int32), Intent(In)  :: ncv
    Integer(Int32), Intent(Out) :: iter
    Real(c_double) :: u1(1), n1(1), h1(1), t1(1), tau1(1)
    Integer(c_int) :: it1(1)
    If ( rho_eq /= 1d0 ) Stop 'Error: solve_u_eqwm supports rho_eq = 1 only'
    u1 = umag_les; n1 = mu_lam; h1 = h_wm; t1 = tau_wall_in
    Call cf_check(cf_eqwm_batch(1, u1, n1, h1, t1, ncv, stretch_factor, 1d-7, 100, tau1, it1), &
         'solve_u_eqwm')
    tau_wall = tau1(1)
    iter = it1(1)
  End Subroutine solve_u_eqwm
End Module EQWM

Module wallmodel
  Use global
  Implicit None
  Integer(Int32) :: n_nan_eqwm, n_iter_eqwm
Contains
  Subroutine compute_wall_model(U_, V_, W_, nu_t_)
    Real(Int64), Dimension(nx, nyg,nzg), Intent(In) :: U_
    Real(Int64), Dimension(nxg, ny,nzg), Intent(In) :: V_
    Real(Int64), Dimension(nxg,nyg, nz), Intent(In) :: W_
    Real(Int64), Dimension(nxg,nyg,nzg), Intent(In) :: nu_t_
    Integer(c_int) :: a, b
    Call cf_check(cf_compute_wall_model(ctx, a, b), 'compute_wall_model')
    n_nan_eqwm = a
    n_iter_eqwm = b
  End Subroutine compute_wall_model
End Module wallmodel

Module time_integration
  Use global
  Implicit None
Contains
  ! compute_time_step_RK3: the whole step (eddy viscosity, wall model,
  ! three substeps with projections, mass-flow correction) on the device
  Subroutine compute_time_step_RK3
    Call cf_check(cf_rk3_step(ctx), 'compute_time_step_RK3')
    Call sync_scalars_from_device
  End Subroutine compute_time_step_RK3

  Subroutine compute_dt
    Real(c_double) :: d
    Call cf_check(cf_compute_dt(ctx, d), 'compute_dt')
    dt = d
  End Subroutine compute_dt

  ! main.f90:67-74 for n steps without returning to the host in between
  Subroutine advance_steps(n)
    Integer(Int32), Intent(In) :: n
    Call cf_check(cf_run_steps(ctx, n, use_graph), 'cf_run_steps')
    Call sync_scalars_from_device
  End Subroutine advance_steps
End Module time_integration
