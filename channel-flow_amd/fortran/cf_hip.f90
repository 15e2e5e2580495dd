!------------------------------------------------------------------!
! ISO_C_BINDING interfaces to the HIP time-step library            !
! (include/channel_flow.h).  Every entry point returns 0 on        !
! success; cf_check turns a failure into Stop with the library's   !
! message, the reference's error convention.                       !
!------------------------------------------------------------------!
Module cf_hip

  Use, Intrinsic :: iso_c_binding
  Implicit None

  Type, Bind(C) :: cf_params_t
     Integer(c_int) :: nx, ny, nz, precision
     Real(c_double) :: cfl, nu, dPdx, dPdz, omega_z
     Integer(c_int) :: x_mass_cte, y_mass_cte, les_model, iwall_model
     Integer(c_int) :: dirichlet_nu_t, jref_eqwm
     Real(c_double) :: cs
     Integer(c_int) :: eqwm_ncv, eqwm_max_iter
     Real(c_double) :: eqwm_stretch, eqwm_tol, tau_mean_eqwm
     Integer(c_int) :: wall_surrogate, device
     Real(c_double) :: u_centerline
  End Type cf_params_t

  Interface
     Function cf_last_error() Bind(C, name='cf_last_error') Result(p)
       Import :: c_ptr
       Type(c_ptr) :: p
     End Function cf_last_error
     Function cf_create(prm, x, y, z, ctx) Bind(C, name='cf_create') Result(rc)
       Import :: c_int, c_double, c_ptr, cf_params_t
       Type(cf_params_t), Intent(In) :: prm
       Real(c_double), Intent(In) :: x(*), y(*), z(*)
       Type(c_ptr), Intent(Out) :: ctx
       Integer(c_int) :: rc
     End Function cf_create
     Function cf_destroy(ctx) Bind(C, name='cf_destroy') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int) :: rc
     End Function cf_destroy
     Function cf_set_state(ctx, U, V, W) Bind(C, name='cf_set_state') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Integer(c_int) :: rc
     End Function cf_set_state
     Function cf_get_state(ctx, U, V, W) Bind(C, name='cf_get_state') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: U(*), V(*), W(*)
       Integer(c_int) :: rc
     End Function cf_get_state
     Function cf_get_pressure(ctx, P) Bind(C, name='cf_get_pressure') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: P(*)
       Integer(c_int) :: rc
     End Function cf_get_pressure
     Function cf_compute_pressure(ctx, P) Bind(C, name='cf_compute_pressure') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: P(*)
       Integer(c_int) :: rc
     End Function cf_compute_pressure
     Function cf_get_nu_t(ctx, a) Bind(C, name='cf_get_nu_t') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: a(*)
       Integer(c_int) :: rc
     End Function cf_get_nu_t
     Function cf_get_avg_nu_t(ctx, a) Bind(C, name='cf_get_avg_nu_t') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: a(*)
       Integer(c_int) :: rc
     End Function cf_get_avg_nu_t
     Function cf_get_wall(ctx, ax, az, ti) Bind(C, name='cf_get_wall') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: ax(*), az(*), ti(*)
       Integer(c_int) :: rc
     End Function cf_get_wall
     Function cf_set_tau_init(ctx, ti) Bind(C, name='cf_set_tau_init') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: ti(*)
       Integer(c_int) :: rc
     End Function cf_set_tau_init
     Function cf_set_nu_t(ctx, a) Bind(C, name='cf_set_nu_t') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: a(*)
       Integer(c_int) :: rc
     End Function cf_set_nu_t
     Function cf_set_scalars(ctx, s) Bind(C, name='cf_set_scalars') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: s(7)
       Integer(c_int) :: rc
     End Function cf_set_scalars
     Function cf_get_scalars(ctx, s) Bind(C, name='cf_get_scalars') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: s(7)
       Integer(c_int) :: rc
     End Function cf_get_scalars
     Function cf_init_mass_flow(ctx) Bind(C, name='cf_init_mass_flow') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int) :: rc
     End Function cf_init_mass_flow
     Function cf_apply_boundary_conditions(ctx) Bind(C, name='cf_apply_boundary_conditions') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int) :: rc
     End Function cf_apply_boundary_conditions
     Function cf_compute_dt(ctx, dt) Bind(C, name='cf_compute_dt') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: dt
       Integer(c_int) :: rc
     End Function cf_compute_dt
     Function cf_rk3_step(ctx) Bind(C, name='cf_rk3_step') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int) :: rc
     End Function cf_rk3_step
     Function cf_run_steps(ctx, n, g) Bind(C, name='cf_run_steps') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: n, g
       Integer(c_int) :: rc
     End Function cf_run_steps
     Function cf_projection_step(ctx, rk) Bind(C, name='cf_projection_step') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: rk
       Integer(c_int) :: rc
     End Function cf_projection_step
     Function cf_compute_rhs(ctx, U, V, W, ru, rv, rw) Bind(C, name='cf_compute_rhs') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Real(c_double), Intent(Out) :: ru(*), rv(*), rw(*)
       Integer(c_int) :: rc
     End Function cf_compute_rhs
     Function cf_compute_eddy_viscosity(ctx) Bind(C, name='cf_compute_eddy_viscosity') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int) :: rc
     End Function cf_compute_eddy_viscosity
     Function cf_compute_wall_model(ctx, nn, ni) Bind(C, name='cf_compute_wall_model') Result(rc)
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int), Intent(Out) :: nn, ni
       Integer(c_int) :: rc
     End Function cf_compute_wall_model
     Function cf_mean_mass_flow_v(ctx, q) Bind(C, name='cf_mean_mass_flow_v') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: q
       Integer(c_int) :: rc
     End Function cf_mean_mass_flow_v
     Function cf_mean_centerline_u(ctx, q) Bind(C, name='cf_mean_centerline_u') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: q
       Integer(c_int) :: rc
     End Function cf_mean_centerline_u
     Function cf_mean_mass_flow_u(ctx, q) Bind(C, name='cf_mean_mass_flow_u') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: q
       Integer(c_int) :: rc
     End Function cf_mean_mass_flow_u
     Function cf_check_divergence(ctx, ms, ma) Bind(C, name='cf_check_divergence') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: ms, ma
       Integer(c_int) :: rc
     End Function cf_check_divergence
     Function cf_statistics(ctx, out) Bind(C, name='cf_statistics') Result(rc)
       Import :: c_int, c_double, c_ptr
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: out(*)
       Integer(c_int) :: rc
     End Function cf_statistics
     Function cf_eqwm_batch(n, u, nu, h, tin, ncv, st, tol, mit, tau, it) &
          Bind(C, name='cf_eqwm_batch') Result(rc)
       Import :: c_int, c_double
       Integer(c_int), Value :: n, ncv, mit
       Real(c_double), Intent(In) :: u(*), nu(*), h(*), tin(*)
       Real(c_double), Value :: st, tol
       Real(c_double), Intent(Out) :: tau(*)
       Integer(c_int), Intent(Out) :: it(*)
       Integer(c_int) :: rc
     End Function cf_eqwm_batch
  End Interface

Contains

  ! Stop with the library's message on a nonzero return code
  Subroutine cf_check(rc, what)
    Integer(c_int), Intent(In) :: rc
    Character(*),   Intent(In) :: what
    Type(c_ptr) :: p
    Character(kind=c_char), Pointer :: s(:)
    Character(512) :: msg
    Integer :: n
    If ( rc == 0 ) Return
    msg = ''
    p = cf_last_error()
    If ( c_associated(p) ) Then
       Call c_f_pointer(p, s, [512])
       n = 0
       Do While ( n < 512 )
          If ( s(n+1) == c_null_char ) Exit
          n = n + 1
          msg(n:n) = s(n)
       End Do
    End If
    Write(*,*) 'Error in ', Trim(what), ': ', Trim(msg)
    Error Stop 'Error: HIP library call failed'
  End Subroutine cf_check

End Module cf_hip
